// Drives the C++20 facade (cppwrf_amd/csrc/dyn_facade.hpp) the way a reference caller drives dyn::RK3: caller-owned
// State / Tendencies / BaseState buffers, RK3::step(State&, Tendencies&, const BaseState&), and a LateralBC object
// with the boundary data of a specified-boundary case.  Input files (written by tests/test_gpu_facade.py):
//   <dir>/spec.bin    the mi_case_spec of the case
//   <dir>/params.bin  the mi_params of the domain
//   <dir>/bdy.bin     optional: records {int32 var, int32 side, int64 n, n doubles values, n doubles tendencies}
// Output: <dir>/out.bin, for each of U V W T PH MU P AL the dense values read back from the caller's own buffers.
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../cppwrf_amd/csrc/dyn_facade.hpp"

template <class T>
static T read_pod(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  T v{};
  f.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!f) throw std::runtime_error("short file " + path);
  return v;
}

int main(int argc, char** argv) {
  namespace dyn = cppwrf_amd::dyn;
  if (argc < 3) return 2;
  const std::string dir = argv[1];
  const int nsteps = std::stoi(argv[2]);
  try {
    const mi_case_spec spec = read_pod<mi_case_spec>(dir + "/spec.bin");
    const mi_params p = read_pod<mi_params>(dir + "/params.bin");
    mi_case* cs = nullptr;
    dyn::check(mi_case_create(&spec, &cs));
    mi_levels lv{};
    dyn::check(mi_case_levels(cs, &lv));
    dyn::Context ctx(0);
    dyn::LargeStepParams lp;
    lp.top_lid = p.top_lid != 0;
    lp.h_adv_order = p.h_adv_order;
    lp.v_adv_order = p.v_adv_order;
    lp.w_damping = p.w_damping;
    lp.w_crit_cfl = p.w_crit_cfl;
    dyn::AcousticParams ap;
    ap.epssm = p.epssm;
    ap.smdiv = p.smdiv;
    ap.emdiv = p.emdiv;
    ap.top_lid = lp.top_lid;
    ap.damp_opt = p.damp_opt;
    ap.hypsometric_opt = p.hypsometric_opt;
    ap.zdamp = p.zdamp;
    ap.dampcoef = p.dampcoef;
    dyn::ScalarParams sp;
    sp.moist_adv_opt = p.moist_adv_opt;
    sp.scalar_adv_opt = p.scalar_adv_opt;
    dyn::DomainOptions opt;
    opt.n_moist = p.n_moist;
    opt.n_scalars = p.n_scalars;
    opt.mapped = p.mapped != 0;
    opt.coriolis = p.coriolis != 0;
    opt.terrain = p.terrain != 0;
    opt.diff_6th_opt = p.diff_6th_opt;
    opt.diff_6th_factor = p.diff_6th_factor;
    opt.precision = p.precision;
    dyn::RK3 rk(ctx, spec.grid, lp, ap, p.time_step_sound, lv, sp, opt);
    dyn::check(mi_case_load(cs, rk.handle()));  // the case's initial fields, then the caller's containers
    dyn::State st(rk);
    dyn::Tendencies tend(rk);
    dyn::BaseState base(rk);
    std::vector<double> u0 = st.field(MI_U).download();
    std::unique_ptr<dyn::LateralBC> bc;
    if (std::FILE* f = std::fopen((dir + "/bdy.bin").c_str(), "rb")) {
      bc = std::make_unique<dyn::LateralBC>(spec.grid, spec.grid.dt);
      int32_t var = 0, side = 0;
      int64_t n = 0;
      while (std::fread(&var, 4, 1, f) == 1 && std::fread(&side, 4, 1, f) == 1 && std::fread(&n, 8, 1, f) == 1) {
        std::vector<double> v(static_cast<size_t>(n)), t(static_cast<size_t>(n));
        if (std::fread(v.data(), 8, size_t(n), f) != size_t(n) || std::fread(t.data(), 8, size_t(n), f) != size_t(n)) return 7;
        if (var >= MI_BDY_Q) bc->set_scalar(var - MI_BDY_Q, dyn::LateralBC::Side(side), v, t);
        else bc->set(dyn::LateralBC::Var(var), dyn::LateralBC::Side(side), v, t);
      }
      std::fclose(f);
      rk.set_lateral_bc(bc.get());
    }
    // a value written into a caller's buffer between steps is the one the next step sees
    u0[0] += 0.0;
    st.field(MI_U).fill_host(u0);
    for (int n = 0; n < nsteps; ++n) rk.step(st, tend, base);
    rk.sync();
    if (bc && !(bc->dtbc > 0)) return 8;  // the LateralBC's clock advanced with the steps
    std::FILE* o = std::fopen((dir + "/out.bin").c_str(), "wb");
    if (!o) return 3;
    for (int id : {MI_U, MI_V, MI_W, MI_T, MI_PH, MI_MU, MI_P, MI_AL}) {
      const std::vector<double> h = st.field(id).download();
      std::fwrite(h.data(), 8, h.size(), o);
    }
    std::fclose(o);
    mi_case_destroy(cs);
    std::printf("facade: %d steps through caller-owned buffers%s\n", nsteps, bc ? " with a LateralBC" : "");
    return 0;
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
}
