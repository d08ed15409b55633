// Compiles the C++20 facade (cppwrf_amd/csrc/dyn_facade.hpp) the way a reference caller would include it, links
// the library, and exercises what runs without a GPU: the sound-step rule, parameter validation, and the error
// path of a context on a machine with no device (std::runtime_error with the library's message).
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../cppwrf_amd/csrc/dyn_facade.hpp"

// a column physics plugin as a reference suite would derive it (compiled here; run by the GPU tests' mirror)
struct Heating : cppwrf_amd::dyn::ColumnPhysics {
  int calls = 0;
  void part1(cppwrf_amd::dyn::RK3&) override { ++calls; }
  void part2(cppwrf_amd::dyn::RK3& rk) override {
    auto st = rk.state();
    std::vector<double> t(st.size(MI_T_TENDF), 1e-3);
    st.set(MI_T_TENDF, t);
  }
  void after_rk(cppwrf_amd::dyn::RK3&) override { ++calls; }
  bool th_forcing() override { return true; }
};

int main() {
  namespace dyn = cppwrf_amd::dyn;
  Heating heating;
  if (heating.th_forcing() != true || heating.qv_forcing() != false) return 13;
  // reference RK3::auto_sound_steps: max(2*(int(300*dt/min(dx,dy) - 0.01) + 1), 4)
  if (dyn::RK3::auto_sound_steps(6.0, 1000.0, 1000.0) != 4) return 2;
  if (dyn::RK3::auto_sound_steps(90.0, 15000.0, 15000.0) != 4) return 3;
  if (dyn::RK3::auto_sound_steps(60.0, 3000.0, 3000.0) != 12) return 4;
  // reference RK3::adaptive_sound_steps: the spacing divided by the largest map factor
  if (dyn::RK3::adaptive_sound_steps(10.0, 1000.0, 1000.0, 1.2) != 8) return 12;
  // dense ABI sizes of the staggered fields (x-y-z and an x-z slice, whose v has no extra row)
  const mi_grid g3{16, 12, 10, MI_BC_PERIODIC, MI_BC_PERIODIC, 1000.0, 1000.0, 6.0};
  const mi_grid g2{16, 1, 10, MI_BC_PERIODIC, MI_BC_PERIODIC, 1000.0, 1000.0, 6.0};
  if (dyn::field_size(MI_U, g3) != 17 * 12 * 10 || dyn::field_size(MI_V, g3) != 16 * 13 * 10) return 7;
  if (dyn::field_size(MI_PH, g3) != 16 * 12 * 11 || dyn::field_size(MI_MU, g3) != 16 * 12) return 8;
  if (dyn::field_size(MI_MSFUX, g3) != 17 * 12 || dyn::field_size(MI_MSFVY, g3) != 16 * 13) return 9;
  if (dyn::field_size(MI_Q + 3, g3) != 16 * 12 * 10 || dyn::field_size(MI_CQW, g3) != 16 * 12 * 11) return 10;
  if (dyn::field_size(MI_V, g2) != 16 * 10 || dyn::field_size(MI_MSFVX, g2) != 16) return 11;
  if (dyn::field_size(MI_RU_TENDF, g3) != 17 * 12 * 10 || dyn::field_size(MI_RW_TENDF, g3) != 16 * 12 * 11 ||
      dyn::field_size(MI_TH_FORCING, g3) != 16 * 12 * 10)
    return 14;
  // the LateralBC object holds its data host side until an RK3 takes it (set_lateral_bc)
  dyn::LateralBC bc(g3, 6.0);
  if (bc.has(dyn::LateralBC::U) || bc.spec_zone() != 1 || bc.width() != 5) return 15;
  bc.set(dyn::LateralBC::U, dyn::LateralBC::XS, std::vector<double>(12 * 10 * 5, 1.0), std::vector<double>(12 * 10 * 5, 0.0));
  if (!bc.has(dyn::LateralBC::U)) return 16;
  bc.advance_time(6.0);
  if (bc.dtbc != 6.0) return 17;
  try {
    dyn::LateralBC bad(g3, 6.0, 6);
    return 18;
  } catch (const std::runtime_error&) {
  }
  try {
    dyn::Context ctx(0);
    // a GPU is present: a domain with an unsupported advection order must be refused
    mi_grid g{16, 16, 10, MI_BC_PERIODIC, MI_BC_PERIODIC, 1000.0, 1000.0, 6.0};
    dyn::LargeStepParams lp;
    lp.h_adv_order = 4;
    dyn::AcousticParams ap;
    mi_levels lv{};
    try {
      dyn::RK3 rk(ctx, g, lp, ap, 4, lv);
    } catch (const std::runtime_error& e) {
      std::printf("refused: %s\n", e.what());
      return 0;
    }
    return 5;
  } catch (const std::runtime_error& e) {
    std::printf("no device: %s\n", e.what());
    return std::string(e.what()).find("cppwrf_amd") == 0 ? 0 : 6;
  }
}
