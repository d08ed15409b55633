// The domain object behind the C ABI: device memory of one forecast, the launch schedule of
// RK3::step (reference src/dyn/rk3.cpp:86-170) and its capture as a HIP graph.
//
// The schedule keeps the reference's points of halo refresh exactly (every halo read returns what the
// reference's halo holds at that moment), because the FP64 results must be bit-identical.  Adjacent
// halo refreshes with no kernel between them are merged into one launch pair (x pass, then y pass):
// every field's x pass still precedes its y pass, as the reference's batched fill (halo.cpp:133-190).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cppwrf_amd.h"
#include "errors.hpp"
#include "kernels_fp64.hpp"

namespace {

using namespace mi_dev;
using mi_detail::Error;

thread_local std::string g_last_error;

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(e == hipErrorOutOfMemory ? MI_ENOMEM : MI_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

enum class St : char { M, U, V, W, M2, U2, V2 };

struct FieldDesc {
  double** slot;  // where the kernels' Fields keeps its pointer
  St st;
  int id;         // MI_* id or -1
};

}  // namespace

namespace mi_detail {
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace mi_detail

struct mi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
};

struct mi_domain {
  mi_ctx* ctx = nullptr;
  mi_grid grid{};
  mi_params prm{};
  int n_sound = 0;
  mi_fp64::Args a{};
  int64_t n3 = 0, n2 = 0;  // doubles per 3-D / 2-D allocation
  int64_t off = 0;         // offset of point (0,0,0) in an allocation
  int nxa = 0, nya = 0;
  std::vector<void*> allocs;
  int64_t bytes = 0;
  std::vector<FieldDesc> fields;
  double *pb_buf = nullptr, *alb_buf = nullptr, *phb_buf = nullptr, *t_init_buf = nullptr;
  int* flags = nullptr;  // device finiteness flags
  // the scalars and their buffers; qpar flips with every step (three stage swaps)
  mi_fp64::Scalars sc{};
  int n_sc = 0, qpar = 0;
  // one graph per scalar-buffer parity: the step's launches name the buffers (none without scalars)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool use_graph = true;
  std::string prof_name;
  mi_fp64::KernelTimer timer{};
  int colcfg[mi_fp64::kNumColKernels] = {0, 0, 0, 0};  // block shape of each column kernel
  std::vector<double*> backup;                          // the carried state, for the autotuner

  // launches of fn, with the kernel timer active when a kernel is selected
  template <class F>
  void timed(F&& fn) {
    if (prof_name.empty()) {
      fn();
      return;
    }
    timer.name = prof_name.c_str();
    mi_fp64::g_timer = &timer;
    try {
      fn();
    } catch (...) {
      mi_fp64::g_timer = nullptr;
      throw;
    }
    mi_fp64::g_timer = nullptr;
  }

  ~mi_domain() {
    for (hipEvent_t e : timer.ev) (void)hipEventDestroy(e);
    drop_graphs();
    for (void* p : allocs) (void)hipFree(p);
  }
  void drop_graphs() {
    for (int q = 0; q < 2; ++q) {
      if (exec[q]) (void)hipGraphExecDestroy(exec[q]);
      if (graph[q]) (void)hipGraphDestroy(graph[q]);
      exec[q] = nullptr;
      graph[q] = nullptr;
    }
  }

  double* alloc(int64_t n) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, size_t(n) * sizeof(double)), "hipMalloc");
    allocs.push_back(p);
    bytes += n * int64_t(sizeof(double));
    hip_check(hipMemsetAsync(p, 0, size_t(n) * sizeof(double), ctx->stream), "hipMemsetAsync");
    return static_cast<double*>(p);
  }

  // ------------------------------------------------------------------ geometry of a staggering
  int nxs(St s) const { return grid.nx + ((s == St::U || s == St::U2) ? 1 : 0); }
  int nys(St s) const { return grid.ny + (((s == St::V || s == St::V2) && grid.ny > 1) ? 1 : 0); }
  int nzs(St s) const {
    if (s == St::M2 || s == St::U2 || s == St::V2) return 1;
    return grid.nz + (s == St::W ? 1 : 0);
  }
  bool face_x(St s) const { return s == St::U || s == St::U2; }
  bool face_y(St s) const { return (s == St::V || s == St::V2) && grid.ny > 1; }

  const FieldDesc* desc_of(const double* p) const {
    for (const FieldDesc& f : fields)
      if (*f.slot == p) return &f;
    return nullptr;
  }

  // ------------------------------------------------------------------ halo batches
  struct HaloReq {
    double* p;
    int px, py;  // -1: default parity of the staggering
  };
  void halo(std::initializer_list<HaloReq> reqs) {
    mi_fp64::HaloSet s{};
    s.nkmax = 1;
    for (const HaloReq& r : reqs) {
      const FieldDesc* d = desc_of(r.p);
      if (!d) throw Error(MI_EINVAL, "internal: halo of an unknown field");
      if (s.n == mi_fp64::kHaloMax) throw Error(MI_EINVAL, "internal: too many fields in one halo batch");
      const int n = s.n++;
      s.p[n] = r.p;
      s.nxs[n] = nxs(d->st);
      s.nys[n] = nys(d->st);
      s.nzs[n] = nzs(d->st);
      s.face_x[n] = face_x(d->st);
      s.face_y[n] = face_y(d->st);
      s.px[n] = (unsigned char)(r.px >= 0 ? r.px : (face_x(d->st) ? 1 : 0));
      s.py[n] = (unsigned char)(r.py >= 0 ? r.py : (face_y(d->st) ? 1 : 0));
      if (s.nzs[n] > s.nkmax) s.nkmax = s.nzs[n];
    }
    mi_fp64::launch_halo(a, s, ctx->stream);
  }

  // ------------------------------------------------------------------ the phases of a step
  // RK3::step's _2 -> _1 copies (rk3.cpp:94-98).  The copies include the halos, which are fresh, so
  // the reference's refresh of mu_1's halo rewrites the values already there and is left out.
  void step_begin() {
    mi_fp64::launch_copy_step(a, n3, n2, ctx->stream);
    mi_fp64::launch_copy_scalars(a, sc, n_sc, n3, ctx->stream);  // q_1 = q_2 (rk3.cpp:99)
  }

  // LargeStep::prep (large_step.cpp:697-733).  Its first halo refresh (u, v, w, t, ph, p, al, mu of the
  // stage) is left out: every path that changes those fields (the end of a stage, mi_domain_set_field)
  // refreshes their halos itself, so at this point it would rewrite identical values.  The refresh of
  // its outputs (mut, ru, rv, rw, alt, php, muu, muv, ww; large_step.cpp:718-720) is written by the
  // kernel itself: each output point stores its halo images.  The same holds for every refresh of the
  // small steps below.
  void rk_prep() {
    mi_fp64::launch_large_prep(a, colcfg[mi_fp64::kColLargePrep], ctx->stream);
    if (a.q.moist) mi_fp64::launch_cq(a, sc, ctx->stream);  // calc_cq (large_step.cpp:724-733)
  }

  // LargeStep::tendencies (large_step.cpp:798-861); at the first stage, with a diffusion driver,
  // zero_tendf and the sixth-order diffusion into the held tendencies first
  void rk_tendency(int rk) {
    if (rk == 1 && a.q.use_tendf) mi_fp64::launch_sixth_dyn(a, ctx->stream);
    mi_fp64::launch_tendencies(a, ctx->stream);
  }

  // SmallStep::run_stage (small_step.cpp:490-607); with prhophi, calc_p_rho_phi (small_step.cpp:609-617)
  // is fused into the finishing kernel
  void small_steps(int rk, bool prhophi) {
    auto& f = a.f;
    const double dt = grid.dt, dts_full = dt / n_sound;
    const double dt_rk = rk == 1 ? dt / 3.0 : (rk == 2 ? 0.5 * dt : dt);
    const double dts = rk == 1 ? dt_rk : dts_full;
    const int ns = rk == 1 ? 1 : (rk == 2 ? n_sound / 2 : n_sound);
    hipStream_t st = ctx->stream;
    if (rk == 1) mi_fp64::launch_fill(f.mudf - off, n2, 0.0, st);
    mi_fp64::launch_ss_prep(a, rk == 1, st);
    const double cof_h = 0.5 * dts * kG * (1.0 + prm.epssm);
    // halos of u'', v'', ph'', t_save, u_save, v_save, mu'', muts, p, al (small_step.cpp:526-528): images
    // written by the two kernels above
    mi_fp64::launch_ss_column(a, colcfg[mi_fp64::kColSsColumn], cof_h * cof_h, st);
    for (int it = 1; it <= ns; ++it) {
      mi_fp64::launch_uv(a, dts, st);  // with the halos of u'', v''
      // with the halos of mu'', mudf, muts, ph'', p, al, and of ru_m, rv_m, ww_m (their refresh after
      // the loop; the sums' halo images are rewritten at every substep, the last one's stay)
      mi_fp64::launch_acoustic(a, colcfg[mi_fp64::kColAcoustic], dts, it, ns, st);
    }
    // with the halos of u, v, w, t, ph, ww, mu (and p, al); the microphysics heating leaves theta_m at
    // the last stage with moisture (small_step.cpp:602)
    const double hd_dt = (rk == 3 && a.q.moist) ? dts * double(ns) : 0.0;
    mi_fp64::launch_finish(a, colcfg[mi_fp64::kColFinish], prhophi, hd_dt, st);
  }

  // ScalarTransport::run_stage (scalars.cpp:436-461): the plain species, then the positive-definite ones
  // (last stage, adv_opt 1); each species' new value becomes its q_2 (the buffers swap), halos included
  void scalar_transport(int rk) {
    if (n_sc == 0) return;
    const double dt_rk = rk == 1 ? grid.dt / 3.0 : (rk == 2 ? 0.5 * grid.dt : grid.dt);
    int plain[MI_MAX_SCALARS], pd[MI_MAX_SCALARS], np = 0, nd = 0;
    for (int n = 0; n < n_sc; ++n) {
      const bool is_pd = rk == 3 && (n < prm.n_moist ? prm.moist_adv_opt : prm.scalar_adv_opt) == 1;
      if (is_pd) pd[nd++] = n;
      else plain[np++] = n;
    }
    // rk_scalar_tend at the first stage: the step's tendency of every scalar, sixth-order diffusion included
    if (rk == 1 && a.q.use_tendf) mi_fp64::launch_sixth_scalars(a, sc, n_sc, dt_rk, ctx->stream);
    mi_fp64::launch_scalars(a, sc, plain, np, false, dt_rk, ctx->stream);
    mi_fp64::launch_scalars(a, sc, pd, nd, true, dt_rk, ctx->stream);
    for (int n = 0; n < n_sc; ++n) std::swap(sc.q2[n], sc.qn[n]);
    qpar ^= 1;
  }

  // SmallStep::calc_p_rho_phi (small_step.cpp:609-617), for the phase-by-phase entry points
  void calc_p_rho_phi() { mi_fp64::launch_p_rho_phi(a, ctx->stream); }  // with the halos of p, al

  void step_launches() {
    step_begin();
    for (int rk = 1; rk <= 3; ++rk) {
      rk_prep();
      rk_tendency(rk);
      small_steps(rk, true);
      scalar_transport(rk);
    }
    if (a.q.terrain) mi_fp64::launch_set_w_surface(a, ctx->stream);  // rk3.cpp:163
  }

  // the fields a step reads from the previous one (everything else is recomputed within a step)
  std::vector<double*> carried() {
    std::vector<double*> c = {a.f.u2, a.f.v2, a.f.w2, a.f.t2, a.f.ph2, a.f.mu2, a.f.p, a.f.al};
    for (int n = 0; n < n_sc; ++n) c.push_back(sc.q2[n]);
    return c;
  }
  mi_fp64::Scalars sc_saved{};
  int qpar_saved = 0;
  void save_or_restore(bool save) {
    if (save) {
      sc_saved = sc;
      qpar_saved = qpar;
    } else {  // the buffers as they were named at the save
      sc = sc_saved;
      qpar = qpar_saved;
    }
    std::vector<double*> c = carried();
    if (backup.empty())
      for (double* p : c) backup.push_back(alloc(p == a.f.mu2 ? n2 : n3));
    for (size_t q = 0; q < c.size(); ++q) {
      const int64_t n = c[q] == a.f.mu2 ? n2 : n3;
      double* live = c[q] - off;
      hip_check(hipMemcpyAsync(save ? backup[q] : live, save ? live : backup[q], size_t(n) * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpyAsync");
    }
  }

  // Time every block shape of each column kernel on this domain and keep the fastest; the state is
  // restored afterwards, so the forecast continues exactly as if no step had been taken.
  void autotune(int trials) {
    if (trials < 1) trials = 1;
    const std::string keep = prof_name;
    save_or_restore(true);
    for (int s = 0; s < 1; ++s) timed([&] { step_launches(); });  // warm-up
    for (int kern = 0; kern < mi_fp64::kNumColKernels; ++kern) {
      double best_t = 1e300;
      int best = colcfg[kern];
      for (int cfg = 0; cfg < mi_fp64::kNumColCfgs; ++cfg) {
        if (!mi_fp64::col_cfg_fits(kern, cfg, grid.nz)) continue;
        colcfg[kern] = cfg;
        prof_name = mi_fp64::kColKernelNames[kern];
        timer.used = 0;
        timer.bytes = 0;
        for (int s = 0; s < trials; ++s) timed([&] { step_launches(); });
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        double tot = 0;
        for (size_t q = 0; q + 1 < timer.used; q += 2) {
          float ms = 0;
          hip_check(hipEventElapsedTime(&ms, timer.ev[q], timer.ev[q + 1]), "hipEventElapsedTime");
          tot += ms;
        }
        if (tot < best_t) {
          best_t = tot;
          best = cfg;
        }
      }
      colcfg[kern] = best;
    }
    prof_name = keep;
    timer.used = 0;
    timer.bytes = 0;
    save_or_restore(false);
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    drop_graphs();  // the graphs hold the old shapes
  }

  void step(int nsteps) {
    if (nsteps <= 0) return;
    if (!use_graph || !prof_name.empty()) {
      for (int s = 0; s < nsteps; ++s) timed([&] { step_launches(); });
      hip_check(hipGetLastError(), "kernel launch");
      return;
    }
    for (int s = 0; s < nsteps; ++s) {
      const int par = qpar;
      if (!exec[par]) {
        // capturing runs the host side of step_launches (the buffer swaps) without running the step, so
        // the host state is put back before the launch below performs the step
        const mi_fp64::Scalars keep = sc;
        hip_check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        step_launches();
        hipGraph_t gr = nullptr;
        hip_check(hipStreamEndCapture(ctx->stream, &gr), "hipStreamEndCapture");
        graph[par] = gr;
        hip_check(hipGraphInstantiate(&exec[par], gr, nullptr, nullptr, 0), "hipGraphInstantiate");
        sc = keep;
        qpar = par;
      }
      hip_check(hipGraphLaunch(exec[par], ctx->stream), "hipGraphLaunch");
      // the host side of the step: three stage swaps of every species' buffers
      for (int r = 0; r < (n_sc > 0 ? 3 : 0); ++r) {
        for (int n = 0; n < n_sc; ++n) std::swap(sc.q2[n], sc.qn[n]);
        qpar ^= 1;
      }
    }
  }
};

namespace {

void build_domain(mi_domain& d, mi_ctx* ctx, const mi_grid& g, const mi_params& p, const mi_levels& lv) {
  if (g.nx < 1 || g.ny < 1 || g.nz < 3) throw Error(MI_EINVAL, "grid: need nx >= 1, ny >= 1, nz >= 3");
  if (g.bc_x != MI_BC_PERIODIC && g.bc_x != MI_BC_SYMMETRIC) throw Error(MI_EINVAL, "bc_x: periodic or symmetric");
  if (g.bc_y != MI_BC_PERIODIC && g.bc_y != MI_BC_SYMMETRIC) throw Error(MI_EINVAL, "bc_y: periodic or symmetric");
  if (g.bc_x == MI_BC_PERIODIC && g.nx < 3) throw Error(MI_EINVAL, "periodic x needs nx >= 3");
  if (g.ny > 1 && g.bc_y == MI_BC_PERIODIC && g.ny < 3) throw Error(MI_EINVAL, "periodic y needs ny >= 3");
  // a symmetric fill with fewer than 4 points mirrors halo slots into one another (order-dependent)
  if (g.bc_x == MI_BC_SYMMETRIC && g.nx < 4) throw Error(MI_EINVAL, "symmetric x needs nx >= 4");
  if (g.ny > 1 && g.bc_y == MI_BC_SYMMETRIC && g.ny < 4) throw Error(MI_EINVAL, "symmetric y needs ny >= 4");
  if (!(g.dx > 0) || !(g.dt > 0) || (g.ny > 1 && !(g.dy > 0))) throw Error(MI_EINVAL, "grid: dx, dy, dt must be > 0");
  if (p.h_adv_order != 5 && p.h_adv_order != 6) throw Error(MI_EINVAL, "h_adv_order must be 5 or 6");
  if (p.v_adv_order != 3 && p.v_adv_order != 4) throw Error(MI_EINVAL, "v_adv_order must be 3 or 4");
  if (p.damp_opt != 0 && p.damp_opt != 3) throw Error(MI_EINVAL, "damp_opt must be 0 or 3");
  if (p.hypsometric_opt != 1) throw Error(MI_EINVAL, "hypsometric_opt must be 1");
  if (p.precision != MI_FP64) throw Error(MI_EINVAL, "precision: only MI_FP64 is available");
  if (p.damp_opt == 3 && !(p.zdamp > 0)) throw Error(MI_EINVAL, "zdamp must be > 0");
  if (p.n_moist < 0 || p.n_moist > MI_MAX_MOIST) throw Error(MI_EINVAL, "n_moist must be in 0..8");
  if (p.n_scalars < p.n_moist || p.n_scalars > MI_MAX_SCALARS) throw Error(MI_EINVAL, "n_scalars must be in n_moist..16");
  if ((p.moist_adv_opt != 0 && p.moist_adv_opt != 1) || (p.scalar_adv_opt != 0 && p.scalar_adv_opt != 1))
    throw Error(MI_EINVAL, "moist_adv_opt and scalar_adv_opt must be 0 or 1");
  if ((p.mapped != 0 && p.mapped != 1) || (p.coriolis != 0 && p.coriolis != 1))
    throw Error(MI_EINVAL, "mapped and coriolis must be 0 or 1");
  if (p.diff_6th_opt < 0 || p.diff_6th_opt > 2) throw Error(MI_EINVAL, "diff_6th_opt must be 0, 1 or 2");
  if (p.terrain != 0 && p.terrain != 1) throw Error(MI_EINVAL, "terrain must be 0 or 1");
  d.ctx = ctx;
  d.grid = g;
  d.prm = p;
  d.n_sound = p.time_step_sound > 0 ? p.time_step_sound : mi_auto_sound_steps(g.dt, g.dx, g.dy);
  if (d.n_sound < 2 || d.n_sound % 2) throw Error(MI_EINVAL, "time_step_sound must be even and >= 2");
  if (!mi_fp64::col_cfg_fits(mi_fp64::kColAcoustic, mi_fp64::kNumColCfgs - 1, g.nz)) throw Error(MI_EINVAL, "nz must be at most 255");
  if (const char* e = std::getenv("MI_GRAPHS")) d.use_graph = std::strcmp(e, "0") != 0;

  const bool two_d = g.ny == 1;
  Geo& G = d.a.g;
  G.nx = g.nx;
  G.ny = g.ny;
  G.nz = g.nz;
  G.nyv = two_d ? 1 : g.ny + 1;
  G.hy = two_d ? 0 : kHalo;
  G.bcx = g.bc_x;
  G.bcy = g.bc_y;
  G.two_d = two_d;
  d.nxa = g.nx + 1;
  d.nya = G.nyv;
  const int64_t pitch = (int64_t(kLpad) + d.nxa + kHalo + 31) / 32 * 32;
  const int64_t rows = d.nya + 2 * G.hy;
  G.sj = two_d ? 0 : pitch;
  G.sk = pitch * rows;
  d.off = int64_t(kLpad) + int64_t(G.hy) * G.sj;
  d.n2 = G.sk;
  d.n3 = G.sk * int64_t(g.nz + 1);

  auto& f = d.a.f;
  auto add3 = [&](double*& slot, St s, int id) {
    slot = d.alloc(d.n3) + d.off;
    d.fields.push_back(FieldDesc{&slot, s, id});
  };
  auto add2 = [&](double*& slot, St s, int id) {
    slot = d.alloc(d.n2) + d.off;
    d.fields.push_back(FieldDesc{&slot, s, id});
  };
  add3(f.u1, St::U, -1); add3(f.u2, St::U, MI_U);
  add3(f.v1, St::V, -1); add3(f.v2, St::V, MI_V);
  add3(f.w1, St::W, -1); add3(f.w2, St::W, MI_W);
  add3(f.t1, St::M, -1); add3(f.t2, St::M, MI_T);
  add3(f.ph1, St::W, -1); add3(f.ph2, St::W, MI_PH);
  add2(f.mu1, St::M2, -1); add2(f.mu2, St::M2, MI_MU);
  add3(f.p, St::M, MI_P); add3(f.al, St::M, MI_AL); add3(f.alt, St::M, MI_ALT); add3(f.php, St::M, MI_PHP);
  add3(f.ww, St::W, MI_WW);
  add2(f.mut, St::M2, -1); add2(f.muu, St::U2, -1); add2(f.muv, St::V2, -1);
  add2(f.muts, St::M2, MI_MUTS); add2(f.muus, St::U2, -1); add2(f.muvs, St::V2, -1);
  add3(f.ru, St::U, -1); add3(f.rv, St::V, -1); add3(f.rw, St::W, -1);
  add3(f.ru_m, St::U, MI_RU_M); add3(f.rv_m, St::V, MI_RV_M); add3(f.ww_m, St::W, MI_WW_M);
  add3(f.ru_tend, St::U, MI_RU_TEND); add3(f.rv_tend, St::V, MI_RV_TEND); add3(f.rw_tend, St::W, MI_RW_TEND);
  add3(f.t_tend, St::M, MI_T_TEND); add3(f.ph_tend, St::W, MI_PH_TEND); add2(f.mu_tend, St::M2, -1);
  add3(f.u_save, St::U, -1); add3(f.v_save, St::V, -1); add3(f.w_save, St::W, -1); add3(f.t_save, St::M, -1);
  add3(f.ph_save, St::W, -1); add3(f.ww1, St::W, -1); add3(f.c2a, St::M, -1); add3(f.pm1, St::M, -1);
  add3(f.ca, St::W, -1); add3(f.calpha, St::W, -1); add3(f.cgamma, St::W, -1);
  add2(f.mu_save, St::M2, -1); add2(f.mudf, St::M2, -1); add2(f.muave, St::M2, -1);
  add2(f.mub, St::M2, MI_MUB);
  add2(f.msftx, St::M2, MI_MSFTX); add2(f.msfty, St::M2, MI_MSFTY);
  add2(f.msfux, St::U2, MI_MSFUX); add2(f.msfuy, St::U2, MI_MSFUY);
  add2(f.msfvx, St::V2, MI_MSFVX); add2(f.msfvy, St::V2, MI_MSFVY); add2(f.msfvx_inv, St::V2, MI_MSFVX_INV);
  for (double* m : {f.msftx, f.msfty, f.msfux, f.msfuy, f.msfvx, f.msfvy, f.msfvx_inv})
    mi_fp64::launch_fill(m - d.off, d.n2, 1.0, ctx->stream);
  // Coriolis and grid rotation: f = e = sina = 0, cosa = 1 unless set
  add2(f.fcor, St::M2, MI_F); add2(f.ecor, St::M2, MI_E); add2(f.sina, St::M2, MI_SINA); add2(f.cosa, St::M2, MI_COSA);
  mi_fp64::launch_fill(f.cosa - d.off, d.n2, 1.0, ctx->stream);
  if (p.diff_6th_opt != 0) {  // the held tendencies of a diffusion driver
    add3(f.ru_tendf, St::U, -1); add3(f.rv_tendf, St::V, -1); add3(f.rw_tendf, St::W, -1); add3(f.t_tendf, St::M, -1);
  }
  // moisture and scalars
  if (p.n_moist > 0) {
    add3(f.cqu, St::U, MI_CQU); add3(f.cqv, St::V, MI_CQV); add3(f.cqw, St::W, MI_CQW);
    add3(f.hdia, St::M, MI_H_DIABATIC);
  }
  d.n_sc = p.n_scalars;
  d.sc.n_moist = p.n_moist;
  for (int n = 0; n < d.n_sc; ++n) {
    add3(d.sc.q1[n], St::M, -1);
    add3(d.sc.q2[n], St::M, MI_Q + n);
    add3(d.sc.qn[n], St::M, -1);
    add3(d.sc.qtf[n], St::M, MI_Q_TENDF + n);
    add3(d.sc.scale[n], St::M, -1);
    if (p.diff_6th_opt != 0) add3(d.sc.qtw[n], St::M, -1);
  }
  // base 3-D fields: profiles (flat terrain, horizontally uniform base state) or full arrays (terrain)
  const int nz = g.nz;
  if (p.terrain) {
    add3(d.pb_buf, St::M, MI_PB); add3(d.alb_buf, St::M, MI_ALB); add3(d.phb_buf, St::W, MI_PHB);
    add3(d.t_init_buf, St::M, MI_T_INIT);
    add2(f.ht, St::M2, MI_HT);
    f.pb = BField{d.pb_buf, 1, G.sj, G.sk};
    f.alb = BField{d.alb_buf, 1, G.sj, G.sk};
    f.phb = BField{d.phb_buf, 1, G.sj, G.sk};
  } else {
    d.pb_buf = d.alloc(nz + 1);
    d.alb_buf = d.alloc(nz + 1);
    d.phb_buf = d.alloc(nz + 1);
    d.t_init_buf = d.alloc(nz + 1);
    f.pb = BField{d.pb_buf, 0, 0, 1};
    f.alb = BField{d.alb_buf, 0, 0, 1};
    f.phb = BField{d.phb_buf, 0, 0, 1};
  }
  {
    void* fl = nullptr;
    hip_check(hipMalloc(&fl, 64 * sizeof(int)), "hipMalloc");
    d.allocs.push_back(fl);
    d.flags = static_cast<int*>(fl);
  }
  // level coefficients
  double* lvb = d.alloc(int64_t(14) * (nz + 1));
  std::vector<double> h(size_t(14) * (nz + 1), 0.0);
  const double* src[14] = {lv.dnw, lv.rdnw, lv.dn, lv.rdn, lv.fnm, lv.fnp, lv.c1h, lv.c2h, lv.c3h, lv.c4h, lv.c1f, lv.c2f, lv.c3f, lv.c4f};
  const int len[14] = {nz, nz, nz + 1, nz + 1, nz + 1, nz + 1, nz, nz, nz, nz, nz + 1, nz + 1, nz + 1, nz + 1};
  for (int q = 0; q < 14; ++q) {
    if (!src[q]) throw Error(MI_EINVAL, "levels: null coefficient array");
    std::memcpy(&h[size_t(q) * (nz + 1)], src[q], size_t(len[q]) * sizeof(double));
  }
  hip_check(hipMemcpyAsync(lvb, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  Lv& c = d.a.c;
  const double** dst[14] = {&c.dnw, &c.rdnw, &c.dn, &c.rdn, &c.fnm, &c.fnp, &c.c1h, &c.c2h, &c.c3h, &c.c4h, &c.c1f, &c.c2f, &c.c3f, &c.c4f};
  for (int q = 0; q < 14; ++q) *dst[q] = lvb + size_t(q) * (nz + 1);
  c.cf1 = lv.cf1; c.cf2 = lv.cf2; c.cf3 = lv.cf3; c.cfn = lv.cfn; c.cfn1 = lv.cfn1;
  c.rdx = lv.rdx; c.rdy = lv.rdy; c.p_top = lv.p_top;
  for (int kern = 0; kern < mi_fp64::kNumColKernels; ++kern) d.colcfg[kern] = mi_fp64::col_cfg_default(kern, g.nx, g.ny, g.nz);
  if (d.n3 >= (int64_t(1) << 31)) throw Error(MI_EINVAL, "grid too large: a field must hold fewer than 2^31 points");
  // run parameters
  mi_fp64::Params& q = d.a.q;
  q.hup = p.h_adv_order == 5 ? 1.0 : 0.0;
  q.vup = p.v_adv_order == 3 ? 1.0 : 0.0;
  q.epssm = p.epssm;
  q.smdiv = p.smdiv;
  q.emdx = p.emdiv * (1. / lv.rdx);
  q.emdy = two_d ? 0.0 : p.emdiv * (1. / lv.rdy);
  q.wd_dt = p.w_damping == 1 ? g.dt : 0.0;
  q.w_crit = p.w_crit_cfl;
  q.dampcoef = p.dampcoef;
  q.zdamp = p.zdamp;
  q.top_lid = p.top_lid;
  q.damp_opt = p.damp_opt;
  q.moist = p.n_moist > 0 ? 1 : 0;
  q.mapped = p.mapped;
  q.cor = (p.mapped != 0 || p.coriolis != 0) ? 1 : 0;
  q.use_tendf = p.diff_6th_opt != 0 ? 1 : 0;
  q.diff_6th_opt = p.diff_6th_opt;
  q.diff_6th_factor = p.diff_6th_factor;
  q.dt = g.dt;
  q.terrain = p.terrain;
}

FieldDesc* find(mi_domain* d, int id) {
  for (FieldDesc& f : d->fields)
    if (f.id == id) return &f;
  return nullptr;
}

}  // namespace

extern "C" {

const char* mi_last_error(void) { return g_last_error.c_str(); }
int mi_abi_version(void) { return MI_ABI_VERSION; }

int mi_auto_sound_steps(double dt, double dx, double dy) {
  const double spacing = dx < dy ? dx : dy;
  const int n = 2 * (int(300.0 * dt / spacing - 0.01) + 1);
  return n > 4 ? n : 4;
}

int mi_ctx_create(int device, void* stream, mi_ctx** out) {
  return mi_detail::guard([&] {
    if (!out) throw Error(MI_EINVAL, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(MI_ENODEV, "no HIP device");
    if (device < 0 || device >= n) throw Error(MI_EINVAL, "device index out of range");
    hip_check(hipSetDevice(device), "hipSetDevice");
    auto c = std::make_unique<mi_ctx>();
    c->device = device;
    if (stream) {
      c->stream = static_cast<hipStream_t>(stream);
    } else {
      hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
      c->own_stream = true;
    }
    *out = c.release();
  });
}

int mi_ctx_destroy(mi_ctx* c) {
  return mi_detail::guard([&] {
    if (!c) return;
    if (c->own_stream) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipStreamDestroy(c->stream);
    }
    delete c;
  });
}

int mi_domain_create(mi_ctx* ctx, const mi_grid* g, const mi_params* p, const mi_levels* lv, mi_domain** out) {
  return mi_detail::guard([&] {
    if (!ctx || !g || !p || !lv || !out) throw Error(MI_EINVAL, "null argument");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    auto d = std::make_unique<mi_domain>();
    build_domain(*d, ctx, *g, *p, *lv);
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    *out = d.release();
  });
}

int mi_domain_destroy(mi_domain* d) {
  return mi_detail::guard([&] {
    if (!d) return;
    (void)hipStreamSynchronize(d->ctx->stream);
    delete d;
  });
}

int mi_domain_n_sound(const mi_domain* d) { return d ? d->n_sound : MI_EINVAL; }
int64_t mi_domain_device_bytes(const mi_domain* d) { return d ? d->bytes : MI_EINVAL; }

int mi_domain_set_field(mi_domain* d, int id, const double* host, int64_t n) {
  return mi_detail::guard([&] {
    if (!d || !host) throw Error(MI_EINVAL, "null argument");
    hipStream_t st = d->ctx->stream;
    const int nz = d->grid.nz;
    const bool base3 = id == MI_PB || id == MI_ALB || id == MI_PHB || id == MI_T_INIT;
    // a base 3-D profile on a domain with terrain: repeated over every column, halos included
    if (base3 && d->prm.terrain && n == nz + (id == MI_PHB ? 1 : 0)) {
      FieldDesc* fd = find(d, id);
      std::vector<double> h(static_cast<size_t>(d->n3), 0.0);
      for (int64_t e = 0; e < d->n3; ++e) h[size_t(e)] = host[std::min<int64_t>(e / d->a.g.sk, n - 1)];
      hip_check(hipMemcpyAsync(*fd->slot - d->off, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st),
                "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
      return;
    }
    // base 3-D profiles
    if (base3 && !d->prm.terrain) {
      const int nzs = nz + (id == MI_PHB ? 1 : 0);
      if (n != nzs) throw Error(MI_EINVAL, "base 3-D fields take a profile of nz (phb: nz+1) values");
      double* dst = id == MI_PB ? d->pb_buf : id == MI_ALB ? d->alb_buf : id == MI_PHB ? d->phb_buf : d->t_init_buf;
      hip_check(hipMemcpyAsync(dst, host, size_t(n) * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
      return;
    }
    FieldDesc* fd = find(d, id);
    if (!fd) throw Error(MI_EINVAL, "field id " + std::to_string(id) + " cannot be set");
    const bool base2 = id == MI_MUB || (id >= MI_MSFTX && id <= MI_HT);
    const int nxs = d->nxs(fd->st), nys = d->nys(fd->st), nzs = d->nzs(fd->st);
    const int64_t alloc_n = (fd->st == St::M2 || fd->st == St::U2 || fd->st == St::V2) ? d->n2 : d->n3;
    double* base = *fd->slot - d->off;
    if (base2 && n == 1) {  // horizontally uniform value, halos included
      mi_fp64::launch_fill(base, alloc_n, host[0], st);
      hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
      return;
    }
    if (n != int64_t(nxs) * nys * nzs) throw Error(MI_EINVAL, "field size mismatch");
    std::vector<double> h(static_cast<size_t>(alloc_n), 0.0);
    hip_check(hipMemcpyAsync(h.data(), base, h.size() * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    const Geo& G = d->a.g;
    for (int k = 0; k < nzs; ++k)
      for (int j = 0; j < nys; ++j)
        for (int i = 0; i < nxs; ++i) h[size_t(d->off + G.at(i, j, k))] = host[(int64_t(k) * nys + j) * nxs + i];
    hip_check(hipMemcpyAsync(base, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    d->halo({{*fd->slot, -1, -1}});
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  });
}

int mi_domain_get_field(mi_domain* d, int id, double* host, int64_t n) {
  return mi_detail::guard([&] {
    if (!d || !host) throw Error(MI_EINVAL, "null argument");
    FieldDesc* fd = find(d, id);
    if (!fd) throw Error(MI_EINVAL, "field id " + std::to_string(id) + " cannot be read");
    const int nxs = d->nxs(fd->st), nys = d->nys(fd->st), nzs = d->nzs(fd->st);
    if (n != int64_t(nxs) * nys * nzs) throw Error(MI_EINVAL, "field size mismatch");
    const int64_t alloc_n = (fd->st == St::M2 || fd->st == St::U2 || fd->st == St::V2) ? d->n2 : d->n3;
    std::vector<double> h(static_cast<size_t>(alloc_n));
    hipStream_t st = d->ctx->stream;
    hip_check(hipMemcpyAsync(h.data(), *fd->slot - d->off, h.size() * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    const Geo& G = d->a.g;
    for (int k = 0; k < nzs; ++k)
      for (int j = 0; j < nys; ++j)
        for (int i = 0; i < nxs; ++i) host[(int64_t(k) * nys + j) * nxs + i] = h[size_t(d->off + G.at(i, j, k))];
  });
}

int mi_step(mi_domain* d, int nsteps) {
  return mi_detail::guard([&] {
    if (!d) throw Error(MI_EINVAL, "null argument");
    if (nsteps < 0) throw Error(MI_EINVAL, "nsteps must be >= 0");
    hip_check(hipSetDevice(d->ctx->device), "hipSetDevice");
    d->step(nsteps);
  });
}

int mi_domain_sync(mi_domain* d) {
  return mi_detail::guard([&] {
    if (!d) throw Error(MI_EINVAL, "null argument");
    hip_check(hipStreamSynchronize(d->ctx->stream), "hipStreamSynchronize");
  });
}

int mi_step_begin(mi_domain* d) {
  return mi_detail::guard([&] {
    if (!d) throw Error(MI_EINVAL, "null argument");
    d->step_begin();
    hip_check(hipGetLastError(), "kernel launch");
  });
}
int mi_rk_prep(mi_domain* d, int rk) {
  return mi_detail::guard([&] {
    if (!d || rk < 1 || rk > 3) throw Error(MI_EINVAL, "bad argument");
    d->rk_prep();
    hip_check(hipGetLastError(), "kernel launch");
  });
}
int mi_rk_tendency(mi_domain* d, int rk) {
  return mi_detail::guard([&] {
    if (!d || rk < 1 || rk > 3) throw Error(MI_EINVAL, "bad argument");
    d->rk_tendency(rk);
    hip_check(hipGetLastError(), "kernel launch");
  });
}
int mi_small_steps(mi_domain* d, int rk) {
  return mi_detail::guard([&] {
    if (!d || rk < 1 || rk > 3) throw Error(MI_EINVAL, "bad argument");
    d->small_steps(rk, false);
    hip_check(hipGetLastError(), "kernel launch");
  });
}
int mi_scalar_transport(mi_domain* d, int rk) {
  return mi_detail::guard([&] {
    if (!d || rk < 1 || rk > 3) throw Error(MI_EINVAL, "bad argument");
    d->scalar_transport(rk);
    hip_check(hipGetLastError(), "kernel launch");
  });
}
int mi_calc_p_rho_phi(mi_domain* d, int rk) {
  return mi_detail::guard([&] {
    if (!d || rk < 1 || rk > 3) throw Error(MI_EINVAL, "bad argument");
    d->calc_p_rho_phi();
    hip_check(hipGetLastError(), "kernel launch");
  });
}

int mi_domain_autotune(mi_domain* d, int trials) {
  return mi_detail::guard([&] {
    if (!d) throw Error(MI_EINVAL, "null argument");
    hip_check(hipSetDevice(d->ctx->device), "hipSetDevice");
    d->autotune(trials);
  });
}

int mi_domain_set_tuning(mi_domain* d, const int* shapes, int n) {
  return mi_detail::guard([&] {
    if (!d || !shapes) throw Error(MI_EINVAL, "null argument");
    for (int q = 0; q < n && q < mi_fp64::kNumColKernels; ++q)
      if (!mi_fp64::col_cfg_fits(q, shapes[q], d->grid.nz)) throw Error(MI_EINVAL, "block shape does not fit this nz");
    for (int q = 0; q < n && q < mi_fp64::kNumColKernels; ++q) d->colcfg[q] = shapes[q];
    d->drop_graphs();
  });
}

int mi_domain_get_tuning(const mi_domain* d, int* shapes, int n) {
  return mi_detail::guard([&] {
    if (!d || !shapes) throw Error(MI_EINVAL, "null argument");
    for (int q = 0; q < n && q < mi_fp64::kNumColKernels; ++q) shapes[q] = d->colcfg[q];
  });
}

int mi_profile_kernel(mi_domain* d, const char* name) {
  return mi_detail::guard([&] {
    if (!d) throw Error(MI_EINVAL, "null argument");
    d->prof_name = name ? name : "";
    d->timer.used = 0;
    d->timer.bytes = 0;
  });
}

int mi_profile_read(mi_domain* d, double* total_ms, int* launches, double* alg_bytes) {
  return mi_detail::guard([&] {
    if (!d) throw Error(MI_EINVAL, "null argument");
    hip_check(hipStreamSynchronize(d->ctx->stream), "hipStreamSynchronize");
    double tot = 0;
    for (size_t q = 0; q + 1 < d->timer.used; q += 2) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, d->timer.ev[q], d->timer.ev[q + 1]), "hipEventElapsedTime");
      tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = int(d->timer.used / 2);
    if (alg_bytes) *alg_bytes = d->timer.bytes;
    d->timer.used = 0;
    d->timer.bytes = 0;
  });
}

int mi_check_finite(mi_domain* d, int* bad_field) {
  int rc = MI_OK;
  const int g = mi_detail::guard([&] {
    if (!d) throw Error(MI_EINVAL, "null argument");
    hipStream_t st = d->ctx->stream;
    hip_check(hipMemsetAsync(d->flags, 0, 64 * sizeof(int), st), "hipMemsetAsync");
    mi_fp64::launch_finite(d->a, d->flags, st);
    int h[64];
    hip_check(hipMemcpyAsync(h, d->flags, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    int bad = -1;
    for (int q = 0; q < 64 && bad < 0; ++q)
      if (h[q]) bad = q;
    if (bad_field) *bad_field = bad;
    if (bad >= 0) {
      mi_detail::set_last_error("non-finite value in field " + std::to_string(bad));
      rc = MI_ENONFINITE;
    }
  });
  return g != MI_OK ? g : rc;
}

}  // extern "C"
