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
#include <type_traits>
#include <vector>

#include "../../include/cppwrf_amd.h"
#include "domain_api.hpp"
#include "errors.hpp"
#include "kernels_fp64.hpp"

namespace MI_NS {
namespace {

using namespace mi_dev;
using mi_detail::Error;

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(e == hipErrorOutOfMemory ? MI_ENOMEM : MI_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

enum class St : char { M, U, V, W, M2, U2, V2 };

struct FieldDesc {
  void** slot;  // where the kernels' Fields keeps its pointer
  St st;
  int id;       // MI_* id or -1
  int esize;    // bytes per element: 8 (FP64) or 4 (FP32 storage)
};

// the parameters that follow the large step: w_damp's dt, the sixth-order coefficient's dt and, with specified
// boundaries, lbc_fcx_gcx (lateral_bc.cpp:198-211; Fortran loop = b + 1 over spec_zone + 1 .. spec_zone + relax_zone)
void set_dt_params(Params& q, double dt, bool w_damping, bool spec) {
  q.wd_dt = w_damping ? dt : 0.0;
  q.dt = dt;
  if (spec)
    for (int loop = 2; loop <= 5; ++loop) {
      const double num = double(1 + 4 - loop), den = double(4 - 1);
      q.fcx[loop - 1] = 0.1 / dt * num / den;
      q.gcx[loop - 1] = 1.0 / dt / 50.0 * num / den;
    }
}

}  // namespace

struct Domain final : mi_domain {
  mi_ctx* ctx = nullptr;
  mi_grid grid{};
  mi_params prm{};
  int n_sound_ = 0;
  Args a{};
  int64_t n3 = 0, n2 = 0;  // doubles per 3-D / 2-D allocation
  int64_t off = 0;         // offset of point (0,0,0) in an allocation
  int nxa = 0, nya = 0;
  std::vector<void*> allocs;
  int64_t bytes = 0;
  std::vector<FieldDesc> fields;
  double *pb_buf = nullptr, *alb_buf = nullptr, *phb_buf = nullptr, *t_init_buf = nullptr;
  int* flags = nullptr;  // device finiteness flags
  // the scalars and their buffers; qpar flips with every step (three stage swaps)
  Scalars sc{};
  int n_sc = 0, qpar = 0;
  // one graph per scalar-buffer parity: the step's launches name the buffers (none without scalars)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  // the domain's own stream for capturing its step: capture is a state of the stream, so work that another
  // thread enqueues on the context's stream meanwhile (a domain built on the same context) would be recorded
  // into this graph and replayed at every step, after that domain is gone too
  hipStream_t cap_stream = nullptr;
  bool use_graph = true;
  std::string prof_name;
  KernelTimer timer{};
  hipStream_t launch_stream = nullptr;  // while the step is captured: cap_stream, or a batch's branch stream
  hipStream_t stream() const { return launch_stream ? launch_stream : ctx->stream; }
  int colcfg[kNumColKernels] = {0, 0, 0, 0};  // block shape of each column kernel
  // specified boundaries: the data buffer of each variable (U V T PH MU, then the scalars) and which
  // scalars have data (the others take flow_dep_bdy)
  bool spec = false;
  double* bdy_buf[5 + MI_MAX_SCALARS] = {};
  int64_t bdy_len[5 + MI_MAX_SCALARS] = {};
  bool q_has_bdy[MI_MAX_SCALARS] = {};
  std::vector<char*> backup;                            // the carried state, for the autotuner

  // launches of fn, with the kernel timer active when a kernel is selected
  template <class F>
  void timed(F&& fn) {
    if (prof_name.empty()) {
      fn();
      return;
    }
    timer.name = prof_name.c_str();
    g_timer = &timer;
    try {
      fn();
    } catch (...) {
      g_timer = nullptr;
      throw;
    }
    g_timer = nullptr;
  }

  ~Domain() override {
    for (hipEvent_t e : timer.ev) (void)hipEventDestroy(e);
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (void* p : allocs) (void)hipFree(p);
  }
  int launch_version = 0;  // bumped whenever the step's launches change (shapes, boundary data present)
  void drop_graphs() {
    ++launch_version;
    for (int q = 0; q < 2; ++q) {
      if (exec[q]) (void)hipGraphExecDestroy(exec[q]);
      if (graph[q]) (void)hipGraphDestroy(graph[q]);
      exec[q] = nullptr;
      graph[q] = nullptr;
    }
  }

  template <class T = double>
  T* alloc(int64_t n) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, size_t(n) * sizeof(T)), "hipMalloc");
    allocs.push_back(p);
    bytes += n * int64_t(sizeof(T));
    hip_check(hipMemsetAsync(p, 0, size_t(n) * sizeof(T), ctx->stream), "hipMemsetAsync");
    return static_cast<T*>(p);
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

  const FieldDesc* desc_of(const void* p) const {
    for (const FieldDesc& f : fields)
      if (*f.slot == p) return &f;
    return nullptr;
  }

  // ------------------------------------------------------------------ halo batches
  struct HaloReq {
    void* p;
    int px, py;  // -1: default parity of the staggering
  };
  // The refresh of the halos of a list of fields (fill_halos, src/core/halo.cpp:170-205): one launch per
  // up to 16 fields; a symmetric OddFace field goes to a batch of its own, which takes the two-pass form.
  void halo(std::initializer_list<HaloReq> reqs) { halo(reqs.begin(), int(reqs.size())); }
  void halo(const HaloReq* reqs, int nreq) {
    HaloSet set[2]{};
    for (int q = 0; q < nreq; ++q) {
      const HaloReq& r = reqs[q];
      const FieldDesc* d = desc_of(r.p);
      if (!d) throw Error(MI_EINVAL, "internal: halo of an unknown field");
      const unsigned char px = (unsigned char)(r.px >= 0 ? r.px : (face_x(d->st) ? 1 : 0));
      const unsigned char py = (unsigned char)(r.py >= 0 ? r.py : (face_y(d->st) ? 1 : 0));
      const bool odd = (grid.bc_x == MI_BC_SYMMETRIC && face_x(d->st) && px == 2) ||
                       (grid.bc_y == MI_BC_SYMMETRIC && face_y(d->st) && py == 2);
      HaloSet& s = set[odd ? 1 : 0];
      if (s.n == kHaloMax) {
        launch_halo(a, s, stream());
        s = HaloSet{};
      }
      const int n = s.n++;
      s.p[n] = r.p;
      s.f32[n] = d->esize == 4 ? 1 : 0;
      s.nxs[n] = nxs(d->st);
      s.nys[n] = nys(d->st);
      s.nzs[n] = nzs(d->st);
      s.face_x[n] = face_x(d->st);
      s.face_y[n] = face_y(d->st);
      s.px[n] = px;
      s.py[n] = py;
      if (s.nzs[n] > s.nkmax) s.nkmax = s.nzs[n];
    }
    for (HaloSet& s : set)
      if (s.n > 0) launch_halo(a, s, stream());
  }
  // the halos of one scalar buffer of the species sp[0..n)
  void halo_scalars(real* const* bufs, const int* sp, int n) {
    HaloReq r[MI_MAX_SCALARS];
    for (int q = 0; q < n; ++q) r[q] = HaloReq{bufs[sp[q]], -1, -1};
    halo(r, n);
  }

  // ------------------------------------------------------------------ the phases of a step
  // RK3::step's _2 -> _1 copies (rk3.cpp:94-98).  The copies include the halos, which are fresh, so
  // the reference's refresh of mu_1's halo rewrites the values already there and is left out.
  //
  // With specified boundaries the step's clock advances first (rk3.cpp:93), and mu_1's halo is refreshed:
  // mu_2's is not fresh after spec_bdy_final (large_step.cpp:700-706), so the copy brings a stale one.
  void step_begin() {
    if (spec) launch_bdy_time(a, stream());
    launch_copy_step(a, n3, n2, stream());
    launch_copy_scalars(a, sc, n_sc, n3, stream());  // q_1 = q_2 (rk3.cpp:99)
    if (spec) halo({{a.f.mu1, -1, -1}});
  }
  // the boundary data of the dry variables (all) or of the scalars with data
  BdySet bdy_dry() const {
    BdySet s{};
    for (int v = 0; v < 5; ++v) s.v[s.n++] = BdyVar{v, 0, bdy_buf[v]};
    return s;
  }
  BdySet bdy_scalars() const {
    BdySet s{};
    for (int n = 0; n < n_sc; ++n)
      if (q_has_bdy[n]) s.v[s.n++] = BdyVar{kBdyQ, n, bdy_buf[5 + n]};
    return s;
  }

  // LargeStep::prep (large_step.cpp:697-733).  Its first halo refresh (u, v, w, t, ph, p, al, mu of the
  // stage) is left out: every path that changes those fields (the end of a stage, mi_domain_set_field)
  // refreshes their halos itself, so at this point it would rewrite identical values.  The kernel fuses
  // the reference's three launches, and its outputs' halos (mut, ru, rv, rw, alt, php, muu, muv, ww;
  // large_step.cpp:711, 719-723) are refreshed in one batch after it: no output of the fused kernel
  // is read at a halo point inside it.  The same holds for every fused kernel of the small steps below.
  //
  // With specified boundaries the first stage does refresh u, v, w, t, ph, p, al (not mu): spec_bdy_final
  // changed the spec zone after the last refresh.
  void rk_prep(int rk) {
    if (spec && rk == 1) halo({{a.f.u2, -1, -1}, {a.f.v2, -1, -1}, {a.f.w2, -1, -1}, {a.f.t2, -1, -1},
                               {a.f.ph2, -1, -1}, {a.f.p, -1, -1}, {a.f.al, -1, -1}});
    launch_large_prep(a, colcfg[kColLargePrep], stream());
    halo({{a.f.mut, -1, -1}, {a.f.ru, -1, -1}, {a.f.rv, -1, -1}, {a.f.rw, -1, -1}, {a.f.alt, -1, -1},
          {a.f.php, -1, -1}, {a.f.muu, 2, 0}, {a.f.muv, 0, 1}, {a.f.ww, -1, -1}});
    if (a.q.moist) launch_cq(a, sc, stream());  // calc_cq (large_step.cpp:724-733)
  }

  // LargeStep::tendencies (large_step.cpp:798-861); at the first stage, with a diffusion driver,
  // zero_tendf and the sixth-order diffusion into the held tendencies first
  //
  // With specified boundaries: mu_tend zeroed every stage (large_step.cpp:815-818), the relaxation of rows
  // 1..3 into the held tendencies at the first stage (:825) and the spec zone's tendencies last (:859).
  //
  // With a physics plugin (hooks on) the held tendencies were zeroed before part1; part2 sets the physics
  // tendencies first (large_step.cpp:820-821) and the diffusion adds to them.
  void rk_tendency(int rk, bool hooks = true) {
    if (spec) launch_fill2d(a, a.f.mu_tend - off, n2, 0.0, stream());
    if (rk == 1 && a.q.karr) {  // DiffusionDriver::prepare (rk3.cpp:113): the step's K, then its halos
      if (a.q.km4) launch_smag2d(a, stream());
      if (a.q.damp_opt == 1) launch_dampkm(a, stream());
      halo({{a.f.xkmh, -1, -1}, {a.f.xkhh, -1, -1}});
    }
    if (rk == 1 && a.q.use_tendf) {
      if (phys_on(hooks)) {
        call_phys(phys.part2);
        launch_sixth_dyn(a, kSixthKeep, stream());
      } else {
        launch_sixth_dyn(a, kSixthZero, stream());
      }
      if (a.q.damp_opt == 2) launch_rayleigh(a, stream());  // the last term of DiffusionDriver::dynamics
    }
    if (rk == 1 && spec) launch_bdy(a, sc, bdy_dry(), kBdyRelax, stream());
    launch_tendencies(a, stream());
    if (spec) launch_bdy(a, sc, bdy_dry(), kBdySpec, stream());
  }

  // SmallStep::run_stage (small_step.cpp:490-607); with prhophi, calc_p_rho_phi (small_step.cpp:609-617)
  // is fused into the finishing kernel.  In four parts, which a batch interleaves with its members' (ss_begin, then
  // per substep advance_uv, after_uv, the column kernel, after_acoustic, then ss_end).
  int stage_ns(int rk) const { return rk == 1 ? 1 : (rk == 2 ? n_sound_ / 2 : n_sound_); }
  double stage_dts(int rk) const { return rk == 1 ? grid.dt / 3.0 : grid.dt / n_sound_; }
  void small_steps(int rk, bool prhophi) {
    const double dts = stage_dts(rk);
    const int ns = stage_ns(rk);
    ss_begin(rk);
    for (int it = 1; it <= ns; ++it) {
      launch_uv(a, dts, stream());
      after_uv();
      launch_acoustic(a, colcfg[kColAcoustic], dts, it, ns, stream());
      after_acoustic(rk, it);
    }
    ss_end(rk, prhophi);
  }
  void ss_begin(int rk) {
    auto& f = a.f;
    const double dts = stage_dts(rk);
    hipStream_t st = stream();
    if (rk == 1) launch_fill2d(a, f.mudf - off, n2, 0.0, st);
    launch_ss_prep(a, rk == 1, st);
    const double cof_h = 0.5 * dts * kG * (1.0 + prm.epssm);
    launch_ss_column(a, colcfg[kColSsColumn], cof_h * cof_h, dts, st);
    // the halos of u'', v'', ph'', t_save, u_save, v_save, mu'', muts, p, al (small_step.cpp:517-522), in
    // one batch after the two kernels: the column kernel reads its inputs in its own column only.  With
    // specified boundaries no substep kernel reads them (as the substep refreshes below)
    if (!spec)
      halo({{f.u2, -1, -1}, {f.v2, -1, -1}, {f.ph2, -1, -1}, {f.t_save, -1, -1}, {f.u_save, -1, -1},
            {f.v_save, -1, -1}, {f.mu2, -1, -1}, {f.muts, -1, -1}, {f.p, -1, -1}, {f.al, -1, -1}});
  }
  // With specified boundaries the substep refreshes below are left out: no kernel of the substeps reads those halo
  // points (the spec zone's faces and columns are advanced from their own values, the interior's stencils reach
  // the spec zone only), and the stage's last refreshes write them before any reader.
  bool needs_after_uv() const { return !spec; }
  void after_uv() {
    auto& f = a.f;
    if (spec) {
    } else if (!a.g.two_d) {
      halo({{f.u2, -1, -1}, {f.v2, -1, -1}});  // small_step.cpp:540-541
    } else {
      halo({{f.u2, -1, -1}});
    }
  }
  // after the substep's column kernel: the spec zone's columns, and the refreshes of mu'', mudf, muts and of ph'', p,
  // al (small_step.cpp:572, 593)
  void after_acoustic(int rk, int it) {
    auto& f = a.f;
    if (spec) launch_acoustic_ring(a, stage_dts(rk), it, stage_ns(rk), stream());
    if (!spec)
      halo({{f.mu2, -1, -1}, {f.mudf, -1, -1}, {f.muts, -1, -1}, {f.ph2, -1, -1}, {f.p, -1, -1}, {f.al, -1, -1}});
  }
  void ss_end(int rk, bool prhophi) {
    auto& f = a.f;
    hipStream_t st = stream();
    const double dts = stage_dts(rk);
    const int ns = stage_ns(rk);
    if (spec) launch_bdy_zero_grad(a, st);  // zero_grad_bdy of the spec zone's w (lateral_bc.cpp:374-387)
    halo({{f.muts, -1, -1}, {f.ru_m, -1, -1}, {f.rv_m, -1, -1}, {f.ww_m, -1, -1}});  // small_step.cpp:597
    // the microphysics heating leaves theta_m at the last stage with moisture (small_step.cpp:602)
    const double hd_dt = (rk == 3 && a.q.moist) ? dts * double(ns) : 0.0;
    launch_finish(a, colcfg[kColFinish], prhophi, hd_dt, st);
    // small_step.cpp:606 (and 616: p, al of the fused calc_p_rho_phi)
    if (prhophi)
      halo({{f.u2, -1, -1}, {f.v2, -1, -1}, {f.w2, -1, -1}, {f.t2, -1, -1}, {f.ph2, -1, -1}, {f.ww, -1, -1},
            {f.mu2, -1, -1}, {f.p, -1, -1}, {f.al, -1, -1}});
    else
      halo({{f.u2, -1, -1}, {f.v2, -1, -1}, {f.w2, -1, -1}, {f.t2, -1, -1}, {f.ph2, -1, -1}, {f.ww, -1, -1},
            {f.mu2, -1, -1}});
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
    // rk_scalar_tend at the first stage: the step's tendency of every scalar, sixth-order diffusion and the
    // boundary relaxation included (scalars.cpp:446-455); the spec zone's tendency every stage
    if (rk == 1 && a.q.use_tendf) launch_sixth_scalars(a, sc, n_sc, dt_rk, stream());
    if (rk == 1 && spec) launch_bdy(a, sc, bdy_scalars(), kBdyRelax, stream());
    if (spec) launch_bdy(a, sc, bdy_scalars(), kBdySpec, stream());
    launch_scalars(a, sc, plain, np, 0, dt_rk, stream());
    if (nd > 0) {  // scalars.cpp:375-418, with the halo refreshes of q_1 and of the limiter's scale
      launch_pd_geom(a, dt_rk, stream());
      launch_scalars(a, sc, pd, nd, 1, dt_rk, stream());
      halo_scalars(sc.q1, pd, nd);
      launch_scalars(a, sc, pd, nd, 2, dt_rk, stream());
      halo_scalars(sc.scale, pd, nd);
      launch_scalars(a, sc, pd, nd, 3, dt_rk, stream());
    }
    for (int n = 0; n < n_sc; ++n) std::swap(sc.q2[n], sc.qn[n]);
    qpar ^= 1;
    int all[MI_MAX_SCALARS];
    for (int n = 0; n < n_sc; ++n) all[n] = n;
    if (spec) {  // flow_dep_bdy of the species without data (scalars.cpp:420-429)
      int fd[MI_MAX_SCALARS], nf = 0;
      for (int n = 0; n < n_sc; ++n)
        if (!q_has_bdy[n]) fd[nf++] = n;
      launch_bdy_flow_dep(a, sc, fd, nf, stream());
    }
    halo_scalars(sc.q2, all, n_sc);  // scalars.cpp:458-460
  }

  // SmallStep::calc_p_rho_phi_interior (small_step.cpp:619-633): the columns spec_zone in from each non-periodic
  // side; with the scalar pow every point gets the bits calc_p_rho_phi gave it (the reference's differences
  // come from libmvec's lanes only)
  void calc_p_rho_phi_interior(int sz) override {
    if (sz < 0) throw Error(MI_EINVAL, "spec_zone must be >= 0");
    const bool px = grid.bc_x == MI_BC_PERIODIC, py = grid.bc_y == MI_BC_PERIODIC || a.g.two_d;
    launch_p_rho_phi_range(a, px ? 0 : sz, px ? grid.nx : grid.nx - sz, py ? 0 : sz, py ? grid.ny : grid.ny - sz, stream());
    halo({{a.f.p, -1, -1}, {a.f.al, -1, -1}});
    hip_check(hipGetLastError(), "kernel launch");
  }
  // SmallStep::calc_p_rho_phi (small_step.cpp:609-617), for the phase-by-phase entry points
  void calc_p_rho_phi() {
    launch_p_rho_phi(a, stream());
    halo({{a.f.p, -1, -1}, {a.f.al, -1, -1}});
  }

  // the end of RK3::step (rk3.cpp:159-167): spec_bdy_final, then the surface w of terrain, which refreshes
  // the u, v halos first (set_w_surface, rk3.cpp:194; those refreshes change nothing without a
  // LateralBC).  The reference also reruns calc_p_rho_phi on the interior after spec_bdy_final: its
  // inputs there are unchanged and the scalar pow gives back the same bits, so it is left out.
  // PhysicsCoupling::step after the RK loop (rk3.cpp:151-158): the microphysics, the halos of what it changed
  // (physics_coupling.cpp:450-455), then calc_p_rho_phi, without the spec zone with specified boundaries
  void microphysics() {
    launch_microphysics(a, sc, stream());
    std::vector<HaloReq> hs = {{a.f.t2, -1, -1}, {a.f.hdia, -1, -1}};
    for (int n = 0; n < prm.n_moist; ++n) hs.push_back({sc.q2[n], -1, -1});
    halo(hs.data(), int(hs.size()));
    if (spec) calc_p_rho_phi_interior(1);
    else calc_p_rho_phi();
  }
  void step_end(bool hooks = true) {
    if (phys_on(hooks)) call_phys(phys.after_rk);  // phy_prep_part2 (rk3.cpp:150)
    if (a.q.mp) microphysics();
    step_end_dyn();
    if (phys_on(hooks)) call_phys(phys.after_all_rk_steps);  // rk3.cpp:166
  }
  void step_end_dyn() {
    if (spec) {
      launch_bdy(a, sc, bdy_dry(), kBdyFinal, stream());
      launch_bdy(a, sc, bdy_scalars(), kBdyFinal, stream());
    }
    if (a.q.terrain) {
      if (spec) halo({{a.f.u2, -1, -1}, {a.f.v2, -1, -1}});
      launch_set_w_surface(a, stream());
      halo({{a.f.w2, -1, -1}});  // rk3.cpp:199
    }
  }

  // RK3::step (rk3.cpp:86-170); with callbacks (host-driven steps only) the stage hooks of the reference:
  // on_first_rk_prep after stage 1's prep (rk3.cpp:105), on_cfl after each stage's tendencies (rk3.cpp:115-119)
  // A physics plugin's hooks run with the other hooks only: the autotuner's trial steps and graph captures
  // (which a domain with a plugin never replays) leave them out.
  void step_launches(bool hooks = false) {
    step_begin();
    for (int rk = 1; rk <= 3; ++rk) {
      rk_prep(rk);
      if (hooks && rk == 1 && prep_fn) {
        hip_check(hipStreamSynchronize(stream()), "hipStreamSynchronize");
        prep_fn(prep_user);
      }
      if (rk == 1) physics_part1(hooks);
      rk_tendency(rk, hooks);
      if (hooks && cfl_fn) {
        double vert = 0, horiz = 0;
        cfl_max(&vert, &horiz);
        cfl_fn(cfl_user, vert, horiz);
      }
      small_steps(rk, true);
      scalar_transport(rk);
    }
    step_end(hooks);
  }
  // ------------------------------------------------------------------ column physics plugin
  mi_column_physics phys{};
  bool has_phys = false;
  bool held_base = false;  // held tendencies without a plugin: a diffusion or specified boundaries
  bool phys_on(bool hooks) const { return hooks && has_phys; }
  void call_phys(mi_phys_fn fn) {
    if (!fn) return;
    hip_check(hipStreamSynchronize(stream()), "hipStreamSynchronize");
    fn(phys.user, this);
    hip_check(hipStreamSynchronize(stream()), "hipStreamSynchronize");
  }
  // first_rk_step_part1 (rk3.cpp:106-111): init_zero_tendency of every held tendency, scalars included, then part1
  void physics_part1(bool hooks = true) {
    if (!phys_on(hooks)) return;
    launch_sixth_dyn(a, kSixthZeroOnly, stream());  // ru, rv, rw, t_tendf (and ph_tendf)
    for (int n = 0; n < n_sc; ++n)
      hip_check(hipMemsetAsync(sc.qtf[n] - off, 0, size_t(n3) * sizeof(real), stream()), "hipMemsetAsync");
    call_phys(phys.part1);
  }
  void set_column_physics(const mi_column_physics* cp) override {
    if (cp && cp->qv_forcing && prm.n_moist == 0) throw Error(MI_EINVAL, "qv_forcing needs moisture (n_moist > 0)");
    if (cp) ensure_held();
    if (cp && cp->th_forcing && !thf_buf) thf_buf = add3(thf_buf, St::M, MI_TH_FORCING);
    if (cp && cp->qv_forcing && !qvf_buf) qvf_buf = add3(qvf_buf, St::M, MI_QV_FORCING);
    has_phys = cp != nullptr;
    a.q.qtf_on = has_phys ? 1 : 0;
    phys = cp ? *cp : mi_column_physics{};
    a.q.use_tendf = (held_base || has_phys) ? 1 : 0;
    if (a.q.use_tendf) {  // every buffer the kernels read or write with held tendencies must exist
      bool ok = a.f.ru_tendf && a.f.rv_tendf && a.f.rw_tendf && a.f.t_tendf;
      for (int n = 0; n < n_sc; ++n) ok = ok && sc.qtw[n];
      if (!ok) throw Error(MI_EHIP, "held tendencies not allocated");
    }
    a.f.thf = (cp && cp->th_forcing) ? thf_buf : nullptr;
    a.f.qvf = (cp && cp->qv_forcing) ? qvf_buf : nullptr;
    drop_graphs();
  }
  real* thf_buf = nullptr;
  real* qvf_buf = nullptr;
  // the held tendencies and the scalars' step tendencies, allocated once they are first needed
  void ensure_held() {
    if (!a.f.ru_tendf) {
      add3(a.f.ru_tendf, St::U, MI_RU_TENDF);
      add3(a.f.rv_tendf, St::V, MI_RV_TENDF);
      add3(a.f.rw_tendf, St::W, MI_RW_TENDF);
      add3(a.f.t_tendf, St::M, MI_T_TENDF);
    }
    for (int n = 0; n < n_sc; ++n)
      if (!sc.qtw[n]) add3(sc.qtw[n], St::M, -1);
  }
  void set_diffusion(int opt, double factor) override {
    if (opt < 0 || opt > 2) throw Error(MI_EINVAL, "diff_6th_opt must be 0, 1 or 2");
    if (opt != 0) ensure_held();
    prm.diff_6th_opt = opt;
    prm.diff_6th_factor = factor;
    a.q.diff_6th_opt = opt;
    a.q.diff_6th_factor = factor;
    held_base = opt != 0 || prm.diff_opt != 0 || prm.damp_opt == 2 || spec;
    a.q.use_tendf = (held_base || has_phys) ? 1 : 0;
    drop_graphs();
  }
  template <class T>
  T* add3(T*& slot, St s, int id) {
    slot = alloc<T>(n3) + off;
    fields.push_back(FieldDesc{reinterpret_cast<void**>(&slot), s, id, int(sizeof(T))});
    return slot;
  }
  mi_cfl_fn cfl_fn = nullptr;
  void* cfl_user = nullptr;
  mi_stage_fn prep_fn = nullptr;
  void* prep_user = nullptr;
  void set_cfl_callback(mi_cfl_fn fn, void* user) override {
    cfl_fn = fn;
    cfl_user = user;
  }
  void set_prep_callback(mi_stage_fn fn, void* user) override {
    prep_fn = fn;
    prep_user = user;
  }
  bool has_callbacks() const override { return cfl_fn != nullptr || prep_fn != nullptr || has_phys; }
  // RK3::set_dt with LateralBC::set_dt and the diffusion's dt (the adaptive time step)
  void set_dt(double dt, int ns) override {
    if (!(dt > 0)) throw Error(MI_EINVAL, "dt must be > 0");
    if (ns < 2 || ns % 2) throw Error(MI_EINVAL, "time_step_sound must be even and >= 2");
    grid.dt = dt;
    n_sound_ = ns;
    set_dt_params(a.q, dt, prm.w_damping == 1, spec);
    drop_graphs();
  }

  // the fields a step reads from the previous one (everything else is recomputed within a step)
  struct Carried {
    void* p;        // point (0,0,0)
    int64_t bytes;  // the whole allocation
    int64_t esize;  // its element size
  };
  // the state a step carries to the next: the prognostic fields, the scalars and, with the microphysics, the
  // heating it leaves for the next step's dynamics and the accumulated precipitation
  std::vector<Carried> carried() {
    const int64_t e3 = int64_t(sizeof(real)), b3 = n3 * e3, b2 = n2 * int64_t(sizeof(double));
    std::vector<Carried> c = {{a.f.u2, b3, e3}, {a.f.v2, b3, e3}, {a.f.w2, b3, e3}, {a.f.t2, b3, e3},
                              {a.f.ph2, b3, e3}, {a.f.mu2, b2, 8}, {a.f.p, b3, e3}, {a.f.al, b3, e3}};
    for (int n = 0; n < n_sc; ++n) c.push_back({sc.q2[n], b3, e3});
    if (a.q.mp) {
      c.push_back({a.f.hdia, b3, e3});
      c.push_back({a.f.rainnc, b2, 8});
    }
    return c;
  }
  Scalars sc_saved{};
  int qpar_saved = 0;
  void save_or_restore(bool save) {
    if (save) {
      sc_saved = sc;
      qpar_saved = qpar;
    } else {  // the buffers as they were named at the save
      sc = sc_saved;
      qpar = qpar_saved;
    }
    std::vector<Carried> c = carried();
    if (backup.empty())
      for (const Carried& x : c) backup.push_back(alloc<char>(x.bytes));
    for (size_t q = 0; q < c.size(); ++q) {
      char* live = static_cast<char*>(c[q].p) - off * c[q].esize;
      hip_check(hipMemcpyAsync(save ? backup[q] : live, save ? live : backup[q], size_t(c[q].bytes),
                               hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpyAsync");
    }
  }

  // ------------------------------------------------------------------ checkpoint and resume
  // One host buffer: a header that names the domain's shape and storage, then every carried allocation (halos
  // included) in carried() order, then the boundary clock.
  struct StateHeader {
    char magic[8];
    int32_t version, nx, ny, nz, real_bytes, n_sc, n_fields, spec;
    int64_t n3, n2, payload;
  };
  int64_t state_bytes() override {
    int64_t b = int64_t(sizeof(StateHeader)) + int64_t(sizeof(double));
    for (const Carried& c : carried()) b += c.bytes;
    return b;
  }
  StateHeader state_header() {
    StateHeader h{};
    const char m[8] = {'M', 'I', 'S', 'T', 'A', 'T', 'E', '1'};
    for (int q = 0; q < 8; ++q) h.magic[q] = m[q];
    h.version = 1;
    h.nx = grid.nx;
    h.ny = grid.ny;
    h.nz = grid.nz;
    h.real_bytes = int32_t(sizeof(real));
    h.n_sc = n_sc;
    h.n_fields = int32_t(carried().size());
    h.spec = spec ? 1 : 0;
    h.n3 = n3;
    h.n2 = n2;
    h.payload = state_bytes();
    return h;
  }
  char* carried_base(const Carried& c) { return static_cast<char*>(c.p) - off * c.esize; }
  void save_state(void* host, int64_t bytes) override {
    if (!host || bytes < state_bytes()) throw Error(MI_EINVAL, "state buffer too small: mi_domain_state_bytes gives the size");
    const StateHeader h = state_header();
    char* o = static_cast<char*>(host);
    std::memcpy(o, &h, sizeof(h));
    o += sizeof(h);
    for (const Carried& c : carried()) {
      hip_check(hipMemcpyAsync(o, carried_base(c), size_t(c.bytes), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
      o += c.bytes;
    }
    double t = spec ? bdy_get_time() : 0.0;
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    std::memcpy(o, &t, sizeof(t));
  }
  void load_state(const void* host, int64_t bytes) override {
    if (!host || bytes < int64_t(sizeof(StateHeader))) throw Error(MI_EINVAL, "state buffer too small");
    StateHeader h{};
    std::memcpy(&h, host, sizeof(h));
    const StateHeader mine = state_header();
    if (std::memcmp(h.magic, mine.magic, 8) != 0 || h.version != mine.version)
      throw Error(MI_EINVAL, "not a state buffer of this library");
    if (h.nx != mine.nx || h.ny != mine.ny || h.nz != mine.nz || h.real_bytes != mine.real_bytes || h.n_sc != mine.n_sc ||
        h.n_fields != mine.n_fields || h.spec != mine.spec || h.n3 != mine.n3 || h.n2 != mine.n2)
      throw Error(MI_EINVAL, "the state buffer belongs to a domain of another shape, precision or scalar count");
    if (h.payload != mine.payload) throw Error(MI_EINVAL, "the state buffer's header does not match this domain");
    if (bytes < h.payload) throw Error(MI_EINVAL, "state buffer truncated");
    const char* o = static_cast<const char*>(host) + sizeof(h);
    for (const Carried& c : carried()) {
      hip_check(hipMemcpyAsync(carried_base(c), o, size_t(c.bytes), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
      o += c.bytes;
    }
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    double t = 0;
    std::memcpy(&t, o, sizeof(t));
    if (spec) bdy_set_time(t);
  }

  // Time every block shape of each column kernel on this domain and keep the fastest; the state is
  // restored afterwards, so the forecast continues exactly as if no step had been taken.
  void autotune(int trials) override {
    if (trials < 1) trials = 1;
    const std::string keep = prof_name;
    const double dtbc = spec ? bdy_get_time() : 0.0;
    save_or_restore(true);
    for (int s = 0; s < 1; ++s) timed([&] { step_launches(); });  // warm-up
    for (int kern = 0; kern < kNumColKernels; ++kern) {
      double best_t = 1e300;
      int best = colcfg[kern];
      for (int cfg = 0; cfg < kNumColShapes; ++cfg) {
        if (!col_cfg_fits(kern, cfg, grid.nz)) continue;
        colcfg[kern] = cfg;
        prof_name = kColKernelNames[kern];
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
    if (spec) bdy_set_time(dtbc);
    drop_graphs();  // the graphs hold the old shapes
  }

  void step(int nsteps) override {
    if (nsteps <= 0) return;
    if (has_callbacks()) {
      for (int s = 0; s < nsteps; ++s) timed([&] { step_launches(true); });
      hip_check(hipGetLastError(), "kernel launch");
      return;
    }
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
        const Scalars keep = sc;
        if (!cap_stream)
          hip_check(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking), "hipStreamCreate");  // default priority
        hip_check(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        launch_stream = cap_stream;
        try {
          step_launches();
        } catch (...) {  // end the capture, so the stream takes launches again, and put the host state back
          launch_stream = nullptr;
          hipGraph_t junk = nullptr;
          (void)hipStreamEndCapture(cap_stream, &junk);
          if (junk) (void)hipGraphDestroy(junk);
          sc = keep;
          qpar = par;
          throw;
        }
        launch_stream = nullptr;
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(cap_stream, &gr);
        sc = keep;
        qpar = par;
        hip_check(ec, "hipStreamEndCapture");
        graph[par] = gr;
        hip_check(hipGraphInstantiate(&exec[par], gr, nullptr, nullptr, 0), "hipGraphInstantiate");
      }
      hip_check(hipGraphLaunch(exec[par], ctx->stream), "hipGraphLaunch");
      // the host side of the step: three stage swaps of every species' buffers
      for (int r = 0; r < (n_sc > 0 ? 3 : 0); ++r) {
        for (int n = 0; n < n_sc; ++n) std::swap(sc.q2[n], sc.qn[n]);
        qpar ^= 1;
      }
    }
  }

  // ------------------------------------------------------------------ the interface behind the C ABI
  mi_ctx* context() const override { return ctx; }
  int n_sound() const override { return n_sound_; }
  int64_t device_bytes() const override { return bytes; }
  double* base_profile(int id) const;
  void set_field(int id, const double* host, int64_t n) override;
  void get_field(int id, double* host, int64_t n) override;
  void field_view(int id, mi_field_view* out) override;
  void field_layout(int id, mi_field_layout* out) const override;
  void field_bind(int id, void* dev, int64_t bytes) override;
  void phase(MiPhase which, int rk) override;
  void set_tuning(const int* shapes, int n) override;
  void get_tuning(int* shapes, int n) const override;
  void profile_kernel(const char* name) override;
  void profile_read(double* total_ms, int* launches, double* alg_bytes) override;
  int check_finite(int* bad_field) override;
  void cfl_max(double* vert, double* horiz) override;
  void cfl_max_noalloc(double* vert, double* horiz) override;
  void bdy_set(int var, int side, const double* values, const double* tends, int64_t n) override;
  void bdy_set_time(double dtbc) override {
    if (!spec) throw Error(MI_EINVAL, "the domain has no specified boundaries");
    hip_check(hipMemcpyAsync(a.f.bdy_time, &dtbc, sizeof(double), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  }
  double bdy_get_time() override {
    if (!spec) throw Error(MI_EINVAL, "the domain has no specified boundaries");
    double t = 0;
    hip_check(hipMemcpyAsync(&t, a.f.bdy_time, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return t;
  }
  // a batch: this domain's step launched on stream s (under capture), the host state left as it was
  void capture_step(hipStream_t s) override {
    const Scalars keep = sc;
    const int kq = qpar;
    launch_stream = s;
    try {
      step_launches();
    } catch (...) {
      launch_stream = nullptr;
      sc = keep;
      qpar = kq;
      throw;
    }
    launch_stream = nullptr;
    sc = keep;
    qpar = kq;
  }
  // the host side of one step run elsewhere (a batch graph): three stage swaps of the scalar buffers
  void advance_host() override {
    for (int r = 0; r < (n_sc > 0 ? 3 : 0); ++r) {
      for (int n = 0; n < n_sc; ++n) std::swap(sc.q2[n], sc.qn[n]);
      qpar ^= 1;
    }
  }
  // the key of a batch's captured step: the scalar-buffer parity and the version of the launch schedule
  int buffer_parity() const override { return qpar + 2 * launch_version; }
  int storage_bytes() const override { return int(sizeof(real)); }
  bool fusable() const override { return !has_callbacks() && batch_fits(a); }
};

namespace {

void build_domain(Domain& d, mi_ctx* ctx, const mi_grid& g, const mi_params& p, const mi_levels& lv) {
  if (g.nx < 1 || g.ny < 1 || g.nz < 3) throw Error(MI_EINVAL, "grid: need nx >= 1, ny >= 1, nz >= 3");
  if (g.bc_x < MI_BC_PERIODIC || g.bc_x > MI_BC_OPEN) throw Error(MI_EINVAL, "bc_x: periodic, symmetric, specified or open");
  if (g.bc_y < MI_BC_PERIODIC || g.bc_y > MI_BC_OPEN) throw Error(MI_EINVAL, "bc_y: periodic, symmetric, specified or open");
  if ((g.bc_x == MI_BC_SPECIFIED) != (g.bc_y == MI_BC_SPECIFIED))
    throw Error(MI_EINVAL, "specified boundaries on both x and y or on neither");
  // the relaxation zone of 5 rows on each side, and the 3-row margin of the sixth-order diffusion
  if (g.bc_x == MI_BC_SPECIFIED && (g.ny == 1 || g.nx < 10 || g.ny < 10))
    throw Error(MI_EINVAL, "specified boundaries need a 3-D grid of at least 10 x 10 columns");
  // an open side degrades the stencils over its 3 outer rows (the radiation face, 2nd, 4th order)
  if ((g.bc_x == MI_BC_OPEN && g.nx < 7) || (g.ny > 1 && g.bc_y == MI_BC_OPEN && g.ny < 7))
    throw Error(MI_EINVAL, "an open direction needs at least 7 columns");
  if (g.bc_x == MI_BC_PERIODIC && g.nx < 3) throw Error(MI_EINVAL, "periodic x needs nx >= 3");
  if (g.ny > 1 && g.bc_y == MI_BC_PERIODIC && g.ny < 3) throw Error(MI_EINVAL, "periodic y needs ny >= 3");
  // a symmetric fill with fewer than 4 points mirrors halo slots into one another (order-dependent)
  if (g.bc_x == MI_BC_SYMMETRIC && g.nx < 4) throw Error(MI_EINVAL, "symmetric x needs nx >= 4");
  if (g.ny > 1 && g.bc_y == MI_BC_SYMMETRIC && g.ny < 4) throw Error(MI_EINVAL, "symmetric y needs ny >= 4");
  if (!(g.dx > 0) || !(g.dt > 0) || (g.ny > 1 && !(g.dy > 0))) throw Error(MI_EINVAL, "grid: dx, dy, dt must be > 0");
  if (p.h_adv_order != 5 && p.h_adv_order != 6) throw Error(MI_EINVAL, "h_adv_order must be 5 or 6");
  if (p.v_adv_order != 3 && p.v_adv_order != 4) throw Error(MI_EINVAL, "v_adv_order must be 3 or 4");
  if (p.damp_opt < 0 || p.damp_opt > 3) throw Error(MI_EINVAL, "damp_opt must be 0, 1, 2 or 3");
  if (p.hypsometric_opt != 1) throw Error(MI_EINVAL, "hypsometric_opt must be 1");
  if (p.precision != MI_FP64 && p.precision != MI_FP32_BULK) throw Error(MI_EINVAL, "precision: MI_FP64 or MI_FP32_BULK");
  if (p.damp_opt != 0 && !(p.zdamp > 0)) throw Error(MI_EINVAL, "zdamp must be > 0");
  if (p.damp_opt == 2 && g.nz < 2) throw Error(MI_EINVAL, "damp_opt 2 needs nz >= 2");
  if (p.n_moist < 0 || p.n_moist > MI_MAX_MOIST) throw Error(MI_EINVAL, "n_moist must be in 0..8");
  if (p.n_scalars < p.n_moist || p.n_scalars > MI_MAX_SCALARS) throw Error(MI_EINVAL, "n_scalars must be in n_moist..16");
  if ((p.moist_adv_opt != 0 && p.moist_adv_opt != 1) || (p.scalar_adv_opt != 0 && p.scalar_adv_opt != 1))
    throw Error(MI_EINVAL, "moist_adv_opt and scalar_adv_opt must be 0 or 1");
  if ((p.mapped != 0 && p.mapped != 1) || (p.coriolis != 0 && p.coriolis != 1))
    throw Error(MI_EINVAL, "mapped and coriolis must be 0 or 1");
  if (p.diff_6th_opt < 0 || p.diff_6th_opt > 2) throw Error(MI_EINVAL, "diff_6th_opt must be 0, 1 or 2");
  if (p.terrain != 0 && p.terrain != 1) throw Error(MI_EINVAL, "terrain must be 0 or 1");
  if (p.diff_opt == 2) throw Error(MI_EINVAL, "diff_opt 2 (diffusion in physical space) is not built");
  if (p.diff_opt < 0 || p.diff_opt > 1) throw Error(MI_EINVAL, "diff_opt must be 0 or 1");
  if (p.diff_opt == 1 && p.km_opt != 1 && p.km_opt != 4)
    throw Error(MI_EINVAL, "km_opt must be 1 (constant K) or 4 (2-D Smagorinsky); the TKE and 3-D Smagorinsky "
                           "options 2 and 3 are not built");
  if (p.diff_opt == 1 && !(p.khdif >= 0 && p.kvdif >= 0 && p.pr_inv >= 0 && p.c_s >= 0))
    throw Error(MI_EINVAL, "khdif, kvdif, pr_inv and c_s must be >= 0");
  if (p.diff_opt == 1 && p.km_opt == 4 && g.nz < 3) throw Error(MI_EINVAL, "km_opt 4 needs nz >= 3");
  if (p.mp_physics != 0 && p.mp_physics != 1)
    throw Error(MI_EINVAL, "mp_physics must be 0 or 1 (Kessler); the reference's ice schemes are not built");
  if (p.mp_physics == 1 && (p.n_moist < 3 || g.nz > kKesslerMaxLevels || g.nz < 2))
    throw Error(MI_EINVAL, "Kessler needs the moist species qv, qc, qr (n_moist >= 3) and nz <= 128");
  d.ctx = ctx;
  d.grid = g;
  d.prm = p;
  d.n_sound_ = p.time_step_sound > 0 ? p.time_step_sound : mi_auto_sound_steps(g.dt, g.dx, g.dy);
  if (d.n_sound_ < 2 || d.n_sound_ % 2) throw Error(MI_EINVAL, "time_step_sound must be even and >= 2");
  if (!col_cfg_fits(kColAcoustic, kColCfgTallest, g.nz)) throw Error(MI_EINVAL, "nz must be at most 255");
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
  // i fastest, then k, then j (WRF's i-k-j memory order): the nz+1 rows of one column group are adjacent, so a
  // column kernel's wavefront (a few columns at several levels) touches one page per array, not one per level
  const int64_t pitch = (int64_t(kLpad) + d.nxa + kHalo + 31) / 32 * 32;
  const int64_t rows = d.nya + 2 * G.hy;
  G.sk = pitch;
  G.sj = two_d ? 0 : pitch * int64_t(g.nz + 1);
  d.off = int64_t(kLpad) + int64_t(G.hy) * G.sj;
  d.n3 = pitch * int64_t(g.nz + 1) * rows;
  d.n2 = d.n3;  // a 2-D field is indexed with the same j stride: it uses one row of each j's block

  auto& f = d.a.f;
  // 3-D fields of the storage type of their Fields member (real, or double for the FP64 islands)
  auto add3 = [&](auto*& slot, St s, int id) {
    using T = std::remove_reference_t<decltype(*slot)>;
    slot = d.template alloc<T>(d.n3) + d.off;
    d.fields.push_back(FieldDesc{reinterpret_cast<void**>(&slot), s, id, int(sizeof(T))});
  };
  auto add2 = [&](double*& slot, St s, int id) {
    slot = d.alloc(d.n2) + d.off;
    d.fields.push_back(FieldDesc{reinterpret_cast<void**>(&slot), s, id, int(sizeof(double))});
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
  add3(f.ky, St::W, -1);  // the scratch of the substep k-march
  add3(f.kr, St::W, -1);
  add2(f.mu_save, St::M2, -1); add2(f.mudf, St::M2, -1); add2(f.muave, St::M2, -1);
  add2(f.mub, St::M2, MI_MUB);
  add2(f.msftx, St::M2, MI_MSFTX); add2(f.msfty, St::M2, MI_MSFTY);
  add2(f.msfux, St::U2, MI_MSFUX); add2(f.msfuy, St::U2, MI_MSFUY);
  add2(f.msfvx, St::V2, MI_MSFVX); add2(f.msfvy, St::V2, MI_MSFVY); add2(f.msfvx_inv, St::V2, MI_MSFVX_INV);
  for (double* m : {f.msftx, f.msfty, f.msfux, f.msfuy, f.msfvx, f.msfvy, f.msfvx_inv})
    launch_fill(m - d.off, d.n2, 1.0, ctx->stream);
  // Coriolis and grid rotation: f = e = sina = 0, cosa = 1 unless set
  add2(f.fcor, St::M2, MI_F); add2(f.ecor, St::M2, MI_E); add2(f.sina, St::M2, MI_SINA); add2(f.cosa, St::M2, MI_COSA);
  launch_fill(f.cosa - d.off, d.n2, 1.0, ctx->stream);
  d.spec = g.bc_x == MI_BC_SPECIFIED;
  // a diffusion driver (DiffusionConfig::on(), diffusion_driver.hpp:50) or a LateralBC
  const bool held = p.diff_6th_opt != 0 || p.diff_opt != 0 || p.damp_opt == 2 || d.spec;
  if (held) {
    add3(f.ru_tendf, St::U, MI_RU_TENDF); add3(f.rv_tendf, St::V, MI_RV_TENDF); add3(f.rw_tendf, St::W, MI_RW_TENDF);
    add3(f.t_tendf, St::M, MI_T_TENDF);
  }
  if (d.spec) {
    add3(f.ph_tendf, St::W, -1);
    f.bdy_time = d.alloc(1);
    // each variable's eight slabs (lateral_bc.cpp:219-233): values and tendencies of xs, xe, ys, ye
    const int64_t nz = g.nz, w = MI_SPEC_BDY_WIDTH;
    for (int v = 0; v < 5 + p.n_scalars; ++v) {
      const int kind = v < 5 ? v : 5;
      const int64_t nk = kind == kBdyPH ? nz + 1 : (kind == kBdyMU ? 1 : nz);
      const int64_t sx = (g.ny + (kind == kBdyV ? 1 : 0)) * nk * w, sy = (g.nx + (kind == kBdyU ? 1 : 0)) * nk * w;
      d.bdy_len[v] = 2 * (2 * sx + 2 * sy);
      d.bdy_buf[v] = d.alloc(d.bdy_len[v]);
    }
  }
  // moisture and scalars
  if (p.n_moist > 0) {
    add3(f.cqu, St::U, MI_CQU); add3(f.cqv, St::V, MI_CQV); add3(f.cqw, St::W, MI_CQW);
    add3(f.hdia, St::M, MI_H_DIABATIC);
  }
  d.n_sc = p.n_scalars;
  d.sc.n_moist = p.n_moist;
  if (p.n_scalars > 0) {  // the limiter's face geometry
    add2(f.pd_dx, St::U2, -1); add2(f.pd_dxdt, St::U2, -1); add2(f.pd_dy, St::V2, -1); add2(f.pd_dydt, St::V2, -1);
    f.pd_dz = d.alloc(2 * (int64_t(g.nz) + 1));
  }
  for (int n = 0; n < d.n_sc; ++n) {
    add3(d.sc.q1[n], St::M, -1);
    add3(d.sc.q2[n], St::M, MI_Q + n);
    add3(d.sc.qn[n], St::M, -1);
    add3(d.sc.qtf[n], St::M, MI_Q_TENDF + n);
    add3(d.sc.scale[n], St::M, -1);
    if (held) add3(d.sc.qtw[n], St::M, -1);
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
  f.t_init = p.terrain ? BField{d.t_init_buf, 1, G.sj, G.sk} : BField{d.t_init_buf, 0, 0, 1};
  if (p.mp_physics) {  // moist_physics_prep's fields and the precipitation
    add3(f.th_phy, St::M, -1); add3(f.mrho, St::M, -1); add3(f.pii, St::M, -1); add3(f.zmid, St::M, -1);
    add3(f.dz8w, St::M, -1); add3(f.qvd, St::M, -1); add3(f.qcd, St::M, -1);
    add2(f.rainnc, St::M2, MI_RAINNC); add2(f.rainncv, St::M2, MI_RAINNCV);
  }
  if (p.diff_opt == 1 && (p.km_opt == 4 || p.damp_opt == 1)) {  // the step's eddy viscosities (0 where not set)
    add3(f.xkmh, St::M, MI_XKMH); add3(f.xkhh, St::M, MI_XKHH);
    if (p.km_opt == 4) add3(f.d12, St::U, -1);
  }
  // what the substeps of a stage read unchanged: advance_w's damping weight, the coupled saves, its coefficients
  if (p.damp_opt == 3) add3(f.dampwt, St::W, -1);
  add3(f.ucs, St::U, -1); add3(f.vcs, St::V, -1); add3(f.cwc, St::M, -1);
  if (p.n_moist > 0) add3(f.cq1, St::W, -1);
  add2(f.rmu, St::U2, -1); add2(f.msfuy_inv, St::U2, -1); add2(f.rmv, St::V2, -1); add2(f.msfty_inv, St::M2, -1);
  f.t_base = d.alloc(nz + 1);
  f.z_base = d.alloc(nz + 1);
  f.u_base = d.alloc(nz + 1);
  f.v_base = d.alloc(nz + 1);
  f.qv_base = d.alloc(nz + 1);
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
  for (int kern = 0; kern < kNumColKernels; ++kern) d.colcfg[kern] = col_cfg_default(kern, g.nx, g.ny, g.nz);
  if (d.n3 >= (int64_t(1) << 31)) throw Error(MI_EINVAL, "grid too large: a field must hold fewer than 2^31 points");
  // run parameters
  Params& q = d.a.q;
  q.hup = p.h_adv_order == 5 ? 1.0 : 0.0;
  q.vup = p.v_adv_order == 3 ? 1.0 : 0.0;
  q.epssm = p.epssm;
  q.smdiv = p.smdiv;
  q.emdx = p.emdiv * (1. / lv.rdx);
  q.emdy = two_d ? 0.0 : p.emdiv * (1. / lv.rdy);
  q.wd_dt = 0.0;
  q.w_crit = p.w_crit_cfl;
  q.dampcoef = p.dampcoef;
  q.zdamp = p.zdamp;
  q.top_lid = p.top_lid;
  q.damp_opt = p.damp_opt;
  q.moist = p.n_moist > 0 ? 1 : 0;
  q.mapped = p.mapped;
  q.cor = (p.mapped != 0 || p.coriolis != 0) ? 1 : 0;
  q.use_tendf = held ? 1 : 0;
  d.held_base = held;
  q.spec = d.spec ? 1 : 0;
  q.diff_6th_opt = p.diff_6th_opt;
  q.diff_6th_factor = p.diff_6th_factor;
  q.mp = p.mp_physics;
  q.mp_lim = p.mp_tend_lim > 0 ? p.mp_tend_lim : 10.0;
  q.diff_opt = p.diff_opt;
  if (p.diff_opt == 1) {  // DiffusionDriver's constants (model.cpp:309-314, km_kh.hpp:35-45)
    const double prandtl = 1.0 / (p.pr_inv > 0 ? p.pr_inv : 3.0);
    q.xkmh = p.khdif;
    q.xkhh = p.khdif / prandtl;
    q.kvdif = p.kvdif;
    q.kvdq = p.kvdif / prandtl;
    q.km4 = p.km_opt == 4 ? 1 : 0;
    q.karr = (p.km_opt == 4 || p.damp_opt == 1) ? 1 : 0;
    q.prandtl = prandtl;
    q.c_s = p.c_s > 0 ? p.c_s : 0.25;
    q.dx = g.dx;
    q.dy = g.dy;
  }
  set_dt_params(q, g.dt, p.w_damping == 1, d.spec);
  q.terrain = p.terrain;
}

FieldDesc* find(Domain* d, int id) {
  for (FieldDesc& f : d->fields)
    if (f.id == id) return &f;
  return nullptr;
}

}  // namespace

namespace {

// host <-> device copies of one field's allocation, converting between the ABI's double and the storage type
void copy_in(void* dev, const std::vector<double>& h, int esize, hipStream_t st) {
  if (esize == 8) {
    hip_check(hipMemcpyAsync(dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  } else {
    std::vector<float> f(h.begin(), h.end());
    hip_check(hipMemcpyAsync(dev, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");  // f goes out of scope
  }
}
void copy_out(std::vector<double>& h, const void* dev, int esize, hipStream_t st) {
  if (esize == 8) {
    hip_check(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  } else {
    std::vector<float> f(h.size());
    hip_check(hipMemcpyAsync(f.data(), dev, f.size() * sizeof(float), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (size_t q = 0; q < f.size(); ++q) h[q] = f[q];
  }
}

}  // namespace

double* Domain::base_profile(int id) const {
  return id == MI_U_BASE ? a.f.u_base : id == MI_V_BASE ? a.f.v_base : id == MI_QV_BASE ? a.f.qv_base
         : id == MI_T_BASE ? a.f.t_base : id == MI_Z_BASE ? a.f.z_base : nullptr;
}

void Domain::set_field(int id, const double* host, int64_t n) {
  hipStream_t st = ctx->stream;
  const int nz = grid.nz;
  if (double* bp = base_profile(id)) {  // the diffusion's 1-D base profiles
    if (n != nz) throw Error(MI_EINVAL, "base profiles take nz values");
    hip_check(hipMemcpyAsync(bp, host, size_t(n) * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return;
  }
  const bool base3 = id == MI_PB || id == MI_ALB || id == MI_PHB || id == MI_T_INIT;
  // a base 3-D profile on a domain with terrain: repeated over every column, halos included
  if (base3 && prm.terrain && n == nz + (id == MI_PHB ? 1 : 0)) {
    FieldDesc* fd = find(this, id);
    std::vector<double> h(static_cast<size_t>(n3), 0.0);
    for (int64_t e = 0; e < n3; ++e) {
      const int64_t lev = a.g.sj ? (e / a.g.sk) % (nz + 1) : e / a.g.sk;  // row e / sk of j block e / sj
      h[size_t(e)] = host[std::min<int64_t>(lev, n - 1)];
    }
    copy_in(static_cast<char*>(*fd->slot) - off * fd->esize, h, fd->esize, st);
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return;
  }
  // base 3-D profiles
  if (base3 && !prm.terrain) {
    const int nzs = nz + (id == MI_PHB ? 1 : 0);
    if (n != nzs) throw Error(MI_EINVAL, "base 3-D fields take a profile of nz (phb: nz+1) values");
    double* dst = id == MI_PB ? pb_buf : id == MI_ALB ? alb_buf : id == MI_PHB ? phb_buf : t_init_buf;
    hip_check(hipMemcpyAsync(dst, host, size_t(n) * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return;
  }
  FieldDesc* fd = find(this, id);
  if (!fd) throw Error(MI_EINVAL, "field id " + std::to_string(id) + " cannot be set");
  const bool base2 = id == MI_MUB || (id >= MI_MSFTX && id <= MI_HT);
  const int fnx = nxs(fd->st), fny = nys(fd->st), fnz = nzs(fd->st);
  const int64_t alloc_n = (fd->st == St::M2 || fd->st == St::U2 || fd->st == St::V2) ? n2 : n3;
  void* base = static_cast<char*>(*fd->slot) - off * fd->esize;
  if (base2 && n == 1) {  // horizontally uniform value, halos included
    launch_fill(static_cast<double*>(base), alloc_n, host[0], st);
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return;
  }
  if (n != int64_t(fnx) * fny * fnz) throw Error(MI_EINVAL, "field size mismatch");
  std::vector<double> h(static_cast<size_t>(alloc_n), 0.0);
  copy_out(h, base, fd->esize, st);
  const Geo& G = a.g;
  for (int k = 0; k < fnz; ++k)
    for (int j = 0; j < fny; ++j)
      for (int i = 0; i < fnx; ++i) h[size_t(off + G.at(i, j, k))] = host[(int64_t(k) * fny + j) * fnx + i];
  copy_in(base, h, fd->esize, st);
  halo({{*fd->slot, -1, -1}});
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

void Domain::field_view(int id, mi_field_view* out) {
  FieldDesc* fd = find(this, id);
  if (!fd) throw Error(MI_EINVAL, "field id " + std::to_string(id) + " has no view");
  const bool two = fd->st == St::M2 || fd->st == St::U2 || fd->st == St::V2;
  out->data = *fd->slot;
  out->elem_bytes = fd->esize;
  out->nx = nxs(fd->st);
  out->ny = nys(fd->st);
  out->nz = two ? 1 : nzs(fd->st);
  out->stride_k = two ? 0 : a.g.sk;
  out->stride_j = a.g.sj;
}

// the fields a caller may give storage of its own: the carried state, held tendencies, 2-D base fields and
// map factors, and the 3-D base fields of a domain with terrain (not the scalars, whose buffers swap)
static bool bindable(int id) {
  return (id >= MI_U && id <= MI_AL) || (id >= MI_RU_TENDF && id <= MI_T_TENDF) || id == MI_MUB ||
         (id >= MI_MSFTX && id <= MI_HT) || id == MI_PB || id == MI_ALB || id == MI_PHB || id == MI_T_INIT;
}

void Domain::field_layout(int id, mi_field_layout* out) const {
  const FieldDesc* fd = nullptr;
  for (const FieldDesc& f : fields)
    if (f.id == id) fd = &f;
  if (!fd || !bindable(id)) throw Error(MI_EINVAL, "field id " + std::to_string(id) + " has no bindable storage");
  const bool two = fd->st == St::M2 || fd->st == St::U2 || fd->st == St::V2;
  out->elem_bytes = fd->esize;
  out->nx = nxs(fd->st);
  out->ny = nys(fd->st);
  out->nz = two ? 1 : nzs(fd->st);
  out->stride_k = two ? 0 : a.g.sk;
  out->stride_j = a.g.sj;
  out->offset = off;
  out->alloc_elems = two ? n2 : n3;
}

void Domain::field_bind(int id, void* dev, int64_t nbytes) {
  mi_field_layout lay{};
  field_layout(id, &lay);
  FieldDesc* fd = find(this, id);
  const int64_t need = lay.alloc_elems * lay.elem_bytes;
  hipStream_t st = ctx->stream;
  char* old_base = static_cast<char*>(*fd->slot) - off * fd->esize;
  char* new_base = static_cast<char*>(dev);
  if (dev) {
    if (nbytes < need) throw Error(MI_EINVAL, "bound buffer too small: " + std::to_string(need) + " bytes needed");
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, dev) != hipSuccess || (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
      (void)hipGetLastError();
      throw Error(MI_EINVAL, "bound buffer is not device memory");
    }
    if (at.device != ctx->device) throw Error(MI_EINVAL, "bound buffer is on another device");
  } else {  // storage of the domain's own again, with the bound values
    if (std::find(allocs.begin(), allocs.end(), static_cast<void*>(old_base)) != allocs.end()) return;  // not bound
    new_base = static_cast<char*>(static_cast<void*>(alloc<char>(need)));
    hip_check(hipMemcpyAsync(new_base, old_base, size_t(need), hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
  }
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  auto own = std::find(allocs.begin(), allocs.end(), static_cast<void*>(old_base));
  if (own != allocs.end() && static_cast<void*>(old_base) != static_cast<void*>(new_base)) {  // the domain's own
    hip_check(hipFree(old_base), "hipFree");
    allocs.erase(own);
    bytes -= need;
  }
  *fd->slot = new_base + off * fd->esize;
  // the base fields of a terrain domain are also read through their BField views
  if (pb_buf && id == MI_PB) a.f.pb.p = pb_buf;
  if (alb_buf && id == MI_ALB) a.f.alb.p = alb_buf;
  if (phb_buf && id == MI_PHB) a.f.phb.p = phb_buf;
  if (t_init_buf && id == MI_T_INIT) a.f.t_init.p = t_init_buf;
  halo({{*fd->slot, -1, -1}});
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  drop_graphs();  // the captured steps name the old buffer
}

void Domain::get_field(int id, double* host, int64_t n) {
  if (double* bp = base_profile(id)) {
    if (n != grid.nz) throw Error(MI_EINVAL, "base profiles hold nz values");
    hip_check(hipMemcpyAsync(host, bp, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return;
  }
  FieldDesc* fd = find(this, id);
  if (!fd) throw Error(MI_EINVAL, "field id " + std::to_string(id) + " cannot be read");
  const int fnx = nxs(fd->st), fny = nys(fd->st), fnz = nzs(fd->st);
  if (n != int64_t(fnx) * fny * fnz) throw Error(MI_EINVAL, "field size mismatch");
  const int64_t alloc_n = (fd->st == St::M2 || fd->st == St::U2 || fd->st == St::V2) ? n2 : n3;
  std::vector<double> h(static_cast<size_t>(alloc_n));
  copy_out(h, static_cast<char*>(*fd->slot) - off * fd->esize, fd->esize, ctx->stream);
  const Geo& G = a.g;
  for (int k = 0; k < fnz; ++k)
    for (int j = 0; j < fny; ++j)
      for (int i = 0; i < fnx; ++i) host[(int64_t(k) * fny + j) * fnx + i] = h[size_t(off + G.at(i, j, k))];
}

void Domain::phase(MiPhase which, int rk) {
  switch (which) {
    case kPhaseStepBegin: step_begin(); break;
    case kPhaseRkPrep:
      rk_prep(rk);
      if (rk == 1) physics_part1();
      break;
    case kPhaseRkTendency: rk_tendency(rk); break;
    case kPhaseSmallSteps: small_steps(rk, false); break;
    case kPhaseScalars: scalar_transport(rk); break;
    case kPhaseCalcPRhoPhi: calc_p_rho_phi(); break;
    case kPhaseStepEnd: step_end(); break;
  }
  hip_check(hipGetLastError(), "kernel launch");
}

void Domain::set_tuning(const int* shapes, int n) {
  for (int q = 0; q < n && q < kNumColKernels; ++q)
    if (!col_cfg_fits(q, shapes[q], grid.nz)) throw Error(MI_EINVAL, "block shape does not fit this nz");
  for (int q = 0; q < n && q < kNumColKernels; ++q) colcfg[q] = shapes[q];
  drop_graphs();
}

void Domain::get_tuning(int* shapes, int n) const {
  for (int q = 0; q < n && q < kNumColKernels; ++q) shapes[q] = colcfg[q];
}

void Domain::profile_kernel(const char* name) {
  prof_name = name ? name : "";
  timer.used = 0;
  timer.bytes = 0;
}

void Domain::profile_read(double* total_ms, int* launches, double* alg_bytes) {
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  double tot = 0;
  for (size_t q = 0; q + 1 < timer.used; q += 2) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, timer.ev[q], timer.ev[q + 1]), "hipEventElapsedTime");
    tot += ms;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = int(timer.used / 2);
  if (alg_bytes) *alg_bytes = timer.bytes;
  timer.used = 0;
  timer.bytes = 0;
}

int Domain::check_finite(int* bad_field) {
  hipStream_t st = ctx->stream;
  hip_check(hipMemsetAsync(flags, 0, 64 * sizeof(int), st), "hipMemsetAsync");
  launch_finite(a, flags, st);
  int h[64];
  hip_check(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  int bad = -1;
  for (int q = 0; q < 64 && bad < 0; ++q)
    if (h[q]) bad = q;
  if (bad_field) *bad_field = bad;
  return bad;
}

void Domain::cfl_max(double* vert, double* horiz) {
  hipStream_t st = ctx->stream;
  unsigned long long* out = reinterpret_cast<unsigned long long*>(flags + 32);  // two 8-byte slots of the flag buffer
  hip_check(hipMemsetAsync(out, 0, 2 * sizeof(unsigned long long), st), "hipMemsetAsync");
  if (grid.nz >= 2) launch_cfl(a, out, st);
  unsigned long long h[2];
  hip_check(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  double v[2];
  std::memcpy(v, h, sizeof(v));
  if (vert) *vert = v[0];
  if (horiz) *horiz = v[1];
}

// the same maxima through the library's file-scope scratch array (no allocation of any kind on this path)
void Domain::cfl_max_noalloc(double* vert, double* horiz) {
  hipStream_t st = ctx->stream;
  unsigned long long* sp = cfl_scratch(ctx->device);
  if (!sp) throw Error(MI_EHIP, "hipGetSymbolAddress of the CFL scratch array failed");
  unsigned long long h[kCflScratchSlots] = {0, 0};
  if (grid.nz >= 2) {
    launch_cfl_scratch(a, sp, st);
    hip_check(hipMemcpyAsync(h, sp, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  }
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  double v[2];
  std::memcpy(v, h, sizeof(v));
  if (vert) *vert = v[0];
  if (horiz) *horiz = v[1];
}

void Domain::bdy_set(int var, int side, const double* values, const double* tends, int64_t n) {
  if (!spec) throw Error(MI_EINVAL, "the domain has no specified boundaries");
  if (side < MI_BDY_XS || side > MI_BDY_YE) throw Error(MI_EINVAL, "boundary side must be MI_BDY_XS .. MI_BDY_YE");
  int v = -1;
  if (var >= MI_BDY_U && var <= MI_BDY_MU) v = var;
  else if (var >= MI_BDY_Q && var < MI_BDY_Q + n_sc) v = 5 + (var - MI_BDY_Q);
  if (v < 0) throw Error(MI_EINVAL, "boundary variable must be MI_BDY_U .. MI_BDY_MU or MI_BDY_Q + a scalar");
  const int kind = v < 5 ? v : 5;
  const int64_t nz = grid.nz, w = MI_SPEC_BDY_WIDTH;
  const int64_t nk = kind == kBdyPH ? nz + 1 : (kind == kBdyMU ? 1 : nz);
  const int64_t sx = (grid.ny + (kind == kBdyV ? 1 : 0)) * nk * w, sy = (grid.nx + (kind == kBdyU ? 1 : 0)) * nk * w;
  const int64_t len = side < 2 ? sx : sy;
  if (n != len) throw Error(MI_EINVAL, "boundary slab size mismatch: expected " + std::to_string(len));
  const int64_t o = side < 2 ? side * sx : 2 * sx + (side - 2) * sy, half = 2 * sx + 2 * sy;
  hipStream_t st = ctx->stream;
  hip_check(hipMemcpyAsync(bdy_buf[v] + o, values, size_t(n) * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(bdy_buf[v] + half + o, tends, size_t(n) * sizeof(double), hipMemcpyHostToDevice, st),
            "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  if (v >= 5 && !q_has_bdy[v - 5]) {
    q_has_bdy[v - 5] = true;  // the step's launches change (relax and spec instead of flow_dep)
    drop_graphs();
  }
}

mi_domain* make_domain(mi_ctx* ctx, const mi_grid& g, const mi_params& p, const mi_levels& lv) {
  auto d = std::make_unique<Domain>();
  build_domain(*d, ctx, g, p, lv);
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return d.release();
}

// ============================================================================= fused batch steps
// A batch whose members share this arithmetic and the acoustic step count steps as one schedule: the phases every
// member runs on a branch of its own (the stage's prep, tendencies, acoustic prologue and finish, the scalars, the
// halo refreshes), and each acoustic substep's two kernels (advance_uv, the column kernel) as one launch over every
// member's blocks.  The device buffer holds the members' Args, then six tables (advance_uv and the column kernel
// of each stage; n + 1 entries each).
size_t fused_bytes(int n) { return size_t(n) * sizeof(Args) + 6 * size_t(n + 1) * sizeof(BatchMember); }

void fused_prepare(mi_domain* const* ms, int n, void* dev, hipStream_t st, FusedInfo& info) {
  if (n > kMaxBatch) throw Error(MI_EINVAL, "internal: too many members for a fused batch");
  std::vector<const Args*> as(static_cast<size_t>(n));
  std::vector<Args> args(static_cast<size_t>(n));
  info.nzmax = 0;
  info.ns = static_cast<Domain*>(ms[0])->n_sound_;
  for (int q = 0; q < n; ++q) {
    Domain* d = static_cast<Domain*>(ms[q]);
    args[size_t(q)] = d->a;
    as[size_t(q)] = &d->a;
    info.nzmax = std::max(info.nzmax, d->grid.nz);
  }
  std::vector<BatchMember> tabs(6 * size_t(n + 1));
  for (int rk = 1; rk <= 3; ++rk) {
    std::vector<double> dts(static_cast<size_t>(n));
    for (int q = 0; q < n; ++q) dts[size_t(q)] = static_cast<Domain*>(ms[q])->stage_dts(rk);
    info.blocks_uv[rk - 1] = batch_table_uv(as.data(), dts.data(), n, &tabs[size_t(rk - 1) * size_t(n + 1)]);
    info.blocks_ac[rk - 1] = batch_table_acoustic(as.data(), dts.data(), n, &tabs[size_t(rk + 2) * size_t(n + 1)]);
  }
  char* p = static_cast<char*>(dev);
  hip_check(hipMemcpyAsync(p, args.data(), args.size() * sizeof(Args), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(p + args.size() * sizeof(Args), tabs.data(), tabs.size() * sizeof(BatchMember),
                           hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

void fused_capture(mi_domain* const* ms, int n, hipStream_t s0, const hipStream_t* st, hipEvent_t fork,
                   const hipEvent_t* join, const void* dev, const FusedInfo& info) {
  std::vector<Domain*> d(static_cast<size_t>(n));
  std::vector<Scalars> keep_sc(static_cast<size_t>(n));
  std::vector<int> keep_q(static_cast<size_t>(n));
  for (int q = 0; q < n; ++q) {
    d[size_t(q)] = static_cast<Domain*>(ms[q]);
    keep_sc[size_t(q)] = d[size_t(q)]->sc;
    keep_q[size_t(q)] = d[size_t(q)]->qpar;
  }
  const Args* dev_args = static_cast<const Args*>(dev);
  const BatchMember* tabs = reinterpret_cast<const BatchMember*>(static_cast<const char*>(dev) + size_t(n) * sizeof(Args));
  // every member's part of a phase on its own branch, forked from and joined back to s0
  auto branch = [&](auto&& fn) {
    hip_check(hipEventRecord(fork, s0), "hipEventRecord");
    for (int q = 0; q < n; ++q) {
      hip_check(hipStreamWaitEvent(st[q], fork, 0), "hipStreamWaitEvent");
      d[size_t(q)]->launch_stream = st[q];
      fn(*d[size_t(q)]);
      d[size_t(q)]->launch_stream = nullptr;
      hip_check(hipEventRecord(join[q], st[q]), "hipEventRecord");
      hip_check(hipStreamWaitEvent(s0, join[q], 0), "hipStreamWaitEvent");
    }
  };
  bool any_after_uv = false;
  for (Domain* m : d) any_after_uv = any_after_uv || m->needs_after_uv();
  try {
    branch([](Domain& m) { m.step_begin(); });
    for (int rk = 1; rk <= 3; ++rk) {
      const int ns = rk == 1 ? 1 : (rk == 2 ? info.ns / 2 : info.ns);
      branch([rk](Domain& m) {
        m.rk_prep(rk);
        m.rk_tendency(rk, false);
        m.ss_begin(rk);
      });
      for (int it = 1; it <= ns; ++it) {
        launch_uv_batch(dev_args, tabs + size_t(rk - 1) * size_t(n + 1), n, info.blocks_uv[rk - 1], s0);
        if (any_after_uv) branch([](Domain& m) { m.after_uv(); });
        launch_acoustic_batch(dev_args, tabs + size_t(rk + 2) * size_t(n + 1), n, info.blocks_ac[rk - 1], info.nzmax, it, ns,
                              s0);
        branch([rk, it](Domain& m) {
          launch_open_sums(m.a, it, m.stage_ns(rk), m.stream());
          m.after_acoustic(rk, it);
        });
      }
      branch([rk](Domain& m) {
        m.ss_end(rk, true);
        m.scalar_transport(rk);
      });
    }
    branch([](Domain& m) { m.step_end(false); });
  } catch (...) {
    for (int q = 0; q < n; ++q) {
      d[size_t(q)]->launch_stream = nullptr;
      d[size_t(q)]->sc = keep_sc[size_t(q)];
      d[size_t(q)]->qpar = keep_q[size_t(q)];
    }
    throw;
  }
  // the host side of the captured step runs at each replay (advance_host): put the buffer names back
  for (int q = 0; q < n; ++q) {
    d[size_t(q)]->sc = keep_sc[size_t(q)];
    d[size_t(q)]->qpar = keep_q[size_t(q)];
  }
}

}  // namespace MI_NS
