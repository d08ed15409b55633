// C++20 facade with the reference's src/dyn names over the C ABI of include/cppwrf_amd.h.
//
// A caller of the reference's dyn::RK3 (src/dyn/rk3.hpp:34-78) keeps its call pattern:
//
//   cppwrf_amd::dyn::Grid g{...};
//   cppwrf_amd::dyn::RK3 rk(ctx, g, large_step_params, acoustic_params, time_step_sound, levels);
//   rk.state().set(MI_U, u_host);  ...                 // State / BaseState fill (fill_host)
//   for (int n = 0; n < nsteps; ++n) rk.step();        // RK3::step
//   rk.state().get(MI_U, u_host);                       // download
//
// The State, Tendencies and BaseState live on the device inside the domain handle; the facade only
// moves host arrays in and out.  Errors come back as std::runtime_error, as the reference reports
// bad parameters (src/dyn/rk3.cpp:50).
#pragma once

#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cppwrf_amd.h"

namespace cppwrf_amd::dyn {

inline void check(int rc) {
  if (rc != MI_OK) throw std::runtime_error(std::string("cppwrf_amd: ") + mi_last_error());
}

// reference LargeStepParams (src/dyn/large_step.hpp:14-22)
struct LargeStepParams {
  bool top_lid = false;
  int h_adv_order = 5, v_adv_order = 3, w_damping = 0;
  double w_crit_cfl = 1;
};
// reference AcousticParams (src/dyn/rk3.hpp:20-27)
struct AcousticParams {
  double epssm = 0.1, smdiv = 0.1, emdiv = 0.01;
  bool top_lid = false;
  int damp_opt = 0, hypsometric_opt = 1;
  double zdamp = 5000, dampcoef = 0.2;
};

class Context {
 public:
  explicit Context(int device = 0, void* stream = nullptr) { check(mi_ctx_create(device, stream, &h_)); }
  ~Context() { mi_ctx_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  mi_ctx* get() const { return h_; }

 private:
  mi_ctx* h_ = nullptr;
};

// reference ScalarParams (src/dyn/scalars.hpp:14-19); the advection orders are the dynamics' own
struct ScalarParams {
  int moist_adv_opt = 1, scalar_adv_opt = 1;
};
// What the reference's BaseState and drivers carry beside the parameter structs: the number of moist
// species and tracers in State::q_2 (state.hpp:53-60), BaseState::mapped / coriolis (state.hpp:93-107),
// a terrain-following base state, the sixth-order diffusion of DiffusionDriver (diffusion/sixth_order.hpp)
// and the storage precision of the bulk 3-D fields.
struct DomainOptions {
  int n_moist = 0, n_scalars = 0;
  bool mapped = false, coriolis = false, terrain = false;
  int diff_6th_opt = 0;
  double diff_6th_factor = 0.12;
  int precision = MI_FP64;
};

// number of values of a field id in the dense layout of the C ABI
inline size_t field_size(int field, const mi_grid& g) {
  const size_t nx = size_t(g.nx), ny = size_t(g.ny), nz = size_t(g.nz), vy = g.ny > 1 ? 1 : 0;
  switch (field) {
    case MI_U: case MI_RU_M: case MI_RU_TEND: case MI_CQU: case MI_RU_TENDF: return (nx + 1) * ny * nz;
    case MI_V: case MI_RV_M: case MI_RV_TEND: case MI_CQV: case MI_RV_TENDF: return nx * (ny + vy) * nz;
    case MI_W: case MI_PH: case MI_WW: case MI_WW_M: case MI_RW_TEND: case MI_PH_TEND: case MI_CQW: case MI_PHB:
    case MI_RW_TENDF:
      return nx * ny * (nz + 1);
    case MI_MU: case MI_MUTS: case MI_MUB: case MI_MSFTX: case MI_MSFTY: case MI_F: case MI_E: case MI_SINA:
    case MI_COSA: case MI_HT:
      return nx * ny;
    case MI_MSFUX: case MI_MSFUY: return (nx + 1) * ny;
    case MI_MSFVX: case MI_MSFVY: case MI_MSFVX_INV: return nx * (ny + vy);
    default: return nx * ny * nz;  // mass points: t, p, al, alt, php, tendencies, h_diabatic, scalars
  }
}

// The device-resident State + BaseState of the domain (reference src/dyn/state.hpp:30-119)
class StateView {
 public:
  explicit StateView(mi_domain* d, const mi_grid& g) : d_(d), g_(g) {}
  size_t size(int field) const { return field_size(field, g_); }
  void set(int field, const std::vector<double>& h) { check(mi_domain_set_field(d_, field, h.data(), int64_t(h.size()))); }
  std::vector<double> get(int field) const {
    std::vector<double> h(size(field));
    check(mi_domain_get_field(d_, field, h.data(), int64_t(h.size())));
    return h;
  }
  // reference State::q_2 of species n (n_moist moist species first, then the tracers)
  void set_q(int n, const std::vector<double>& h) { set(MI_Q + n, h); }
  std::vector<double> q(int n) const { return get(MI_Q + n); }

 private:
  mi_domain* d_;
  mi_grid g_;
};

class RK3;
// reference ColumnPhysics (src/dyn/column_physics.hpp:21-34): the column physics suite's hooks into the step.
// The State and Tendencies are the domain's (rk.state()); part2 sets the physics tendencies into MI_RU_TENDF,
// MI_RV_TENDF, MI_T_TENDF and MI_Q_TENDF + n.  th_forcing() / qv_forcing() = true keep the advective forcings
// MI_TH_FORCING / MI_QV_FORCING (the reference returns the fields to fill).
class ColumnPhysics {
 public:
  virtual ~ColumnPhysics() = default;
  virtual void part1(RK3& rk) = 0;
  virtual void part2(RK3& rk) = 0;
  virtual void after_rk(RK3& rk) = 0;
  virtual void after_all_rk_steps(RK3&) {}
  virtual bool th_forcing() { return false; }
  virtual bool qv_forcing() { return false; }
};

// reference dyn::RK3 (src/dyn/rk3.hpp:34-78)
class RK3 {
 public:
  RK3(const Context& ctx, const mi_grid& g, const LargeStepParams& lp, const AcousticParams& ap, int time_step_sound,
      const mi_levels& levels, const ScalarParams& sp = {}, const DomainOptions& opt = {})
      : g_(g) {
    if (lp.top_lid != ap.top_lid) throw std::runtime_error("top_lid must agree between LargeStepParams and AcousticParams");
    mi_params p{};
    p.h_adv_order = lp.h_adv_order;
    p.v_adv_order = lp.v_adv_order;
    p.top_lid = lp.top_lid ? 1 : 0;
    p.w_damping = lp.w_damping;
    p.w_crit_cfl = lp.w_crit_cfl;
    p.epssm = ap.epssm;
    p.smdiv = ap.smdiv;
    p.emdiv = ap.emdiv;
    p.damp_opt = ap.damp_opt;
    p.hypsometric_opt = ap.hypsometric_opt;
    p.zdamp = ap.zdamp;
    p.dampcoef = ap.dampcoef;
    p.time_step_sound = time_step_sound;
    p.precision = opt.precision;
    p.n_moist = opt.n_moist;
    p.n_scalars = opt.n_scalars;
    p.moist_adv_opt = sp.moist_adv_opt;
    p.scalar_adv_opt = sp.scalar_adv_opt;
    p.mapped = opt.mapped ? 1 : 0;
    p.coriolis = opt.coriolis ? 1 : 0;
    p.diff_6th_opt = opt.diff_6th_opt;
    p.terrain = opt.terrain ? 1 : 0;
    p.diff_6th_factor = opt.diff_6th_factor;
    check(mi_domain_create(ctx.get(), &g, &p, &levels, &d_));
  }
  ~RK3() { mi_domain_destroy(d_); }
  RK3(const RK3&) = delete;
  RK3& operator=(const RK3&) = delete;

  // reference RK3::auto_sound_steps (src/dyn/rk3.cpp:69-72)
  static int auto_sound_steps(double dt, double dx, double dy) { return mi_auto_sound_steps(dt, dx, dy); }
  // reference RK3::adaptive_sound_steps (src/dyn/rk3.cpp:74-77)
  static int adaptive_sound_steps(double dt, double dx, double dy, double max_msft) {
    return mi_adaptive_sound_steps(dt, dx, dy, max_msft);
  }
  int n_sound() const { return mi_domain_n_sound(d_); }
  // reference RK3::set_dt (src/dyn/rk3.cpp:79-84), with the boundary weights and the diffusion's dt
  void set_dt(double dt, int n_sound) {
    check(mi_domain_set_dt(d_, dt, n_sound));
    g_.dt = dt;
  }
  // checkpoint and resume: the carried state as an opaque byte buffer (no counterpart in src/dyn)
  std::vector<unsigned char> save_state() {
    const int64_t n = mi_domain_state_bytes(d_);
    if (n < 0) check(int(n));
    std::vector<unsigned char> b(static_cast<size_t>(n));
    check(mi_domain_save_state(d_, b.data(), n));
    return b;
  }
  void load_state(const std::vector<unsigned char>& b) { check(mi_domain_load_state(d_, b.data(), int64_t(b.size()))); }
  // reference RK3::on_cfl and on_first_rk_prep (rk3.hpp:46-47, 62-64): while either is set, steps run launch by
  // launch and call them at the reference's points
  std::function<void(double, double)> on_cfl;
  std::function<void()> on_first_rk_prep;
  // reference RK3::step (src/dyn/rk3.cpp:86-170); nsteps > 1 runs them back to back on the device
  void step(int nsteps = 1) {
    check(mi_domain_set_cfl_callback(d_, on_cfl ? &RK3::cfl_thunk : nullptr, this));
    check(mi_domain_set_first_rk_prep_callback(d_, on_first_rk_prep ? &RK3::prep_thunk : nullptr, this));
    check(mi_step(d_, nsteps));
  }
  // reference Field3::view (field.hpp:16-44): device pointer and strides of a field, for device-side plugins
  mi_field_view view(int field) {
    mi_field_view v{};
    check(mi_domain_field_view(d_, field, &v));
    return v;
  }
  // reference RK3::set_diffusion (rk3.hpp:53-62) with the sixth-order diffusion: opt 0 detaches
  void set_diffusion(int diff_6th_opt, double diff_6th_factor = 0.12) {
    check(mi_domain_set_diffusion(d_, diff_6th_opt, diff_6th_factor));
  }
  // reference RK3::set_column_physics: the plugin is not owned (rk3.hpp:53-62); nullptr detaches
  void set_column_physics(ColumnPhysics* cp) {
    cp_ = cp;
    if (!cp) {
      check(mi_domain_set_column_physics(d_, nullptr));
      return;
    }
    mi_column_physics s{};
    s.user = this;
    s.part1 = [](void* u, mi_domain*) { auto* r = static_cast<RK3*>(u); r->cp_->part1(*r); };
    s.part2 = [](void* u, mi_domain*) { auto* r = static_cast<RK3*>(u); r->cp_->part2(*r); };
    s.after_rk = [](void* u, mi_domain*) { auto* r = static_cast<RK3*>(u); r->cp_->after_rk(*r); };
    s.after_all_rk_steps = [](void* u, mi_domain*) { auto* r = static_cast<RK3*>(u); r->cp_->after_all_rk_steps(*r); };
    s.th_forcing = cp->th_forcing() ? 1 : 0;
    s.qv_forcing = cp->qv_forcing() ? 1 : 0;
    check(mi_domain_set_column_physics(d_, &s));
  }
  void sync() { check(mi_domain_sync(d_)); }
  StateView state() { return StateView(d_, g_); }
  mi_domain* handle() const { return d_; }
  // w_damp's vertical and horizontal Courant maxima of the current fields (the values the reference
  // passes to RK3::on_cfl, rk3.hpp:46-47)
  std::pair<double, double> cfl_max() {
    double v = 0, h = 0;
    check(mi_cfl_max(d_, &v, &h));
    return {v, h};
  }
  // the reference runner's finiteness check (src/cases/domain_runner.hpp:70-71); -1 when all are finite
  int first_nonfinite() {
    int bad = -1;
    const int rc = mi_check_finite(d_, &bad);
    if (rc != MI_ENONFINITE) check(rc);
    return bad;
  }

 private:
  static void cfl_thunk(void* self, double vert, double horiz) { static_cast<RK3*>(self)->on_cfl(vert, horiz); }
  static void prep_thunk(void* self) { static_cast<RK3*>(self)->on_first_rk_prep(); }
  mi_grid g_;
  mi_domain* d_ = nullptr;
  ColumnPhysics* cp_ = nullptr;
};

// Several domains stepped together, one HIP graph per step (the reference's tiles of main.cpp:30-42)
class Batch {
 public:
  Batch(const Context& ctx, const std::vector<RK3*>& members) {
    std::vector<mi_domain*> h;
    for (RK3* m : members) h.push_back(m->handle());
    check(mi_batch_create(ctx.get(), h.data(), int(h.size()), &b_));
  }
  ~Batch() { mi_batch_destroy(b_); }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  void step(int nsteps = 1) { check(mi_batch_step(b_, nsteps)); }

 private:
  mi_batch* b_ = nullptr;
};

}  // namespace cppwrf_amd::dyn
