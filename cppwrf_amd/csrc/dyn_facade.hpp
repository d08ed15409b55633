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
// or, with caller-owned device buffers as the reference's containers own their Field3s:
//
//   cppwrf_amd::dyn::State st(rk); cppwrf_amd::dyn::Tendencies tend(rk); cppwrf_amd::dyn::BaseState base(rk);
//   st.field(MI_U).fill_host(u_host); ...
//   rk.set_lateral_bc(&bc);                             // LateralBC data, not owned
//   for (int n = 0; n < nsteps; ++n) rk.step(st, tend, base);  // RK3::step(State&, Tendencies&, const BaseState&)
//
// Errors come back as std::runtime_error, as the reference reports bad parameters (src/dyn/rk3.cpp:50).
#pragma once

#include <functional>
#include <map>
#include <memory>
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
// a terrain-following base state, the diffusion of DiffusionDriver (its DiffusionConfig's diff_opt 1 with
// km_opt 1 or 4, khdif, kvdif, c_s and 1 / prandtl, diffusion_driver.hpp:23-51; and diff_6th_opt,
// diffusion/sixth_order.hpp)
// and the storage precision of the bulk 3-D fields.
struct DomainOptions {
  int n_moist = 0, n_scalars = 0;
  bool mapped = false, coriolis = false, terrain = false;
  int diff_6th_opt = 0;
  double diff_6th_factor = 0.12;
  int precision = MI_FP64;
  int diff_opt = 0, km_opt = 1;
  double khdif = 0, kvdif = 0, pr_inv = 3.0, c_s = 0.25;
  // RK3::enable_physics(PhysicsParams) (rk3.hpp:34-78, physics_coupling.hpp:23-37): mp_physics 1 (Kessler) after the
  // RK loop, with mp_tend_lim
  int mp_physics = 0;
  double mp_tend_lim = 10.0;
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
    case MI_COSA: case MI_HT: case MI_RAINNC: case MI_RAINNCV:
      return nx * ny;
    case MI_XKMH: case MI_XKHH: return nx * ny * nz;
    case MI_MSFUX: case MI_MSFUY: return (nx + 1) * ny;
    case MI_U_BASE: case MI_V_BASE: case MI_QV_BASE: case MI_T_BASE: case MI_Z_BASE: return nz;
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

// reference Field3 (src/core/field.hpp:76-143): a device buffer its owner (the caller) allocates, in the layout the
// domain asks for (mi_domain_field_layout); RK3::step(State&, Tendencies&, const BaseState&) makes it the field's
// storage (mi_field_bind).  Move-only, freed with its owner.
class Field {
 public:
  Field() = default;
  Field(mi_ctx* c, int id, const mi_field_layout& l) : c_(c), id_(id), l_(l) { check(mi_device_alloc(c, bytes(), &p_)); }
  ~Field() {
    if (p_) mi_device_free(c_, p_);
  }
  Field(Field&& o) noexcept { swap(o); }
  Field& operator=(Field&& o) noexcept {
    swap(o);
    return *this;
  }
  Field(const Field&) = delete;
  Field& operator=(const Field&) = delete;
  int id() const { return id_; }
  void* data() const { return p_; }
  int64_t bytes() const { return l_.alloc_elems * l_.elem_bytes; }
  const mi_field_layout& layout() const { return l_; }
  size_t points() const { return size_t(l_.nx) * size_t(l_.ny) * size_t(l_.nz); }
  // reference Field3::fill_host / download: dense (k, j, i) values of the logical points (the halos are the
  // domain's to refresh, which it does when the buffer is bound or changed)
  void fill_host(const std::vector<double>& h) {
    if (h.size() != points()) throw std::runtime_error("cppwrf_amd: fill_host size mismatch");
    std::vector<unsigned char> buf(static_cast<size_t>(bytes()));
    check(mi_device_download(c_, buf.data(), p_, bytes()));
    for_each([&](size_t e, size_t q) { put(buf.data(), e, h[q]); });
    check(mi_device_upload(c_, p_, buf.data(), bytes()));
    ++version_;
  }
  std::vector<double> download() const {
    std::vector<unsigned char> buf(static_cast<size_t>(bytes()));
    check(mi_device_download(c_, buf.data(), p_, bytes()));
    std::vector<double> h(points());
    for_each([&](size_t e, size_t q) { h[q] = get(buf.data(), e); });
    return h;
  }
  uint64_t version() const { return version_; }

 private:
  template <class F>
  void for_each(F&& f) const {
    size_t q = 0;
    for (int k = 0; k < l_.nz; ++k)
      for (int j = 0; j < l_.ny; ++j)
        for (int i = 0; i < l_.nx; ++i, ++q) f(size_t(l_.offset + i + int64_t(j) * l_.stride_j + int64_t(k) * l_.stride_k), q);
  }
  void put(unsigned char* b, size_t e, double v) const {
    if (l_.elem_bytes == 8) reinterpret_cast<double*>(b)[e] = v;
    else reinterpret_cast<float*>(b)[e] = float(v);
  }
  double get(const unsigned char* b, size_t e) const {
    return l_.elem_bytes == 8 ? reinterpret_cast<const double*>(b)[e] : double(reinterpret_cast<const float*>(b)[e]);
  }
  void swap(Field& o) noexcept {
    std::swap(c_, o.c_);
    std::swap(id_, o.id_);
    std::swap(l_, o.l_);
    std::swap(p_, o.p_);
    std::swap(version_, o.version_);
  }
  mi_ctx* c_ = nullptr;
  int id_ = -1;
  mi_field_layout l_{};
  void* p_ = nullptr;
  uint64_t version_ = 1;
};

class RK3;

// A set of caller-owned fields of one domain, made in the domain's layouts and holding its current values
// (reference State, Tendencies and BaseState, src/dyn/state.hpp:30-119; the scalars stay in the domain, whose two
// buffers of a species swap at every stage).
class FieldSet {
 public:
  Field& field(int id) {
    for (Field& f : f_)
      if (f.id() == id) return f;
    throw std::runtime_error("cppwrf_amd: no field " + std::to_string(id) + " in this set");
  }
  bool has(int id) const {
    for (const Field& f : f_)
      if (f.id() == id) return true;
    return false;
  }
  std::vector<Field>& fields() { return f_; }
  const std::vector<Field>& fields() const { return f_; }

 protected:
  void make(RK3& rk, std::initializer_list<int> ids);
  std::vector<Field> f_;
};
// the prognostic state the step carries: u, v, w, theta, phi', mu', p', alpha'
class State : public FieldSet {
 public:
  explicit State(RK3& rk) { make(rk, {MI_U, MI_V, MI_W, MI_T, MI_PH, MI_MU, MI_P, MI_AL}); }
};
// the held (stage-1) tendencies, when the domain holds them (a diffusion, specified boundaries or a plugin)
class Tendencies : public FieldSet {
 public:
  explicit Tendencies(RK3& rk) { make(rk, {MI_RU_TENDF, MI_RV_TENDF, MI_RW_TENDF, MI_T_TENDF}); }
};
// the base state, map factors, Coriolis and terrain (3-D base fields only with terrain)
class BaseState : public FieldSet {
 public:
  explicit BaseState(RK3& rk) {
    make(rk, {MI_MUB, MI_PB, MI_ALB, MI_PHB, MI_T_INIT, MI_MSFTX, MI_MSFTY, MI_MSFUX, MI_MSFUY, MI_MSFVX, MI_MSFVY,
              MI_MSFVX_INV, MI_F, MI_E, MI_SINA, MI_COSA, MI_HT});
  }
};

// reference LateralBC (src/dyn/lateral_bc.hpp:28-91): the boundary data of a real-data run for the current
// boundary interval, host side; RK3::set_lateral_bc attaches it (not owned), and each step uploads what changed
// and advances dtbc as the reference's step does.  Width 5 (spec_zone 1, relax_zone 4), not nested.
class LateralBC {
 public:
  enum Var { U = MI_BDY_U, V = MI_BDY_V, T = MI_BDY_T, PH = MI_BDY_PH, MU = MI_BDY_MU };
  enum Side { XS = MI_BDY_XS, XE = MI_BDY_XE, YS = MI_BDY_YS, YE = MI_BDY_YE };
  LateralBC(const mi_grid& g, double dt, int width = MI_SPEC_BDY_WIDTH, int spec_zone = 1, int relax_zone = 4)
      : g_(g), dt_(dt) {
    if (width != MI_SPEC_BDY_WIDTH || spec_zone != 1 || relax_zone != 4)
      throw std::runtime_error("cppwrf_amd: LateralBC supports width 5 (spec_zone 1, relax_zone 4)");
  }
  int spec_zone() const { return 1; }
  int width() const { return MI_SPEC_BDY_WIDTH; }
  // the relaxation weights follow the domain's large step (RK3::set_dt passes it on)
  void set_dt(double dt) { dt_ = dt; }
  void set(Var v, Side s, const std::vector<double>& values, const std::vector<double>& tends) {
    put(int(v), int(s), values, tends);
  }
  void set_scalar(int n, Side s, const std::vector<double>& values, const std::vector<double>& tends) {
    put(MI_BDY_Q + n, int(s), values, tends);
  }
  bool has(Var v) const {
    for (const auto& kv : data_)
      if (kv.first.first == int(v)) return true;
    return false;
  }
  double dtbc = 0;  // seconds since the start of the current boundary interval
  void advance_time(double dt) { dtbc += dt; }

 private:
  friend class RK3;
  struct Slab {
    std::vector<double> values, tends;
    uint64_t version = 0;
  };
  void put(int var, int side, const std::vector<double>& values, const std::vector<double>& tends) {
    if (values.size() != tends.size()) throw std::runtime_error("cppwrf_amd: boundary values and tendencies differ in size");
    Slab& sl = data_[{var, side}];
    sl.values = values;
    sl.tends = tends;
    sl.version = ++clock_;
  }
  mi_grid g_;
  double dt_;
  uint64_t clock_ = 0;
  std::map<std::pair<int, int>, Slab> data_;
};

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
      : g_(g), c_(ctx.get()) {
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
    p.diff_opt = opt.diff_opt;
    p.km_opt = opt.km_opt;
    p.khdif = opt.khdif;
    p.kvdif = opt.kvdif;
    p.pr_inv = opt.pr_inv;
    p.mp_physics = opt.mp_physics;
    p.mp_tend_lim = opt.mp_tend_lim;
    p.c_s = opt.c_s;
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
    if (bc_) bc_->set_dt(dt);
  }
  mi_ctx* context() const { return c_; }
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
    if (!bc_) {
      check(mi_step(d_, nsteps));
      return;
    }
    for (int n = 0; n < nsteps; ++n) {  // the boundary clock lives in the LateralBC, as the reference's
      push_boundary();
      check(mi_step(d_, 1));
      check(mi_bdy_get_time(d_, &bc_->dtbc));
    }
  }
  // reference RK3::step(State&, Tendencies&, const BaseState&) (src/dyn/rk3.cpp:86-170): one step of the caller's
  // containers, whose buffers become (or stay) the domain's storage of their fields
  void step(State& s, Tendencies& t, const BaseState& b) {
    bind(s);
    bind(t);
    bind(b);
    step(1);
  }
  // reference RK3::set_lateral_bc (rk3.hpp:66-67): specified boundaries from bc's data; not owned
  void set_lateral_bc(LateralBC* bc) {
    bc_ = bc;
    pushed_.clear();
    if (bc_) push_boundary();
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
  void bind(const FieldSet& fs) {
    for (const Field& f : fs.fields()) {
      auto it = bound_.find(f.id());
      if (it != bound_.end() && it->second.first == f.data() && it->second.second == f.version()) continue;
      check(mi_field_bind(d_, f.id(), f.data(), f.bytes()));
      bound_[f.id()] = {f.data(), f.version()};
    }
  }
  void push_boundary() {
    for (const auto& kv : bc_->data_) {
      auto it = pushed_.find(kv.first);
      if (it != pushed_.end() && it->second == kv.second.version) continue;
      const auto& sl = kv.second;
      check(mi_bdy_set(d_, kv.first.first, kv.first.second, sl.values.data(), sl.tends.data(), int64_t(sl.values.size())));
      pushed_[kv.first] = sl.version;
    }
    check(mi_bdy_set_time(d_, bc_->dtbc));
  }
  mi_grid g_;
  mi_ctx* c_ = nullptr;
  mi_domain* d_ = nullptr;
  ColumnPhysics* cp_ = nullptr;
  LateralBC* bc_ = nullptr;
  std::map<int, std::pair<void*, uint64_t>> bound_;
  std::map<std::pair<int, int>, uint64_t> pushed_;
};

// the containers hold the domain's current values of their fields (those the domain has)
inline void FieldSet::make(RK3& rk, std::initializer_list<int> ids) {
  for (int id : ids) {
    mi_field_layout l{};
    if (mi_domain_field_layout(rk.handle(), id, &l) != MI_OK) continue;  // not a field of this domain
    Field f(rk.context(), id, l);
    std::vector<double> h(f.points());
    check(mi_domain_get_field(rk.handle(), id, h.data(), int64_t(h.size())));
    f.fill_host(h);
    f_.push_back(std::move(f));
  }
}

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
