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

#include <stdexcept>
#include <string>
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

// The device-resident State + BaseState of the domain (reference src/dyn/state.hpp:30-119)
class StateView {
 public:
  explicit StateView(mi_domain* d, const mi_grid& g) : d_(d), g_(g) {}
  size_t size(int field) const {
    const bool u = field == MI_U || field == MI_RU_M || field == MI_RU_TEND;
    const bool v = field == MI_V || field == MI_RV_M || field == MI_RV_TEND;
    const bool w = field == MI_W || field == MI_PH || field == MI_WW || field == MI_WW_M || field == MI_RW_TEND ||
                   field == MI_PH_TEND || field == MI_PHB;
    const bool two = field == MI_MU || field == MI_MUTS || field == MI_MUB;
    const size_t nxs = size_t(g_.nx + (u ? 1 : 0)), nys = size_t(g_.ny + (v && g_.ny > 1 ? 1 : 0));
    const size_t nzs = two ? 1 : size_t(g_.nz + (w ? 1 : 0));
    return nxs * nys * nzs;
  }
  void set(int field, const std::vector<double>& h) { check(mi_domain_set_field(d_, field, h.data(), int64_t(h.size()))); }
  std::vector<double> get(int field) const {
    std::vector<double> h(size(field));
    check(mi_domain_get_field(d_, field, h.data(), int64_t(h.size())));
    return h;
  }

 private:
  mi_domain* d_;
  mi_grid g_;
};

// reference dyn::RK3 (src/dyn/rk3.hpp:34-78)
class RK3 {
 public:
  RK3(const Context& ctx, const mi_grid& g, const LargeStepParams& lp, const AcousticParams& ap, int time_step_sound,
      const mi_levels& levels, int precision = MI_FP64)
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
    p.precision = precision;
    check(mi_domain_create(ctx.get(), &g, &p, &levels, &d_));
  }
  ~RK3() { mi_domain_destroy(d_); }
  RK3(const RK3&) = delete;
  RK3& operator=(const RK3&) = delete;

  // reference RK3::auto_sound_steps (src/dyn/rk3.cpp:69-72)
  static int auto_sound_steps(double dt, double dx, double dy) { return mi_auto_sound_steps(dt, dx, dy); }
  int n_sound() const { return mi_domain_n_sound(d_); }
  // reference RK3::step (src/dyn/rk3.cpp:86-170); nsteps > 1 runs them back to back on the device
  void step(int nsteps = 1) { check(mi_step(d_, nsteps)); }
  void sync() { check(mi_domain_sync(d_)); }
  StateView state() { return StateView(d_, g_); }
  mi_domain* handle() const { return d_; }

 private:
  mi_grid g_;
  mi_domain* d_ = nullptr;
};

}  // namespace cppwrf_amd::dyn
