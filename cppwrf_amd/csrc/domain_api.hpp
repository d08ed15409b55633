// The precision-independent interface of a domain behind the C ABI.  One implementation per arithmetic
// (domain.hip compiled for FP64 and for FP32-bulk storage); the extern "C" functions of capi.hip reach
// them through this interface.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

#include "../../include/cppwrf_amd.h"

// A context lives until it is destroyed and no domain or batch made on it remains (each holds a
// reference), so the order in which a caller releases objects cannot free a stream still in use.
struct mi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::atomic<int> refs{1};  // the caller's, plus one per live domain or batch
};
void mi_ctx_release(mi_ctx* c);  // drop one reference; the last one frees the context

// phases of RK3::step for the phase-by-phase entry points
enum MiPhase {
  kPhaseStepBegin, kPhaseRkPrep, kPhaseRkTendency, kPhaseSmallSteps, kPhaseScalars, kPhaseCalcPRhoPhi, kPhaseStepEnd
};

struct mi_domain {
  virtual ~mi_domain() = default;
  virtual mi_ctx* context() const = 0;
  virtual int n_sound() const = 0;
  virtual int64_t device_bytes() const = 0;
  virtual void set_field(int id, const double* host, int64_t n) = 0;
  virtual void get_field(int id, double* host, int64_t n) = 0;
  virtual void field_view(int id, mi_field_view* out) = 0;
  virtual void field_layout(int id, mi_field_layout* out) const = 0;
  virtual void field_bind(int id, void* dev, int64_t bytes) = 0;
  virtual void step(int nsteps) = 0;
  virtual void phase(MiPhase which, int rk) = 0;
  virtual void autotune(int trials) = 0;
  virtual void set_tuning(const int* shapes, int n) = 0;
  virtual void get_tuning(int* shapes, int n) const = 0;
  virtual void profile_kernel(const char* name) = 0;
  virtual void profile_read(double* total_ms, int* launches, double* alg_bytes) = 0;
  virtual int check_finite(int* bad_field) = 0;  // -1 or the first field id with a non-finite value
  virtual void cfl_max(double* vert, double* horiz) = 0;
  virtual void cfl_max_noalloc(double* vert, double* horiz) = 0;
  virtual void set_dt(double dt, int n_sound) = 0;
  virtual int64_t state_bytes() = 0;
  virtual void save_state(void* host, int64_t bytes) = 0;
  virtual void load_state(const void* host, int64_t bytes) = 0;
  virtual void set_cfl_callback(mi_cfl_fn fn, void* user) = 0;
  virtual void set_prep_callback(mi_stage_fn fn, void* user) = 0;
  virtual bool has_callbacks() const = 0;
  virtual void set_column_physics(const mi_column_physics* cp) = 0;
  virtual void calc_p_rho_phi_interior(int spec_zone) = 0;
  virtual void set_diffusion(int opt, double factor) = 0;
  // specified lateral boundaries (LateralBC::set / set_scalar, dtbc)
  virtual void bdy_set(int var, int side, const double* values, const double* tends, int64_t n) = 0;
  virtual void bdy_set_time(double dtbc) = 0;
  virtual double bdy_get_time() = 0;
  // batches: launch one step onto stream s (the caller captures it), the host bookkeeping of a step that
  // ran elsewhere, and the parity of the scalar buffers a captured step names
  virtual void capture_step(hipStream_t s) = 0;
  virtual void advance_host() = 0;
  virtual int buffer_parity() const = 0;
  // fused batch steps: the storage size of the bulk fields (4 or 8) and whether this domain can join one (no hooks,
  // its levels fit the batched column kernel)
  virtual int storage_bytes() const = 0;
  virtual bool fusable() const = 0;
};

// what a fused batch step launches with: the acoustic step count, the largest nz, each stage's block counts
struct FusedInfo {
  int ns = 0, nzmax = 0;
  int blocks_uv[3] = {0, 0, 0}, blocks_ac[3] = {0, 0, 0};
};

// one implementation per arithmetic: domains, and the fused batch steps of members of that arithmetic (fused_bytes
// of device memory for n members; fused_prepare fills it, before a capture; fused_capture launches one step)
#define MI_DOMAIN_NS_API                                                                                          \
  mi_domain* make_domain(mi_ctx* ctx, const mi_grid& g, const mi_params& p, const mi_levels& lv);              \
  size_t fused_bytes(int n);                                                                                     \
  void fused_prepare(mi_domain* const* ms, int n, void* dev, hipStream_t st, FusedInfo& info);                  \
  void fused_capture(mi_domain* const* ms, int n, hipStream_t s0, const hipStream_t* st, hipEvent_t fork,        \
                     const hipEvent_t* join, const void* dev, const FusedInfo& info);
namespace mi_fp64 {
MI_DOMAIN_NS_API
constexpr int kMaxFusedMembers = 64;  // the batched kernels' member tables
}
namespace mi_fp32 {
MI_DOMAIN_NS_API
}
