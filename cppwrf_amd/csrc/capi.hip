// The extern "C" boundary of include/cppwrf_amd.h: argument checks, error capture (no exception crosses
// the ABI) and dispatch to the domain implementation of the requested arithmetic.
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/cppwrf_amd.h"
#include "domain_api.hpp"
#include "errors.hpp"

namespace {

using mi_detail::Error;

thread_local std::string g_last_error;

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(e == hipErrorOutOfMemory ? MI_ENOMEM : MI_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

mi_domain* checked(mi_domain* d) {
  if (!d) throw Error(MI_EINVAL, "null argument");
  hip_check(hipSetDevice(d->context()->device), "hipSetDevice");
  return d;
}

int run_phase(mi_domain* d, MiPhase which, int rk) {
  return mi_detail::guard([&] {
    if (which != kPhaseStepBegin && which != kPhaseStepEnd && (rk < 1 || rk > 3))
      throw Error(MI_EINVAL, "rk must be 1, 2 or 3");
    checked(d)->phase(which, rk);
  });
}

}  // namespace

// A batch: one graph per combination of the members' scalar-buffer parities (the captured launches name
// the buffers), each member's step a branch forked from and joined back to the context's stream.
struct mi_batch {
  mi_ctx* ctx = nullptr;
  std::vector<mi_domain*> m;
  std::vector<hipStream_t> st;
  // the stream the batch's graphs are captured on: the context's own stream may take other threads' work
  // meanwhile (domains built on the context), which a capture on it would record into the graph
  hipStream_t cap = nullptr;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  struct Graph {
    std::vector<int> key;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
  };
  std::vector<Graph> graphs;
  int schedule = MI_BATCH_BRANCHES;
  void* fused_dev = nullptr;  // the fused schedule's device tables (members' Args, block tables)
  size_t fused_cap = 0;
  FusedInfo fused_info;

  ~mi_batch() {
    for (Graph& gr : graphs) destroy(gr);
    if (fused_dev) (void)hipFree(fused_dev);
    for (hipEvent_t e : join) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    for (hipStream_t s : st) (void)hipStreamDestroy(s);
    if (cap) (void)hipStreamDestroy(cap);
  }
  bool fusable() const {
    if (m.empty() || m.size() > size_t(mi_fp64::kMaxFusedMembers)) return false;
    for (mi_domain* d : m)
      if (!d->fusable() || d->storage_bytes() != m[0]->storage_bytes() || d->n_sound() != m[0]->n_sound()) return false;
    return true;
  }
  std::vector<int> key() const {
    std::vector<int> k;
    for (mi_domain* d : m) k.push_back(d->buffer_parity());
    return k;
  }
  static void destroy(Graph& gr) {
    if (gr.x) (void)hipGraphExecDestroy(gr.x);
    if (gr.g) (void)hipGraphDestroy(gr.g);
    gr.x = nullptr;
    gr.g = nullptr;
  }
  hipGraphExec_t exec_for(const std::vector<int>& k) {
    for (Graph& gr : graphs)
      if (gr.key == k) return gr.x;
    // a member's launch schedule changed (set_dt, set_tuning, ...): the graphs of its older schedules are
    // never replayed again, so only those of the current schedules (one per buffer parity) are kept
    std::vector<Graph> keep;
    for (Graph& gr : graphs) {
      bool current = gr.key.size() == k.size();
      for (size_t q = 0; current && q < k.size(); ++q) current = (gr.key[q] >> 1) == (k[q] >> 1);
      if (current) keep.push_back(gr);
      else destroy(gr);
    }
    graphs.swap(keep);
    hipStream_t s0 = ctx->stream;
    // MI_BATCH_FUSED: each acoustic substep's kernels as one launch over all members (every member can join:
    // one arithmetic, one acoustic step count, no hooks); MI_BATCH_BRANCHES: each member's step a branch of its own
    const bool fused = schedule == MI_BATCH_FUSED && fusable();
    FusedInfo& info = fused_info;
    if (fused && graphs.empty()) {  // the tables of the members' current schedules (any buffer parity)
      const int n = int(m.size());
      const size_t need = m[0]->storage_bytes() == 8 ? mi_fp64::fused_bytes(n) : mi_fp32::fused_bytes(n);
      if (need > fused_cap) {
        if (fused_dev) hip_check(hipFree(fused_dev), "hipFree");
        fused_dev = nullptr;
        fused_cap = 0;
        hip_check(hipMalloc(&fused_dev, need), "hipMalloc");
        fused_cap = need;
      }
      // the tables are filled before the capture (graphs read them at every replay; a new capture refills them)
      hip_check(hipStreamSynchronize(s0), "hipStreamSynchronize");
      if (m[0]->storage_bytes() == 8) mi_fp64::fused_prepare(m.data(), n, fused_dev, s0, info);
      else mi_fp32::fused_prepare(m.data(), n, fused_dev, s0, info);
    }
    hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      if (fused) {
        if (m[0]->storage_bytes() == 8)
          mi_fp64::fused_capture(m.data(), int(m.size()), cap, st.data(), fork, join.data(), fused_dev, info);
        else
          mi_fp32::fused_capture(m.data(), int(m.size()), cap, st.data(), fork, join.data(), fused_dev, info);
      } else {
        hip_check(hipEventRecord(fork, cap), "hipEventRecord");
        for (size_t q = 0; q < m.size(); ++q) {
          hip_check(hipStreamWaitEvent(st[q], fork, 0), "hipStreamWaitEvent");
          m[q]->capture_step(st[q]);
          hip_check(hipEventRecord(join[q], st[q]), "hipEventRecord");
          hip_check(hipStreamWaitEvent(cap, join[q], 0), "hipStreamWaitEvent");
        }
      }
    } catch (...) {
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(cap, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      throw;
    }
    Graph gr;
    gr.key = k;
    hip_check(hipStreamEndCapture(cap, &gr.g), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&gr.x, gr.g, nullptr, nullptr, 0), "hipGraphInstantiate");
    graphs.push_back(gr);
    return gr.x;
  }
};

namespace mi_detail {
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace mi_detail

void mi_ctx_release(mi_ctx* c) {
  if (c->refs.fetch_sub(1) != 1) return;
  if (c->own_stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

extern "C" {

const char* mi_last_error(void) { return g_last_error.c_str(); }
int mi_abi_version(void) { return MI_ABI_VERSION; }

// reference RK3::auto_sound_steps (src/dyn/rk3.cpp:69-72)
int mi_auto_sound_steps(double dt, double dx, double dy) {
  const double spacing = dx < dy ? dx : dy;
  const int n = 2 * (int(300.0 * dt / spacing - 0.01) + 1);
  return n > 4 ? n : 4;
}

// reference RK3::adaptive_sound_steps (src/dyn/rk3.cpp:74-77), in double
int mi_adaptive_sound_steps(double dt, double dx, double dy, double max_msft) {
  const double spacing = dx < dy ? dx : dy;
  const int n = 2 * (int(300.0 * dt / (spacing / max_msft) - 0.01) + 1);
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
    if (c) mi_ctx_release(c);
  });
}

int mi_domain_create(mi_ctx* ctx, const mi_grid* g, const mi_params* p, const mi_levels* lv, mi_domain** out) {
  return mi_detail::guard([&] {
    if (!ctx || !g || !p || !lv || !out) throw Error(MI_EINVAL, "null argument");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (p->precision == MI_FP64) *out = mi_fp64::make_domain(ctx, *g, *p, *lv);
    else if (p->precision == MI_FP32_BULK) *out = mi_fp32::make_domain(ctx, *g, *p, *lv);
    else throw Error(MI_EINVAL, "precision must be MI_FP64 or MI_FP32_BULK");
    ctx->refs.fetch_add(1);
  });
}

int mi_domain_destroy(mi_domain* d) {
  return mi_detail::guard([&] {
    if (!d) return;
    mi_ctx* c = d->context();
    (void)hipStreamSynchronize(c->stream);
    delete d;
    mi_ctx_release(c);
  });
}

int mi_domain_n_sound(const mi_domain* d) { return d ? d->n_sound() : MI_EINVAL; }
int64_t mi_domain_device_bytes(const mi_domain* d) { return d ? d->device_bytes() : MI_EINVAL; }

int mi_domain_set_field(mi_domain* d, int id, const double* host, int64_t n) {
  return mi_detail::guard([&] {
    if (!host) throw Error(MI_EINVAL, "null argument");
    checked(d)->set_field(id, host, n);
  });
}

int mi_domain_field_view(mi_domain* d, int id, mi_field_view* out) {
  return mi_detail::guard([&] {
    if (!out) throw Error(MI_EINVAL, "null argument");
    checked(d)->field_view(id, out);
  });
}

int mi_domain_field_layout(const mi_domain* d, int id, mi_field_layout* out) {
  return mi_detail::guard([&] {
    if (!d || !out) throw Error(MI_EINVAL, "null argument");
    d->field_layout(id, out);
  });
}

int mi_field_bind(mi_domain* d, int id, void* dev, int64_t bytes) {
  return mi_detail::guard([&] { checked(d)->field_bind(id, dev, bytes); });
}

int mi_device_alloc(mi_ctx* c, int64_t bytes, void** out) {
  return mi_detail::guard([&] {
    if (!c || !out || bytes <= 0) throw Error(MI_EINVAL, "bad argument");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    *out = nullptr;
    hip_check(hipMalloc(out, size_t(bytes)), "hipMalloc");
  });
}

int mi_device_free(mi_ctx* c, void* p) {
  return mi_detail::guard([&] {
    if (!c) throw Error(MI_EINVAL, "null argument");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (p) hip_check(hipFree(p), "hipFree");
  });
}

int mi_device_upload(mi_ctx* c, void* dst, const void* src, int64_t n) {
  return mi_detail::guard([&] {
    if (!c || !dst || !src || n < 0) throw Error(MI_EINVAL, "bad argument");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hip_check(hipMemcpyAsync(dst, src, size_t(n), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

int mi_device_download(mi_ctx* c, void* dst, const void* src, int64_t n) {
  return mi_detail::guard([&] {
    if (!c || !dst || !src || n < 0) throw Error(MI_EINVAL, "bad argument");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hip_check(hipMemcpyAsync(dst, src, size_t(n), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

int mi_domain_get_field(mi_domain* d, int id, double* host, int64_t n) {
  return mi_detail::guard([&] {
    if (!host) throw Error(MI_EINVAL, "null argument");
    checked(d)->get_field(id, host, n);
  });
}

int mi_step(mi_domain* d, int nsteps) {
  return mi_detail::guard([&] {
    if (nsteps < 0) throw Error(MI_EINVAL, "nsteps must be >= 0");
    checked(d)->step(nsteps);
  });
}

int mi_domain_sync(mi_domain* d) {
  return mi_detail::guard([&] { hip_check(hipStreamSynchronize(checked(d)->context()->stream), "hipStreamSynchronize"); });
}

int mi_step_begin(mi_domain* d) { return run_phase(d, kPhaseStepBegin, 0); }
int mi_rk_prep(mi_domain* d, int rk) { return run_phase(d, kPhaseRkPrep, rk); }
int mi_rk_tendency(mi_domain* d, int rk) { return run_phase(d, kPhaseRkTendency, rk); }
int mi_small_steps(mi_domain* d, int rk) { return run_phase(d, kPhaseSmallSteps, rk); }
int mi_scalar_transport(mi_domain* d, int rk) { return run_phase(d, kPhaseScalars, rk); }
int mi_calc_p_rho_phi(mi_domain* d, int rk) { return run_phase(d, kPhaseCalcPRhoPhi, rk); }
int mi_calc_p_rho_phi_interior(mi_domain* d, int spec_zone) {
  return mi_detail::guard([&] { checked(d)->calc_p_rho_phi_interior(spec_zone); });
}
int mi_step_end(mi_domain* d) { return run_phase(d, kPhaseStepEnd, 0); }

int mi_bdy_set(mi_domain* d, int var, int side, const double* values, const double* tends, int64_t n) {
  return mi_detail::guard([&] {
    if (!values || !tends) throw Error(MI_EINVAL, "null boundary data");
    checked(d)->bdy_set(var, side, values, tends, n);
  });
}
int mi_bdy_set_time(mi_domain* d, double dtbc) {
  return mi_detail::guard([&] { checked(d)->bdy_set_time(dtbc); });
}
int mi_bdy_get_time(mi_domain* d, double* dtbc) {
  return mi_detail::guard([&] {
    if (!dtbc) throw Error(MI_EINVAL, "null argument");
    *dtbc = checked(d)->bdy_get_time();
  });
}

int mi_domain_autotune(mi_domain* d, int trials) {
  return mi_detail::guard([&] { checked(d)->autotune(trials); });
}

int mi_domain_set_tuning(mi_domain* d, const int* shapes, int n) {
  return mi_detail::guard([&] {
    if (!shapes) throw Error(MI_EINVAL, "null argument");
    checked(d)->set_tuning(shapes, n);
  });
}

int mi_domain_get_tuning(const mi_domain* d, int* shapes, int n) {
  return mi_detail::guard([&] {
    if (!d || !shapes) throw Error(MI_EINVAL, "null argument");
    d->get_tuning(shapes, n);
  });
}

int mi_profile_kernel(mi_domain* d, const char* name) {
  return mi_detail::guard([&] { checked(d)->profile_kernel(name); });
}

int mi_profile_read(mi_domain* d, double* total_ms, int* launches, double* alg_bytes) {
  return mi_detail::guard([&] { checked(d)->profile_read(total_ms, launches, alg_bytes); });
}

int mi_batch_create(mi_ctx* ctx, mi_domain* const* members, int n, mi_batch** out) {
  return mi_detail::guard([&] {
    if (!ctx || !members || !out || n < 1) throw Error(MI_EINVAL, "bad argument");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    auto b = std::make_unique<mi_batch>();
    b->ctx = ctx;
    for (int q = 0; q < n; ++q) {
      if (!members[q]) throw Error(MI_EINVAL, "null member");
      if (members[q]->context()->device != ctx->device) throw Error(MI_EINVAL, "members must be on the batch's device");
      for (int r = 0; r < q; ++r)
        if (members[r] == members[q]) throw Error(MI_EINVAL, "a domain appears twice in a batch");
      b->m.push_back(members[q]);
      hipStream_t s = nullptr;
      hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");  // default priority
      b->st.push_back(s);
      hipEvent_t e = nullptr;
      hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      b->join.push_back(e);
    }
    hip_check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking), "hipStreamCreate");  // default priority
    ctx->refs.fetch_add(1);
    *out = b.release();
  });
}

int mi_batch_destroy(mi_batch* b) {
  return mi_detail::guard([&] {
    if (!b) return;
    mi_ctx* c = b->ctx;
    (void)hipStreamSynchronize(c->stream);
    delete b;
    mi_ctx_release(c);
  });
}

int mi_batch_step(mi_batch* b, int nsteps) {
  return mi_detail::guard([&] {
    if (!b || nsteps < 0) throw Error(MI_EINVAL, "bad argument");
    hip_check(hipSetDevice(b->ctx->device), "hipSetDevice");
    // the members' own streams may still hold work: the batch's stream waits for it
    for (mi_domain* d : b->m)
      if (d->context()->stream != b->ctx->stream) {
        hip_check(hipEventRecord(b->fork, d->context()->stream), "hipEventRecord");
        hip_check(hipStreamWaitEvent(b->ctx->stream, b->fork, 0), "hipStreamWaitEvent");
      }
    for (mi_domain* d : b->m)
      if (d->has_callbacks()) throw Error(MI_EINVAL, "a batch member has stage callbacks: step it alone");
    for (int s = 0; s < nsteps; ++s) {
      hipGraphExec_t x = b->exec_for(b->key());
      hip_check(hipGraphLaunch(x, b->ctx->stream), "hipGraphLaunch");
      for (mi_domain* d : b->m) d->advance_host();
    }
    // and each member's stream waits for the batch, so work queued on a member afterwards follows it
    hip_check(hipEventRecord(b->fork, b->ctx->stream), "hipEventRecord");
    for (mi_domain* d : b->m)
      if (d->context()->stream != b->ctx->stream)
        hip_check(hipStreamWaitEvent(d->context()->stream, b->fork, 0), "hipStreamWaitEvent");
  });
}

int mi_batch_set_schedule(mi_batch* b, int schedule) {
  return mi_detail::guard([&] {
    if (!b) throw Error(MI_EINVAL, "null batch");
    if (schedule != MI_BATCH_BRANCHES && schedule != MI_BATCH_FUSED)
      throw Error(MI_EINVAL, "schedule: MI_BATCH_BRANCHES or MI_BATCH_FUSED");
    if (schedule == MI_BATCH_FUSED && !b->fusable())
      throw Error(MI_EINVAL, "fused schedule: every member needs the same arithmetic and acoustic step count, no "
                             "callbacks, nz within the batched column kernel, and at most 64 members");
    if (schedule == b->schedule) return;
    hip_check(hipStreamSynchronize(b->ctx->stream), "hipStreamSynchronize");
    for (mi_batch::Graph& gr : b->graphs) mi_batch::destroy(gr);
    b->graphs.clear();
    b->schedule = schedule;
  });
}

int mi_batch_graph_count(const mi_batch* b) { return b ? int(b->graphs.size()) : MI_EINVAL; }

int mi_cfl_max(mi_domain* d, double* vert, double* horiz) {
  return mi_detail::guard([&] { checked(d)->cfl_max(vert, horiz); });
}

int mi_cfl_max_noalloc(mi_domain* d, double* vert, double* horiz) {
  return mi_detail::guard([&] { checked(d)->cfl_max_noalloc(vert, horiz); });
}

int mi_domain_set_dt(mi_domain* d, double dt, int n_sound) {
  return mi_detail::guard([&] { checked(d)->set_dt(dt, n_sound); });
}

int64_t mi_domain_state_bytes(mi_domain* d) {
  int64_t b = MI_EINVAL;
  const int rc = mi_detail::guard([&] { b = checked(d)->state_bytes(); });
  return rc == MI_OK ? b : rc;
}

int mi_domain_save_state(mi_domain* d, void* host, int64_t bytes) {
  return mi_detail::guard([&] { checked(d)->save_state(host, bytes); });
}

int mi_domain_load_state(mi_domain* d, const void* host, int64_t bytes) {
  return mi_detail::guard([&] { checked(d)->load_state(host, bytes); });
}

int mi_domain_set_cfl_callback(mi_domain* d, mi_cfl_fn fn, void* user) {
  return mi_detail::guard([&] { checked(d)->set_cfl_callback(fn, user); });
}

int mi_domain_set_first_rk_prep_callback(mi_domain* d, mi_stage_fn fn, void* user) {
  return mi_detail::guard([&] { checked(d)->set_prep_callback(fn, user); });
}

int mi_domain_set_column_physics(mi_domain* d, const mi_column_physics* cp) {
  return mi_detail::guard([&] { checked(d)->set_column_physics(cp); });
}

int mi_domain_set_diffusion(mi_domain* d, int opt, double factor) {
  return mi_detail::guard([&] { checked(d)->set_diffusion(opt, factor); });
}

int mi_check_finite(mi_domain* d, int* bad_field) {
  int bad = -1;
  const int rc = mi_detail::guard([&] { bad = checked(d)->check_finite(bad_field); });
  if (rc != MI_OK) return rc;
  if (bad >= 0) {
    mi_detail::set_last_error("non-finite value in field " + std::to_string(bad));
    return MI_ENONFINITE;
  }
  return MI_OK;
}

}  // extern "C"
