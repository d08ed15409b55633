/* cppwrf_amd: the RK3 split-explicit dynamics step of WRF-ARW on AMD MI355X
 * (gfx950), behind a C ABI.
 *
 * The reference has no C ABI: its operator interface is the C++ src/dyn
 * surface, dyn::RK3 (reference src/dyn/rk3.hpp:34-78) with
 *   RK3(const Grid&, const LargeStepParams&, const AcousticParams&,
 *       int time_step_sound, const ScalarParams&)
 *   void step(State&, Tendencies&, const BaseState&)
 * and the containers Grid (src/core/grid.hpp:21-55), State / Tendencies /
 * BaseState (src/dyn/state.hpp:30-119).  This header is the replacement
 * boundary: a domain handle owns the device-resident State and BaseState of
 * one forecast, the caller hands the initial fields over as plain host arrays
 * and steps it.  The C++ facade cppwrf_amd/csrc/dyn_facade.hpp gives the
 * reference's class names and methods on top of these functions.
 *
 * Conventions
 *  - Every function returns MI_OK (0) or a negative MI_E* code; no exception
 *    crosses the ABI.  mi_last_error() gives the message of the calling
 *    thread's last failure.
 *  - Arrays are double, dense, logical points only, i fastest, then j, then k:
 *    element (i, j, k) of a field with nxs x nys x nzs points is
 *    a[(k * nys + j) * nxs + i].  Staggering as the reference
 *    (src/core/grid.hpp:3-7): U has nx+1 points in x, V ny+1 in y (ny == 1:
 *    one), W nz+1 levels.
 *  - The caller owns every pointer it passes; the library copies what it
 *    needs before the call returns.
 *  - One domain must not be used by two host threads at once.  Different
 *    domains may be stepped from different threads, also domains of one
 *    context: a domain may be created on a context while another thread
 *    steps the context's domains (graphs are captured on a stream of the
 *    domain's or batch's own, never on the context's stream).
 */
#ifndef CPPWRF_AMD_H
#define CPPWRF_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_ABI_VERSION 5

/* status codes */
#define MI_OK 0
#define MI_EINVAL (-1)     /* bad argument (reference: std::runtime_error for bad parameters, rk3.cpp:50) */
#define MI_EHIP (-2)       /* HIP runtime error (reference: CUDA errors, backend.cpp:23-25) */
#define MI_ENOMEM (-3)     /* device allocation failed (reference: backend.cpp:312-331) */
#define MI_ENONFINITE (-4) /* a non-finite value was found by mi_check_finite */
#define MI_ENODEV (-5)     /* no GPU, or the library was built without device code */

/* lateral boundary conditions, in the order of the reference's enum BC (src/core/grid.hpp:19).  SPECIFIED on
 * both directions or neither (a 3-D grid of at least 10 x 10 columns, data through mi_bdy_*); OPEN (WRF
 * open_xs/xe, open_ys/ye: radiation faces and inflow terms) on either or both, at least 7 columns across */
enum { MI_BC_PERIODIC = 0, MI_BC_SYMMETRIC = 1, MI_BC_SPECIFIED = 2, MI_BC_OPEN = 3 };

/* arithmetic of a domain: plain FP64 (bit-identical to the reference CPU FP64 build whose calc_p_rho_phi takes
 * glibc's scalar pow on every point, i.e. built with CPPWRF_NO_LIBMVEC) or FP32-bulk: the
 * 3-D prognostic, tendency, scalar and work fields stored in FP32 (half the HBM traffic), with FP64 kept
 * for the base state, the 2-D column masses, the full geopotential php, the advance_w tridiagonal solve,
 * the column mass-divergence sums and calc_p_rho; fields cross the ABI as double either way */
enum { MI_FP64 = 0, MI_FP32_BULK = 1 };

/* reference Grid (src/core/grid.hpp:21-31) */
typedef struct mi_grid {
  int32_t nx, ny, nz; /* mass points; ny == 1 is an x-z slice (the j stride is 0) */
  int32_t bc_x, bc_y; /* MI_BC_* */
  double dx, dy;      /* m */
  double dt;          /* large step, s */
} mi_grid;

/* reference LargeStepParams (src/dyn/large_step.hpp:14-22) and AcousticParams (src/dyn/rk3.hpp:20-27) */
typedef struct mi_params {
  int32_t h_adv_order;     /* 5 (upwind) or 6 (centred) */
  int32_t v_adv_order;     /* 3 or 4 */
  int32_t top_lid;         /* 0: constant-pressure top (WRF default) */
  int32_t w_damping;       /* 1: damp w where the vertical Courant number exceeds 1 */
  double w_crit_cfl;
  double epssm, smdiv, emdiv;
  /* damping in the top zdamp metres (dampcoef its strength): 0 off; 1 the horizontal eddy viscosities raised to a
   * damping profile (with diff_opt 1; reference cal_dampkm, src/dyn/diffusion/km_kh.hpp:485-568); 2 Rayleigh
   * relaxation of u, v, w and theta towards MI_U_BASE, MI_V_BASE, 0 and MI_T_BASE interpolated to each point's
   * height (MI_Z_BASE), into the held tendencies at the first RK stage (rk_rayleigh_damp,
   * src/dyn/diffusion/rayleigh_damp.hpp); 3 implicit Rayleigh damping of w in the acoustic steps */
  int32_t damp_opt;
  int32_t hypsometric_opt; /* 1 */
  double zdamp, dampcoef;
  int32_t time_step_sound; /* acoustic steps per large step (even); 0: the reference's automatic rule */
  int32_t precision;       /* MI_FP64 or MI_FP32_BULK */
  /* advected scalars (reference State::q_1 / q_2 and ScalarParams, src/dyn/state.hpp:53-60,
   * src/dyn/scalars.hpp:14-19): the first n_moist are moist species (qv first), which couple to the
   * dynamics through calc_cq and theta_m; the rest are passive tracers */
  int32_t n_moist;         /* 0..MI_MAX_MOIST */
  int32_t n_scalars;       /* n_moist..MI_MAX_SCALARS */
  int32_t moist_adv_opt;   /* 0: advect_scalar; 1: positive-definite advect_scalar_pd at the last stage */
  int32_t scalar_adv_opt;  /* the same for the passive tracers */
  /* map projection and Coriolis (reference BaseState::mapped / coriolis, src/dyn/state.hpp:93-107):
   * mapped = 1 adds the map-factor curvature terms (vxgm) to the u and v tendencies; coriolis = 1 the
   * Coriolis terms with the fields MI_F, MI_E, MI_SINA, MI_COSA (both as the reference's MapC::cor) */
  int32_t mapped;
  int32_t coriolis;
  /* sixth-order horizontal hyperdiffusion (reference sixth_order_diffusion, src/dyn/diffusion/sixth_order.hpp):
   * 0 off, 1 linear, 2 monotonic flux-limited; factor = the non-dimensional diff_6th_factor.  Applied at the
   * first RK stage into the held tendencies of u, v, w, theta and every scalar */
  int32_t diff_6th_opt;
  /* 1: terrain (reference BaseState not flat): the base state pb, alb, phb, t_init is given in full
   * (nxs*nys*nzs values), the terrain height MI_HT sets the lower boundary w (advance_w's surface w''
   * and set_w_surface at the end of each step, src/dyn/small_step.cpp:339-353, rk3.cpp:173-200) */
  int32_t terrain;
  double diff_6th_factor;
  /* second-order diffusion along eta (reference DiffusionDriver with diff_opt 1, src/dyn/diffusion_driver.cpp:92-193,
   * 297-327; src/dyn/diffusion/big_step_diffusion.hpp): diff_opt 0 off or 1 (2, physical space, is not built);
   * km_opt 1 (constant K: khdif, kvdif, m2/s) or 4 (the horizontal K from the 2-D Smagorinsky closure of the
   * step's horizontal deformation, c_s its constant, 0 takes the reference's default 0.25; km_kh.hpp:47-102,
   * deform_div.hpp:20-170; the vertical K stays kvdif); the TKE and 3-D Smagorinsky options 2 and 3 are not
   * built.  pr_inv = 1 / prandtl for the scalars' K (0 takes the reference's default 3).  Applied at the first
   * RK stage into the held tendencies of u, v, w, theta (vertical K of theta 3 kvdif, as WRF's rk_tendency)
   * and every scalar, before the sixth-order diffusion.  The base profiles MI_U_BASE, MI_V_BASE, MI_QV_BASE
   * enter the vertical diffusion of u, v and qv. */
  int32_t diff_opt;
  int32_t km_opt;
  double khdif, kvdif, pr_inv;
  /* microphysics after the RK loop (reference RK3::enable_physics with PhysicsCoupling, src/dyn/rk3.cpp:151-158,
   * src/dyn/physics_coupling.cpp:262-455): mp_physics 0 off or 1 (Kessler warm rain, moist species 0..2 = qv,
   * qc, qr; the reference's ice schemes are not built); mp_tend_lim the limit of the heating rate, K/s (0 takes
   * the reference's default 10).  Each step ends with moist_physics_prep, Kessler, moist_physics_finish (the
   * heating kept in MI_H_DIABATIC for the next step) and calc_p_rho_phi. */
  int32_t mp_physics;
  double mp_tend_lim;
  double c_s; /* the Smagorinsky constant of km_opt 4 (appended in ABI 5) */
} mi_params;

#define MI_MAX_MOIST 8    /* reference KCq::kMax (src/dyn/large_step.cpp:128) */
#define MI_MAX_SCALARS 16

/* reference BaseState vertical coefficients (src/dyn/state.hpp:109-116, set by BaseState::set_levels,
 * src/dyn/state.cpp:86-116).  Arrays of nz (mass levels) or nz+1 (w levels) values. */
typedef struct mi_levels {
  const double *dnw, *rdnw;              /* nz */
  const double *dn, *rdn, *fnm, *fnp;    /* nz+1 (index 0 and nz unused) */
  const double *c1h, *c2h, *c3h, *c4h;   /* nz */
  const double *c1f, *c2f, *c3f, *c4f;   /* nz+1 */
  double cf1, cf2, cf3, cfn, cfn1;
  double rdx, rdy; /* 1/dx and 1/dy (0 for an x-z slice, as the reference BaseState) */
  double p_top;
} mi_levels;

/* field identifiers for mi_domain_set_field / mi_domain_get_field */
enum {
  /* state carried from step to step (reference State::*_2, p, al) */
  MI_U = 0, MI_V = 1, MI_W = 2, MI_T = 3, MI_PH = 4, MI_MU = 5, MI_P = 6, MI_AL = 7,
  /* diagnostics of the last stage (read only) */
  MI_WW = 8, MI_RU_M = 9, MI_RV_M = 10, MI_WW_M = 11, MI_ALT = 12, MI_PHP = 13, MI_MUTS = 14,
  MI_RU_TEND = 15, MI_RV_TEND = 16, MI_RW_TEND = 17, MI_T_TEND = 18, MI_PH_TEND = 19,
  /* calc_cq coefficients of the last stage (read only; with moisture): U, V and W points */
  MI_CQU = 20, MI_CQV = 21, MI_CQW = 22,
  /* microphysics heating of the last step, theta_m K/s (input; zero unless set; used with moisture:
   * rk_addtend_dry and small_step_finish, src/dyn/large_step.cpp:557, small_step.cpp:437-440) */
  MI_H_DIABATIC = 23,
  /* held tendencies of the first stage (reference Tendencies::ru_tendf, rv_tendf, rw_tendf, t_tendf, coupled;
   * src/dyn/state.hpp:85-88): allocated when a diffusion, specified boundaries or a column physics plugin is
   * attached.  At the first stage of every step they are zeroed, then take the plugin's part2 (set here),
   * the sixth-order diffusion and the boundary relaxation, in that order (large_step.cpp:811-825) */
  MI_RU_TENDF = 24, MI_RV_TENDF = 25, MI_RW_TENDF = 26, MI_T_TENDF = 27,
  /* the advective forcings of a column physics plugin that asks for them (read only; WRF RTHFTEN, RQVFTEN:
   * the advection tendency of theta and of the vapour times msfty, of the last stage;
   * large_step.cpp:506, 543-555, scalars.cpp:147-150, 260-287) */
  MI_TH_FORCING = 28, MI_QV_FORCING = 29,
  /* base state (write once, before the first step).  3-D fields take nzs values for a horizontally
   * uniform profile or nxs*nys*nzs values; 2-D fields take 1 value or nx*ny values. */
  MI_MUB = 32, MI_PB = 33, MI_ALB = 34, MI_PHB = 35, MI_T_INIT = 36,
  MI_MSFTX = 40, MI_MSFTY = 41, MI_MSFUX = 42, MI_MSFUY = 43, MI_MSFVX = 44, MI_MSFVY = 45,
  MI_MSFVX_INV = 46,
  /* Coriolis parameters 2 Omega sin(lat), 2 Omega cos(lat) and the grid rotation (2-D mass points) */
  MI_F = 47, MI_E = 48, MI_SINA = 49, MI_COSA = 50,
  MI_HT = 51, /* terrain height, m (2-D mass points; with terrain) */
  /* the 1-D base profiles of the diffusion (reference BaseState::u_base, v_base, qv_base, src/dyn/state.hpp:100,
   * 112; the ideal cases take u and v of column (0, 0), src/cases/model.cpp:149-177): nz values each, zero
   * unless set */
  MI_U_BASE = 52, MI_V_BASE = 53, MI_QV_BASE = 54,
  /* the microphysics' accumulated and last-step grid-scale precipitation, mm (2-D mass points; with mp_physics) */
  MI_RAINNC = 55, MI_RAINNCV = 56,
  /* the horizontal eddy viscosities of the step (3-D mass points, read only; with km_opt 4 or damp_opt 1): xkmh of the
   * momenta and xkhh = xkmh / prandtl of theta and the scalars, m2/s.  km_opt 4 leaves 0 on the outermost
   * rows of open and specified sides, as the reference */
  MI_XKMH = 57, MI_XKHH = 58,
  /* the 1-D base profiles of damp_opt 2 (reference BaseState::t_base, z_base; the ideal cases take theta of
   * t_init and the height of the mass levels of column (0, 0), src/cases/model.cpp:149-177): nz values each */
  MI_T_BASE = 59, MI_Z_BASE = 60,
  /* scalar n (0 <= n < n_scalars) at mass points: MI_Q + n is its state (q_2), MI_Q_TENDF + n its held
   * tendency (input, mass-coupled, held over the step as the reference's q_tendf) */
  MI_Q = 64, MI_Q_TENDF = 96
};

typedef struct mi_ctx mi_ctx;       /* a GPU and the HIP stream its domains launch on */
typedef struct mi_domain mi_domain; /* one forecast: device-resident State, BaseState, work fields */

/* message of the calling thread's last failure ("" if none) */
const char* mi_last_error(void);
int mi_abi_version(void);

/* reference RK3::auto_sound_steps (src/dyn/rk3.cpp:69-72): max(2*(int(300 dt/min(dx,dy) - 0.01)+1), 4) */
int mi_auto_sound_steps(double dt, double dx, double dy);
/* RK3::adaptive_sound_steps (rk3.hpp:40-41, rk3.cpp:74-77): the acoustic step count of an adaptive time step,
 * max(2*(int(300*dt/(min(dx,dy)/max_msft) - 0.01) + 1), 4), evaluated in double as the reference's FP64 build */
int mi_adaptive_sound_steps(double dt, double dx, double dy, double max_msft);

/* A context on HIP device `device`.  stream: a hipStream_t to launch on, or NULL for a stream the
 * context creates (reference: the thread-local launch stream, src/core/backend.cpp:73-81). */
int mi_ctx_create(int device, void* stream, mi_ctx** out);
int mi_ctx_destroy(mi_ctx* ctx);

/* reference RK3::RK3 with State/Tendencies/BaseState allocation (src/dyn/rk3.cpp:48-56,
 * src/dyn/state.cpp:10-23): allocates every device field of the domain, zero-filled */
int mi_domain_create(mi_ctx* ctx, const mi_grid* g, const mi_params* p, const mi_levels* lv, mi_domain** out);
int mi_domain_destroy(mi_domain* d);
/* acoustic steps per large step in use (reference RK3::n_sound, rk3.hpp:44) */
int mi_domain_n_sound(const mi_domain* d);
/* device bytes the domain holds */
int64_t mi_domain_device_bytes(const mi_domain* d);

/* Copy a host field into the domain (reference Field::fill_host, src/core/field.hpp:113-119, followed by
 * the halo fill of src/core/halo.cpp).  n = number of values (see the field ids above). */
int mi_domain_set_field(mi_domain* d, int field, const double* host, int64_t n);
/* Copy a field's logical points to the host (stream-ordered: waits for the domain's queued work). */
int mi_domain_get_field(mi_domain* d, int field, double* host, int64_t n);

/* Zero-copy device access for device-side plugins (reference Field3::view(), src/core/field.hpp:16-44: a
 * pointer and int strides): the device address of point (0,0,0) of a field, the element size (8: double; 4:
 * float, the FP32-bulk storage of the 3-D bulk fields) and the strides in elements: i contiguous, then k
 * (stride_k), then j (stride_j; 0 for an x-z slice).  The 3 halo points on each horizontal side are
 * addressable; a 2-D field has nz = 1 and stride_k = 0.  Work on the view goes on the context's stream (or
 * synchronises with it).  A scalar's view (MI_Q + n) names the buffer of the moment: the two buffers of a
 * scalar swap at every stage, so take it again after any step or stage.  Base fields stored as profiles (flat
 * terrain) have no view (MI_EINVAL). */
typedef struct mi_field_view {
  void* data;
  int32_t elem_bytes;
  int32_t nx, ny, nz; /* logical points of the field's staggering */
  int64_t stride_k, stride_j;
} mi_field_view;
int mi_domain_field_view(mi_domain* d, int field, mi_field_view* out);

/* Caller-owned device buffers (the reference's State / Tendencies / BaseState own their Field3 buffers and RK3
 * only steps them, src/dyn/rk3.hpp:34-78, src/core/field.hpp:76-143).  mi_domain_field_layout gives the layout
 * a field's buffer must have: element size, logical points, strides (elements), the offset of point (0,0,0)
 * from the start of the allocation and the allocation's size in elements (halos included).
 * mi_field_bind makes a caller's device allocation of at least alloc_elems * elem_bytes bytes, laid out so,
 * the storage of the field: the domain reads and writes it from then on and never frees it (the caller keeps
 * it alive while the domain may step, and may read or write it between steps on the context's stream).  Its
 * contents become the field's values (the halo points are refreshed from the interior at the bind).  Binding
 * NULL gives the field storage of the domain's own again, with the bound buffer's values copied in.
 * Bindable: the carried state (MI_U .. MI_AL), the held tendencies, the 2-D base fields and map factors, and
 * the 3-D base fields of a domain with terrain.  Not the scalars (their two buffers swap at every stage) nor
 * the base profiles of a flat domain (MI_EINVAL). */
typedef struct mi_field_layout {
  int32_t elem_bytes;
  int32_t nx, ny, nz;
  int64_t stride_k, stride_j;
  int64_t offset;      /* elements from the start of the allocation to point (0,0,0) */
  int64_t alloc_elems; /* elements of the allocation, halos included */
} mi_field_layout;
int mi_domain_field_layout(const mi_domain* d, int field, mi_field_layout* out);
int mi_field_bind(mi_domain* d, int field, void* dev_alloc, int64_t bytes);
/* device memory for such buffers, on the context's device (hipMalloc / hipFree) */
int mi_device_alloc(mi_ctx* c, int64_t bytes, void** out);
int mi_device_free(mi_ctx* c, void* p);
/* host <-> device copies of n bytes on the context's stream, complete when they return */
int mi_device_upload(mi_ctx* c, void* dst, const void* src, int64_t n);
int mi_device_download(mi_ctx* c, void* dst, const void* src, int64_t n);

/* reference RK3::step (src/dyn/rk3.cpp:86-170): advance nsteps large steps.  Asynchronous on the
 * context's stream; mi_domain_sync waits. */
int mi_step(mi_domain* d, int nsteps);
int mi_domain_sync(mi_domain* d);

/* The phases of one step, for tests (same launches as mi_step, in its order):
 *   mi_step_begin      the boundary clock (rk3.cpp:93) and the _2 -> _1 copies (rk3.cpp:94-98)
 *   mi_rk_prep         LargeStep::prep (large_step.cpp:697-733)
 *   mi_rk_tendency     LargeStep::tendencies (large_step.cpp:798-861)
 *   mi_small_steps     SmallStep::run_stage (small_step.cpp:490-607)
 *   mi_scalar_transport ScalarTransport::run_stage (scalars.cpp:436-461)
 *   mi_calc_p_rho_phi  SmallStep::calc_p_rho_phi (small_step.cpp:609-617)
 *   mi_step_end        spec_bdy_final and the surface w of terrain (rk3.cpp:159-167)
 * rk = 1, 2, 3. */
int mi_step_begin(mi_domain* d);
int mi_rk_prep(mi_domain* d, int rk);
int mi_rk_tendency(mi_domain* d, int rk);
int mi_small_steps(mi_domain* d, int rk);
int mi_scalar_transport(mi_domain* d, int rk);
int mi_calc_p_rho_phi(mi_domain* d, int rk);
/* reference SmallStep::calc_p_rho_phi_interior (small_step.cpp:619-633): p and al again on the columns spec_zone
 * in from each non-periodic side (solve_em after spec_bdy_final and after the microphysics; rk3.cpp:155-161) */
int mi_calc_p_rho_phi_interior(mi_domain* d, int spec_zone);
int mi_step_end(mi_domain* d);

/* ---------------------------------------------------------------- specified lateral boundaries
 * Real-data boundaries, bc_x = bc_y = MI_BC_SPECIFIED (reference LateralBC, src/dyn/lateral_bc.hpp:28-91,
 * lateral_bc.cpp:67-457; WRF spec_bdy_width 5 = spec_zone 1 + relax_zone 4).  The outermost row of every
 * side (the spec zone) takes its tendency and, at the end of each step, its value from boundary data; the
 * next three rows are relaxed toward the data (fcx/gcx ramp of lbc_fcx_gcx, lateral_bc.cpp:198-211).
 *
 * Boundary data of one variable and side: values and tendencies for the current boundary interval, each
 * npos * nk * 5 doubles in (pos, k, b) order, pos fastest: pos = j on the x sides MI_BDY_XS/XE (ny values,
 * ny + 1 for V), pos = i on the y sides MI_BDY_YS/YE (nx, nx + 1 for U); k over the variable's levels (nz,
 * nz + 1 for PH, 1 for MU); b = 0 the outermost row.  The values are coupled as a WRF boundary file holds
 * them: u (c1h muu + c2h) / msfuy, v (c1h muv + c2h) / msfvx, t (c1h mut + c2h), ph (c1f mut + c2f),
 * q (c1h mut + c2h), and mu' itself.  At a step the boundary value is values + dtbc * tendencies, dtbc
 * the seconds since the start of the interval, which each step advances by dt before anything else;
 * set it back with mi_bdy_set_time when the next interval's data is loaded.  Variables without data hold
 * zeros; scalars without data take flow-dependent values in the spec zone (flow_dep_bdy). */
enum { MI_BDY_U = 0, MI_BDY_V = 1, MI_BDY_T = 2, MI_BDY_PH = 3, MI_BDY_MU = 4, MI_BDY_Q = 16 };
enum { MI_BDY_XS = 0, MI_BDY_XE = 1, MI_BDY_YS = 2, MI_BDY_YE = 3 };
#define MI_SPEC_BDY_WIDTH 5
/* reference LateralBC::set / set_scalar (lateral_bc.cpp:240-275); var MI_BDY_Q + n for scalar n */
int mi_bdy_set(mi_domain* d, int var, int side, const double* values, const double* tends, int64_t n);
/* reference LateralBC::dtbc (lateral_bc.hpp:58-59) */
int mi_bdy_set_time(mi_domain* d, double dtbc);
int mi_bdy_get_time(mi_domain* d, double* dtbc);

/* The domain maxima of w_damp's vertical and horizontal Courant numbers from the current fields
 * (reference LargeStep::cfl_max, src/dyn/large_step.cpp:862-876: |ww / (c1f mut + c2f) rdnw dt| and
 * max(|u rdx msfux dt|, |v rdy msfvy dt|) over levels 1..nz-1; a NaN never wins).  Called between
 * mi_rk_tendency and mi_small_steps it gives the reference's on_cfl values of that stage. */
int mi_cfl_max(mi_domain* d, double* vert, double* horiz);
/* The same maxima without allocating anything: no counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The two running
 * maxima live in a file-scope device array of the library (2 x 8 bytes, one copy per GPU, held until the process
 * ends), cleared by a kernel at every call; the results equal mi_cfl_max's.  Calls that overlap on one GPU share
 * that array and race. */
int mi_cfl_max_noalloc(mi_domain* d, double* vert, double* horiz);

/* Adaptive time step.  mi_domain_set_dt is RK3::set_dt (rk3.cpp:79-84) together with what an adaptive-step
 * driver changes beside it (WRF adapt_timestep): the specified boundaries' relaxation weights
 * (LateralBC::set_dt, lateral_bc.cpp:198-211) and the sixth-order diffusion's dt.  The next steps use the
 * large step dt (> 0) and n_sound acoustic steps (even, >= 2); the boundary clock is kept. */
int mi_domain_set_dt(mi_domain* d, double dt, int n_sound);
/* RK3::on_cfl (rk3.hpp:46-47, rk3.cpp:115-119): called after each stage's rk_tendency with LargeStep::cfl_max's
 * maxima (vertical, horizontal).  RK3::on_first_rk_prep (rk3.hpp:62-64, rk3.cpp:105): called in stage 1 right
 * after rk_step_prep.  NULL removes a callback.  While either is set, mi_step runs its steps launch by launch (not
 * as a replayed graph), synchronising the stream before each call; a batch refuses such a member. */
/* Checkpoint and resume.  No counterpart in src/dyn: the reference restarts from NetCDF restart files
 * (tests/robustness/run_test.sh:40-51); here the state a step carries to the next (u, v, w, theta, ph, mu, p, al
 * and every scalar, halos included, and the boundary clock) goes to one opaque host buffer of
 * mi_domain_state_bytes bytes.  A domain created with the same grid, parameters, base state and boundary data and
 * loaded from the buffer continues bit for bit; a buffer of another shape, precision or scalar count is refused
 * (MI_EINVAL).  mi_domain_state_bytes returns a negative status on error. */
int64_t mi_domain_state_bytes(mi_domain* d);
int mi_domain_save_state(mi_domain* d, void* host, int64_t bytes);
int mi_domain_load_state(mi_domain* d, const void* host, int64_t bytes);
typedef void (*mi_cfl_fn)(void* user, double vert, double horiz);
typedef void (*mi_stage_fn)(void* user);
int mi_domain_set_cfl_callback(mi_domain* d, mi_cfl_fn fn, void* user);
int mi_domain_set_first_rk_prep_callback(mi_domain* d, mi_stage_fn fn, void* user);

/* Column physics plugin (reference ColumnPhysics, src/dyn/column_physics.hpp:21-34, attached by
 * RK3::set_column_physics): host functions called at WRF's points of the step (rk3.cpp:100-166,
 * large_step.cpp:820-823), each with the stream synchronised, so that they may read and write fields with
 * mi_domain_get_field / mi_domain_set_field:
 *   part1              stage 1, after rk_step_prep and the first-stage hook, once the held tendencies
 *                      (MI_*_TENDF and every MI_Q_TENDF) are zeroed: first_rk_step_part1
 *   part2              stage 1, first thing of the tendencies: sets the physics tendencies into MI_RU_TENDF,
 *                      MI_RV_TENDF, MI_T_TENDF, MI_Q_TENDF + n (update_phy_ten); the diffusion and the
 *                      boundary relaxation are added after it
 *   after_rk           after the third stage, before the specified boundaries' final values (phy_prep_part2)
 *   after_all_rk_steps the end of the step (after_all_rk_steps)
 * th_forcing / qv_forcing = 1 keep MI_TH_FORCING / MI_QV_FORCING (the latter needs n_moist > 0).  Any pointer
 * may be NULL.  While a plugin is attached the held tendencies are used (as with a diffusion or specified
 * boundaries), the caller's MI_Q_TENDF holds only for the step whose part2 sets it, and steps run launch by
 * launch as with the callbacks above (mi_batch_step refuses the domain).  The per-phase entry points call
 * part1 from mi_rk_prep(d, 1), part2 from mi_rk_tendency(d, 1) and the last two from mi_step_end.  cp == NULL
 * detaches the plugin. */
typedef void (*mi_phys_fn)(void* user, mi_domain* d);
typedef struct mi_column_physics {
  void* user;
  mi_phys_fn part1, part2, after_rk, after_all_rk_steps;
  int32_t th_forcing, qv_forcing;
} mi_column_physics;
int mi_domain_set_column_physics(mi_domain* d, const mi_column_physics* cp);

/* Attach, change or detach the sixth-order diffusion between steps (reference RK3::set_diffusion with a
 * DiffusionDriver whose sixth-order option is on, rk3.hpp:53-62, diffusion_driver.cpp:266-327): opt 0 off, 1
 * linear, 2 monotonic; factor the non-dimensional diff_6th_factor.  The same as creating the domain with these
 * mi_params values, from the next step on. */
int mi_domain_set_diffusion(mi_domain* d, int diff_6th_opt, double diff_6th_factor);

/* Batches (ensembles, serving): the steps of several domains of one context advance together.  Each
 * step of the batch is one HIP graph holding every member's step as an independent branch, so the
 * members' kernels overlap on the device (reference: the tiles of main.cpp:30-42 stepped one thread per
 * patch; here one graph launch for all).  Members keep their own state and may differ in size and
 * options; a member stepped on its own in between is picked up (the batch re-captures). */
typedef struct mi_batch mi_batch;
int mi_batch_create(mi_ctx* ctx, mi_domain* const* members, int n, mi_batch** out);
int mi_batch_destroy(mi_batch* b);
int mi_batch_step(mi_batch* b, int nsteps);
/* How a batch step is laid out in its graph.  MI_BATCH_BRANCHES (the default): each member's step is a
 * branch of its own.  MI_BATCH_FUSED: advance_uv and the column kernel of every acoustic substep run as one
 * launch over all members (a table of the members' arguments and block ranges), the rest of each member's
 * step in branches between them; every member needs the same arithmetic and acoustic step count, no
 * callbacks, and at most 64 members (MI_EINVAL otherwise).  Both give each member the bits of its step alone;
 * which is faster depends on the mix (measured by bench.py).  No counterpart in the reference. */
enum { MI_BATCH_BRANCHES = 0, MI_BATCH_FUSED = 1 };
int mi_batch_set_schedule(mi_batch* b, int schedule);
/* The graphs a batch holds: at most one per buffer parity of the members' current launch schedules (a
 * schedule change such as mi_domain_set_dt on a member frees the graphs of the old one). */
int mi_batch_graph_count(const mi_batch* b);

/* MI_ENONFINITE if any point of u, v, w, t, ph, mu is NaN or infinite (the reference's runner check,
 * src/cases/domain_runner.hpp:70-71); *bad_field receives the first such field id (or -1). */
int mi_check_finite(mi_domain* d, int* bad_field);

/* Autotuner (no counterpart in the reference): times each block shape (columns per block x threads per
 * column) of the column kernels on this domain for `trials` steps and keeps the fastest.  The carried
 * state is saved first and restored after, so the forecast's trajectory is unchanged, and every shape
 * gives bit-identical results.  mi_domain_get_tuning returns the chosen shape index of each column kernel
 * (prep, acoustic prologue, acoustic substep, finish) into shapes[0..n). */
int mi_domain_autotune(mi_domain* d, int trials);
int mi_domain_get_tuning(const mi_domain* d, int* shapes, int n);
/* sets the shapes explicitly (index 0..12 of the library's table of columns x threads per column: 16x16,
 * 16x32, 32x16, 8x32, 8x64, 4x128, 2x256, 64x4, 64x8, 64x16, 32x8, 16x64, 4x64; the substep kernel needs at
 * least nz+1 threads per column, or index 13: one thread per column marching k; MI_EINVAL if a shape does not
 * fit nz) */
int mi_domain_set_tuning(mi_domain* d, const int* shapes, int n);

/* Per-kernel timing, for measurement (no counterpart in the reference; its CPPWRF_KTIME timer,
 * src/core/backend.hpp:419-425, synchronises around every launch instead).  While a kernel name is
 * selected, mi_step launches without a graph and brackets every launch of that kernel with a pair of
 * HIP events on the domain's stream.  mi_profile_read waits for the stream, returns the summed event
 * time, the number of timed launches and their algorithmic bytes (distinct 3-D arrays read and written
 * x 8 bytes x points), and clears the counters.  name NULL or "" stops the timing. */
int mi_profile_kernel(mi_domain* d, const char* name);
int mi_profile_read(mi_domain* d, double* total_ms, int* launches, double* alg_bytes);

/* ---------------------------------------------------------------- synthetic initial states
 * Host-only builders of the idealized states the tests and benchmarks use: the reference's case
 * recipes (src/cases/cases.cpp:17-287: sounding, eta levels, hydrostatic base state) and the 3-D
 * warm bubble of this library's workloads.  No GPU is touched. */
enum {
  MI_CASE_REST_STATE = 0,  /* reference configs/rest_state.ini recipe */
  MI_CASE_WARM_BUBBLE = 1, /* reference warm_bubble (x-z Bryan & Fritsch 2002) */
  MI_CASE_IGW = 2,         /* reference igw (Skamarock & Klemp 1994, f = 0) */
  MI_CASE_BUBBLE3D = 3,    /* 3-D dry bubble in an N = 0.01 atmosphere at rest, seeded noise */
  MI_CASE_HILLS = 4,       /* real-data-shaped: seeded hills, US standard atmosphere, jets, map factors,
                              Coriolis, hybrid coordinate, optional qv; needs terrain, mapped, coriolis
                              (and n_moist >= 1 when moist) in mi_params */
  MI_CASE_STRAKA = 5,      /* reference straka (x-z density current, Straka et al. 1993) */
  MI_CASE_SQUALL2D = 6     /* reference squall2d (x-z moist bubble in the Weisman & Klemp 1982 sounding) */
};
typedef struct mi_case_spec {
  int32_t kind; /* MI_CASE_* */
  mi_grid grid;
  double ztop;          /* m */
  uint64_t seed;        /* noise seed (bubble3d) */
  double noise_amp;     /* K, half-width of the uniform theta noise (bubble3d) */
  double bubble_dtheta; /* K (bubble3d: +2) */
  double bubble_radius; /* m (bubble3d: 4000) */
  double bubble_zc;     /* m (bubble3d: 2000) */
  double u0;            /* uniform wind (rest_state's case.u0) */
  /* hills: a real-data-shaped domain (terrain, US standard atmosphere, jets, map factors, Coriolis) */
  int32_t n_hills;      /* seeded Gaussian hills (periodic images) */
  int32_t moist;        /* 1: qv from RH 80 % at the ground to 10 % at 10 km (scalar 0) */
  double hill_height;   /* m, heights uniform in [0, hill_height] */
  double hill_wmin, hill_wmax; /* m, e-folding widths uniform in [hill_wmin, hill_wmax] */
  double jet_min, jet_max;     /* m/s, u and v jet maxima uniform in [jet_min, jet_max], plus +-1 m/s noise */
  double map_amp;       /* map factors 1 +- map_amp (0.05) */
  double etac;          /* hybrid coordinate: B(eta) = 0 above eta = etac (WRF hybrid_opt 2); 0: sigma */
  double p_top;         /* Pa (hills; the model top is where the sounding reaches p_top) */
  /* idealized kinds: a moist full state with qv(z) = qv0 exp(-z / qv_hscale) (reference case.qv0 and
   * case.qv_hscale, cases.cpp:184-188; 0: dry), in hydrostatic balance with its vapour load (scalar 0) */
  double qv0;
  double qv_hscale;     /* m (0: 2500) */
} mi_case_spec;
typedef struct mi_case mi_case;
int mi_case_create(const mi_case_spec* spec, mi_case** out);
int mi_case_destroy(mi_case* c);
/* the level coefficients; the pointers stay valid until mi_case_destroy */
int mi_case_levels(const mi_case* c, mi_levels* out);
/* the initial value of a state or base field (same ids and sizes as mi_domain_get_field; base
 * 3-D fields as full nxs*nys*nzs arrays) */
int mi_case_get_field(const mi_case* c, int field, double* host, int64_t n);
/* load every initial field of a case into a domain of the same grid */
int mi_case_load(const mi_case* c, mi_domain* d);

#ifdef __cplusplus
}
#endif
#endif /* CPPWRF_AMD_H */
