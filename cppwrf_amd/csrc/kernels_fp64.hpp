// The dynamics kernels: declarations and the argument block shared by all of them.
//
// The source is compiled once per arithmetic: MI_REAL = double in namespace mi_fp64 (plain FP64, bit-
// identical to the reference CPU FP64 build) and MI_REAL = float in namespace mi_fp32 (FP32-bulk: the
// 3-D prognostic, tendency and work fields stored in FP32).  Whatever the storage, the base state, the
// 2-D column masses, php, the tridiagonal coefficients and every LDS column buffer (the mass-divergence
// sums, the omega integral, the Thomas solve of advance_w, p'' and al'' of calc_p_rho) stay FP64.
#pragma once

#include <vector>

#include "dev_common.hpp"

#ifndef MI_REAL
#define MI_REAL double
#define MI_NS mi_fp64
#endif

namespace MI_NS {

using real = MI_REAL;  // storage of the 3-D bulk fields
// arithmetic of the bulk kernels (advection fluxes, the positive-definite limiter, the scalar updates): double
// in FP64, float in FP32-bulk; the column solves, column sums and p/al keep double in both
using acc = MI_REAL;

using mi_dev::BField;
using mi_dev::Geo;
using mi_dev::Lv;

// Every device array of a domain (pointers to point (0,0,0) of each field) plus the run parameters.
// Passed by value to every kernel: kernels read the members they use from the kernel-argument segment.
struct Fields {
  // state (reference State, src/dyn/state.hpp:30-78)
  real *u1, *u2, *v1, *v2, *w1, *w2, *t1, *t2, *ph1, *ph2;
  double *mu1, *mu2;
  real *p, *al, *alt, *ww;
  double* php;  // the full geopotential: its horizontal differences need FP64
  double *mut, *muu, *muv, *muts, *muus, *muvs;
  real *ru, *rv, *rw, *ru_m, *rv_m, *ww_m;
  // tendencies (src/dyn/state.hpp:80-90)
  real *ru_tend, *rv_tend, *rw_tend, *t_tend, *ph_tend;
  double* mu_tend;
  // acoustic-step work fields (src/dyn/small_step.hpp:43-49)
  real *u_save, *v_save, *w_save, *t_save, *ph_save, *ww1, *c2a, *pm1;
  double *ca, *calpha, *cgamma;  // the tridiagonal coefficients of advance_w
  double *mu_save, *mudf, *muave;
  // base state: 2-D fields stored in full with halos, 3-D through BField
  double *mub, *msftx, *msfty, *msfux, *msfuy, *msfvx, *msfvy, *msfvx_inv;
  BField pb, alb, phb;
  // moisture (src/dyn/state.hpp:72-74): calc_cq coefficients and the microphysics heating
  real *cqu, *cqv, *cqw, *hdia;
  // Coriolis and grid rotation (state.hpp:105-106); held tendencies of the first stage (state.hpp:85-88)
  double *fcor, *ecor, *sina, *cosa, *ht;
  real *ru_tendf, *rv_tendf, *rw_tendf, *t_tendf;
  // specified boundaries: the held tendency of ph (reference Tendencies::ph_tendf) and the seconds since
  // the start of the boundary interval (LateralBC::dtbc), kept on the device so a replayed step advances it
  real* ph_tendf;
  double* bdy_time;
  // the positive-definite limiter's face widths and their ratios to dt (2-D; pd_dz: dz, then dz / dt)
  double *pd_dx, *pd_dxdt, *pd_dy, *pd_dydt, *pd_dz;
  // the advective forcings a column physics plugin asks for (WRF RTHFTEN, RQVFTEN); null when not wanted
  real *thf, *qvf;
  // the substep k-march's scratch: the forward sweep's values and the explicit phi'' of the substep (FP64)
  double *ky, *kr;
  // the diffusion's base state: theta's t_init (3-D), the 1-D profiles u_base, v_base, qv_base (nz values)
  BField t_init;
  double *u_base, *v_base, *qv_base;
  // km_opt 4: the horizontal eddy viscosities of the step at mass points (xkmh of the momenta, xkhh of theta and
  // the scalars) and the deformation defor12 at the vorticity points ((nx+1) x (ny+1), u-staggered storage whose
  // row ny lies in the y halo)
  real *xkmh, *xkhh, *d12;
  // what the acoustic substeps of a stage read unchanged, set by the stage's kernels: damp_opt 3's weight at the w
  // levels; the coupled saves (c1h mu + c2h) u_save / msfuy and (c1h mu + c2h) v_save / msfvx; advance_w's
  // c2a rdnw / (c1h mut + c2h) at the mass levels; 1 / (1 + cqw) with moisture
  real *dampwt, *ucs, *vcs, *cwc, *cq1;
  // and advance_uv's map-factor ratios (2-D): msfux / msfuy, 1 / msfuy at u faces, msfvy / msfvx at v faces
  double *rmu, *msfuy_inv, *rmv;
  double* msfty_inv;  // 1 / msfty (mass points), for the acoustic substeps
  // damp_opt 2: the base profiles of theta and of the mass levels' height (nz values)
  double *t_base, *z_base;
  // the microphysics after the RK loop: moist_physics_prep's fields (dry theta, density, Exner function, height
  // and thickness of the layer, the qv and qc before the scheme) and the grid-scale precipitation (2-D)
  real *th_phy, *mrho, *pii, *zmid, *dz8w, *qvd, *qcd;
  double *rainnc, *rainncv;
};

// The advected scalars of a domain (State::q_1 / q_2, Tendencies::q_tendf) and the work fields of their
// transport.  q2[n] / qn[n] are the two buffers of species n; the host swaps them after every stage.
struct Scalars {
  real* q1[MI_MAX_SCALARS];
  real* q2[MI_MAX_SCALARS];
  real* qtf[MI_MAX_SCALARS];
  real* qn[MI_MAX_SCALARS];
  real* scale[MI_MAX_SCALARS];  // advect_scalar_pd's outflow scale of each species
  real* qtw[MI_MAX_SCALARS];    // the step's tendency: the caller's plus diffusion and boundary relaxation
  int n_moist;
};

// Specified lateral boundaries: the boundary data of one variable, one device buffer of eight slabs (values
// of the sides xs, xe, ys, ye, then their tendencies), each (pos, k, b) with pos fastest (reference
// LateralBC, src/dyn/lateral_bc.hpp:17-19, 77-84)
enum BdyKind : int { kBdyU = 0, kBdyV = 1, kBdyT = 2, kBdyPH = 3, kBdyMU = 4, kBdyQ = 5 };
struct BdyVar {
  int kind;  // BdyKind
  int q;     // the scalar (kBdyQ)
  const double* base;
};
constexpr int kBdyMax = 5 + MI_MAX_SCALARS;
struct BdySet {
  int n;
  BdyVar v[kBdyMax];
};
// what a ring launch does: the spec zone's tendency from the data (spec_bdytend), the relaxation of rows
// 1..3 (relax_bdytend), the spec zone's values at the end of a step (spec_bdy_final)
enum BdyMode : int { kBdySpec = 0, kBdyRelax = 1, kBdyFinal = 2 };

struct Params {
  double hup, vup;   // 1 for 5th / 3rd-order upwind, 0 for 6th / 4th-order centred
  double epssm, smdiv, emdx, emdy;
  double wd_dt, w_crit;  // w_damp: the large step (0: off) and the critical Courant number
  double dampcoef, zdamp;
  int top_lid, damp_opt;
  int moist;  // moist species present: calc_cq coefficients and the heating terms are used
  int mapped, cor;     // MapC::mapped (curvature with vxgm), MapC::cor (Coriolis terms)
  int use_tendf;       // held tendencies added (a diffusion driver is attached)
  int qtf_on;          // a column physics plugin sets the scalars' q_tendf (else zero_tendf leaves it 0)
  int diff_6th_opt;    // 1 linear, 2 monotonic
  double diff_6th_factor;
  int diff_opt;        // 1: second-order diffusion along eta
  int km4;             // km_opt 4: the horizontal K from the 2-D Smagorinsky closure
  int karr;            // the horizontal K read from xkmh / xkhh (km_opt 4 or damp_opt 1), else the constants below
  double xkmh, xkhh, kvdif, kvdq;  // its eddy viscosities: khdif, khdif / prandtl, kvdif, kvdif / prandtl
  double c_s, prandtl, dx, dy;     // km_opt 4's constant, 1 / pr_inv and the grid spacing
  int mp;              // microphysics after the RK loop: 1 Kessler
  double mp_lim;       // the heating-rate limit, K/s
  double dt;           // the large step
  int terrain;         // the surface w follows the terrain (ht)
  int spec;            // specified lateral boundaries (both directions)
  double fcx[5], gcx[5];  // their relaxation weights (lbc_fcx_gcx)
};

struct Args {
  Geo g;
  Lv c;
  Fields f;
  Params q;
};

// halo fill of up to kHaloMax fields per launch; parity 0 even, 1 odd, 2 odd with the wall face flipped
constexpr int kHaloMax = 16;
struct HaloSet {
  int n;
  int nkmax;
  void* p[kHaloMax];
  unsigned char f32[kHaloMax];  // element type: 1 float, 0 double
  int nxs[kHaloMax], nys[kHaloMax], nzs[kHaloMax];
  unsigned char face_x[kHaloMax], face_y[kHaloMax], px[kHaloMax], py[kHaloMax];
};

// per-kernel event timing: while g_timer points to a timer, every launch of the kernel named
// g_timer->name is bracketed by a pair of events on its stream
struct KernelTimer {
  const char* name;
  std::vector<hipEvent_t> ev;  // start, end, start, end, ...
  size_t used = 0;
  double bytes = 0;            // algorithmic bytes of the timed launches
};
extern thread_local KernelTimer* g_timer;

void launch_halo(const Args& a, const HaloSet& s, hipStream_t st);
void launch_copy_step(const Args& a, int64_t n_alloc3, int64_t n_alloc2, hipStream_t st);
void launch_fill(double* p, int64_t n, double v, hipStream_t st);
// the same over a 2-D field's points (level 0 of every row) of an allocation of n_alloc values
void launch_fill2d(const Args& a, double* p, int64_t n_alloc, double v, hipStream_t st);
// column kernels and their block shapes (columns x threads per column)
enum ColKernel { kColLargePrep = 0, kColSsColumn = 1, kColAcoustic = 2, kColFinish = 3, kNumColKernels = 4 };
constexpr int kNumColCfgs = 13;
// 64 or 32 columns per block: a wavefront reads 64 (32) consecutive points of one level (two levels); 4x64: four
// waves per block, so a barrier of the substep kernel holds fewer waves
constexpr int kColCfgs[kNumColCfgs][2] = {{16, 16}, {16, 32}, {32, 16}, {8, 32}, {8, 64}, {4, 128}, {2, 256},
                                          {64, 4}, {64, 8}, {64, 16}, {32, 8}, {16, 64}, {4, 64}};
constexpr int kColCfgTallest = 6;  // 2 x 256: the most levels per column (the substep kernel's nz limit)
// one more shape of the substep kernel: one thread per column marching k (k_acoustic_march)
constexpr int kColMarch = kNumColCfgs;
constexpr int kNumColShapes = kColMarch + 1;
constexpr const char* kColKernelNames[kNumColKernels] = {"large_prep", "ss_column", "acoustic", "finish"};
bool col_cfg_fits(int kernel, int cfg, int nz);
int col_cfg_default(int kernel, int nx, int ny, int nz);

void launch_large_prep(const Args& a, int cfg, hipStream_t st);
void launch_tendencies(const Args& a, hipStream_t st);
void launch_ss_prep(const Args& a, bool first, hipStream_t st);
void launch_ss_column(const Args& a, int cfg, double cof, double dts, hipStream_t st);
void launch_uv(const Args& a, double dts, hipStream_t st);
void launch_acoustic(const Args& a, int cfg, double dts, int it, int ns, hipStream_t st);
void launch_open_sums(const Args& a, int it, int ns, hipStream_t st);
// A batch's acoustic substeps in one launch per kernel over every member (the members' Args in a device array):
// one table entry per member (its blocks' first index in the launch, its block grid, its stage dts) and one more
// holding the launch's block count.  batch_table_* fill the tables on the host and return the block count.
struct BatchMember {
  int i0, i1, j0, j1, k0, k1;
  int start, gx, gy;
  double dts;
};
constexpr int kMaxBatch = 64;
int batch_table_uv(const Args* const* as, const double* dts, int n, BatchMember* out);
int batch_table_acoustic(const Args* const* as, const double* dts, int n, BatchMember* out);
bool batch_fits(const Args& a);  // the batched column kernel's levels per block hold the member's
void launch_uv_batch(const Args* dev_args, const BatchMember* dev_tab, int n, int blocks, hipStream_t st);
void launch_acoustic_batch(const Args* dev_args, const BatchMember* dev_tab, int n, int blocks, int nzmax, int it, int ns,
                           hipStream_t st);
void launch_finish(const Args& a, int cfg, bool prhophi, double hd_dt, hipStream_t st);
// the held tendencies of u, v, w, theta at the first stage: kSixthZero zeroes them and adds the sixth-order
// diffusion; kSixthKeep adds it to what they hold (a physics plugin's part2); kSixthZeroOnly only zeroes
enum SixthMode : int { kSixthZero = 0, kSixthKeep = 1, kSixthZeroOnly = 2 };
void launch_sixth_dyn(const Args& a, int mode, hipStream_t st);
// specified boundaries (reference LateralBC): ring launches over the variables of bs
void launch_bdy(const Args& a, const Scalars& s, const BdySet& bs, BdyMode mode, hipStream_t st);
void launch_bdy_zero_grad(const Args& a, hipStream_t st);
void launch_acoustic_ring(const Args& a, double dts, int it, int ns, hipStream_t st);
void launch_bdy_flow_dep(const Args& a, const Scalars& s, const int* sp, int n, hipStream_t st);
void launch_bdy_time(const Args& a, hipStream_t st);
// the same for the scalars, into their step tendencies (first stage, after the acoustic steps)
void launch_sixth_scalars(const Args& a, const Scalars& s, int n, double dt_rk, hipStream_t st);
// set_w_surface (fill = false) after a step, with terrain
void launch_set_w_surface(const Args& a, hipStream_t st);
// w_damp's Courant-number maxima into out[0] (vertical), out[1] (horizontal) as non-negative bit patterns
void launch_cfl(const Args& a, unsigned long long* out, hipStream_t st);
// the same into the file-scope scratch array of this file (no allocation): cfl_scratch gives the array's
// address on a device (nullptr if hipGetSymbolAddress fails), launch_cfl_scratch zeroes its slots with a kernel
// and runs the maxima into them
constexpr int kCflScratchSlots = 2;
unsigned long long* cfl_scratch(int device);
void launch_cfl_scratch(const Args& a, unsigned long long* scratch_ptr, hipStream_t st);
// calc_cq (KCq) of the stage state
void launch_cq(const Args& a, const Scalars& s, hipStream_t st);
// the transport of species sp[0..n): rk_update_scalar (pd = 0) or rk_update_scalar_pd (pd = 1)
void launch_scalars(const Args& a, const Scalars& s, const int* sp, int n, int step, double dt_rk, hipStream_t st);
void launch_pd_geom(const Args& a, double dt, hipStream_t st);
void launch_copy_scalars(const Args& a, const Scalars& s, int n, int64_t n3, hipStream_t st);
void launch_p_rho_phi(const Args& a, hipStream_t st);
// the microphysics after the RK loop (moist_physics_prep, Kessler, moist_physics_finish); the halos and the final
// calc_p_rho_phi are the caller's
constexpr int kKesslerMaxLevels = 128;
void launch_smag2d(const Args& a, hipStream_t st);
void launch_dampkm(const Args& a, hipStream_t st);
void launch_rayleigh(const Args& a, hipStream_t st);
void launch_microphysics(const Args& a, const Scalars& s, hipStream_t st);
// the same on the columns i0 <= i < i1, j0 <= j < j1
void launch_p_rho_phi_range(const Args& a, int i0, int i1, int j0, int j1, hipStream_t st);
void launch_finite(const Args& a, int* flags, hipStream_t st);

}  // namespace MI_NS
