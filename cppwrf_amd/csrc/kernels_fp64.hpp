// FP64 dynamics kernels: declarations and the argument block shared by all of them.
#pragma once

#include "dev_common.hpp"

namespace mi_fp64 {

using mi_dev::BField;
using mi_dev::Geo;
using mi_dev::Lv;

// Every device array of a domain (pointers to point (0,0,0) of each field) plus the run parameters.
// Passed by value to every kernel: kernels read the members they use from the kernel-argument segment.
struct Fields {
  // state (reference State, src/dyn/state.hpp:30-78)
  double *u1, *u2, *v1, *v2, *w1, *w2, *t1, *t2, *ph1, *ph2, *mu1, *mu2;
  double *p, *al, *alt, *php, *ww;
  double *mut, *muu, *muv, *muts, *muus, *muvs;
  double *ru, *rv, *rw, *ru_m, *rv_m, *ww_m;
  // tendencies (src/dyn/state.hpp:80-90)
  double *ru_tend, *rv_tend, *rw_tend, *t_tend, *ph_tend, *mu_tend;
  // acoustic-step work fields (src/dyn/small_step.hpp:43-49)
  double *u_save, *v_save, *w_save, *t_save, *ph_save, *ww1, *c2a, *pm1, *t_2save, *ca, *calpha, *cgamma;
  double *mu_save, *mudf, *muave, *w_rhs, *t2a, *dvdxi;
  // base state: 2-D fields stored in full with halos, 3-D through BField
  double *mub, *msftx, *msfty, *msfux, *msfuy, *msfvx, *msfvy, *msfvx_inv;
  BField pb, alb, phb;
};

struct Params {
  double hup, vup;   // 1 for 5th / 3rd-order upwind, 0 for 6th / 4th-order centred
  double epssm, smdiv, emdx, emdy;
  double wd_dt, w_crit;  // w_damp: the large step (0: off) and the critical Courant number
  double dampcoef, zdamp;
  int top_lid, damp_opt;
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
  double* p[kHaloMax];
  int nxs[kHaloMax], nys[kHaloMax], nzs[kHaloMax];
  unsigned char face_x[kHaloMax], face_y[kHaloMax], px[kHaloMax], py[kHaloMax];
};

void launch_halo(const Args& a, const HaloSet& s, hipStream_t st);
void launch_copy_step(const Args& a, int64_t n_alloc3, int64_t n_alloc2, hipStream_t st);
void launch_fill(double* p, int64_t n, double v, hipStream_t st);
void launch_rk_prep_mu(const Args& a, hipStream_t st);
void launch_rk_prep_couple(const Args& a, hipStream_t st);
void launch_rk_prep_ww(const Args& a, hipStream_t st);
void launch_tendencies(const Args& a, hipStream_t st);
void launch_ss_prep(const Args& a, bool first, hipStream_t st);
void launch_p_rho(const Args& a, bool damp, hipStream_t st);
void launch_coef_w(const Args& a, double cof, hipStream_t st);
void launch_uv(const Args& a, double dts, hipStream_t st);
void launch_mu_t(const Args& a, double dts, int it, int ns, hipStream_t st);
void launch_w(const Args& a, double dts, hipStream_t st);
void launch_finish(const Args& a, hipStream_t st);
void launch_p_rho_phi(const Args& a, hipStream_t st);
void launch_finite(const Args& a, int* flags, hipStream_t st);

}  // namespace mi_fp64
