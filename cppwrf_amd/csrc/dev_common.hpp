// Device-side building blocks of the dynamics kernels: field views, the domain geometry, the level
// coefficients and the advection flux operators.
//
// Layout in HBM.  Every 3-D field of a domain has the same strides: a row of P doubles (P a multiple
// of 32, i.e. 256 bytes, with point i = 0 at element 8 so a wavefront's 64 consecutive i start on a
// 64-byte boundary), ny+1+2*hy rows per level and nz+1 levels.  U, V, W and mass fields differ only in
// how many of those points they use (staggering as the reference, src/core/grid.hpp:3-7), so one
// index i + j*sj + k*sk serves every field and kernels need a single set of strides.  An x-z slice
// (ny == 1) has sj == 0: j +- 1 reads the same row, as in the reference.  2-D fields use one level of
// the same plane.  Horizontally uniform base-state fields are 1-D profiles read through BField
// (strides 0 in i and j).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cppwrf_amd.h"
#include <stdint.h>

namespace mi_dev {

constexpr double kG = 9.81;
constexpr double kRd = 287.0;
constexpr double kCp = 7.0 * kRd / 2.0;
constexpr double kCpovcv = kCp / (kCp - kRd);
constexpr double kP1000 = 100000.0;
constexpr double kT0 = 300.0;
constexpr double kRcp = kRd / kCp;
constexpr double kRvovrd = 461.6 / kRd;  // r_v / r_d
constexpr double H = 0.5, Q = 0.25;
constexpr double kReradius = 1. / 6370.0e03;
constexpr int kHalo = 3;
constexpr int kLpad = 8;

// Element offsets are 32-bit (every allocation holds fewer than 2^31 doubles; the domain checks it):
// half the address registers of 64-bit index arithmetic, which matters for the occupancy of the
// column kernels.
using idx_t = int;

struct Geo {
  int nx, ny, nz;
  int nyv;       // V points in y: ny + 1, or 1 for an x-z slice
  idx_t sj, sk;  // strides of j and k (sj == 0 for an x-z slice)
  int hy;        // halo rows in y (0 for an x-z slice)
  int bcx, bcy;  // MI_BC_*
  bool two_d;
  __host__ __device__ idx_t at(int i, int j, int k) const { return i + j * sj + k * sk; }
  __host__ __device__ idx_t at2(int i, int j) const { return i + j * sj; }
  // the lateral boundary kinds (reference grid.hpp:45-53, kernels_common.hpp:33-44): specified (both
  // directions), open (WRF open_xs/xe, open_ys/ye) and flat = specified or open (the outermost row copied
  // outward, the advection orders degraded next to the edge)
  __host__ __device__ bool spec() const { return bcx == MI_BC_SPECIFIED; }
  __host__ __device__ bool ox() const { return bcx == MI_BC_OPEN; }
  __host__ __device__ bool oy() const { return bcy == MI_BC_OPEN && !two_d; }
  __host__ __device__ bool flx() const { return bcx == MI_BC_SPECIFIED || bcx == MI_BC_OPEN; }
  __host__ __device__ bool fly() const { return (bcy == MI_BC_SPECIFIED || bcy == MI_BC_OPEN) && !two_d; }
};
__host__ __device__ inline bool flat_bc(int bc) { return bc == MI_BC_SPECIFIED || bc == MI_BC_OPEN; }
constexpr double kCb = 25.;  // phase speed of the open (radiative) boundaries (reference constants.hpp:19)

// base-state field: full (si = 1, sj, sk of the domain) or a horizontally uniform profile (si = sj = 0)
struct BField {
  const double* p;
  idx_t si, sj, sk;
  __device__ double operator()(int i, int j, int k) const { return p[i * si + j * sj + k * sk]; }
};

// vertical coefficients (device copies of the level arrays)
struct Lv {
  const double *dnw, *rdnw, *dn, *rdn, *fnm, *fnp, *c1h, *c2h, *c3h, *c4h, *c1f, *c2f, *c3f, *c4f;
  double cf1, cf2, cf3, cfn, cfn1, rdx, rdy, p_top;
};

// The flux operators are templates on the arithmetic type T: double in the plain FP64 mode (the reference's
// expressions, bit for bit) and float in the bulk of the FP32-bulk mode.  by60 / by12 are the operators'
// divisions by 60 and 12 and qdiv a division of two values: IEEE divisions in FP64; in FP32 a multiply by
// the rounded reciprocal of the constant and by the hardware reciprocal (1 ulp) of a value, which keeps the
// division sequence out of the FP32-bulk kernels.
__device__ __forceinline__ double by60(double x) { return x / 60.0; }
__device__ __forceinline__ float by60(float x) { return x * (1.0f / 60.0f); }
__device__ __forceinline__ double by12(double x) { return x / 12.0; }
__device__ __forceinline__ float by12(float x) { return x * (1.0f / 12.0f); }
__device__ __forceinline__ double qdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <class T>
__device__ __forceinline__ T sgn1(T a) { return a >= T(0) ? T(1) : T(-1); }

// 5th-order upwind flux of the reference (src/dyn/kernels_common.hpp:13-17); up = 0 gives the 6th-order
// centred flux
template <class T>
__device__ __forceinline__ T flux5(T qm3, T qm2, T qm1, T q0, T qp1, T qp2, T ua, T up) {
  return by60(T(37) * (q0 + qm1) - T(8) * (qp1 + qm2) + (qp2 + qm3)) -
         by60(up * sgn1(ua) * ((qp2 - qm3) - T(5) * (qp1 - qm2) + T(10) * (q0 - qm1)));
}
// 3rd-order upwind flux (kernels_common.hpp:19-22); up = 0 gives the 4th-order centred flux
template <class T>
__device__ __forceinline__ T flux3(T qm2, T qm1, T q0, T qp1, T ua, T up) {
  return by12(T(7) * (q0 + qm1) - (qp1 + qm2)) + by12(up * sgn1(ua) * ((qp1 - qm2) - T(3) * (q0 - qm1)));
}

// advect_scalar_pd's own statement functions (kernels_common.hpp:23-33): folded constants, so their
// roundoff differs from flux5 / flux3
template <class T>
__device__ __forceinline__ T flux5_pd(T qm3, T qm2, T qm1, T q0, T qp1, T qp2, T ua, T up) {
  const T f6 = T(37. / 60.) * (q0 + qm1) - T(2. / 15.) * (qp1 + qm2) + T(1. / 60.) * (qp2 + qm3);
  return f6 - up * sgn1(ua) * T(1. / 60.) * ((qp2 - qm3) - T(5.) * (qp1 - qm2) + T(10.) * (q0 - qm1));
}
template <class T>
__device__ __forceinline__ T flux3_pd(T qm2, T qm1, T q0, T qp1, T ua, T up) {
  const T f4 = T(7. / 12.) * (q0 + qm1) - T(1. / 12.) * (qp1 + qm2);
  return f4 + up * sgn1(ua) * T(1. / 12.) * ((qp1 - qm2) - T(3.) * (q0 - qm1));
}
// first-order upwind flux with Courant number cr (kernels_common.hpp:83-88)
template <class T>
__device__ __forceinline__ T flux_upwind(T qm1, T q0, T cr) {
  const T a = cr < T(0) ? -cr : cr;
  const T lo = cr + a, hi = cr - a;
  return T(0.5) * (lo < T(1) ? lo : T(1)) * qm1 + T(0.5) * (hi > T(-1) ? hi : T(-1)) * q0;
}

}  // namespace mi_dev
