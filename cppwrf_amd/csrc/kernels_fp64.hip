// FP64 dynamics kernels for gfx950.
//
// Each kernel computes, point by point, the same expression as the reference CPU FP64 build, operand
// for operand (the results are compared bit for bit); the file and line of the reference kernel it
// matches is given at each.  Built with -ffp-contract=off: no multiply-add is fused, as the reference.
// Transcendental calls go to libm_port.hpp, which reproduces the host glibc's pow and sin.
//
// Launch shapes: point kernels run one thread per point with 64 consecutive i per wavefront (one
// 512-byte coalesced row segment per load); column kernels (vertical sums, the tridiagonal solve) run
// one thread per column and march in k, every load still i-contiguous across the wavefront.
#include "kernels_fp64.hpp"
#include "libm_port.hpp"

namespace mi_fp64 {

using namespace mi_dev;

namespace {

constexpr int kBx = 64;

struct Range {
  int i0, i1, j0, j1, k0, k1;
};

inline dim3 grid_of(const Range& r, int by) {
  return dim3(unsigned((r.i1 - r.i0 + kBx - 1) / kBx), unsigned((r.j1 - r.j0 + by - 1) / by), unsigned(r.k1 - r.k0));
}

#define MI_POINT_INDEX(r)                          \
  const int i = (r).i0 + int(blockIdx.x) * kBx + int(threadIdx.x); \
  const int j = (r).j0 + int(blockIdx.y) * int(blockDim.y) + int(threadIdx.y); \
  const int k = (r).k0 + int(blockIdx.z);          \
  if (i >= (r).i1 || j >= (r).j1) return;

#define MI_COLUMN_INDEX(nx, ny)                    \
  const int i = int(blockIdx.x) * kBx + int(threadIdx.x); \
  const int j = int(blockIdx.y);                   \
  if (i >= (nx) || j >= (ny)) return;

template <class K, class... T>
void launch3(K kern, const Range& r, hipStream_t st, T... args) {
  if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) return;
  const int by = (r.j1 - r.j0) == 1 ? 1 : 4;
  hipLaunchKernelGGL(kern, grid_of(r, by), dim3(kBx, by, 1), 0, st, r, args...);
}
template <class K, class... T>
void launch_cols(K kern, int nx, int ny, hipStream_t st, T... args) {
  hipLaunchKernelGGL(kern, dim3(unsigned((nx + kBx - 1) / kBx), unsigned(ny), 1), dim3(kBx, 1, 1), 0, st, args...);
}

// ----------------------------------------------------------------------------- halo fill
// One halo slot of one direction (reference semantics src/core/halo.cpp:52-84): slot a = 0..7 maps to
// i = 0, -1, -2, -3, n, n+1, n+2, n+3 on face-staggered points and -1, -2, -3, n, n+1, n+2 otherwise.
__device__ __forceinline__ void halo_slot(double* base, int64_t sn, int a, int n, bool face, bool periodic, int par) {
  const int nl = face ? 4 : 3;
  const int i = a < nl ? (face ? -a : -1 - a) : n + (a - nl);
  int src;
  if (periodic) src = ((i % n) + n) % n;
  else if (face) src = i <= 0 ? -i : 2 * n - i;
  else src = i < 0 ? -1 - i : 2 * n - 1 - i;
  double v = base[int64_t(src) * sn];
  if (!periodic && src == i) {
    if (par == 2) base[int64_t(i) * sn] = -v;
    return;
  }
  if (!periodic && par != 0) v = -v;
  base[int64_t(i) * sn] = v;
}

// x pass: thread (slot, row j, level k, field); rows of the field's interior
__global__ void k_halo_x(Geo g, HaloSet s) {
  const int a = int(threadIdx.x) & 7;
  const int j = int(blockIdx.x) * 32 + (int(threadIdx.x) >> 3);
  const int fk = int(blockIdx.y);
  const int f = fk / s.nkmax, k = fk - f * s.nkmax;
  if (f >= s.n || k >= s.nzs[f] || j >= s.nys[f]) return;
  const bool face = s.face_x[f] != 0;
  if (a >= (face ? 8 : 6)) return;
  double* base = s.p[f] + g.at(0, j, k);
  halo_slot(base, 1, a, g.nx, face, g.bcx == MI_BC_PERIODIC, s.px[f]);
}
// y pass: thread (x index over the whole row incl. halos, slot, level, field)
__global__ void k_halo_y(Geo g, HaloSet s) {
  const int i = -kHalo + int(blockIdx.x) * 64 + int(threadIdx.x);
  const int a = int(threadIdx.y);
  const int fk = int(blockIdx.y);
  const int f = fk / s.nkmax, k = fk - f * s.nkmax;
  if (f >= s.n || k >= s.nzs[f] || i >= s.nxs[f] + kHalo) return;
  const bool face = s.face_y[f] != 0;
  if (a >= (face ? 8 : 6)) return;
  double* base = s.p[f] + g.at(i, 0, k);
  halo_slot(base, g.sj, a, g.ny, face, g.bcy == MI_BC_PERIODIC, s.py[f]);
}

// ----------------------------------------------------------------------------- copies
// RK3::step's _2 -> _1 copies of whole allocations, halos included (reference rk3.cpp:94-98)
__global__ void k_copy_step(Fields f, const double* base3, const double* base2, int64_t n3, int64_t n2,
                            int64_t off3_u, int64_t off3_v, int64_t off3_w, int64_t off3_t, int64_t off3_ph,
                            int64_t off2_mu) {
  (void)base3;
  (void)base2;
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int which = int(blockIdx.y);
  if (which < 5) {
    if (e >= n3) return;
    double *dst, *src;
    int64_t off;
    switch (which) {
      case 0: dst = f.u1; src = f.u2; off = off3_u; break;
      case 1: dst = f.v1; src = f.v2; off = off3_v; break;
      case 2: dst = f.w1; src = f.w2; off = off3_w; break;
      case 3: dst = f.t1; src = f.t2; off = off3_t; break;
      default: dst = f.ph1; src = f.ph2; off = off3_ph; break;
    }
    (dst - off)[e] = (src - off)[e];
  } else {
    if (e >= n2) return;
    (f.mu1 - off2_mu)[e] = (f.mu2 - off2_mu)[e];
  }
}

__global__ void k_fill(double* p, int64_t n, double v) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (e < n) p[e] = v;
}

// ----------------------------------------------------------------------------- rk_step_prep
// KMu: full column masses at the stage (reference src/dyn/large_step.cpp:49-66)
__global__ void k_mu_uv(Range r, Args a) {
  MI_POINT_INDEX(r)
  (void)k;
  const Geo& g = a.g;
  const Fields& f = a.f;
  const int nx = g.nx, ny = g.ny;
  auto mu = [&](int ii, int jj) { return f.mu2[g.at2(ii, jj)]; };
  auto mub = [&](int ii, int jj) { return f.mub[g.at2(ii, jj)]; };
  if (i < nx && j < ny) f.mut[g.at2(i, j)] = mu(i, j) + mub(i, j);
  if (j < ny) f.muu[g.at2(i, j)] = H * (mu(i, j) + mu(i - 1, j) + mub(i, j) + mub(i - 1, j));
  if (i < nx && j < g.nyv) f.muv[g.at2(i, j)] = H * (mu(i, j) + mu(i, j - 1) + mub(i, j) + mub(i, j - 1));
}

// KPrep3: couple_momentum, calc_alt, calc_php (large_step.cpp:68-81)
__global__ void k_couple(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  if (j < ny && k < nz) f.ru[x] = f.u2[x] * (c.c1h[k] * f.muu[x2] + c.c2h[k]) / f.msfuy[x2];
  if (i < nx && j < g.nyv && k < nz) f.rv[x] = f.v2[x] * (c.c1h[k] * f.muv[x2] + c.c2h[k]) * f.msfvx_inv[x2];
  if (i < nx && j < ny) f.rw[x] = f.w2[x] * (c.c1f[k] * f.mut[x2] + c.c2f[k]) / f.msfty[x2];
  if (i < nx && j < ny && k < nz) {
    f.alt[x] = f.al[x] + f.alb(i, j, k);
    f.php[x] = H * (f.phb(i, j, k) + f.phb(i, j, k + 1) + f.ph2[x] + f.ph2[x + g.sk]);
  }
}

// KWW calc_ww_cp (large_step.cpp:83-108): column sum of the mass divergence in k order, then the
// running eta-velocity integral
__global__ void k_ww(Args a) {
  const Geo& g = a.g;
  MI_COLUMN_INDEX(g.nx, g.ny)
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  auto mu_u = [&](int ii, int jj) {
    return H * (f.mu2[g.at2(ii, jj)] + f.mub[g.at2(ii, jj)] + f.mu2[g.at2(ii - 1, jj)] + f.mub[g.at2(ii - 1, jj)]);
  };
  auto mu_v = [&](int ii, int jj) {
    return H * (f.mu2[g.at2(ii, jj)] + f.mub[g.at2(ii, jj)] + f.mu2[g.at2(ii, jj - 1)] + f.mub[g.at2(ii, jj - 1)]);
  };
  const double mu1 = mu_u(i + 1, j), mu0 = mu_u(i, j), mv1 = mu_v(i, j + 1), mv0 = mu_v(i, j);
  const double msftx = f.msftx[g.at2(i, j)], msfuy1 = f.msfuy[g.at2(i + 1, j)], msfuy0 = f.msfuy[g.at2(i, j)];
  const double mvi1 = f.msfvx_inv[g.at2(i, j + 1)], mvi0 = f.msfvx_inv[g.at2(i, j)];
  auto divv = [&](int kk) {
    const int64_t x = g.at(i, j, kk);
    return msftx * c.dnw[kk] *
           (c.rdx * ((c.c1h[kk] * mu1 + c.c2h[kk]) * f.u2[x + 1] / msfuy1 - (c.c1h[kk] * mu0 + c.c2h[kk]) * f.u2[x] / msfuy0) +
            c.rdy * ((c.c1h[kk] * mv1 + c.c2h[kk]) * f.v2[x + g.sj] * mvi1 - (c.c1h[kk] * mv0 + c.c2h[kk]) * f.v2[x] * mvi0));
  };
  double dmdt = 0;
  for (int kk = 0; kk < nz; ++kk) dmdt += divv(kk);
  double w = 0;
  f.ww[g.at(i, j, 0)] = 0;
  for (int kk = 1; kk < nz; ++kk) {
    w = w - c.dnw[kk - 1] * c.c1h[kk - 1] * dmdt - divv(kk - 1);
    f.ww[g.at(i, j, kk)] = w;
  }
  f.ww[g.at(i, j, nz)] = 0;
}

// ----------------------------------------------------------------------------- rk_tendency
// KTendU (large_step.cpp:152-264): advect_u + horizontal pressure gradient + curvature, interior faces
__global__ void k_tend_u(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t sj = g.sj, sk = g.sk;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double* u = f.u2;
  const double hup = a.q.hup, vup = a.q.vup;
  auto fx = [&](int64_t y) {  // face between y-1 and y
    const double vel = H * (f.ru[y] + f.ru[y - 1]);
    return vel * flux5(u[y - 3], u[y - 2], u[y - 1], u[y], u[y + 1], u[y + 2], vel, hup);
  };
  auto fy = [&](int64_t y) {
    const double vel = H * (f.rv[y] + f.rv[y - 1]);
    return vel * flux5(u[y - 3 * sj], u[y - 2 * sj], u[y - sj], u[y], u[y + sj], u[y + 2 * sj], vel, hup);
  };
  auto fz = [&](int kk) {
    if (kk == 0 || kk == nz) return 0.0;
    const int64_t y = g.at(i, j, kk);
    const double vel = H * (f.ww[y - 1] + f.ww[y]);
    if (kk == 1 || kk == nz - 1) return vel * (c.fnm[kk] * u[y] + c.fnp[kk] * u[y - sk]);
    return vel * flux3(u[y - 2 * sk], u[y - sk], u[y], u[y + sk], -vel, vup);
  };
  auto pu = [&](int kk) { const int64_t y = g.at(i, j, kk); return f.p[y - 1] + f.p[y]; };
  auto dpn = [&](int kk) {
    if (kk == 0) return H * (c.cf1 * pu(0) + c.cf2 * pu(1) + c.cf3 * pu(2));
    if (kk == nz) return a.q.top_lid ? H * (c.cfn * pu(nz - 1) + c.cfn1 * pu(nz - 2)) : 0.0;
    return H * (c.fnm[kk] * pu(kk) + c.fnp[kk] * pu(kk - 1));
  };
  const double msfux = f.msfux[x2], msfuy = f.msfuy[x2];
  double t = -(msfux * c.rdx) * (fx(x + 1) - fx(x));
  if (!g.two_d) t -= (msfux * c.rdy) * (fy(x + sj) - fy(x));
  t -= c.rdnw[k] * (fz(k + 1) - fz(k));
  const double rm = msfux / msfuy;
  double dpx = rm * H * c.rdx * (c.c1h[k] * f.muu[x2] + c.c2h[k]) *
               ((f.ph2[x + sk] - f.ph2[x + sk - 1] + f.ph2[x] - f.ph2[x - 1]) + (f.alt[x] + f.alt[x - 1]) * (f.p[x] - f.p[x - 1]) +
                (f.al[x] + f.al[x - 1]) * (f.pb(i, j, k) - f.pb(i - 1, j, k)));
  dpx += rm * c.rdx * (f.php[x] - f.php[x - 1]) *
         (c.rdnw[k] * (dpn(k + 1) - dpn(k)) - H * (c.c1h[k] * f.mu2[x2 - 1] + c.c1h[k] * f.mu2[x2]));
  t -= 1.0 * dpx;
  const double rw4 = Q * (f.rw[x - 1 + sk] + f.rw[x - 1] + f.rw[x + sk] + f.rw[x]);
  t = t - u[x] * kReradius * rw4;
  f.ru_tend[x] = t;
}

// KTendV (large_step.cpp:267-376)
__global__ void k_tend_v(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t sj = g.sj, sk = g.sk;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double* v = f.v2;
  const double hup = a.q.hup, vup = a.q.vup;
  auto fx = [&](int64_t y) {
    const double vel = H * (f.ru[y] + f.ru[y - sj]);
    return vel * flux5(v[y - 3], v[y - 2], v[y - 1], v[y], v[y + 1], v[y + 2], vel, hup);
  };
  auto fy = [&](int64_t y) {
    const double vel = H * (f.rv[y] + f.rv[y - sj]);
    return vel * flux5(v[y - 3 * sj], v[y - 2 * sj], v[y - sj], v[y], v[y + sj], v[y + 2 * sj], vel, hup);
  };
  auto fz = [&](int kk) {
    if (kk == 0 || kk == nz) return 0.0;
    const int64_t y = g.at(i, j, kk);
    const double vel = H * (f.ww[y - sj] + f.ww[y]);
    if (kk == 1 || kk == nz - 1) return vel * (c.fnm[kk] * v[y] + c.fnp[kk] * v[y - sk]);
    return vel * flux3(v[y - 2 * sk], v[y - sk], v[y], v[y + sk], -vel, vup);
  };
  auto pv = [&](int kk) { const int64_t y = g.at(i, j, kk); return f.p[y - sj] + f.p[y]; };
  auto dpn = [&](int kk) {
    if (kk == 0) return H * (c.cf1 * pv(0) + c.cf2 * pv(1) + c.cf3 * pv(2));
    if (kk == nz) return a.q.top_lid ? H * (c.cfn * pv(nz - 1) + c.cfn1 * pv(nz - 2)) : 0.0;
    return H * (c.fnm[kk] * pv(kk) + c.fnp[kk] * pv(kk - 1));
  };
  const double msfvy = f.msfvy[x2], msfvx = f.msfvx[x2];
  const double rm = msfvy / msfvx;
  double t = -(msfvy * c.rdx) * (fx(x + 1) - fx(x));
  t -= (msfvy * c.rdy) * (fy(x + sj) - fy(x));
  t -= rm * c.rdnw[k] * (fz(k + 1) - fz(k));
  double dpy = rm * H * c.rdy * (c.c1h[k] * f.muv[x2] + c.c2h[k]) *
               ((f.ph2[x + sk] - f.ph2[x + sk - sj] + f.ph2[x] - f.ph2[x - sj]) + (f.alt[x] + f.alt[x - sj]) * (f.p[x] - f.p[x - sj]) +
                (f.al[x] + f.al[x - sj]) * (f.pb(i, j, k) - f.pb(i, j - 1, k)));
  dpy += rm * c.rdy * (f.php[x] - f.php[x - sj]) *
         (c.rdnw[k] * (dpn(k + 1) - dpn(k)) - H * (c.c1h[k] * f.mu2[x2 - sj] + c.c1h[k] * f.mu2[x2]));
  t -= 1.0 * dpy;
  const double rw4 = Q * (f.rw[x - sj + sk] + f.rw[x - sj] + f.rw[x + sk] + f.rw[x]);
  t = t - rm * v[x] * kReradius * rw4;
  f.rv_tend[x] = t;
}

// KTendW (large_step.cpp:379-499), w levels 1..nz
__global__ void k_tend_w(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t sj = g.sj, sk = g.sk;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double* w = f.w2;
  const double hup = a.q.hup, vup = a.q.vup;
  auto rux = [&](int64_t y) {  // ru at w level k
    if (k < nz) return c.fnm[k] * f.ru[y] + c.fnp[k] * f.ru[y - sk];
    return (2.0 - c.fnm[nz - 1]) * f.ru[y - sk] - c.fnp[nz - 1] * f.ru[y - 2 * sk];
  };
  auto rvy = [&](int64_t y) {
    if (k < nz) return c.fnm[k] * f.rv[y] + c.fnp[k] * f.rv[y - sk];
    return (2.0 - c.fnm[nz - 1]) * f.rv[y - sk] - c.fnp[nz - 1] * f.rv[y - 2 * sk];
  };
  auto fx = [&](int64_t y) {
    const double vel = rux(y);
    return vel * flux5(w[y - 3], w[y - 2], w[y - 1], w[y], w[y + 1], w[y + 2], vel, hup);
  };
  auto fy = [&](int64_t y) {
    const double vel = rvy(y);
    return vel * flux5(w[y - 3 * sj], w[y - 2 * sj], w[y - sj], w[y], w[y + sj], w[y + 2 * sj], vel, hup);
  };
  auto fz = [&](int m) {  // vertical flux at mass level m
    const int64_t y = g.at(i, j, m);
    if (m == 0 || m == nz - 1) return 0.25 * (f.ww[y + sk] + f.ww[y]) * (w[y + sk] + w[y]);
    const double vel = H * (f.ww[y + sk] + f.ww[y]);
    return vel * flux3(w[y - sk], w[y], w[y + sk], w[y + 2 * sk], -vel, vup);
  };
  const double msftx = f.msftx[x2], msfty = f.msfty[x2];
  double t = -(msftx * c.rdx) * (fx(x + 1) - fx(x));
  if (!g.two_d) t -= (msftx * c.rdy) * (fy(x + sj) - fy(x));
  const double rmsft = 1.0 / msfty;
  if (k < nz) {
    t -= c.rdn[k] * (fz(k) - fz(k - 1));
    const double cqw = 0.0, cq1 = 1.0 / (1.0 + cqw), cq2 = cqw * cq1;
    t += rmsft * kG * (cq1 * c.rdn[k] * (f.p[x] - f.p[x - sk]) - c.c1f[k] * f.mu2[x2] - cq2 * (c.c1f[k] * f.mub[x2] + c.c2f[k]));
    if (a.q.wd_dt != 0.0) {
      const double muf = c.c1f[k] * f.mut[x2] + c.c2f[k];
      double cfl = f.ww[x] / muf * c.rdnw[k] * a.q.wd_dt;
      cfl = cfl < 0.0 ? -cfl : cfl;
      if (cfl > 1.0) t = t - (w[x] < 0.0 ? -1.0 : 1.0) * 0.3 * (cfl - a.q.w_crit) * muf;
    }
    const double ruw = H * (c.fnm[k] * (f.ru[x] + f.ru[x + 1]) + c.fnp[k] * (f.ru[x - sk] + f.ru[x - sk + 1]));
    const double rvw = H * (c.fnm[k] * (f.rv[x] + f.rv[x + sj]) + c.fnp[k] * (f.rv[x - sk] + f.rv[x - sk + sj]));
    const double rmt = msftx / msfty;
    const double uw = H * (c.fnm[k] * (f.u2[x] + f.u2[x + 1]) + c.fnp[k] * (f.u2[x - sk] + f.u2[x - sk + 1]));
    const double vw = H * (c.fnm[k] * (f.v2[x] + f.v2[x + sj]) + c.fnp[k] * (f.v2[x - sk] + f.v2[x - sk + sj]));
    t = t + kReradius * (ruw * uw + rmt * rvw * vw);
  } else {
    t += 2.0 * c.rdn[nz - 1] * fz(nz - 1);
    const double cqw = 0.0, cq1 = 1.0 / (1.0 + cqw), cq2 = cqw * cq1;
    t += rmsft * kG *
         (cq1 * 2.0 * c.rdnw[nz - 1] * (-f.p[x - sk]) - c.c1f[nz] * f.mu2[x2] - cq2 * (c.c1f[nz] * f.mub[x2] + c.c2f[nz]));
  }
  f.rw_tend[x] = t;
}

// KTendT (large_step.cpp:502-565)
__global__ void k_tend_t(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t sj = g.sj, sk = g.sk;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double* t2 = f.t2;
  const double hup = a.q.hup, vup = a.q.vup;
  auto fx = [&](int64_t y) {
    const double vel = f.ru[y];
    return vel * flux5(t2[y - 3], t2[y - 2], t2[y - 1], t2[y], t2[y + 1], t2[y + 2], vel, hup);
  };
  auto fy = [&](int64_t y) {
    const double vel = f.rv[y];
    return vel * flux5(t2[y - 3 * sj], t2[y - 2 * sj], t2[y - sj], t2[y], t2[y + sj], t2[y + 2 * sj], vel, hup);
  };
  auto fz = [&](int kk) {
    if (kk == 0 || kk == nz) return 0.0;
    const int64_t y = g.at(i, j, kk);
    const double vel = f.ww[y];
    if (kk == 1 || kk == nz - 1) return vel * (c.fnm[kk] * t2[y] + c.fnp[kk] * t2[y - sk]);
    return vel * flux3(t2[y - 2 * sk], t2[y - sk], t2[y], t2[y + sk], -vel, vup);
  };
  const double msftx = f.msftx[x2];
  double rr = -(msftx * c.rdx) * (fx(x + 1) - fx(x));
  if (!g.two_d) rr -= (msftx * c.rdy) * (fy(x + sj) - fy(x));
  rr -= c.rdnw[k] * (fz(k + 1) - fz(k));
  f.t_tend[x] = rr;
}

// KTendPh rhs_ph (large_step.cpp:570-668), w levels 1..nz, 6th-order horizontal gradient
__global__ void k_tend_ph(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t sj = g.sj, sk = g.sk;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double* ph = f.ph2;
  const double* u = f.u2;
  const double* v = f.v2;
  auto wdwn = [&](int m) {
    const int64_t y = g.at(i, j, m);
    return H * (f.ww[y + sk] + f.ww[y]) * c.rdnw[m] * (ph[y + sk] - ph[y] + f.phb(i, j, m + 1) - f.phb(i, j, m));
  };
  auto grad6x = [&]() {
    return 45.0 * (ph[x + 1] - ph[x - 1]) - 9.0 * (ph[x + 2] - ph[x - 2]) + (ph[x + 3] - ph[x - 3]) +
           45.0 * (f.phb(i + 1, j, k) - f.phb(i - 1, j, k)) - 9.0 * (f.phb(i + 2, j, k) - f.phb(i - 2, j, k)) +
           (f.phb(i + 3, j, k) - f.phb(i - 3, j, k));
  };
  auto grad6y = [&]() {
    return 45.0 * (ph[x + sj] - ph[x - sj]) - 9.0 * (ph[x + 2 * sj] - ph[x - 2 * sj]) + (ph[x + 3 * sj] - ph[x - 3 * sj]) +
           45.0 * (f.phb(i, j + 1, k) - f.phb(i, j - 1, k)) - 9.0 * (f.phb(i, j + 2, k) - f.phb(i, j - 2, k)) +
           (f.phb(i, j + 3, k) - f.phb(i, j - 3, k));
  };
  const double msfty = f.msfty[x2];
  double t = 0;
  if (k < nz) t = t - (c.fnm[k] * wdwn(k) + c.fnp[k] * wdwn(k - 1));
  t = t + (c.c1f[k] * f.mut[x2] + c.c2f[k]) * kG * f.w2[x] / msfty;
  const double r60 = 1.0 / 60.0;
  const int64_t xt = g.at(i, j, nz - 1);  // top extrapolation reads levels nz-1, nz-2
  if (!g.two_d) {
    const double mvp = c.c1f[k] * f.muv[x2 + sj] + c.c2f[k], mvm = c.c1f[k] * f.muv[x2] + c.c2f[k];
    const double vp = k < nz ? v[x + sj] + v[x + sj - sk] : c.cfn * v[xt + sj] + c.cfn1 * v[xt + sj - sk];
    const double vm = k < nz ? v[x] + v[x - sk] : c.cfn * v[xt] + c.cfn1 * v[xt - sk];
    const double cy = (k < nz ? 0.25 : H) * c.rdy / msfty;
    t = t - cy * ((mvp * vp * f.msfvy[x2 + sj] + mvm * vm * f.msfvy[x2]) * r60 * grad6y());
  }
  const double mup = c.c1f[k] * f.muu[x2 + 1] + c.c2f[k], mum = c.c1f[k] * f.muu[x2] + c.c2f[k];
  const double up = k < nz ? u[x + 1] + u[x + 1 - sk] : c.cfn * u[xt + 1] + c.cfn1 * u[xt + 1 - sk];
  const double um = k < nz ? u[x] + u[x - sk] : c.cfn * u[xt] + c.cfn1 * u[xt - sk];
  const double cx = (k < nz ? 0.25 : H) * c.rdx / msfty;
  t = t - cx * ((mup * up * f.msfux[x2 + 1] + mum * um * f.msfux[x2]) * r60 * grad6x());
  f.ph_tend[x] = t;
}

// ----------------------------------------------------------------------------- small_step_prep
// KPrep (reference src/dyn/small_step.cpp:39-93)
__global__ void k_ss_prep(Range r, Args a, int first) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double mts = f.mub[x2] + f.mu1[x2];
  const double mus_u = first ? f.muu[x2] : H * (mts + f.mub[x2 - 1] + f.mu1[x2 - 1]);
  const double mus_v = first ? f.muv[x2] : H * (mts + f.mub[x2 - g.sj] + f.mu1[x2 - g.sj]);
  if (k == 0) {
    if (i < nx && j < ny) {
      f.muts[x2] = mts;
      const double m2 = f.mu2[x2];
      f.mu_save[x2] = m2;
      f.mu2[x2] = f.mu1[x2] - m2;
    }
    if (j < ny) f.muus[x2] = mus_u;
    if (i < nx && j < g.nyv) f.muvs[x2] = mus_v;
  }
  if (j < ny && k < nz) {
    const double ui = f.u2[x];
    f.u_save[x] = ui;
    f.u2[x] = ((c.c1h[k] * mus_u + c.c2h[k]) * f.u1[x] - (c.c1h[k] * f.muu[x2] + c.c2h[k]) * ui) / f.msfuy[x2];
  }
  if (i < nx && j < g.nyv && k < nz) {
    const double vi = f.v2[x];
    f.v_save[x] = vi;
    f.v2[x] = ((c.c1h[k] * mus_v + c.c2h[k]) * f.v1[x] - (c.c1h[k] * f.muv[x2] + c.c2h[k]) * vi) * f.msfvx_inv[x2];
  }
  if (i < nx && j < ny) {
    if (k < nz) {
      const double ti = f.t2[x];
      f.t_save[x] = ti;
      f.t2[x] = (c.c1h[k] * mts + c.c2h[k]) * f.t1[x] - (c.c1h[k] * f.mut[x2] + c.c2h[k]) * ti;
      f.c2a[x] = kCpovcv * (f.pb(i, j, k) + f.p[x]) / f.alt[x];
    }
    const double wi = f.w2[x], phi = f.ph2[x];
    f.w_save[x] = wi;
    f.w2[x] = ((c.c1f[k] * mts + c.c2f[k]) * f.w1[x] - (c.c1f[k] * f.mut[x2] + c.c2f[k]) * wi) / f.msfty[x2];
    f.ph_save[x] = phi;
    f.ph2[x] = f.ph1[x] - phi;
    f.ww1[x] = (k == 0 || k == nz) ? 0.0 : f.ww[x];
  }
}

// p_rho_point (small_step.cpp:97-114): KPRho (damp = 0) and KWPRho (damp = 1, smdiv damping)
__global__ void k_p_rho(Range r, Args a, int damp) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double m = c.c1h[k] * f.muts[x2] + c.c2h[k];
  const double al = -((1. / m) * (f.alt[x] * (c.c1h[k] * f.mu2[x2]) + c.rdnw[k] * (f.ph2[x + g.sk] - f.ph2[x])));
  f.al[x] = al;
  double pp = f.c2a[x] * (f.alt[x] * (f.t2[x] - c.c1h[k] * f.mu2[x2] * f.t_save[x]) / (m * (kT0 + f.t_save[x])) - al);
  if (damp) {
    const double ptmp = pp;
    pp = pp + a.q.smdiv * (pp - f.pm1[x]);
    f.pm1[x] = ptmp;
  } else {
    f.pm1[x] = pp;
  }
  f.p[x] = pp;
}

// KCoefW calc_coef_w (small_step.cpp:124-148), dry
__global__ void k_coef_w(Args a, double cof) {
  const Geo& g = a.g;
  MI_COLUMN_INDEX(g.nx, g.ny)
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t x0 = g.at(i, j, 0), sk = g.sk;
  const double m = f.mut[g.at2(i, j)];
  const double lid = a.q.top_lid ? 0.0 : 1.0;
  auto mh = [&](int kk) { return c.c1h[kk] * m + c.c2h[kk]; };
  auto mf = [&](int kk) { return c.c1f[kk] * m + c.c2f[kk]; };
  auto c2a = [&](int kk) { return f.c2a[x0 + kk * sk]; };
  f.ca[x0] = 0;
  f.calpha[x0] = 0;
  f.cgamma[x0] = 0;
  f.ca[x0 + sk] = 0;
  f.ca[x0 + nz * sk] = -2.0 * cof * (c.rdnw[nz - 1] * c.rdnw[nz - 1]) * c2a(nz - 1) * lid / (mh(nz - 1) * mf(nz - 1));
  for (int kk = 2; kk < nz; ++kk)
    f.ca[x0 + kk * sk] = -1.0 * cof * c.rdn[kk] * c.rdnw[kk - 1] * c2a(kk - 1) / (mh(kk - 1) * mf(kk - 1));
  double gam = 0.0;  // gamma(k-1)
  for (int kk = 1; kk < nz; ++kk) {
    const double cq1 = 1.0;
    const double b = 1.0 + cq1 * cof * c.rdn[kk] * (c.rdnw[kk] * c2a(kk) / (mh(kk) * mf(kk)) + c.rdnw[kk - 1] * c2a(kk - 1) / (mh(kk - 1) * mf(kk)));
    const double cc = -cq1 * cof * c.rdn[kk] * c.rdnw[kk] * c2a(kk) / (mh(kk) * mf(kk + 1));
    const double al = 1.0 / (b - f.ca[x0 + kk * sk] * gam);
    f.calpha[x0 + kk * sk] = al;
    gam = cc * al;
    f.cgamma[x0 + kk * sk] = gam;
  }
  const double b = 1.0 + 2.0 * cof * (c.rdnw[nz - 1] * c.rdnw[nz - 1]) * c2a(nz - 1) / (mh(nz - 1) * mf(nz));
  f.calpha[x0 + nz * sk] = 1.0 / (b - f.ca[x0 + nz * sk] * gam);
  f.cgamma[x0 + nz * sk] = 0;
}

// ----------------------------------------------------------------------------- advance_uv
// KUV (small_step.cpp:151-190)
__global__ void k_uv(Range r, Args a, double dts) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j), sk = g.sk;
  auto dpn = [&](int64_t d, int kk) {
    auto ps = [&](int m) { const int64_t y = g.at(i, j, m); return f.p[y] + f.p[y - d]; };
    if (kk == 0) return H * (c.cf1 * ps(0) + c.cf2 * ps(1) + c.cf3 * ps(2));
    if (kk == nz) return a.q.top_lid ? H * (c.cf1 * ps(nz - 1) + c.cf2 * ps(nz - 2) + c.cf3 * ps(nz - 3)) : 0.0;
    return H * (c.fnm[kk] * ps(kk) + c.fnp[kk] * ps(kk - 1));
  };
  auto dp = [&](int64_t d, int di, int dj, double mface, double rd, double rm) {
    double dd = rm * H * rd * (c.c1h[k] * mface + c.c2h[k]) *
                (((f.ph2[x + sk] - f.ph2[x + sk - d]) + (f.ph2[x] - f.ph2[x - d])) + (f.alt[x] + f.alt[x - d]) * (f.p[x] - f.p[x - d]) +
                 (f.al[x] + f.al[x - d]) * (f.pb(i, j, k) - f.pb(i - di, j - dj, k)));
    dd += rm * rd * (f.php[x] - f.php[x - d]) *
          (c.rdnw[k] * (dpn(d, k + 1) - dpn(d, k)) - H * (c.c1h[k] * f.mu2[x2 - d] + c.c1h[k] * f.mu2[x2]));
    return dd;
  };
  const int ui0 = g.bcx != MI_BC_PERIODIC ? 1 : 0;
  if (i >= ui0 && i < g.nx) {
    const double mudf_x = -a.q.emdx * (f.mudf[x2] - f.mudf[x2 - 1]) / f.msfuy[x2];
    f.u2[x] = f.u2[x] + dts * f.ru_tend[x] - dts * 1.0 * dp(1, 1, 0, f.muu[x2], c.rdx, f.msfux[x2] / f.msfuy[x2]) + c.c1h[k] * mudf_x;
  }
  const int vj0 = (g.bcy != MI_BC_PERIODIC && !g.two_d) ? 1 : 0;
  if (!g.two_d && j >= vj0 && j < g.ny) {
    const int64_t sj = g.sj;
    const double mudf_y = -a.q.emdy * (f.mudf[x2] - f.mudf[x2 - sj]) * f.msfvx_inv[x2];
    f.v2[x] = f.v2[x] + dts * f.rv_tend[x] - dts * 1.0 * dp(sj, 0, 1, f.muv[x2], c.rdy, f.msfvy[x2] / f.msfvx[x2]) + c.c1h[k] * mudf_y;
  }
}

// ----------------------------------------------------------------------------- advance_mu_t + sumflux
// KMuTDiv (small_step.cpp:194-207)
__global__ void k_mut_div(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j), sj = g.sj;
  const double fx = (f.u2[x + 1] + (c.c1h[k] * f.muu[x2 + 1] + c.c2h[k]) * f.u_save[x + 1] / f.msfuy[x2 + 1]) -
                    (f.u2[x] + (c.c1h[k] * f.muu[x2] + c.c2h[k]) * f.u_save[x] / f.msfuy[x2]);
  double d = c.rdx * fx;
  if (!g.two_d)
    d += c.rdy * ((f.v2[x + sj] + (c.c1h[k] * f.muv[x2 + sj] + c.c2h[k]) * f.v_save[x + sj] * f.msfvx_inv[x2 + sj]) -
                  (f.v2[x] + (c.c1h[k] * f.muv[x2] + c.c2h[k]) * f.v_save[x] * f.msfvx_inv[x2]));
  f.dvdxi[x] = f.msftx[x2] * f.msfty[x2] * d;
}

// KMuTCol (small_step.cpp:209-248): mu, mudf, muts, muave; omega'' and its running sum ww_m
__global__ void k_mut_col(Args a, double dts, int it, int ns) {
  const Geo& g = a.g;
  MI_COLUMN_INDEX(g.nx, g.ny)
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t x0 = g.at(i, j, 0), x2 = g.at2(i, j), sk = g.sk;
  const bool first = it == 1, last = it == ns;
  const double rn = double(ns);
  double dmdt = 0;
  for (int kk = 0; kk < nz; ++kk) dmdt += c.dnw[kk] * f.dvdxi[x0 + kk * sk];
  const double mtend = f.mu_tend[x2];
  const double mu_old = f.mu2[x2];
  const double mu_new = mu_old + dts * (dmdt + mtend);
  f.mu2[x2] = mu_new;
  f.mudf[x2] = dmdt + mtend;
  f.muts[x2] = f.mut[x2] + mu_new;
  f.muave[x2] = H * ((1.0 + a.q.epssm) * mu_new + (1.0 - a.q.epssm) * mu_old);
  const double msfty = f.msfty[x2];
  auto acc_ww = [&](int kk, double wpp) {
    const int64_t y = x0 + kk * sk;
    const double sw = (first ? 0.0 : f.ww_m[y]) + wpp;
    f.ww_m[y] = last ? sw / rn + f.ww1[y] : sw;
  };
  double ww_full = 0;
  f.ww[x0] = 0;
  acc_ww(0, 0.0);
  for (int kk = 0; kk < nz; ++kk) {
    double wpp_hi = 0;
    if (kk + 1 < nz) {
      ww_full = ww_full - c.dnw[kk] * (c.c1h[kk] * dmdt + f.dvdxi[x0 + kk * sk] + c.c1h[kk] * mtend) / msfty;
      wpp_hi = ww_full - f.ww1[x0 + (kk + 1) * sk];
      f.ww[x0 + (kk + 1) * sk] = wpp_hi;
    } else {
      f.ww[x0 + nz * sk] = 0;
    }
    acc_ww(kk + 1, wpp_hi);
  }
}

// KMuTPoint (small_step.cpp:250-276): ru_m / rv_m sums and the theta'' update
__global__ void k_mut_point(Range r, Args a, double dts, int it, int ns) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j), sj = g.sj, sk = g.sk;
  const bool first = it == 1, last = it == ns;
  const double rn = double(ns);
  auto wdtn = [&](int kk) {
    if (kk == 0 || kk == nz) return 0.0;
    const int64_t y = g.at(i, j, kk);
    return f.ww[y] * (c.fnm[kk] * f.t_save[y] + c.fnp[kk] * f.t_save[y - sk]);
  };
  const double su = (first ? 0.0 : f.ru_m[x]) + f.u2[x];
  f.ru_m[x] = last ? su / rn + (c.c1h[k] * f.muu[x2] + c.c2h[k]) * f.u_save[x] / f.msfuy[x2] : su;
  if (!g.two_d) {
    const double sv = (first ? 0.0 : f.rv_m[x]) + f.v2[x];
    f.rv_m[x] = last ? sv / rn + (c.c1h[k] * f.muv[x2] + c.c2h[k]) * f.v_save[x] * f.msfvx_inv[x2] : sv;
  }
  const double tk = f.t2[x];
  f.t_2save[x] = tk;
  const double* ts = f.t_save;
  double hflux = H * c.rdx * (f.u2[x + 1] * (ts[x + 1] + ts[x]) - f.u2[x] * (ts[x] + ts[x - 1]));
  if (!g.two_d) hflux += H * c.rdy * (f.v2[x + sj] * (ts[x + sj] + ts[x]) - f.v2[x] * (ts[x] + ts[x - sj]));
  const double msfty = f.msfty[x2];
  const double tn = tk + msfty * dts * f.t_tend[x];
  f.t2[x] = tn - dts * msfty * (f.msftx[x2] * hflux + c.rdnw[k] * (wdtn(k + 1) - wdtn(k)));
}

// ----------------------------------------------------------------------------- advance_w
// KWRhs (small_step.cpp:312-328): explicit phi'' (rhs) and t_2ave
__global__ void k_w_rhs(Range r, Args a, double dts) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j), sk = g.sk;
  const double epssm = a.q.epssm;
  double R;
  if (k == 0) {
    R = 0;
  } else if (a.q.top_lid && k == nz) {
    R = 0;
  } else {
    auto wdwn = [&](int m) {
      const int64_t y = g.at(i, j, m);
      return H * (f.ww[y] + f.ww[y - sk]) * c.rdnw[m - 1] * (f.ph_save[y] - f.ph_save[y - sk] + f.phb(i, j, m) - f.phb(i, j, m - 1));
    };
    double rr = dts * (f.ph_tend[x] + H * kG * (1.0 - epssm) * f.w2[x]);
    if (k < nz) rr -= dts * (c.fnm[k] * wdwn(k + 1) + c.fnp[k] * wdwn(k));
    R = f.ph2[x] + f.msfty[x2] * rr / (c.c1f[k] * f.mut[x2] + c.c2f[k]);
  }
  f.w_rhs[x] = R;
  if (k < nz) {
    const double tav = H * ((1.0 + epssm) * f.t2[x] + (1.0 - epssm) * f.t_2save[x]);
    f.t2a[x] = (tav + c.c1h[k] * f.muave[x2] * kT0) / ((c.c1h[k] * f.muts[x2] + c.c2h[k]) * (kT0 + f.t_save[x]));
  }
}

// KWExpl (small_step.cpp:331-378): explicit w'', flat lower boundary
__global__ void k_w_expl(Range r, Args a, double dts) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j), sk = g.sk;
  if (k == 0) {
    f.w2[x] = 0;
    return;
  }
  const double m = f.mut[x2], mav = f.muave[x2];
  const double ope = 1.0 + a.q.epssm, ome = 1.0 - a.q.epssm;
  const double msft_inv = 1.0 / f.msfty[x2];
  auto mh = [&](int kk) { return c.c1h[kk] * m + c.c2h[kk]; };
  const double* R = f.w_rhs;
  const double* ph = f.ph2;
  const double r_k = R[x], r_km1 = R[x - sk], ph_k = ph[x], ph_km1 = ph[x - sk];
  if (k < nz) {
    const double r_kp1 = R[x + sk], ph_kp1 = ph[x + sk];
    f.w2[x] = f.w2[x] + dts * f.rw_tend[x] +
              msft_inv * 1.0 *
                  (H * dts * kG * c.rdn[k] *
                   (f.c2a[x] * c.rdnw[k] / mh(k) * (ope * (r_kp1 - r_k) + ome * (ph_kp1 - ph_k)) -
                    f.c2a[x - sk] * c.rdnw[k - 1] / mh(k - 1) * (ope * (r_k - r_km1) + ome * (ph_k - ph_km1)))) +
              dts * kG * msft_inv *
                  (c.rdn[k] * (f.c2a[x] * f.alt[x] * f.t2a[x] - f.c2a[x - sk] * f.alt[x - sk] * f.t2a[x - sk]) - c.c1f[k] * mav);
  } else {
    const int km = nz - 1;
    const int64_t xm = x - sk;
    double wn = f.w2[x] + dts * f.rw_tend[x] +
                msft_inv * (-H * dts * kG / mh(km) * (c.rdnw[km] * c.rdnw[km]) * 2.0 * f.c2a[xm] * (ope * (r_k - r_km1) + ome * (ph_k - ph_km1)) -
                            dts * kG * (2.0 * c.rdnw[km] * f.c2a[xm] * f.alt[xm] * f.t2a[xm] + c.c1f[k] * mav));
    if (a.q.top_lid) wn = 0;
    f.w2[x] = wn;
  }
}

// KWSolve (small_step.cpp:381-407): tridiagonal forward/back substitution, damp_opt = 3 Rayleigh
// damping of w near the top, phi'' from the implicit w
__global__ void k_w_solve(Args a, double dts) {
  const Geo& g = a.g;
  MI_COLUMN_INDEX(g.nx, g.ny)
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const int64_t x0 = g.at(i, j, 0), x2 = g.at2(i, j), sk = g.sk;
  const double m = f.mut[x2], mts = f.muts[x2], ope = 1.0 + a.q.epssm;
  double* w = f.w2 + x0;
  double wprev = w[0];
  for (int kk = 1; kk <= nz; ++kk) {
    const int64_t y = kk * sk;
    wprev = (w[y] - f.ca[x0 + y] * wprev) * f.calpha[x0 + y];
    w[y] = wprev;
  }
  double wnext = w[nz * sk];
  for (int kk = nz - 1; kk >= 1; --kk) {
    const int64_t y = kk * sk;
    wnext = w[y] - f.cgamma[x0 + y] * wnext;
    w[y] = wnext;
  }
  const double dampmag = a.q.damp_opt == 3 ? dts * a.q.dampcoef : 0.0;
  if (dampmag != 0.0) {
    const double pi = 3.141592653589793;  // 4 atan(1), as the compiler folds it
    const double hdepth = a.q.zdamp;
    const double htop = (f.ph_save[x0 + nz * sk] + f.phb(i, j, nz)) / kG, hbot = htop - hdepth;
    for (int kk = nz; kk >= 1; --kk) {
      const int64_t y = kk * sk;
      const double hk = (f.ph_save[x0 + y] + f.phb(i, j, kk)) / kG;
      double dampwt = 0;
      if (hk >= hbot) {
        const double sn = mi_libm::sin(H * pi * (hk - hbot) / hdepth);
        dampwt = dampmag * sn * sn;
      }
      w[y] = (w[y] - dampwt * (c.c1f[kk] * m + c.c2f[kk]) * f.w_save[x0 + y]) / (1.0 + dampwt);
    }
  }
  const double msfty = f.msfty[x2];
  for (int kk = nz; kk >= 1; --kk) {
    const int64_t y = kk * sk;
    f.ph2[x0 + y] = f.w_rhs[x0 + y] + msfty * H * dts * kG * ope * w[y] / (c.c1f[kk] * mts + c.c2f[kk]);
  }
}

// ----------------------------------------------------------------------------- small_step_finish
// KFinish (small_step.cpp:419-446), dry
__global__ void k_finish(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  if (j < ny && k < nz) {
    const double mus_u = H * (f.muts[x2] + f.muts[x2 - 1]);
    f.u2[x] = (f.msfuy[x2] * f.u2[x] + f.u_save[x] * (c.c1h[k] * f.muu[x2] + c.c2h[k])) / (c.c1h[k] * mus_u + c.c2h[k]);
  }
  if (i < nx && j < g.nyv && k < nz) {
    const double mus_v = H * (f.muts[x2] + f.muts[x2 - g.sj]);
    f.v2[x] = (f.msfvx[x2] * f.v2[x] + f.v_save[x] * (c.c1h[k] * f.muv[x2] + c.c2h[k])) / (c.c1h[k] * mus_v + c.c2h[k]);
  }
  if (i < nx && j < ny) {
    const double m = f.mut[x2], mts = f.muts[x2];
    f.w2[x] = (f.msfty[x2] * f.w2[x] + f.w_save[x] * (c.c1f[k] * m + c.c2f[k])) / (c.c1f[k] * mts + c.c2f[k]);
    f.ph2[x] = f.ph2[x] + f.ph_save[x];
    f.ww[x] = f.ww[x] + f.ww1[x];
    if (k < nz) f.t2[x] = (f.t2[x] - 0.0 + f.t_save[x] * (c.c1h[k] * m + c.c2h[k])) / (c.c1h[k] * mts + c.c2h[k]);
  }
}

// KFinishMu (small_step.cpp:448-454)
__global__ void k_finish_mu(Range r, Args a) {
  MI_POINT_INDEX(r)
  (void)k;
  const Geo& g = a.g;
  const Fields& f = a.f;
  const int64_t x2 = g.at2(i, j);
  if (i < g.nx && j < g.ny) f.mu2[x2] = f.mu2[x2] + f.mu_save[x2];
  if (j < g.ny) f.muus[x2] = H * (f.muts[x2] + f.muts[x2 - 1]);
  if (i < g.nx && j < g.nyv) f.muvs[x2] = H * (f.muts[x2] + f.muts[x2 - g.sj]);
}

// ----------------------------------------------------------------------------- calc_p_rho_phi
// KPRhoPhi (small_step.cpp:456-477), hypsometric_opt = 1, glibc pow
__global__ void k_p_rho_phi(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int64_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double al = -1.0 / (c.c1h[k] * f.muts[x2] + c.c2h[k]) * (f.alb(i, j, k) * (c.c1h[k] * f.mu2[x2]) + c.rdnw[k] * (f.ph2[x + g.sk] - f.ph2[x]));
  f.al[x] = al;
  const double base = kRd * (kT0 + f.t2[x]) / (kP1000 * (al + f.alb(i, j, k)));
  f.p[x] = kP1000 * mi_libm::pow(base, kCpovcv) - f.pb(i, j, k);
}

// finiteness of the prognostic fields (flags[fld] = 1 when a NaN or an infinity is found)
__global__ void k_finite(Range r, Args a, int* flags) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const int64_t x = g.at(i, j, k);
  auto bad = [](double v) { return !(v - v == 0.0); };
  if (i < g.nx + 1 && j < g.ny && k < g.nz && bad(f.u2[x])) flags[MI_U] = 1;
  if (i < g.nx && j < g.nyv && k < g.nz && bad(f.v2[x])) flags[MI_V] = 1;
  if (i < g.nx && j < g.ny) {
    if (bad(f.w2[x])) flags[MI_W] = 1;
    if (bad(f.ph2[x])) flags[MI_PH] = 1;
    if (k < g.nz && bad(f.t2[x])) flags[MI_T] = 1;
    if (k == 0 && bad(f.mu2[g.at2(i, j)])) flags[MI_MU] = 1;
  }
}

}  // namespace

// ============================================================================= launchers
void launch_halo(const Args& a, const HaloSet& s, hipStream_t st) {
  if (s.n == 0) return;
  const Geo& g = a.g;
  int nymax = 0, nxmax = 0;
  for (int f = 0; f < s.n; ++f) {
    nymax = s.nys[f] > nymax ? s.nys[f] : nymax;
    nxmax = s.nxs[f] > nxmax ? s.nxs[f] : nxmax;
  }
  const unsigned nfk = unsigned(s.n * s.nkmax);
  hipLaunchKernelGGL(k_halo_x, dim3(unsigned((nymax + 31) / 32), nfk, 1), dim3(256, 1, 1), 0, st, g, s);
  if (!g.two_d) {
    const int ni = nxmax + 2 * kHalo;
    hipLaunchKernelGGL(k_halo_y, dim3(unsigned((ni + 63) / 64), nfk, 1), dim3(64, 8, 1), 0, st, g, s);
  }
}

void launch_copy_step(const Args& a, int64_t n3, int64_t n2, hipStream_t st) {
  // offsets of point (0,0,0) inside the allocations: identical for every field of the domain
  const int64_t off3 = int64_t(kLpad) + int64_t(a.g.hy) * a.g.sj;
  hipLaunchKernelGGL(k_copy_step, dim3(unsigned((n3 + 255) / 256), 6, 1), dim3(256, 1, 1), 0, st, a.f, nullptr, nullptr, n3,
                     n2, off3, off3, off3, off3, off3, off3);
}

void launch_fill(double* p, int64_t n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(unsigned((n + 255) / 256), 1, 1), dim3(256, 1, 1), 0, st, p, n, v);
}

void launch_rk_prep_mu(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_mu_uv, Range{0, g.nx + 1, 0, g.nyv, 0, 1}, st, a);
}
void launch_rk_prep_couple(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_couple, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a);
}
void launch_rk_prep_ww(const Args& a, hipStream_t st) { launch_cols(k_ww, a.g.nx, a.g.ny, st, a); }

void launch_tendencies(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  const int ui0 = g.bcx != MI_BC_PERIODIC ? 1 : 0;
  const int vj0 = (g.bcy != MI_BC_PERIODIC && !g.two_d) ? 1 : 0;
  launch3(k_tend_u, Range{ui0, g.nx, 0, g.ny, 0, g.nz}, st, a);
  if (!g.two_d) launch3(k_tend_v, Range{0, g.nx, vj0, g.ny, 0, g.nz}, st, a);
  launch3(k_tend_w, Range{0, g.nx, 0, g.ny, 1, g.nz + 1}, st, a);
  launch3(k_tend_t, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a);
  launch3(k_tend_ph, Range{0, g.nx, 0, g.ny, 1, g.nz + 1}, st, a);
}

void launch_ss_prep(const Args& a, bool first, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_ss_prep, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a, first ? 1 : 0);
}
void launch_p_rho(const Args& a, bool damp, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_p_rho, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, damp ? 1 : 0);
}
void launch_coef_w(const Args& a, double cof, hipStream_t st) { launch_cols(k_coef_w, a.g.nx, a.g.ny, st, a, cof); }
void launch_uv(const Args& a, double dts, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_uv, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, dts);
}
void launch_mu_t(const Args& a, double dts, int it, int ns, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_mut_div, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a);
  launch_cols(k_mut_col, g.nx, g.ny, st, a, dts, it, ns);
  launch3(k_mut_point, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, dts, it, ns);
}
void launch_w(const Args& a, double dts, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_w_rhs, Range{0, g.nx, 0, g.ny, 0, g.nz + 1}, st, a, dts);
  launch3(k_w_expl, Range{0, g.nx, 0, g.ny, 0, g.nz + 1}, st, a, dts);
  launch_cols(k_w_solve, g.nx, g.ny, st, a, dts);
  launch3(k_p_rho, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, 1);
}
void launch_finish(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_finish, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a);
  launch3(k_finish_mu, Range{0, g.nx + 1, 0, g.nyv, 0, 1}, st, a);
}
void launch_p_rho_phi(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_p_rho_phi, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a);
}
void launch_finite(const Args& a, int* flags, hipStream_t st) {
  const Geo& g = a.g;
  launch3(k_finite, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a, flags);
}

}  // namespace mi_fp64
