// CPU oracle of the RK3 split-explicit dynamics step -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Each function below restates one kernel of the reference CPU FP64 build and cites it.  The
// expressions keep the reference's operand order token for token: the FP64 mode of the HIP library
// must agree with this file bit for bit, and both must agree with the reference.  Compiled with
// -ffp-contract=off (no fused multiply-add), as the reference (CMakeLists.txt:33-37).
//
// Scope: the dynamics step on flat or terrain-following grids (hybrid or sigma coordinate) with map factors,
// Coriolis and periodic, symmetric, open or specified lateral boundaries; h_adv_order 5/6, v_adv_order 3/4,
// damp_opt 0..3, w_damping 0/1, top_lid 0/1, hypsometric_opt 1; moist species coupled through calc_cq and
// theta_m, the transport of moist species and passive tracers (advect_scalar, advect_scalar_pd) with the mass
// fluxes of the acoustic steps; the sixth-order and the diff_opt 1 diffusion (km_opt 1 and 4); Kessler
// microphysics after the RK loop; the column physics hooks.
#include "oracle.h"

#include <emmintrin.h>
#include <omp.h>

#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" __m128d _ZGVbN2vv_pow(__m128d, __m128d);
extern "C" __m128d _ZGVbN2v_exp(__m128d);
extern "C" __m128d _ZGVbN2v_sin(__m128d);
extern "C" __m128d _ZGVbN2v_cos(__m128d);

namespace {

thread_local std::string g_err;

constexpr double kG = 9.81;                   // constants.hpp:5
constexpr double kRd = 287.0;                 // constants.hpp:6
constexpr double kCp = 7.0 * kRd / 2.0;       // constants.hpp:7
constexpr double kCpovcv = kCp / (kCp - kRd); // constants.hpp:11
constexpr double kP1000 = 100000.0;           // constants.hpp:12
constexpr double kT0 = 300.0;                 // constants.hpp:13
constexpr double kRcp = kRd / kCp;            // constants.hpp
constexpr double kRvovrd = 461.6 / kRd;       // r_v / r_d, constants.hpp
constexpr double H = 0.5, Q = 0.25;
constexpr int kMaxLev = 256;                  // levels of a column's work arrays
constexpr double kReradius = 1. / 6370.0e03;  // large_step.cpp:29

enum class St { M, U, V, W };

// A field with the reference's allocation semantics: halo 3 in x, 3 in y (0 for an x-z slice, whose
// j stride is 0), staggering as src/core/grid.hpp:3-7 (field.hpp:44-70, halo.cpp:12-48).
struct F {
  int nxs = 0, nys = 0, nzs = 0, hy = 0;
  long sj = 0, sk = 0, off = 0;
  bool face_x = false, face_y = false;
  std::vector<double> a;
  // extra_k: levels beyond the staggering's own (the diffusion's zx, zy carry nz + 1 at u and v points)
  void init(int nx, int ny, int nz, St s, bool rank3, int extra_k = 0) {
    const bool two_d = ny == 1;
    nxs = nx + (s == St::U ? 1 : 0);
    nys = ny + ((s == St::V && !two_d) ? 1 : 0);
    nzs = rank3 ? nz + (s == St::W ? 1 : 0) + extra_k : 1;
    hy = two_d ? 0 : 3;
    face_x = s == St::U;
    face_y = s == St::V && !two_d;
    const long pitch = nxs + 6;
    sj = two_d ? 0 : pitch;
    const long plane = pitch * (nys + 2 * hy);
    sk = rank3 ? plane : 0;
    off = 3 + hy * sj;
    a.assign(size_t(plane) * size_t(nzs), 0.0);
  }
  double& operator()(int i, int j, int k) { return a[size_t(off + i + j * sj + k * sk)]; }
  double operator()(int i, int j, int k) const { return a[size_t(off + i + j * sj + k * sk)]; }
  double& operator()(int i, int j) { return a[size_t(off + i + j * sj)]; }
  double operator()(int i, int j) const { return a[size_t(off + i + j * sj)]; }
};

enum class Par { Even, Odd, OddFace };

struct Grid {
  int nx, ny, nz, bcx, bcy;
  double dx, dy, dt;
  bool two_d() const { return ny == 1; }
  int nyv() const { return two_d() ? ny : ny + 1; }
  int u_i0() const { return bcx != MI_BC_PERIODIC ? 1 : 0; }  // grid.hpp:41
  int v_j0() const { return (bcy != MI_BC_PERIODIC && !two_d()) ? 1 : 0; }  // grid.hpp:43
};

// One halo line (halo.cpp:52-84): slot a of the direction with n mass points.
inline void halo_point(double* base, long sn, int a, int n, bool face, bool periodic, bool odd, bool odd_face,
                       bool flat) {
  const int nl = face ? 4 : 3;
  const int i = a < nl ? (face ? -a : -1 - a) : n + (a - nl);
  int src;
  if (flat) {  // specified or open: the outermost row copied outward (halo.cpp:67-69)
    const int last = face ? n : n - 1;
    src = i < 0 ? 0 : (i > last ? last : i);
    base[i * sn] = base[src * sn];
    return;
  }
  if (periodic) src = ((i % n) + n) % n;
  else if (face) src = i <= 0 ? -i : 2 * n - i;
  else src = i < 0 ? -1 - i : 2 * n - 1 - i;
  double v = base[src * sn];
  if (!periodic && src == i) {
    if (odd_face) base[i * sn] = -v;
    return;
  }
  if (!periodic && odd) v = -v;
  base[i * sn] = v;
}

// fill_halo (halo.cpp:114-131): x over the interior rows, then y over the full x extent.
void fill(F& f, const Grid& g, Par px, Par py) {
  const int na = f.face_x ? 8 : 6;
  for (int k = 0; k < f.nzs; ++k)
    for (int j = 0; j < f.nys; ++j)
      for (int a = 0; a < na; ++a)
        halo_point(&f(0, j, k), 1, a, g.nx, f.face_x, g.bcx == MI_BC_PERIODIC, px != Par::Even, px == Par::OddFace,
                   g.bcx == MI_BC_SPECIFIED || g.bcx == MI_BC_OPEN);
  if (g.two_d()) return;
  const int nb = f.face_y ? 8 : 6;
  for (int k = 0; k < f.nzs; ++k)
    for (int i = -3; i < f.nxs + 3; ++i)
      for (int a = 0; a < nb; ++a)
        halo_point(&f(i, 0, k), f.sj, a, g.ny, f.face_y, g.bcy == MI_BC_PERIODIC, py != Par::Even, py == Par::OddFace,
                   g.bcy == MI_BC_SPECIFIED || g.bcy == MI_BC_OPEN);
}
// default parities (halo.cpp:224-228): U odd in x, V odd in y, everything else even
void fill(F& f, const Grid& g) { fill(f, g, f.face_x ? Par::Odd : Par::Even, f.face_y ? Par::Odd : Par::Even); }

template <class K>
void pfor(int i0, int i1, int j0, int j1, int k0, int k1, K k) {  // backend.hpp:509-519
#pragma omp parallel for collapse(2) schedule(static)
  for (int kk = k0; kk < k1; ++kk)
    for (int j = j0; j < j1; ++j)
      for (int i = i0; i < i1; ++i) k(i, j, kk);
}
template <class K>
void pfor2(int i0, int i1, int j0, int j1, K k) {  // column_for over a range (backend.hpp:520-528)
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = j0; j < j1; ++j)
    for (int i = i0; i < i1; ++i) k(i, j);
}
template <class K>
void cfor(int nx, int ny, K k) {  // backend.hpp:520-528
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) k(i, j);
}

inline double sgn1(double a) { return a >= 0.0 ? 1.0 : -1.0; }  // kernels_common.hpp:11
// kernels_common.hpp:13-17
inline double flux5(double qm3, double qm2, double qm1, double q0, double qp1, double qp2, double ua, double up) {
  return (37.0 * (q0 + qm1) - 8.0 * (qp1 + qm2) + (qp2 + qm3)) / 60.0 -
         up * sgn1(ua) * ((qp2 - qm3) - 5.0 * (qp1 - qm2) + 10.0 * (q0 - qm1)) / 60.0;
}
// kernels_common.hpp:19-22
inline double flux3(double qm2, double qm1, double q0, double qp1, double ua, double up) {
  return (7.0 * (q0 + qm1) - (qp1 + qm2)) / 12.0 + up * sgn1(ua) * ((qp1 - qm2) - 3.0 * (q0 - qm1)) / 12.0;
}
// advect_scalar_pd's statement functions with folded constants (kernels_common.hpp:23-33)
inline double flux5_pd(double qm3, double qm2, double qm1, double q0, double qp1, double qp2, double ua, double up) {
  const double f6 = (37. / 60.) * (q0 + qm1) - (2. / 15.) * (qp1 + qm2) + (1. / 60.) * (qp2 + qm3);
  return f6 - up * sgn1(ua) * (1. / 60.) * ((qp2 - qm3) - 5. * (qp1 - qm2) + 10. * (q0 - qm1));
}
inline double flux3_pd(double qm2, double qm1, double q0, double qp1, double ua, double up) {
  const double f4 = (7. / 12.) * (q0 + qm1) - (1. / 12.) * (qp1 + qm2);
  return f4 + up * sgn1(ua) * (1. / 12.) * ((qp1 - qm2) - 3. * (q0 - qm1));
}
// first-order upwind flux with Courant number cr (kernels_common.hpp:83-88)
inline double flux_upwind(double qm1, double q0, double cr) {
  const double a = cr < 0.0 ? -cr : cr;
  const double lo = cr + a, hi = cr - a;
  return 0.5 * (lo < 1.0 ? lo : 1.0) * qm1 + 0.5 * (hi > -1.0 ? hi : -1.0) * q0;
}

// kern::ord_face / ord_mass (kernels_common.hpp:45-58): the horizontal advection order next to a
// specified edge, for a flux between mass points f-1 and f (ord_face) or at mass point f-1 of a field
// staggered in that direction (ord_mass); 5 stands for the domain's own order
inline int ord_face(int f, int n, bool on) {
  if (!on) return 5;
  if (f <= 1 || f >= n - 1) return 2;
  if (f == 2 || f == n - 2) return 3;
  return 5;
}
// open boundaries (kernels_common.hpp:60-81): the outermost column (row) of a scalar takes, instead of the flux
// divergence, rdx*(ub*(q_old(inner)-q_old(outer)) + q*(ru(b)-ru(a))) with the inflow mean normal mass flux ub
template <class Q, class R>
inline double open_x_scalar(const Q& q, const Q& qo, const R& ru, double rdx, int i, int j, int k, int nx) {
  if (i == 0) {
    double ub = 0.5 * (ru(0, j, k) + ru(1, j, k));
    ub = ub < 0.0 ? ub : 0.0;
    return rdx * (ub * (qo(1, j, k) - qo(0, j, k)) + q(0, j, k) * (ru(1, j, k) - ru(0, j, k)));
  }
  double ub = 0.5 * (ru(nx - 1, j, k) + ru(nx, j, k));
  ub = ub > 0.0 ? ub : 0.0;
  return rdx * (ub * (qo(nx - 1, j, k) - qo(nx - 2, j, k)) + q(nx - 1, j, k) * (ru(nx, j, k) - ru(nx - 1, j, k)));
}
template <class Q, class R>
inline double open_y_scalar(const Q& q, const Q& qo, const R& rv, double rdy, int i, int j, int k, int ny) {
  if (j == 0) {
    double vb = 0.5 * (rv(i, 0, k) + rv(i, 1, k));
    vb = vb < 0.0 ? vb : 0.0;
    return rdy * (vb * (qo(i, 1, k) - qo(i, 0, k)) + q(i, 0, k) * (rv(i, 1, k) - rv(i, 0, k)));
  }
  double vb = 0.5 * (rv(i, ny - 1, k) + rv(i, ny, k));
  vb = vb > 0.0 ? vb : 0.0;
  return rdy * (vb * (qo(i, ny - 1, k) - qo(i, ny - 2, k)) + q(i, ny - 1, k) * (rv(i, ny, k) - rv(i, ny - 1, k)));
}
constexpr double kCb = 25.;  // phase speed of the open (radiative) boundaries (constants.hpp:19)

inline int ord_mass(int f, int n, bool on) {
  if (!on) return 5;
  if (f <= 1 || f >= n) return 2;
  if (f == 2 || f == n - 1) return 3;
  return 5;
}

struct Lev {
  std::vector<double> dnw, rdnw, dn, rdn, fnm, fnp, c1h, c2h, c3h, c4h, c1f, c2f, c3f, c4f;
  double cf1, cf2, cf3, cfn, cfn1, rdx, rdy, p_top;
};

struct Dom {
  Grid g;
  mi_params p;
  Lev c;
  int n_sound = 0;
  bool libmvec = false;
  // State (state.hpp:30-78)
  F u1, u2, v1, v2, w1, w2, t1, t2, ph1, ph2, mu1, mu2, pp, al, alt, php, ww;
  F mut, muu, muv, muts, muus, muvs, ru, rv, rw, ru_m, rv_m, ww_m;
  // Tendencies (state.hpp:80-90)
  F ru_tend, rv_tend, rw_tend, t_tend, ph_tend, mu_tend;
  // SmallStep work fields (small_step.hpp:43-49)
  F u_save, v_save, w_save, t_save, ph_save, ww1, c2a, pm1, t_2save, ca, calpha, cgamma, mu_save, mudf, muave, w_rhs,
      t2a, dvdxi;
  // BaseState (state.hpp:92-119), stored in full with halos
  F mub, pb, alb, t_init, phb, msftx, msfty, msfux, msfuy, msfvx, msfvy, msfvx_inv;
  // scalars (state.hpp:53-76): q_1, q_2, q_tendf, calc_cq coefficients, h_diabatic
  int n_moist = 0, n_sc = 0;
  std::vector<F> q1, q2, qtf, qnew;
  F cqu, cqv, cqw, hdia;
  // advect_scalar_pd work (scalars.cpp:295-298)
  F sc_scale, dhx, lox, dhy, loy, dhz, loz;
  // Coriolis and grid rotation (state.hpp:105-107); held tendencies of the first stage (state.hpp:85-88)
  F fcor, ecor, sina, cosa, ht;
  F ru_tendf, rv_tendf, rw_tendf, t_tendf, ph_tendf;
  std::vector<F> qtw;  // q_tendf of the step: the caller's tendency plus the sixth-order diffusion
  bool mapped = false, cor = false, use_tendf = false, terrain = false;
  // the diffusion's 1-D base profiles (BaseState::u_base, v_base, qv_base; state.hpp:100, 112) and its horizontal
  // eddy viscosities at mass points: km_opt 1 the constants of KIsotropicKm (km_kh.hpp:35-45: xkmh = khdif, xkhh =
  // khdif / prandtl), km_opt 4 those of KSmag2dKm, set each step from the metrics zx, zy, rdzw and the horizontal
  // deformations defor11, defor22, defor12 (DiffusionDriver::prepare, diffusion_driver.cpp:92-138)
  std::vector<double> u_base, v_base, qv_base, t_base, z_base;
  double kvdif = 0, kvdq = 0, prandtl = 1.0 / 3.0, c_s = 0.25;
  F xkmh, xkhh, zx, zy, rdz, rdzw, d11, d22, d12;
  bool karr = false;  // K set each step (km_opt 4, or damp_opt 1 raising the constants)
  // microphysics after the RK loop (PhysicsCoupling, physics_coupling.hpp:52-94): moist_physics_prep's fields and
  // the precipitation
  int mp = 0;
  double mp_lim = 10.0;
  F th_phy, mrho, pii, zmid, dz8w, qvd, qcd, p_phy, rainnc, rainncv;
  // a column physics plugin (ColumnPhysics, column_physics.hpp:21-34; RK3::set_column_physics): its hooks and
  // the advective forcings it asks for (RTHFTEN, RQVFTEN)
  orc_column_physics phys{};
  bool has_phys = false, held_base = false;
  F thf, qvf;
  bool want_thf = false, want_qvf = false;
  // specified lateral boundaries (LateralBC, lateral_bc.hpp:28-91): spec_zone 1, relax_zone 4
  static constexpr int kBw = 5, kSpecZone = 1, kRelaxZone = 4;
  bool spec = false;
  // open boundaries (WRF open_xs/xe, open_ys/ye; grid.hpp:17-19, kernels_common.hpp:33-44): ox, oy; flx, fly:
  // the degraded advection orders of a flat (specified or open) direction (kern::Spec x, y)
  bool ox = false, oy = false, flx = false, fly = false;
  double fcx[kBw] = {}, gcx[kBw] = {}, dtbc = 0;
  // one variable's boundary data: values and tendencies of the sides xs, xe, ys, ye, (pos, k, b) pos fastest
  struct Bdy {
    bool have = false;
    int nk = 0, npos[4] = {};
    std::vector<double> v[4], t[4];
  };
  Bdy bdy[5];             // MI_BDY_U, V, T, PH, MU
  std::vector<Bdy> qbdy;  // scalars

  explicit Dom(const mi_grid& gg, const mi_params& pp_, const mi_levels& lv) {
    g = Grid{gg.nx, gg.ny, gg.nz, gg.bc_x, gg.bc_y, gg.dx, gg.dy, gg.dt};
    p = pp_;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    auto v1d = [](const double* a, int n) { return std::vector<double>(a, a + n); };
    c.dnw = v1d(lv.dnw, nz); c.rdnw = v1d(lv.rdnw, nz);
    c.dn = v1d(lv.dn, nz + 1); c.rdn = v1d(lv.rdn, nz + 1); c.fnm = v1d(lv.fnm, nz + 1); c.fnp = v1d(lv.fnp, nz + 1);
    c.c1h = v1d(lv.c1h, nz); c.c2h = v1d(lv.c2h, nz); c.c3h = v1d(lv.c3h, nz); c.c4h = v1d(lv.c4h, nz);
    c.c1f = v1d(lv.c1f, nz + 1); c.c2f = v1d(lv.c2f, nz + 1); c.c3f = v1d(lv.c3f, nz + 1); c.c4f = v1d(lv.c4f, nz + 1);
    c.cf1 = lv.cf1; c.cf2 = lv.cf2; c.cf3 = lv.cf3; c.cfn = lv.cfn; c.cfn1 = lv.cfn1;
    c.rdx = lv.rdx; c.rdy = lv.rdy; c.p_top = lv.p_top;
    n_sound = p.time_step_sound > 0 ? p.time_step_sound : mi_auto_sound_steps_(g.dt, g.dx, g.dy);
    auto f3 = [&](F& f, St s) { f.init(nx, ny, nz, s, true); };
    auto f2 = [&](F& f, St s) { f.init(nx, ny, nz, s, false); };
    for (F* f : {&u1, &u2, &ru, &ru_m, &ru_tend, &u_save}) f3(*f, St::U);
    for (F* f : {&v1, &v2, &rv, &rv_m, &rv_tend, &v_save}) f3(*f, St::V);
    for (F* f : {&w1, &w2, &ph1, &ph2, &ww, &rw, &ww_m, &rw_tend, &ph_tend, &w_save, &ph_save, &ww1, &ca, &calpha,
                 &cgamma, &w_rhs, &phb})
      f3(*f, St::W);
    for (F* f : {&t1, &t2, &pp, &al, &alt, &php, &t_tend, &t_save, &c2a, &pm1, &t_2save, &t2a, &dvdxi, &pb, &alb, &t_init})
      f3(*f, St::M);
    for (F* f : {&mu1, &mu2, &mut, &muts, &mu_tend, &mu_save, &mudf, &muave, &mub, &msftx, &msfty}) f2(*f, St::M);
    for (F* f : {&muu, &muus, &msfux, &msfuy}) f2(*f, St::U);
    for (F* f : {&muv, &muvs, &msfvx, &msfvy, &msfvx_inv}) f2(*f, St::V);
    for (F* f : {&msftx, &msfty, &msfux, &msfuy, &msfvx, &msfvy, &msfvx_inv}) std::fill(f->a.begin(), f->a.end(), 1.0);
    n_moist = p.n_moist;
    n_sc = p.n_scalars;
    for (auto* v : {&q1, &q2, &qtf, &qnew}) {
      v->resize(size_t(n_sc));
      for (F& f : *v) f3(f, St::M);
    }
    f3(cqu, St::U); f3(cqv, St::V); f3(cqw, St::W); f3(hdia, St::M);
    f3(sc_scale, St::M); f3(dhx, St::U); f3(lox, St::U); f3(dhy, St::V); f3(loy, St::V); f3(dhz, St::W); f3(loz, St::W);
    for (F* f : {&fcor, &ecor, &sina, &cosa, &ht}) f2(*f, St::M);
    std::fill(cosa.a.begin(), cosa.a.end(), 1.0);
    f3(ru_tendf, St::U); f3(rv_tendf, St::V); f3(rw_tendf, St::W); f3(t_tendf, St::M); f3(ph_tendf, St::W);
    qtw.resize(size_t(n_sc));
    for (F& f : qtw) f3(f, St::M);
    mapped = p.mapped != 0;
    cor = mapped || p.coriolis != 0;  // MapC::from(b, b.mapped || b.coriolis) (large_step.cpp:810)
    spec = g.bcx == MI_BC_SPECIFIED;  // with bcy the same (checked at creation)
    ox = g.bcx == MI_BC_OPEN;
    oy = g.bcy == MI_BC_OPEN && !g.two_d();
    flx = spec || ox;
    fly = (spec || g.bcy == MI_BC_OPEN) && !g.two_d();
    u_base.assign(size_t(nz), 0.0);
    v_base.assign(size_t(nz), 0.0);
    qv_base.assign(size_t(nz), 0.0);
    if (p.diff_opt == 1) {  // DiffusionDriver's params (model.cpp:314: prandtl = 1 / pr_inv)
      const double pr_inv = p.pr_inv > 0 ? p.pr_inv : 3.0;
      prandtl = 1.0 / pr_inv;
      kvdif = p.kvdif;
      kvdq = p.kvdif / prandtl;
      c_s = p.c_s > 0 ? p.c_s : 0.25;
      f3(xkmh, St::M);
      f3(xkhh, St::M);
      karr = p.km_opt == 4 || p.damp_opt == 1;
      if (!karr) {  // the constants everywhere, halos included
        std::fill(xkmh.a.begin(), xkmh.a.end(), p.khdif);
        std::fill(xkhh.a.begin(), xkhh.a.end(), p.khdif / prandtl);
      } else {
        zx.init(nx, ny, nz, St::U, true, 1);
        zy.init(nx, ny, nz, St::V, true, 1);
        for (F* f : {&rdz, &rdzw, &d11, &d22}) f3(*f, St::M);
        f3(d12, St::U);  // (nx+1) x (ny+1): row ny lives in the y halo (diffusion_driver.cpp:30)
      }
    }
    t_base.assign(size_t(nz), 0.0);
    z_base.assign(size_t(nz), 0.0);
    // a DiffusionDriver (DiffusionConfig::on(), diffusion_driver.hpp:50) or a LateralBC (large_step.cpp:828)
    use_tendf = p.diff_6th_opt != 0 || p.diff_opt != 0 || p.damp_opt == 2 || spec;
    mp = p.mp_physics;
    mp_lim = p.mp_tend_lim > 0 ? p.mp_tend_lim : 10.0;
    if (mp) {
      for (F* f : {&th_phy, &mrho, &pii, &zmid, &dz8w, &qvd, &qcd, &p_phy}) f3(*f, St::M);
      f2(rainnc, St::M);
      f2(rainncv, St::M);
    }
    held_base = use_tendf;
    f3(thf, St::M);
    f3(qvf, St::M);
    terrain = p.terrain != 0;          // !b.ht.uniform() (small_step.cpp:583, rk3.cpp:163)
    if (spec) {
      set_fcx();
      const int nkv[5] = {nz, nz, nz, nz + 1, 1};  // U V T PH MU (lateral_bc.cpp:219)
      for (int v = 0; v < 5; ++v) init_bdy(bdy[v], nkv[v], v == 0, v == 1);
      qbdy.resize(size_t(n_sc));
      for (Bdy& b : qbdy) init_bdy(b, nz, false, false);
    }
  }
  // lbc_fcx_gcx (LateralBC::set_dt, lateral_bc.cpp:198-211), Fortran loop = b + 1 over spec_zone + 1 ..
  // spec_zone + relax_zone
  void set_fcx() {
    for (int loop = kSpecZone + 1; loop <= kSpecZone + kRelaxZone; ++loop) {
      const double num = double(kSpecZone + kRelaxZone - loop), den = double(kRelaxZone - 1);
      fcx[loop - 1] = 0.1 / g.dt * num / den;
      gcx[loop - 1] = 1.0 / g.dt / 50.0 * num / den;
    }
  }
  // RK3::set_dt (rk3.cpp:79-84) with LateralBC::set_dt and DiffusionDriver::set_dt (every user of the grid's dt
  // reads it at its use)
  void set_dt(double dt, int ns) {
    g.dt = dt;
    n_sound = ns;
    if (spec) set_fcx();
  }
  void init_bdy(Bdy& b, int nk, bool u, bool v) {  // lateral_bc.cpp:225-233
    b.nk = nk;
    b.npos[0] = b.npos[1] = g.ny + (v ? 1 : 0);  // x sides: pos = j
    b.npos[2] = b.npos[3] = g.nx + (u ? 1 : 0);  // y sides: pos = i
    for (int s_ = 0; s_ < 4; ++s_) {
      b.v[s_].assign(size_t(b.npos[s_]) * size_t(nk) * kBw, 0.0);
      b.t[s_].assign(size_t(b.npos[s_]) * size_t(nk) * kBw, 0.0);
    }
  }
  // Slab / SlabSet (lateral_bc.cpp:21-32): value + dtbc * tendency of side slab sl
  size_t bidx(const Bdy& B, int sl, int pos, int k, int b) const {
    return size_t(pos) + size_t(k) * size_t(B.npos[sl]) + size_t(b) * size_t(B.npos[sl]) * size_t(B.nk);
  }
  double bval(const Bdy& B, int sl, int pos, int k, int b) const {
    const size_t o = bidx(B, sl, pos, k, b);
    return B.v[sl][o] + dtbc * B.t[sl][o];
  }
  static int slab_of_side(int side) { return side < 2 ? side + 2 : side - 2; }  // S,N,W,E -> ys,ye,xs,xe
  // Ring (lateral_bc.cpp:34-58): rows b0 .. b0+nb-1 of the sides S, N (i = b .. ibe-b) and W, E
  // (j = b+1 .. jbe-b-1); ibe, jbe the last index of the staggering ('u': nx, 'v': ny in y)
  template <class K>
  void ring(char var, int b0, int nb, int nk, K kern) {
    const int ibe = var == 'u' ? g.nx : g.nx - 1, jbe = var == 'v' ? g.ny : g.ny - 1;
    for (int side = 0; side < 4; ++side)
      for (int b = b0; b < b0 + nb; ++b)
        for (int k = 0; k < nk; ++k) {
          if (side < 2) {
            for (int i = b; i <= ibe - b; ++i) kern(i, side, b, i, side == 0 ? b : jbe - b, k);
          } else {
            for (int j = b + 1; j <= jbe - b - 1; ++j) kern(j, side, b, side == 2 ? b : ibe - b, j, k);
          }
        }
  }
  // KRelax (lateral_bc.cpp:82-121): fls0 .. fls4 of row b (1..3) against the model field fm as the data
  // holds it; out(i, j, k, b, fls0, fls1 + fls2 + fls3 + fls4 - 4 fls0)
  template <class Fm, class Out>
  void relax(char var, int nk, const Bdy& B, const Fm& fm, const Out& out) {
    const int ibe = var == 'u' ? g.nx : g.nx - 1;
    ring(var, kSpecZone, kRelaxZone - 1, nk, [&](int pos, int side, int b, int i, int j, int k) {
      const int sl = slab_of_side(side);
      auto bv = [&](int p_, int bb) { return bval(B, sl, p_, k, bb); };
      double fls0, fls1, fls2, fls3, fls4;
      fls0 = bv(pos, b) - fm(i, j, k);
      if (side <= 1) {
        const int im1 = i - 1 > 0 ? i - 1 : 0, ip1 = i + 1 < ibe ? i + 1 : ibe;
        const int jin = side == 0 ? j + 1 : j - 1, jout = side == 0 ? j - 1 : j + 1;
        fls1 = bv(im1, b) - fm(im1, j, k);
        fls2 = bv(ip1, b) - fm(ip1, j, k);
        fls3 = bv(i, b - 1) - fm(i, jout, k);
        fls4 = bv(i, b + 1) - fm(i, jin, k);
      } else {
        const int iin = side == 2 ? i + 1 : i - 1, iout = side == 2 ? i - 1 : i + 1;
        fls1 = bv(j - 1, b) - fm(i, j - 1, k);
        fls2 = bv(j + 1, b) - fm(i, j + 1, k);
        fls3 = bv(j, b - 1) - fm(iout, j, k);
        fls4 = bv(j, b + 1) - fm(iin, j, k);
      }
      out(i, j, k, b, fls0, fls1 + fls2 + fls3 + fls4 - 4.0 * fls0);
    });
  }
  // LateralBC::relax_dry (lateral_bc.cpp:292-338): into the held tendencies (u, v times msfuy, msfvx)
  void relax_dry() {
    const int nz = g.nz;
    relax('u', nz, bdy[MI_BDY_U], [&](int i, int j, int k) { return ru(i, j, k); },
          [&](int i, int j, int k, int b, double f0, double sm) { ru_tendf(i, j, k) += (fcx[b] * f0 - gcx[b] * sm) * msfuy(i, j); });
    relax('v', nz, bdy[MI_BDY_V], [&](int i, int j, int k) { return rv(i, j, k); },
          [&](int i, int j, int k, int b, double f0, double sm) { rv_tendf(i, j, k) += (fcx[b] * f0 - gcx[b] * sm) * msfvx(i, j); });
    relax('h', nz + 1, bdy[MI_BDY_PH], [&](int i, int j, int k) { return ph2(i, j, k) * (c.c1f[k] * mut(i, j) + c.c2f[k]); },
          [&](int i, int j, int k, int b, double f0, double sm) { ph_tendf(i, j, k) += fcx[b] * f0 - gcx[b] * sm; });
    relax('t', nz, bdy[MI_BDY_T], [&](int i, int j, int k) { return t2(i, j, k) * (c.c1h[k] * mut(i, j) + c.c2h[k]); },
          [&](int i, int j, int k, int b, double f0, double sm) { t_tendf(i, j, k) += fcx[b] * f0 - gcx[b] * sm; });
    relax('m', 1, bdy[MI_BDY_MU], [&](int i, int j, int) { return mu2(i, j); },
          [&](int i, int j, int, int b, double f0, double sm) { mu_tend(i, j) += fcx[b] * f0 - gcx[b] * sm; });
  }
  // KSpecTend (lateral_bc.cpp:67-79): the spec zone's tendency = the data's
  void spec_tend(char var, int nk, const Bdy& B, F& out) {
    ring(var, 0, kSpecZone, nk, [&](int pos, int side, int b, int i, int j, int k) {
      const int sl = slab_of_side(side);
      out(i, j, k) = B.t[sl][bidx(B, sl, pos, k, b)];
    });
  }
  void spec_dry() {  // LateralBC::spec_dry (lateral_bc.cpp:340-354)
    spec_tend('u', g.nz, bdy[MI_BDY_U], ru_tend);
    spec_tend('v', g.nz, bdy[MI_BDY_V], rv_tend);
    spec_tend('h', g.nz + 1, bdy[MI_BDY_PH], ph_tend);
    spec_tend('t', g.nz, bdy[MI_BDY_T], t_tend);
    spec_tend('m', 1, bdy[MI_BDY_MU], mu_tend);
  }
  // KUpdate (lateral_bc.cpp:123-131): field += dts * tend in the spec zone
  void spec_update(char var, int nk, F& f, const F& t, double dts) {
    ring(var, 0, kSpecZone, nk, [&](int, int, int, int i, int j, int k) { f(i, j, k) += dts * t(i, j, k); });
  }
  // LateralBC::update_ph_w (lateral_bc.cpp:374-387): KUpdatePh (:133-144) and zero_grad_bdy on w (:146-161)
  void update_ph_w(double dts) {
    const int ibe = g.nx - 1, jbe = g.ny - 1;
    ring('h', 0, kSpecZone, g.nz + 1, [&](int, int side, int, int i, int j, int k) {
      const double mu_old = muts(i, j) - dts * mu_tend(i, j);
      const double m_old = c.c1f[k] * mu_old + c.c2f[k], m_new = c.c1f[k] * muts(i, j) + c.c2f[k];
      ph2(i, j, k) = ph2(i, j, k) * m_old / m_new + dts * ph_tend(i, j, k) / m_new + ph_save(i, j, k) * (m_old / m_new - 1.0);
      const int ic = i < 1 ? 1 : (i > ibe - 1 ? ibe - 1 : i);
      const int jc = j < 1 ? 1 : (j > jbe - 1 ? jbe - 1 : j);
      switch (side) {
        case 0: w2(i, j, k) = w2(ic, 1, k); break;
        case 1: w2(i, j, k) = w2(ic, jbe - 1, k); break;
        case 2: w2(i, j, k) = w2(1, jc, k); break;
        default: w2(i, j, k) = w2(ibe - 1, jc, k); break;
      }
    });
  }
  // KFlowDep flow_dep_bdy (lateral_bc.cpp:163-178): the spec zone of a scalar without data
  void flow_dep(F& f) {
    const int ibe = g.nx - 1, jbe = g.ny - 1;
    ring('q', 0, kSpecZone, g.nz, [&](int, int side, int, int i, int j, int k) {
      const int ic = i < 1 ? 1 : (i > ibe - 1 ? ibe - 1 : i);
      const int jc = j < 1 ? 1 : (j > jbe - 1 ? jbe - 1 : j);
      switch (side) {
        case 0: f(i, j, k) = rv_m(i, j, k) < 0.0 ? f(ic, 1, k) : 0.0; break;
        case 1: f(i, j, k) = rv_m(i, j + 1, k) > 0.0 ? f(ic, jbe - 1, k) : 0.0; break;
        case 2: f(i, j, k) = ru_m(i, j, k) < 0.0 ? f(1, jc, k) : 0.0; break;
        default: f(i, j, k) = ru_m(i + 1, j, k) > 0.0 ? f(ibe - 1, jc, k) : 0.0; break;
      }
    });
  }
  // KFinal spec_bdy_final (lateral_bc.cpp:180-194, 420-455): the exact boundary values at t + dt
  template <class Cpl>
  void final_one(char var, int nk, const Bdy& B, F& f, const Cpl& cpl) {
    ring(var, 0, kSpecZone, nk, [&](int pos, int side, int b, int i, int j, int k) {
      f(i, j, k) = cpl(i, j, k, bval(B, slab_of_side(side), pos, k, b));
    });
  }
  void spec_final() {
    const int nz = g.nz;
    final_one('u', nz, bdy[MI_BDY_U], u2,
              [&](int i, int j, int k, double bv) { return msfuy(i, j) * bv / (c.c1h[k] * muus(i, j) + c.c2h[k]); });
    final_one('v', nz, bdy[MI_BDY_V], v2,
              [&](int i, int j, int k, double bv) { return msfvx(i, j) * bv / (c.c1h[k] * muvs(i, j) + c.c2h[k]); });
    final_one('t', nz, bdy[MI_BDY_T], t2, [&](int i, int j, int k, double bv) { return bv / (c.c1h[k] * muts(i, j) + c.c2h[k]); });
    final_one('h', nz + 1, bdy[MI_BDY_PH], ph2,
              [&](int i, int j, int k, double bv) { return bv / (c.c1f[k] * muts(i, j) + c.c2f[k]); });
    final_one('m', 1, bdy[MI_BDY_MU], mu2, [&](int, int, int, double bv) { return bv; });
    for (int n = 0; n < n_sc; ++n)
      if (qbdy[size_t(n)].have)
        final_one('t', nz, qbdy[size_t(n)], q2[size_t(n)],
                  [&](int i, int j, int k, double bv) { return bv / (c.c1h[k] * muts(i, j) + c.c2h[k]); });
  }

  // MapC::vxgm (large_step.cpp:121-124): curvature at mass point (i, j, k) from the uncoupled u, v
  double vxgm(int i, int j, int k) const {
    return H * (u2(i, j, k) + u2(i + 1, j, k)) * (msfvx(i, j + 1) - msfvx(i, j)) * c.rdy -
           H * (v2(i, j, k) + v2(i, j + 1, k)) * (msfuy(i + 1, j) - msfuy(i, j)) * c.rdx;
  }

  // ------------------------------------------------------------------ sixth_order_diffusion
  // (diffusion/sixth_order.hpp:13-210), periodic / symmetric loop limits: u over i = 0..nx, v over
  // j = 0..ny, w levels 1..nz-1, mass points otherwise; coef = factor 0.015625 / (2 dt)
  static double d6(double m3, double m2, double m1, double q0, double p1, double p2) {
    return 10.0 * (q0 - m1) - 5.0 * (p1 - m2) + (p2 - m3);
  }
  template <class Fld, class Mu>
  void sixth(char name, const Fld& fld, F& tend, const Mu& mut_, double dt, const std::vector<double>& c1,
             const std::vector<double>& c2) {
    const int nx = g.nx, ny = g.ny, nz = g.nz, opt = p.diff_6th_opt;
    const double coef = p.diff_6th_factor * 0.015625 / (2.0 * dt);
    auto lim = [&](double dfl, double diff) { return (opt == 2 && dfl * diff <= 0.0) ? 0.0 : dfl; };
    int i0 = 0, j0 = 0, i1 = nx, j1 = ny, k0 = 0;
    if (name == 'u') i1 = nx + 1;
    if (name == 'v') j1 = ny + 1;
    if (name == 'w') k0 = 1;
    if (spec) {  // specified limits (sixth_order.hpp:176-205): three rows in from every side
      i0 = 3; j0 = 3;
      i1 = name == 'u' ? nx - 2 : nx - 3;
      j1 = name == 'v' ? ny - 2 : ny - 3;
    } else if (ox || oy) {  // open sides (sixth_order.hpp:179-204), 'v' with open_xs on all four sides
      if (name == 'u') {
        if (ox) { i0 = 3; i1 = nx - 2; }
        if (oy) { j0 = 3; j1 = ny - 3; }
      } else if (name == 'v') {
        if (ox) { i0 = 3; i1 = nx - 3; j0 = 3; j1 = ny - 2; }
        if (oy) { j0 = 3; j1 = ny - 2; }
      } else {
        if (ox) { i0 = 3; i1 = nx - 3; }
        if (oy) { j0 = 3; j1 = ny - 3; }
      }
      if (g.bcx == MI_BC_PERIODIC) { i0 = 0; i1 = name == 'u' ? nx + 1 : nx; }
    }
    pfor(i0, i1, j0, j1, k0, nz, [&](int i, int j, int k) {
      auto q = [&](int a, int b) { return fld(i + a, j + b, k); };
      const double dxp0 = lim(d6(q(-3, 0), q(-2, 0), q(-1, 0), q(0, 0), q(1, 0), q(2, 0)), q(0, 0) - q(-1, 0));
      const double dxp1 = lim(d6(q(-2, 0), q(-1, 0), q(0, 0), q(1, 0), q(2, 0), q(3, 0)), q(1, 0) - q(0, 0));
      const double dyp0 = lim(d6(q(0, -3), q(0, -2), q(0, -1), q(0, 0), q(0, 1), q(0, 2)), q(0, 0) - q(0, -1));
      const double dyp1 = lim(d6(q(0, -2), q(0, -1), q(0, 0), q(0, 1), q(0, 2), q(0, 3)), q(0, 1) - q(0, 0));
      auto mf = [&](int a, int b) { return c1[k] * mut_(i + a, j + b) + c2[k]; };
      double mx0, mx1, my0, my1, sx, sy;
      if (name == 'u') {
        mx0 = mf(-1, 0);
        mx1 = mf(0, 0);
        my0 = 0.25 * (mf(-1, -1) + mf(0, -1) + mf(-1, 0) + mf(0, 0));
        my1 = 0.25 * (mf(-1, 0) + mf(0, 0) + mf(-1, 1) + mf(0, 1));
        sx = msfux(i, j);
        sy = msfuy(i, j);
      } else if (name == 'v') {
        mx0 = 0.25 * (mf(-1, -1) + mf(0, -1) + mf(-1, 0) + mf(0, 0));
        mx1 = 0.25 * (mf(0, -1) + mf(1, -1) + mf(0, 0) + mf(1, 0));
        my0 = mf(0, -1);
        my1 = mf(0, 0);
        sx = msfvx(i, j);
        sy = msfvy(i, j);
      } else {
        mx0 = 0.5 * (mf(-1, 0) + mf(0, 0));
        mx1 = 0.5 * (mf(0, 0) + mf(1, 0));
        my0 = 0.5 * (mf(0, -1) + mf(0, 0));
        my1 = 0.5 * (mf(0, 0) + mf(0, 1));
        sx = msftx(i, j);
        sy = msfty(i, j);
      }
      const double tx = coef * sx * (mx1 * dxp1 - mx0 * dxp0);
      const double ty = coef * sy * (my1 * dyp1 - my0 * dyp0);
      tend(i, j, k) = tend(i, j, k) + tx + ty;
    });
  }
  // ------------------------------------------------------------------ diff_opt 1 (big_step_diffusion.hpp)
  // horizontal_diffusion (big_step_diffusion.hpp:26-186, 343-389): the second-order fluxes along eta of one field
  // with the eddy viscosity K at mass points, averaged onto each face as the reference; name 'u', 'v', 'w' or 'm'
  // (mass points); loop limits one row in from open and specified sides (a periodic x resets the x limits)
  template <class Fld, class Mu>
  void hdiff(char name, const Fld& q, F& tend, const Mu& mu, const F& K, const std::vector<double>& c1,
             const std::vector<double>& c2) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool fx = ox || spec, fy = oy || spec, do_y = !g.two_d();
    int is = fx ? 1 : 0, js = fy ? 1 : 0, ie, je, k0 = 0;
    if (name == 'u') { ie = fx ? nx : nx + 1; je = fy ? ny - 1 : ny; }
    else if (name == 'v') { ie = fx ? nx - 1 : nx; je = fy ? ny : ny + 1; }
    else { ie = fx ? nx - 1 : nx; je = fy ? ny - 1 : ny; }
    if (name == 'w') k0 = 1;
    if (g.bcx == MI_BC_PERIODIC) { is = 0; ie = name == 'u' ? nx + 1 : nx; }
    const double rdx = c.rdx, rdy = c.rdy;
    pfor(is, ie, js, je, k0, nz, [&](int i, int j, int k) {
      auto m = [&](int a, int b) { return c1[k] * mu(i + a, j + b) + c2[k]; };
      auto k_ = [&](int a, int b, int dk = 0) { return K(i + a, j + b, k + dk); };
      double mkrdxm, mkrdxp, mrdx, mkrdym = 0, mkrdyp = 0, mrdy = 0;
      if (name == 'u') {  // KHorizDiffU1
        mkrdxm = (msftx(i - 1, j) / msfty(i - 1, j)) * m(-1, 0) * k_(-1, 0) * rdx;
        mkrdxp = (msftx(i, j) / msfty(i, j)) * m(0, 0) * k_(0, 0) * rdx;
        mrdx = msfux(i, j) * msfuy(i, j) * rdx;
        if (do_y) {
          mkrdym = ((msfuy(i, j) + msfuy(i, j - 1)) / (msfux(i, j) + msfux(i, j - 1))) * 0.25 *
                   (m(0, 0) + m(0, -1) + m(-1, -1) + m(-1, 0)) * 0.25 *
                   (k_(0, 0) + k_(0, -1) + k_(-1, -1) + k_(-1, 0)) * rdy;
          mkrdyp = ((msfuy(i, j) + msfuy(i, j + 1)) / (msfux(i, j) + msfux(i, j + 1))) * 0.25 *
                   (m(0, 0) + m(0, 1) + m(-1, 1) + m(-1, 0)) * 0.25 * (k_(0, 0) + k_(0, 1) + k_(-1, 1) + k_(-1, 0)) *
                   rdy;
          mrdy = msfux(i, j) * msfuy(i, j) * rdy;
        }
      } else if (name == 'v') {  // KHorizDiffV1 (no mass factor in y, WRF v4.8.0 lines 2854-2856)
        mkrdxm = ((msfvx(i, j) + msfvx(i - 1, j)) / (msfvy(i, j) + msfvy(i - 1, j))) * 0.25 *
                 (m(0, 0) + m(0, -1) + m(-1, -1) + m(-1, 0)) * 0.25 * (k_(0, 0) + k_(0, -1) + k_(-1, -1) + k_(-1, 0)) *
                 rdx;
        mkrdxp = ((msfvx(i, j) + msfvx(i + 1, j)) / (msfvy(i, j) + msfvy(i + 1, j))) * 0.25 *
                 (m(0, 0) + m(0, -1) + m(1, -1) + m(1, 0)) * 0.25 * (k_(0, 0) + k_(0, -1) + k_(1, -1) + k_(1, 0)) * rdx;
        mrdx = msfvx(i, j) * msfvy(i, j) * rdx;
        mkrdym = (msfty(i, j - 1) / msftx(i, j - 1)) * k_(0, -1) * rdy;
        mkrdyp = (msfty(i, j) / msftx(i, j)) * k_(0, 0) * rdy;
        mrdy = msfvx(i, j) * msfvy(i, j) * rdy;
      } else if (name == 'w') {  // KHorizDiffW1: mass factors of the level, K averaged over k and k-1
        mkrdxm = (msfux(i, j) / msfuy(i, j)) * 0.25 * (m(0, 0) + m(-1, 0) + m(0, 0) + m(-1, 0)) * 0.25 *
                 (k_(0, 0) + k_(-1, 0) + k_(0, 0, -1) + k_(-1, 0, -1)) * rdx;
        mkrdxp = (msfux(i + 1, j) / msfuy(i + 1, j)) * 0.25 * (m(1, 0) + m(0, 0) + m(1, 0) + m(0, 0)) * 0.25 *
                 (k_(1, 0) + k_(0, 0) + k_(1, 0, -1) + k_(0, 0, -1)) * rdx;
        mrdx = msftx(i, j) * msfty(i, j) * rdx;
        if (do_y) {
          mkrdym = (msfvy(i, j) * msfvx_inv(i, j)) * 0.25 * (m(0, 0) + m(0, -1) + m(0, 0) + m(0, -1)) * 0.25 *
                   (k_(0, 0) + k_(0, -1) + k_(0, 0, -1) + k_(0, -1, -1)) * rdy;
          mkrdyp = (msfvy(i, j + 1) * msfvx_inv(i, j + 1)) * 0.25 * (m(0, 1) + m(0, 0) + m(0, 1) + m(0, 0)) * 0.25 *
                   (k_(0, 1) + k_(0, 0) + k_(0, 1, -1) + k_(0, 0, -1)) * rdy;
          mrdy = msftx(i, j) * msfty(i, j) * rdy;
        }
      } else {  // KHorizDiffM1
        mkrdxm = (msfux(i, j) / msfuy(i, j)) * 0.5 * (k_(0, 0) + k_(-1, 0)) * 0.5 * (m(0, 0) + m(-1, 0)) * rdx;
        mkrdxp = (msfux(i + 1, j) / msfuy(i + 1, j)) * 0.5 * (k_(1, 0) + k_(0, 0)) * 0.5 * (m(1, 0) + m(0, 0)) * rdx;
        mrdx = msftx(i, j) * msfty(i, j) * rdx;
        if (do_y) {
          mkrdym = (msfvy(i, j) * msfvx_inv(i, j)) * 0.5 * (k_(0, 0) + k_(0, -1)) * 0.5 * (m(0, 0) + m(0, -1)) * rdy;
          mkrdyp = (msfvy(i, j + 1) * msfvx_inv(i, j + 1)) * 0.5 * (k_(0, 1) + k_(0, 0)) * 0.5 * (m(0, 1) + m(0, 0)) *
                   rdy;
          mrdy = msftx(i, j) * msfty(i, j) * rdy;
        }
      }
      const double qx = mrdx * (mkrdxp * (q(i + 1, j, k) - q(i, j, k)) - mkrdxm * (q(i, j, k) - q(i - 1, j, k)));
      // an x-z slice: rdy = 0 makes the y term a zero that leaves the sum's bits unchanged (the x term is added
      // to a tendency that is never -0), so it is left out there (no y neighbours)
      const double qy = do_y ? mrdy * (mkrdyp * (q(i, j + 1, k) - q(i, j, k)) - mkrdym * (q(i, j, k) - q(i, j - 1, k))) : 0.0;
      tend(i, j, k) += do_y ? qx + qy : qx;
    });
  }
  // horizontal_diffusion_3dmp (big_step_diffusion.hpp:188-219, 391-414): theta minus its base state t_init, the
  // differences evaluated left to right as WRF
  template <class Fld>
  void hdiff_3dmp(const Fld& q, F& tend, const F& K) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool fx = ox || spec, fy = oy || spec, do_y = !g.two_d();
    int is = fx ? 1 : 0, js = fy ? 1 : 0, ie = fx ? nx - 1 : nx, je = fy ? ny - 1 : ny;
    if (g.bcx == MI_BC_PERIODIC) { is = 0; ie = nx; }
    const double rdx = c.rdx, rdy = c.rdy;
    pfor(is, ie, js, je, 0, nz, [&](int i, int j, int k) {
      auto m = [&](int a, int b) { return c.c1h[k] * mut(i + a, j + b) + c.c2h[k]; };
      auto k_ = [&](int a, int b) { return K(i + a, j + b, k); };
      const double mkrdxm = (msfux(i, j) / msfuy(i, j)) * 0.5 * (k_(0, 0) + k_(-1, 0)) * 0.5 * (m(0, 0) + m(-1, 0)) * rdx;
      const double mkrdxp =
          (msfux(i + 1, j) / msfuy(i + 1, j)) * 0.5 * (k_(1, 0) + k_(0, 0)) * 0.5 * (m(1, 0) + m(0, 0)) * rdx;
      const double mrdx = msftx(i, j) * msfty(i, j) * rdx;
      auto b = [&](int a, int bb) { return t_init(i + a, j + bb, k); };
      const double dxp = q(i + 1, j, k) - q(i, j, k) - b(1, 0) + b(0, 0);
      const double dxm = q(i, j, k) - q(i - 1, j, k) - b(0, 0) + b(-1, 0);
      const double qx = mrdx * (mkrdxp * dxp - mkrdxm * dxm);
      if (!do_y) {
        tend(i, j, k) += qx;
        return;
      }
      const double mkrdym =
          (msfvy(i, j) * msfvx_inv(i, j)) * 0.5 * (k_(0, 0) + k_(0, -1)) * 0.5 * (m(0, 0) + m(0, -1)) * rdy;
      const double mkrdyp =
          (msfvy(i, j + 1) * msfvx_inv(i, j + 1)) * 0.5 * (k_(0, 1) + k_(0, 0)) * 0.5 * (m(0, 1) + m(0, 0)) * rdy;
      const double mrdy = msftx(i, j) * msfty(i, j) * rdy;
      const double dyp = q(i, j + 1, k) - q(i, j, k) - b(0, 1) + b(0, 0);
      const double dym = q(i, j, k) - q(i, j - 1, k) - b(0, 0) + b(0, -1);
      tend(i, j, k) += qx + mrdy * (mkrdyp * dyp - mkrdym * dym);
    });
  }
  // ------------------------------------------------------------------ km_opt 4: DiffusionDriver::prepare
  // compute_diff_metrics with physical_bc (diff_metrics.hpp:20-138): rdzw at mass points, zx at u points and zy at
  // v points over nz + 1 levels; a periodic side leaves the top level 0 on its boundary faces, a specified or open
  // end side copies the last inner face, other non-periodic boundary faces are 0.  z and rdz are not read by
  // diff_opt 1 and are left out.
  void diff_metrics() {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool px = g.bcx == MI_BC_PERIODIC, py = g.bcy == MI_BC_PERIODIC;
    const bool cx = spec || g.bcx == MI_BC_OPEN, cy = spec || (g.bcy == MI_BC_OPEN && !g.two_d());
    const double rdx = c.rdx, rdy = c.rdy;
    pfor2(0, nx, 0, ny, [&](int i, int j) {
      double zw[kMaxLev + 1] = {};
      for (int k = 0; k <= nz; ++k) zw[k] = (ph2(i, j, k) + phb(i, j, k)) / kG;
      for (int k = 0; k < nz; ++k) rdzw(i, j, k) = 1.0 / (zw[k + 1] - zw[k]);
      rdz(i, j, 0) = 2.0 / (zw[1] - zw[0]);
      for (int k = 1; k < nz; ++k) rdz(i, j, k) = 2.0 / (zw[k + 1] - zw[k - 1]);
    });
    pfor(0, nx + 1, 0, ny, 0, nz + 1, [&](int i, int j, int k) {
      if (px && k == nz && (i == 0 || i == nx)) {
        zx(i, j, k) = 0;
      } else if (!px && (i == 0 || (i == nx && !cx))) {
        zx(i, j, k) = 0;
      } else {
        const int i0 = (i == nx) ? (px ? 0 : nx - 1) : i, im1 = (i == 0) ? nx - 1 : (i == nx && !px ? nx - 2 : i - 1);
        double v = rdx * (phb(i0, j, k) - phb(im1, j, k)) / kG;
        v += rdx * (ph2(i0, j, k) - ph2(im1, j, k)) / kG;
        zx(i, j, k) = v;
      }
    });
    pfor(0, nx, 0, ny + 1, 0, nz + 1, [&](int i, int j, int k) {
      if (py && k == nz && (j == 0 || j == ny)) {
        zy(i, j, k) = 0;
      } else if (!py && (j == 0 || (j == ny && !cy))) {
        zy(i, j, k) = 0;
      } else {
        const int j0 = (j == ny) ? (py ? 0 : ny - 1) : j, jm1 = (j == 0) ? ny - 1 : (j == ny && !py ? ny - 2 : j - 1);
        double v = rdy * (phb(i, j0, k) - phb(i, jm1, k)) / kG;
        v += rdy * (ph2(i, j0, k) - ph2(i, jm1, k)) / kG;
        zy(i, j, k) = v;
      }
    });
  }
  // the vertical average of u / msf (or v / msf) onto the w levels of KDeformDiagAndDiv / KDeform12 (deform_div.hpp:
  // 36-48): h0, h1 the two columns' values at the mass levels, out[0..nz]
  void hat_avg(const double* h0, const double* h1, double* out) const {
    const int nz = g.nz;
    const double cft2 = -0.5 * c.dnw[size_t(nz - 1)] / c.dn[size_t(nz - 1)], cft1 = 1.0 - cft2;
    for (int k = 1; k < nz; ++k) out[k] = 0.5 * (c.fnm[size_t(k)] * (h0[k] + h1[k]) + c.fnp[size_t(k)] * (h0[k - 1] + h1[k - 1]));
    out[0] = 0.5 * (c.cf1 * h0[0] + c.cf2 * h0[1] + c.cf3 * h0[2] + c.cf1 * h1[0] + c.cf2 * h1[1] + c.cf3 * h1[2]);
    out[nz] = 0.5 * (cft1 * (h0[nz - 1] + h1[nz - 1]) + cft2 * (h0[nz - 2] + h1[nz - 2]));
  }
  // cal_deform_and_div's defor11, defor22 (KDeformDiagAndDiv, deform_div.hpp:20-84) over the mass points and defor12
  // (KDeform12, :86-155) over the vorticity points, with its edge copies in the reference's order (:257-311);
  // defor33, defor13, defor23 and div are not read by diff_opt 1 with km_opt 4 and are left out
  void deform() {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool px = g.bcx == MI_BC_PERIODIC, py = g.bcy == MI_BC_PERIODIC;
    const double rdx = c.rdx, rdy = c.rdy;
    pfor2(0, nx, 0, ny, [&](int i, int j) {
      const double mm = msftx(i, j) * msfty(i, j);
      double h0[kMaxLev], h1[kMaxLev], ha[kMaxLev + 1], g0[kMaxLev], g1[kMaxLev], ga[kMaxLev + 1];
      for (int k = 0; k < nz; ++k) {
        h0[k] = u2(i, j, k) / msfuy(i, j);
        h1[k] = u2(i + 1, j, k) / msfuy(i + 1, j);
        g0[k] = v2(i, j, k) / msfvx(i, j);
        g1[k] = v2(i, j + 1, k) / msfvx(i, j + 1);
      }
      hat_avg(h0, h1, ha);
      hat_avg(g0, g1, ga);
      for (int k = 0; k < nz; ++k) {
        const double tzx = 0.25 * (zx(i, j, k) + zx(i + 1, j, k) + zx(i, j, k + 1) + zx(i + 1, j, k + 1));
        const double dudx = mm * (rdx * (h1[k] - h0[k]) - (ha[k + 1] - ha[k]) * tzx * rdzw(i, j, k));
        d11(i, j, k) = 2.0 * dudx;
        const double tzy = 0.25 * (zy(i, j, k) + zy(i, j + 1, k) + zy(i, j, k + 1) + zy(i, j + 1, k + 1));
        const double dvdy = mm * (rdy * (g1[k] - g0[k]) - (ga[k + 1] - ga[k]) * tzy * rdzw(i, j, k));
        d22(i, j, k) = 2.0 * dvdy;
      }
    });
    const bool bxs = !px && (spec || ox), bys = !py && (spec || oy);
    pfor2(bxs ? 1 : 0, bxs ? nx : nx + 1, bys ? 1 : 0, bys ? ny : ny + 1, [&](int i, int j) {
      const int im1 = (i == 0) ? (px ? nx - 1 : 0) : i - 1, i0 = (i == nx) ? (px ? 0 : nx - 1) : i;
      const int jm1 = (j == 0) ? (py ? ny - 1 : 0) : j - 1, j0 = (j == ny) ? (py ? 0 : ny - 1) : j;
      const double mm = 0.25 * (msfux(i, jm1) + msfux(i, j0)) * (msfvy(im1, j) + msfvy(i0, j));
      double hm[kMaxLev], hp[kMaxLev], ha[kMaxLev + 1], gm[kMaxLev], gp[kMaxLev], ga[kMaxLev + 1];
      for (int k = 0; k < nz; ++k) {
        hm[k] = u2(i, jm1, k) / msfux(i, jm1);
        hp[k] = u2(i, j0, k) / msfux(i, j0);
        gm[k] = v2(im1, j, k) / msfvy(im1, j);
        gp[k] = v2(i0, j, k) / msfvy(i0, j);
      }
      hat_avg(hm, hp, ha);
      hat_avg(gm, gp, ga);
      for (int k = 0; k < nz; ++k) {
        const double tzy = 0.25 * (zy(im1, j, k) + zy(i0, j, k) + zy(im1, j, k + 1) + zy(i0, j, k + 1));
        const double tu = (ha[k + 1] - ha[k]) * 0.25 * tzy *
                          (rdzw(i0, j0, k) + rdzw(im1, j0, k) + rdzw(im1, jm1, k) + rdzw(i0, jm1, k));
        const double term_u = mm * (rdy * (hp[k] - hm[k]) - tu);
        const double tzx = 0.25 * (zx(i, jm1, k) + zx(i, j0, k) + zx(i, jm1, k + 1) + zx(i, j0, k + 1));
        const double tv = (ga[k + 1] - ga[k]) * 0.25 * tzx *
                          (rdzw(i0, j0, k) + rdzw(i0, jm1, k) + rdzw(im1, jm1, k) + rdzw(im1, j0, k));
        const double term_v = mm * (rdx * (gp[k] - gm[k]) - tv);
        d12(i, j, k) = term_u + term_v;
      }
    });
    const bool bxe = bxs, bye = bys;
    auto row = [&](int dst, int src) { cfor(1, ny + 1, [&](int, int j) { for (int k = 0; k < nz; ++k) d12(dst, j, k) = d12(src, j, k); }); };
    auto col = [&](int dst, int src) { cfor(nx + 1, 1, [&](int i, int) { for (int k = 0; k < nz; ++k) d12(i, dst, k) = d12(i, src, k); }); };
    if (px) row(nx, 0);
    else if (bxs) row(0, 1);
    if (py) col(ny, 0);
    else if (bys) col(0, 1);
    if (bxe) row(nx, nx - 1);
    if (bye) col(ny, ny - 1);
  }
  // calculate_km_kh with km_opt 4 (km_kh.hpp:586-629): KSmag2dKm (:47-102) with diff_opt 1 over the mass points one
  // row in from open and specified sides (a periodic x all of them); the rows outside keep 0; then fill_halos
  void smag2d() {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool fx = ox || spec, fy = oy || spec;
    int is = fx ? 1 : 0, ie = fx ? nx - 1 : nx;
    if (g.bcx == MI_BC_PERIODIC) { is = 0; ie = nx; }
    const double dx = g.dx, dy = g.dy;
    pfor(is, ie, fy ? 1 : 0, fy ? ny - 1 : ny, 0, nz, [&](int i, int j, int k) {
      const double dd = d11(i, j, k) - d22(i, j, k);
      double def2 = 0.25 * (dd * dd);
      const double d12a = 0.25 * (d12(i, j, k) + d12(i, j + 1, k) + d12(i + 1, j, k) + d12(i + 1, j + 1, k));
      def2 += d12a * d12a;
      const double mlen = std::sqrt(dx / msftx(i, j) * dy / msfty(i, j));
      const double tmp = std::sqrt(def2);
      double kmh = c_s * c_s * mlen * mlen * tmp;
      kmh = kmh < 10.0 * mlen ? kmh : 10.0 * mlen;
      xkmh(i, j, k) = kmh;
      xkhh(i, j, k) = kmh / prandtl;
    });
    fill(xkmh, g);
    fill(xkhh, g);
  }
  // cal_dampkm (km_kh.hpp:485-568, damp_opt 1): in each column (one row in from the relaxation zone of specified
  // sides) the eddy viscosities raised to kmmax cos^2(pi/2 min(1, depth below the top / zdamp)) dampcoef, kmmax =
  // min(dx, dy)^2 / dt of the column, the depth summed down from half the top layer
  void dampkm() {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int is = 0, ie = nx, js = 0, je = ny;
    if (spec) {
      is = std::max(is, kBw - 1);
      ie = std::min(ie, nx - kBw + 1);
      js = std::max(js, kBw - 1);
      je = std::min(je, ny - kBw + 1);
    }
    const int ni = ie - is;
    const double piconst = 3.1415926535897932384626433, degrad = piconst / 180.0, degrad90 = degrad * 90.0;
    const double dx = g.dx, dy = g.dy, dt = g.dt, dc = p.dampcoef, zd = p.zdamp;
    pfor2(is, ie, js, je, [&](int i, int j) {
      double dampk[kMaxLev], dampkv[kMaxLev], deltaz = 0;
      for (int k = nz - 1; k >= 0; --k) {
        const double dx_m = dx / msftx(i, j), dy_m = dy / msfty(i, j);
        const double ds = dx_m < dy_m ? dx_m : dy_m;
        const double kmmax = ds * ds / dt;
        double dz;
        if (k == nz - 1) {
          dz = 1.0 / rdzw(i, j, k);
          deltaz = 0.5 * dz;
        } else {
          dz = 1.0 / rdz(i, j, k);
          deltaz = deltaz + dz;
          dz = 1.0 / rdzw(i, j, k);
        }
        const double kmmvmax = dz * dz / dt;
        const double ratio = deltaz / zd;
        const double tmp = ratio < 1.0 ? ratio : 1.0;
        const double c = vcos(degrad90 * tmp, i - is, ni);
        dampk[k] = c * c * kmmax * dc;
        dampkv[k] = c * c * kmmvmax * dc;
        if (dampk[k] < dampkv[k]) dampkv[k] = dampk[k];
      }
      for (int k = 0; k < nz; ++k) {
        if (dampk[k] > xkmh(i, j, k)) xkmh(i, j, k) = dampk[k];
        if (dampk[k] > xkhh(i, j, k)) xkhh(i, j, k) = dampk[k];
      }
    });
  }
  void diffusion_prepare() {  // DiffusionDriver::prepare at the first RK stage (rk3.cpp:113)
    if (p.diff_opt != 1 || !karr) return;
    diff_metrics();
    if (p.km_opt == 4) {
      deform();
      smag2d();
    } else {  // KIsotropicKm over the mass points
      pfor(0, g.nx, 0, g.ny, 0, g.nz, [&](int i, int j, int k) {
        xkmh(i, j, k) = p.khdif;
        xkhh(i, j, k) = p.khdif / prandtl;
      });
    }
    if (p.damp_opt == 1) dampkm();
    fill(xkmh, g);
    fill(xkhh, g);
  }
  // rk_rayleigh_damp (rayleigh_damp.hpp:14-259, damp_opt 2): in the top zdamp metres u, v, theta relax towards their
  // base profiles interpolated to each point's height (u_base, v_base, t_base over z_base) and w towards 0, with the
  // coefficient sin^2(pi/2 (1 - min(1, depth below the top / zdamp))) dampcoef, into the held tendencies.  The
  // damped levels of a column run from the top down to the one above the first point below the layer (level 1 if
  // none is below); u, v and theta take libmvec's sine lanes over those levels, w the scalar sine.
  void rayleigh() {
    const int nx = g.nx, ny = g.ny, nz = g.nz, ktf = nz - 1;
    const bool px = g.bcx == MI_BC_PERIODIC, py = g.bcy == MI_BC_PERIODIC;
    const double pii = 2.0 * std::asin(1.0), dc = p.dampcoef, zd = p.zdamp;
    // one column: zlev(k) the height of mass level k, base the profile, apply(k, coef, ref) the tendency update
    auto column = [&](double ztop, auto zlev, const std::vector<double>& base, auto apply) {
      double z00[kMaxLev], q00[kMaxLev];
      int k1 = ktf;
      double z = ztop;
      while (k1 >= 0 && z >= (ztop - zd)) {
        z = zlev(k1);
        z00[k1] = z;
        k1 = k1 - 1;
      }
      k1 = k1 + 2;
      for (int k = k1; k <= ktf; ++k) {
        int k2 = ktf;
        while (k2 > 0 && z_base[size_t(k2)] > z00[k]) k2 = k2 - 1;
        const auto b = [&](int q) { return base[size_t(q)]; };
        const auto zb = [&](int q) { return z_base[size_t(q)]; };
        if (k2 + 1 > ktf)
          q00[k] = b(k2) + (b(k2) - b(k2 - 1)) * (z00[k] - zb(k2)) / (zb(k2) - zb(k2 - 1));
        else
          q00[k] = b(k2) + (b(k2 + 1) - b(k2)) * (z00[k] - zb(k2)) / (zb(k2 + 1) - zb(k2));
      }
      const int nk = ktf - k1 + 1;
      for (int k = k1; k <= ktf; ++k) {
        const double ratio = (ztop - z00[k]) / zd;
        const double m = ratio < 1.0 ? ratio : 1.0;
        const double s = vsin(0.5 * pii * (1.0 - m), k - k1, nk);
        apply(k, s * s, q00[k]);
      }
    };
    pfor2(0, nx + 1, 0, ny, [&](int i, int j) {  // KRayleighDampU
      const int i0 = (i == nx) ? (px ? 0 : nx - 1) : i, im1 = (i == 0) ? (px ? nx - 1 : 0) : (i - 1);
      const double ztop = 0.5 * (phb(i0, j, nz) + phb(im1, j, nz) + ph2(i0, j, nz) + ph2(im1, j, nz)) / kG;
      auto zl = [&](int k) {
        return 0.25 * (phb(i0, j, k) + phb(i0, j, k + 1) + phb(im1, j, k) + phb(im1, j, k + 1) + ph2(i0, j, k) +
                       ph2(i0, j, k + 1) + ph2(im1, j, k) + ph2(im1, j, k + 1)) / kG;
      };
      column(ztop, zl, u_base, [&](int k, double dcoef, double u00) {
        ru_tendf(i, j, k) = ru_tendf(i, j, k) - (c.c1h[size_t(k)] * muu(i, j) + c.c2h[size_t(k)]) * (dcoef * dc) *
                                                    (u2(i, j, k) - u00);
      });
    });
    if (!g.two_d()) {  // an x-z slice holds no v tendency
      pfor2(0, nx, 0, ny + 1, [&](int i, int j) {  // KRayleighDampV
        const int j0 = (j == ny) ? (py ? 0 : ny - 1) : j, jm1 = (j == 0) ? (py ? ny - 1 : 0) : (j - 1);
        const double ztop = 0.5 * (phb(i, j0, nz) + phb(i, jm1, nz) + ph2(i, j0, nz) + ph2(i, jm1, nz)) / kG;
        auto zl = [&](int k) {
          return 0.25 * (phb(i, j0, k) + phb(i, j0, k + 1) + phb(i, jm1, k) + phb(i, jm1, k + 1) + ph2(i, j0, k) +
                         ph2(i, j0, k + 1) + ph2(i, jm1, k) + ph2(i, jm1, k + 1)) / kG;
        };
        column(ztop, zl, v_base, [&](int k, double dcoef, double v00) {
          rv_tendf(i, j, k) = rv_tendf(i, j, k) - (c.c1h[size_t(k)] * muv(i, j) + c.c2h[size_t(k)]) * (dcoef * dc) *
                                                      (v2(i, j, k) - v00);
        });
      });
    }
    pfor2(0, nx, 0, ny, [&](int i, int j) {  // KRayleighDampWT
      const double ztop = (phb(i, j, nz) + ph2(i, j, nz)) / kG;
      for (int k = 0; k <= nz; ++k) {
        const double z = (phb(i, j, k) + ph2(i, j, k)) / kG;
        if (z >= (ztop - zd)) {
          const double ratio = (ztop - z) / zd;
          const double m = ratio < 1.0 ? ratio : 1.0;
          const double s = std::sin(0.5 * pii * (1.0 - m));
          rw_tendf(i, j, k) = rw_tendf(i, j, k) - (c.c1f[size_t(k)] * mut(i, j) + c.c2f[size_t(k)]) * (s * s * dc) *
                                                      w2(i, j, k);
        }
      }
      auto zl = [&](int k) { return 0.5 * (phb(i, j, k) + phb(i, j, k + 1) + ph2(i, j, k) + ph2(i, j, k + 1)) / kG; };
      column(ztop, zl, t_base, [&](int k, double dcoef, double t00) {
        t_tendf(i, j, k) = t_tendf(i, j, k) - (c.c1h[size_t(k)] * mut(i, j) + c.c2h[size_t(k)]) * (dcoef * dc) *
                                                  (t2(i, j, k) - t00);
      });
    });
  }
  // vertical_diffusion, _mp, _u, _v (big_step_diffusion.hpp:225-338, 416-500): the column's fluxes at the w levels
  // 1..nz-1 (vflux[k] between levels k and k+1, 0 at the top), then their differences.  'u'/'v' at the faces (muu /
  // muv, alt averaged onto the face, the base profile subtracted, no tendency at the top level); 'w' at the w
  // levels 1..nz-1; 'm' at the mass levels, with `base` (t_init, qv_base) subtracted when given
  template <class Fld>
  void vdiff(char name, const Fld& q, F& tend, double kv, const std::function<double(int, int, int)>* base) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool fx = ox || spec, fy = oy || spec;
    int is = 0, ie = nx, js = 0, je = ny;
    if (name == 'u') { is = fx ? 1 : 0; ie = fx ? nx : nx + 1; }
    if (name == 'v') { js = fy ? 1 : 0; je = fy ? ny : ny + 1; }
    if (name == 'u' && g.bcx == MI_BC_PERIODIC) { is = 0; ie = nx + 1; }
    const double G = kG;
    pfor2(is, ie, js, je, [&](int i, int j) {
      std::vector<double> vf(size_t(nz), 0.0);
      for (int k = 0; k < nz - 1; ++k) {
        if (name == 'u' || name == 'v') {
          const int di = name == 'u' ? -1 : 0, dj = name == 'v' ? -1 : 0;
          const double a = 0.25 * (alt(i, j, k) + alt(i + di, j + dj, k) + alt(i, j, k + 1) + alt(i + di, j + dj, k + 1));
          const std::vector<double>& bp = name == 'u' ? u_base : v_base;
          const double dp = q(i, j, k + 1) - q(i, j, k) - bp[k + 1] + bp[k];
          vf[k] = kv * c.rdn[k + 1] / a * dp;
        } else if (name == 'w') {
          vf[k] = (kv / alt(i, j, k)) * c.rdnw[k] * (q(i, j, k + 1) - q(i, j, k));
        } else {
          const double dq = base ? q(i, j, k + 1) - q(i, j, k) - (*base)(i, j, k + 1) + (*base)(i, j, k)
                                 : q(i, j, k + 1) - q(i, j, k);
          vf[k] = kv * c.rdn[k + 1] / (0.5 * (alt(i, j, k) + alt(i, j, k + 1))) * dq;
        }
      }
      const double vf0 = vf[0];
      if (name == 'u' || name == 'v') {
        const int di = name == 'u' ? -1 : 0, dj = name == 'v' ? -1 : 0;
        auto mm = [&](int k) { return c.c1h[k] * (name == 'u' ? muu(i, j) : muv(i, j)) + c.c2h[k]; };
        tend(i, j, 0) += G * G * c.rdnw[0] / mm(0) / (0.5 * (alt(i + di, j + dj, 0) + alt(i, j, 0))) * (vf[0] - vf0);
        for (int k = 1; k < nz - 1; ++k) {
          const double a = 0.5 * (alt(i + di, j + dj, k) + alt(i, j, k));
          tend(i, j, k) += G * G * c.rdnw[k] / mm(k) / a * (vf[k] - vf[k - 1]);
        }
      } else if (name == 'w') {
        for (int k = 1; k < nz; ++k) {
          const double mu = c.c1f[k] * mut(i, j) + c.c2f[k];
          const double alt_avg = 0.5 * (alt(i, j, k) + alt(i, j, k - 1));
          tend(i, j, k) += c.rdn[k] * G * G / mu / alt_avg * (vf[k] - vf[k - 1]);
        }
      } else {
        const F& mus = name == 's' ? muts : mut;
        tend(i, j, 0) += G * G / (c.c1h[0] * mus(i, j) + c.c2h[0]) / alt(i, j, 0) * c.rdnw[0] * (vf[0] - vf0);
        for (int k = 1; k < nz; ++k) {
          const double mu = c.c1h[k] * mus(i, j) + c.c2h[k];
          tend(i, j, k) += G * G / mu / alt(i, j, k) * c.rdnw[k] * (vf[k] - vf[k - 1]);
        }
      }
    });
  }
  bool moist() const { return n_moist > 0; }
  // kern::MoistCoef (kernels_common.hpp:93-100): 1, 1, 0 when dry
  double mc_cqu(int i, int j, int k) const { return moist() ? cqu(i, j, k) : 1.0; }
  double mc_cqv(int i, int j, int k) const { return moist() ? cqv(i, j, k) : 1.0; }
  double mc_cqw(int i, int j, int k) const { return moist() ? cqw(i, j, k) : 0.0; }
  double mc_cq1(int i, int j, int k) const { return 1.0 / (1.0 + mc_cqw(i, j, k)); }
  static int mi_auto_sound_steps_(double dt, double dx, double dy) {
    const double spacing = dx < dy ? dx : dy;
    const int n = 2 * (int(300.0 * dt / spacing - 0.01) + 1);
    return n > 4 ? n : 4;
  }
  int nyv() const { return g.nyv(); }

  // ------------------------------------------------------------------ LargeStep::prep (large_step.cpp:697-733)
  void rk_prep() {
    const int nx = g.nx, ny = g.ny, nz = g.nz, nyv_ = nyv();
    // mu_2's halo is not refreshed with specified boundaries (large_step.cpp:700-706)
    for (F* f : {&u2, &v2, &w2, &t2, &ph2, &pp, &al}) fill(*f, g);
    if (!spec) fill(mu2, g);
    // KMu (large_step.cpp:49-66): a specified or open edge face takes its column's mass twice
    pfor(0, nx + 1, 0, nyv_, 0, 1, [&](int i, int j, int) {
      if (i < nx && j < ny) mut(i, j) = mu2(i, j) + mub(i, j);
      if (j < ny) {
        if (flx && i == 0) muu(i, j) = H * (mu2(0, j) + mu2(0, j) + mub(0, j) + mub(0, j));
        else if (flx && i == nx) muu(i, j) = H * (mu2(nx - 1, j) + mu2(nx - 1, j) + mub(nx - 1, j) + mub(nx - 1, j));
        else muu(i, j) = H * (mu2(i, j) + mu2(i - 1, j) + mub(i, j) + mub(i - 1, j));
      }
      if (i < nx && j < nyv_) {
        if (fly && j == 0) muv(i, j) = H * (mu2(i, 0) + mu2(i, 0) + mub(i, 0) + mub(i, 0));
        else if (fly && j == ny) muv(i, j) = H * (mu2(i, ny - 1) + mu2(i, ny - 1) + mub(i, ny - 1) + mub(i, ny - 1));
        else muv(i, j) = H * (mu2(i, j) + mu2(i, j - 1) + mub(i, j) + mub(i, j - 1));
      }
    });
    fill(mut, g);
    // KPrep3 (large_step.cpp:68-81)
    pfor(0, nx + 1, 0, nyv_, 0, nz + 1, [&](int i, int j, int k) {
      if (j < ny && k < nz) ru(i, j, k) = u2(i, j, k) * (c.c1h[k] * muu(i, j) + c.c2h[k]) / msfuy(i, j);
      if (i < nx && j < nyv_ && k < nz) rv(i, j, k) = v2(i, j, k) * (c.c1h[k] * muv(i, j) + c.c2h[k]) * msfvx_inv(i, j);
      if (i < nx && j < ny) rw(i, j, k) = w2(i, j, k) * (c.c1f[k] * mut(i, j) + c.c2f[k]) / msfty(i, j);
      if (i < nx && j < ny && k < nz) {
        alt(i, j, k) = al(i, j, k) + alb(i, j, k);
        php(i, j, k) = H * (phb(i, j, k) + phb(i, j, k + 1) + ph2(i, j, k) + ph2(i, j, k + 1));
      }
    });
    for (F* f : {&ru, &rv, &rw, &alt, &php}) fill(*f, g);
    fill(muu, g, Par::OddFace, Par::Even);
    fill(muv, g, Par::Even, Par::Odd);
    // KWW calc_ww_cp (large_step.cpp:83-108)
    auto mu_u = [&](int i, int j) { return H * (mu2(i, j) + mub(i, j) + mu2(i - 1, j) + mub(i - 1, j)); };
    auto mu_v = [&](int i, int j) { return H * (mu2(i, j) + mub(i, j) + mu2(i, j - 1) + mub(i, j - 1)); };
    auto divv = [&](int i, int j, int k) {
      const double mu1_ = mu_u(i + 1, j), mu0 = mu_u(i, j), mv1 = mu_v(i, j + 1), mv0 = mu_v(i, j);
      return msftx(i, j) * c.dnw[k] *
             (c.rdx * ((c.c1h[k] * mu1_ + c.c2h[k]) * u2(i + 1, j, k) / msfuy(i + 1, j) -
                       (c.c1h[k] * mu0 + c.c2h[k]) * u2(i, j, k) / msfuy(i, j)) +
              c.rdy * ((c.c1h[k] * mv1 + c.c2h[k]) * v2(i, j + 1, k) * msfvx_inv(i, j + 1) -
                       (c.c1h[k] * mv0 + c.c2h[k]) * v2(i, j, k) * msfvx_inv(i, j)));
    };
    cfor(nx, ny, [&](int i, int j) {
      double dmdt = 0;
      for (int k = 0; k < nz; ++k) dmdt += divv(i, j, k);
      double w = 0;
      ww(i, j, 0) = 0;
      for (int k = 1; k < nz; ++k) {
        w = w - c.dnw[k - 1] * c.c1h[k - 1] * dmdt - divv(i, j, k - 1);
        ww(i, j, k) = w;
      }
      ww(i, j, nz) = 0;
    });
    fill(ww, g);
    if (moist()) {  // KCq calc_cq (large_step.cpp:127-148, 724-733): per species, both neighbours
      pfor(0, nx + 1, 0, nyv_, 0, nz, [&](int i, int j, int k) {
        if (j < ny) {
          double t = 0;
          for (int s = 0; s < n_moist; ++s) t = t + q2[s](i, j, k) + q2[s](i - 1, j, k);
          cqu(i, j, k) = 1.0 / (1.0 + 0.5 * t);
        }
        if (i < nx && j < nyv_) {
          double t = 0;
          for (int s = 0; s < n_moist; ++s) t = t + q2[s](i, j, k) + q2[s](i, j - 1, k);
          cqv(i, j, k) = 1.0 / (1.0 + 0.5 * t);
        }
        if (i < nx && j < ny && k >= 1) {
          double t = 0;
          for (int s = 0; s < n_moist; ++s) t = t + q2[s](i, j, k) + q2[s](i, j, k - 1);
          cqw(i, j, k) = 0.5 * t;
        }
      });
    }
  }

  // ------------------------------------------------------------------ LargeStep::tendencies (large_step.cpp:798-861)
  void rk_tendency(int rk) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (spec) std::fill(mu_tend.a.begin(), mu_tend.a.end(), 0.0);  // zero_tend of mu_tend, every stage (:815-818)
    // zero_tendf at the first stage without a plugin zeroes the scalars' q_tendf too (large_step.cpp:814,
    // 888-893): a tendency the caller left there has no effect on the step
    if (rk == 1 && !has_phys)
      for (F& f : qtf) std::fill(f.a.begin(), f.a.end(), 0.0);
    if (rk == 1) diffusion_prepare();
    if (rk == 1 && use_tendf) {  // zero_tendf, then DiffusionDriver::dynamics (diffusion_driver.cpp:266-281)
      if (!has_phys) {
        for (F* f : {&ru_tendf, &rv_tendf, &rw_tendf, &t_tendf, &ph_tendf}) std::fill(f->a.begin(), f->a.end(), 0.0);
      } else {  // zeroed before part1; part2 adds the physics tendencies (large_step.cpp:820-821)
        if (phys.part2) phys.part2(phys.user);
        for (F* f : {&ru_tendf, &rv_tendf, &rw_tendf, &t_tendf})
          for (double& v : f->a) v = 0.0 + v;
      }
      auto fu = [&](int i, int j, int k) { return u2(i, j, k); };
      auto fv = [&](int i, int j, int k) { return v2(i, j, k); };
      auto fw = [&](int i, int j, int k) { return w2(i, j, k); };
      auto ft = [&](int i, int j, int k) { return t2(i, j, k); };
      auto mu_t = [&](int i, int j) { return mut(i, j); };
      if (p.diff_opt == 1) {  // diffusion_driver.cpp:159-193: horizontal u, v, w, theta, then vertical
        hdiff('u', fu, ru_tendf, mu_t, xkmh, c.c1h, c.c2h);
        if (!g.two_d()) hdiff('v', fv, rv_tendf, mu_t, xkmh, c.c1h, c.c2h);
        hdiff('w', fw, rw_tendf, mu_t, xkmh, c.c1f, c.c2f);
        hdiff_3dmp(ft, t_tendf, xkhh);
        vdiff('u', fu, ru_tendf, kvdif, nullptr);
        if (!g.two_d()) vdiff('v', fv, rv_tendf, kvdif, nullptr);
        vdiff('w', fw, rw_tendf, kvdif, nullptr);
        const std::function<double(int, int, int)> tb = [&](int i, int j, int k) { return t_init(i, j, k); };
        vdiff('m', ft, t_tendf, 3.0 * kvdif, &tb);  // kvdq = 3 kvdif for theta (WRF rk_tendency)
      }
      if (p.diff_6th_opt != 0) {
        sixth('u', fu, ru_tendf, mu_t, g.dt, c.c1h, c.c2h);
        if (!g.two_d()) sixth('v', fv, rv_tendf, mu_t, g.dt, c.c1h, c.c2h);
        sixth('w', fw, rw_tendf, mu_t, g.dt, c.c1f, c.c2f);
        sixth('m', ft, t_tendf, mu_t, g.dt, c.c1h, c.c2h);
      }
      if (p.damp_opt == 2) rayleigh();  // the last term of DiffusionDriver::dynamics (diffusion_driver.cpp:283-289)
      if (spec) relax_dry();  // relax_bdy_dry (large_step.cpp:825)
    }
    const bool do_y = !g.two_d();
    const bool top_lid = p.top_lid != 0;
    const double hup = p.h_adv_order == 5 ? 1.0 : 0.0, vup = p.v_adv_order == 3 ? 1.0 : 0.0;
    const double rdx = c.rdx, rdy = c.rdy;
    const double wd_dt = p.w_damping == 1 ? g.dt : 0.0, w_crit = p.w_crit_cfl;
    const int nxo = nx, nyo = ny;  // kern::Spec: degraded orders next to specified edges

    // KTendU (large_step.cpp:152-264), interior faces, no Coriolis, unmapped
    {
      auto fx = [&](int i, int j, int k) {
        const int o = ord_mass(i, nxo, flx);
        if (o == 2) {  // specified: the upstream normal velocity at the boundary face (large_step.cpp:160-167)
          if (i <= 1) {
            const double ub = (spec && u2(i, j, k) < 0.0) ? u2(i, j, k) : u2(i - 1, j, k);
            return Q * (ru(i, j, k) + ru(i - 1, j, k)) * (u2(i, j, k) + ub);
          }
          const double ub = (spec && u2(i - 1, j, k) > 0.0) ? u2(i - 1, j, k) : u2(i, j, k);
          return Q * (ru(i, j, k) + ru(i - 1, j, k)) * (u2(i - 1, j, k) + ub);
        }
        double vel = H * (ru(i, j, k) + ru(i - 1, j, k));
        if (o == 3) return vel * flux3(u2(i - 2, j, k), u2(i - 1, j, k), u2(i, j, k), u2(i + 1, j, k), vel, hup);
        return vel * flux5(u2(i - 3, j, k), u2(i - 2, j, k), u2(i - 1, j, k), u2(i, j, k), u2(i + 1, j, k), u2(i + 2, j, k),
                           vel, hup);
      };
      auto fy = [&](int i, int j, int k) {
        const int o = ord_face(j, nyo, fly);
        if (o == 2) return Q * (rv(i, j, k) + rv(i - 1, j, k)) * (u2(i, j, k) + u2(i, j - 1, k));
        double vel = H * (rv(i, j, k) + rv(i - 1, j, k));
        if (o == 3) return vel * flux3(u2(i, j - 2, k), u2(i, j - 1, k), u2(i, j, k), u2(i, j + 1, k), vel, hup);
        return vel * flux5(u2(i, j - 3, k), u2(i, j - 2, k), u2(i, j - 1, k), u2(i, j, k), u2(i, j + 1, k), u2(i, j + 2, k),
                           vel, hup);
      };
      auto fz = [&](int i, int j, int k) {
        if (k == 0 || k == nz) return 0.0;
        double vel = H * (ww(i - 1, j, k) + ww(i, j, k));
        if (k == 1 || k == nz - 1) return vel * (c.fnm[k] * u2(i, j, k) + c.fnp[k] * u2(i, j, k - 1));
        return vel * flux3(u2(i, j, k - 2), u2(i, j, k - 1), u2(i, j, k), u2(i, j, k + 1), -vel, vup);
      };
      auto pu = [&](int i, int j, int k) { return pp(i - 1, j, k) + pp(i, j, k); };
      auto dpn = [&](int i, int j, int k) {
        if (k == 0) return H * (c.cf1 * pu(i, j, 0) + c.cf2 * pu(i, j, 1) + c.cf3 * pu(i, j, 2));
        if (k == nz) return top_lid ? H * (c.cfn * pu(i, j, nz - 1) + c.cfn1 * pu(i, j, nz - 2)) : 0.0;
        return H * (c.fnm[k] * pu(i, j, k) + c.fnp[k] * pu(i, j, k - 1));
      };
      // advect_u open_ys / open_ye (large_step.cpp:193-207): v-flux terms of the outermost rows
      auto open_y_term = [&](int i, int j, int k) {
        const int imin = 0, imax = nx - 1;
        const int ip = i < imax ? i : imax, im = i - 1 > imin ? i - 1 : imin;
        const double mrdy = msfux(i, j) * rdy;
        if (j == 0) {
          const double vw = H * (rv(ip, 0, k) + rv(im, 0, k));
          const double vb = vw < 0.0 ? vw : 0.0;
          const double dvm = rv(ip, 1, k) - rv(ip, 0, k), dvp = rv(im, 1, k) - rv(im, 0, k);
          return mrdy * (vb * (u2(i, 1, k) - u2(i, 0, k)) + H * u2(i, 0, k) * (dvm + dvp));
        }
        const int je = ny;
        const double vw = H * (rv(ip, je, k) + rv(im, je, k));
        const double vb = vw > 0.0 ? vw : 0.0;
        const double dvm = rv(ip, je, k) - rv(ip, je - 1, k), dvp = rv(im, je, k) - rv(im, je - 1, k);
        return mrdy * (vb * (u2(i, je - 1, k) - u2(i, je - 2, k)) + H * u2(i, je - 1, k) * (dvm + dvp));
      };
      // open x: the boundary faces are advanced too (large_step.cpp:805)
      pfor(ox ? 0 : g.u_i0(), ox ? nx + 1 : nx, 0, ny, 0, nz, [&](int i, int j, int k) {
        if (ox && (i == 0 || i == nx)) {
          // open x boundary face (large_step.cpp:211-228): the radiation condition with phase speed cb and,
          // unless open_ys/ye, the vertical advection
          double t;
          if (i == 0) {
            double ub = ru(0, j, k) - kCb * (c.c1h[k] * mut(0, j) + c.c2h[k]);
            ub = ub < 0.0 ? ub : 0.0;
            t = 0.0 - rdx * ub * (u2(1, j, k) - u2(0, j, k));
          } else {
            double ub = ru(i, j, k) + kCb * (c.c1h[k] * mut(i - 1, j) + c.c2h[k]);
            ub = ub > 0.0 ? ub : 0.0;
            t = 0.0 - rdx * ub * (u2(i, j, k) - u2(i - 1, j, k));
          }
          if (!oy) t -= c.rdnw[k] * (fz(i, j, k + 1) - fz(i, j, k));
          if (use_tendf) t += ru_tendf(i, j, k) / msfuy(i, j);
          ru_tend(i, j, k) = t;
          return;
        }
        double t = -(msfux(i, j) * rdx) * (fx(i + 1, j, k) - fx(i, j, k));
        if (do_y) {
          const bool orow = oy && (j == 0 || j == ny - 1);  // open_ys/ye: the outermost rows take the open terms
          if (!orow) t -= (msfux(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
          else t -= open_y_term(i, j, k);
        }
        if (!(oy && flx && (i == 0 || i == nx))) t -= c.rdnw[k] * (fz(i, j, k + 1) - fz(i, j, k));
        const double rm = msfux(i, j) / msfuy(i, j);
        double dpx = rm * H * rdx * (c.c1h[k] * muu(i, j) + c.c2h[k]) *
                     ((ph2(i, j, k + 1) - ph2(i - 1, j, k + 1) + ph2(i, j, k) - ph2(i - 1, j, k)) +
                      (alt(i, j, k) + alt(i - 1, j, k)) * (pp(i, j, k) - pp(i - 1, j, k)) +
                      (al(i, j, k) + al(i - 1, j, k)) * (pb(i, j, k) - pb(i - 1, j, k)));
        dpx += rm * rdx * (php(i, j, k) - php(i - 1, j, k)) *
               (c.rdnw[k] * (dpn(i, j, k + 1) - dpn(i, j, k)) - H * (c.c1h[k] * mu2(i - 1, j) + c.c1h[k] * mu2(i, j)));
        t -= mc_cqu(i, j, k) * dpx;
        const double rv4 = Q * (rv(i - 1, j + 1, k) + rv(i, j + 1, k) + rv(i - 1, j, k) + rv(i, j, k));
        const double rw4 = Q * (rw(i - 1, j, k + 1) + rw(i - 1, j, k) + rw(i, j, k + 1) + rw(i, j, k));
        if (cor)
          t = t + rm * H * (fcor(i, j) + fcor(i - 1, j)) * rv4 -
              H * (ecor(i, j) + ecor(i - 1, j)) * H * (cosa(i, j) + cosa(i - 1, j)) * rw4;
        if (mapped)
          t = t + H * (vxgm(i, j, k) + vxgm(i - 1, j, k)) * rv4 - u2(i, j, k) * kReradius * rw4;
        else
          t = t - u2(i, j, k) * kReradius * rw4;
        if (use_tendf) t += ru_tendf(i, j, k) / msfuy(i, j);
        ru_tend(i, j, k) = t;
      });
    }
    // KTendV (large_step.cpp:267-376), 3-D only
    if (do_y) {
      auto fx = [&](int i, int j, int k) {
        const int o = ord_face(i, nxo, flx);
        if (o == 2) return Q * (ru(i, j, k) + ru(i, j - 1, k)) * (v2(i, j, k) + v2(i - 1, j, k));
        double vel = H * (ru(i, j, k) + ru(i, j - 1, k));
        if (o == 3) return vel * flux3(v2(i - 2, j, k), v2(i - 1, j, k), v2(i, j, k), v2(i + 1, j, k), vel, hup);
        return vel * flux5(v2(i - 3, j, k), v2(i - 2, j, k), v2(i - 1, j, k), v2(i, j, k), v2(i + 1, j, k), v2(i + 2, j, k),
                           vel, hup);
      };
      auto fy = [&](int i, int j, int k) {
        const int o = ord_mass(j, nyo, fly);
        if (o == 2) {  // upstream normal velocity at the boundary face (large_step.cpp:282-289)
          if (j <= 1) {
            const double vb = (spec && v2(i, j, k) < 0.0) ? v2(i, j, k) : v2(i, j - 1, k);
            return Q * (rv(i, j, k) + rv(i, j - 1, k)) * (v2(i, j, k) + vb);
          }
          const double vb = (spec && v2(i, j - 1, k) > 0.0) ? v2(i, j - 1, k) : v2(i, j, k);
          return Q * (rv(i, j, k) + rv(i, j - 1, k)) * (vb + v2(i, j - 1, k));
        }
        double vel = H * (rv(i, j, k) + rv(i, j - 1, k));
        if (o == 3) return vel * flux3(v2(i, j - 2, k), v2(i, j - 1, k), v2(i, j, k), v2(i, j + 1, k), vel, hup);
        return vel * flux5(v2(i, j - 3, k), v2(i, j - 2, k), v2(i, j - 1, k), v2(i, j, k), v2(i, j + 1, k), v2(i, j + 2, k),
                           vel, hup);
      };
      auto fz = [&](int i, int j, int k) {
        if (k == 0 || k == nz) return 0.0;
        double vel = H * (ww(i, j - 1, k) + ww(i, j, k));
        if (k == 1 || k == nz - 1) return vel * (c.fnm[k] * v2(i, j, k) + c.fnp[k] * v2(i, j, k - 1));
        return vel * flux3(v2(i, j, k - 2), v2(i, j, k - 1), v2(i, j, k), v2(i, j, k + 1), -vel, vup);
      };
      auto pv = [&](int i, int j, int k) { return pp(i, j - 1, k) + pp(i, j, k); };
      auto dpn = [&](int i, int j, int k) {
        if (k == 0) return H * (c.cf1 * pv(i, j, 0) + c.cf2 * pv(i, j, 1) + c.cf3 * pv(i, j, 2));
        if (k == nz) return top_lid ? H * (c.cfn * pv(i, j, nz - 1) + c.cfn1 * pv(i, j, nz - 2)) : 0.0;
        return H * (c.fnm[k] * pv(i, j, k) + c.fnp[k] * pv(i, j, k - 1));
      };
      // advect_v open_xs / open_xe (large_step.cpp:307-323): u-flux terms of the outermost columns
      auto open_x_term = [&](int i, int j, int k) {
        const int jmin = 0, jmax = ny - 1;
        const int jp = j < jmax ? j : jmax, jm = j - 1 > jmin ? j - 1 : jmin;
        if (i == 0) {
          const double mrdx = msfvy(0, j) * rdx;
          const double uw = H * (ru(0, jp, k) + ru(0, jm, k));
          const double ub = uw < 0.0 ? uw : 0.0;
          const double dup = ru(1, jp, k) - ru(0, jp, k), dum = ru(1, jm, k) - ru(0, jm, k);
          return mrdx * (ub * (v2(1, j, k) - v2(0, j, k)) + H * v2(0, j, k) * (dup + dum));
        }
        const int ie = nx;
        const double mrdx = msfvy(ie - 1, j) * rdx;
        const double uw = H * (ru(ie, jp, k) + ru(ie, jm, k));
        const double ub = uw > 0.0 ? uw : 0.0;
        const double dup = ru(ie, jp, k) - ru(ie - 1, jp, k), dum = ru(ie, jm, k) - ru(ie - 1, jm, k);
        return mrdx * (ub * (v2(ie - 1, j, k) - v2(ie - 2, j, k)) + H * v2(ie - 1, j, k) * (dup + dum));
      };
      pfor(0, nx, oy ? 0 : g.v_j0(), oy ? ny + 1 : ny, 0, nz, [&](int i, int j, int k) {
        if (oy && (j == 0 || j == ny)) {  // open y boundary face: radiation only (large_step.cpp:327-340)
          double t;
          if (j == 0) {
            double vb = rv(i, 0, k) - kCb * (c.c1h[k] * mut(i, 0) + c.c2h[k]);
            vb = vb < 0.0 ? vb : 0.0;
            t = 0.0 - rdy * vb * (v2(i, 1, k) - v2(i, 0, k));
          } else {
            double vb = rv(i, j, k) + kCb * (c.c1h[k] * mut(i, j - 1) + c.c2h[k]);
            vb = vb > 0.0 ? vb : 0.0;
            t = 0.0 - rdy * vb * (v2(i, j, k) - v2(i, j - 1, k));
          }
          if (use_tendf) t += rv_tendf(i, j, k) * msfvx_inv(i, j);
          rv_tend(i, j, k) = t;
          return;
        }
        const double rm = msfvy(i, j) / msfvx(i, j);
        double t;
        if (ox && (i == 0 || i == nx - 1)) {  // open x: the y flux, then the open term instead of the x flux
          t = -(msfvy(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
          t -= open_x_term(i, j, k);
        } else {
          t = -(msfvy(i, j) * rdx) * (fx(i + 1, j, k) - fx(i, j, k));
          t -= (msfvy(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
        }
        t -= rm * c.rdnw[k] * (fz(i, j, k + 1) - fz(i, j, k));
        double dpy = rm * H * rdy * (c.c1h[k] * muv(i, j) + c.c2h[k]) *
                     ((ph2(i, j, k + 1) - ph2(i, j - 1, k + 1) + ph2(i, j, k) - ph2(i, j - 1, k)) +
                      (alt(i, j, k) + alt(i, j - 1, k)) * (pp(i, j, k) - pp(i, j - 1, k)) +
                      (al(i, j, k) + al(i, j - 1, k)) * (pb(i, j, k) - pb(i, j - 1, k)));
        dpy += rm * rdy * (php(i, j, k) - php(i, j - 1, k)) *
               (c.rdnw[k] * (dpn(i, j, k + 1) - dpn(i, j, k)) - H * (c.c1h[k] * mu2(i, j - 1) + c.c1h[k] * mu2(i, j)));
        t -= mc_cqv(i, j, k) * dpy;
        const double ru4 = Q * (ru(i, j, k) + ru(i + 1, j, k) + ru(i, j - 1, k) + ru(i + 1, j - 1, k));
        const double rw4 = Q * (rw(i, j - 1, k + 1) + rw(i, j - 1, k) + rw(i, j, k + 1) + rw(i, j, k));
        if (cor)
          t = t - rm * H * (fcor(i, j) + fcor(i, j - 1)) * ru4 +
              rm * H * (ecor(i, j) + ecor(i, j - 1)) * H * (sina(i, j) + sina(i, j - 1)) * rw4;
        if (mapped)
          t = t - H * (vxgm(i, j, k) + vxgm(i, j - 1, k)) * ru4 - rm * v2(i, j, k) * kReradius * rw4;
        else
          t = t - rm * v2(i, j, k) * kReradius * rw4;
        if (use_tendf) t += rv_tendf(i, j, k) * msfvx_inv(i, j);
        rv_tend(i, j, k) = t;
      });
    }
    // KTendW (large_step.cpp:379-499), levels 1..nz
    {
      auto rux = [&](int i, int j, int k) {
        if (k < nz) return c.fnm[k] * ru(i, j, k) + c.fnp[k] * ru(i, j, k - 1);
        return (2.0 - c.fnm[nz - 1]) * ru(i, j, nz - 1) - c.fnp[nz - 1] * ru(i, j, nz - 2);
      };
      auto rvy = [&](int i, int j, int k) {
        if (k < nz) return c.fnm[k] * rv(i, j, k) + c.fnp[k] * rv(i, j, k - 1);
        return (2.0 - c.fnm[nz - 1]) * rv(i, j, nz - 1) - c.fnp[nz - 1] * rv(i, j, nz - 2);
      };
      auto fx = [&](int i, int j, int k) {
        double vel = rux(i, j, k);
        const int o = ord_face(i, nxo, flx);
        if (o == 2) return H * vel * (w2(i, j, k) + w2(i - 1, j, k));
        if (o == 3) return vel * flux3(w2(i - 2, j, k), w2(i - 1, j, k), w2(i, j, k), w2(i + 1, j, k), vel, hup);
        return vel * flux5(w2(i - 3, j, k), w2(i - 2, j, k), w2(i - 1, j, k), w2(i, j, k), w2(i + 1, j, k), w2(i + 2, j, k),
                           vel, hup);
      };
      auto fy = [&](int i, int j, int k) {
        double vel = rvy(i, j, k);
        const int o = ord_face(j, nyo, fly);
        if (o == 2) return H * vel * (w2(i, j, k) + w2(i, j - 1, k));
        if (o == 3) return vel * flux3(w2(i, j - 2, k), w2(i, j - 1, k), w2(i, j, k), w2(i, j + 1, k), vel, hup);
        return vel * flux5(w2(i, j - 3, k), w2(i, j - 2, k), w2(i, j - 1, k), w2(i, j, k), w2(i, j + 1, k), w2(i, j + 2, k),
                           vel, hup);
      };
      auto fz = [&](int i, int j, int m) {
        if (m == 0 || m == nz - 1) return 0.25 * (ww(i, j, m + 1) + ww(i, j, m)) * (w2(i, j, m + 1) + w2(i, j, m));
        double vel = H * (ww(i, j, m + 1) + ww(i, j, m));
        return vel * flux3(w2(i, j, m - 1), w2(i, j, m), w2(i, j, m + 1), w2(i, j, m + 2), -vel, vup);
      };
      // advect_w open_xs/xe (open_ys/ye) terms of the outermost columns (rows), large_step.cpp:416-453
      auto open_x_term = [&](int i, int j, int k) {
        const int a = i == 0 ? 0 : nx - 1, b = i == 0 ? 1 : nx;
        double uw, dru;
        if (k < nz) {
          uw = H * (c.fnm[k] * (ru(a, j, k) + ru(b, j, k)) + c.fnp[k] * (ru(a, j, k - 1) + ru(b, j, k - 1)));
          dru = c.fnm[k] * (ru(b, j, k) - ru(a, j, k)) + c.fnp[k] * (ru(b, j, k - 1) - ru(a, j, k - 1));
        } else {
          const double fm = 2.0 - c.fnm[nz - 1], fp = c.fnp[nz - 1];
          uw = H * (fm * (ru(a, j, nz - 1) + ru(b, j, nz - 1)) - fp * (ru(a, j, nz - 2) + ru(b, j, nz - 2)));
          dru = fm * (ru(b, j, nz - 1) - ru(a, j, nz - 1)) - fp * (ru(b, j, nz - 2) - ru(a, j, nz - 2));
        }
        if (i == 0) {
          const double ub = uw < 0.0 ? uw : 0.0;
          return rdx * (ub * (w2(1, j, k) - w2(0, j, k)) + w2(0, j, k) * dru);
        }
        const double ub = uw > 0.0 ? uw : 0.0;
        return rdx * (ub * (w2(i, j, k) - w2(i - 1, j, k)) + w2(i, j, k) * dru);
      };
      auto open_y_term = [&](int i, int j, int k) {
        const int a = j == 0 ? 0 : ny - 1, b = j == 0 ? 1 : ny;
        double vw, drv;
        if (k < nz) {
          vw = H * (c.fnm[k] * (rv(i, a, k) + rv(i, b, k)) + c.fnp[k] * (rv(i, a, k - 1) + rv(i, b, k - 1)));
          drv = c.fnm[k] * (rv(i, b, k) - rv(i, a, k)) + c.fnp[k] * (rv(i, b, k - 1) - rv(i, a, k - 1));
        } else {
          const double fm = 2.0 - c.fnm[nz - 1], fp = c.fnp[nz - 1];
          vw = H * (fm * (rv(i, a, nz - 1) + rv(i, b, nz - 1)) - fp * (rv(i, a, nz - 2) + rv(i, b, nz - 2)));
          drv = fm * (rv(i, b, nz - 1) - rv(i, a, nz - 1)) - fp * (rv(i, b, nz - 2) - rv(i, a, nz - 2));
        }
        if (j == 0) {
          const double vb = vw < 0.0 ? vw : 0.0;
          return rdy * (vb * (w2(i, 1, k) - w2(i, 0, k)) + w2(i, 0, k) * drv);
        }
        const double vb = vw > 0.0 ? vw : 0.0;
        return rdy * (vb * (w2(i, j, k) - w2(i, j - 1, k)) + w2(i, j, k) * drv);
      };
      pfor(0, nx, 0, ny, 1, nz + 1, [&](int i, int j, int k) {
        double t;
        const bool ocol = ox && (i == 0 || i == nx - 1), orow = oy && (j == 0 || j == ny - 1);
        if (ocol || orow) {  // WRF order: y flux, x flux, open x terms, open y terms
          t = 0.0;
          if (do_y && !orow) t = t - (msftx(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
          if (!ocol) t = t - (msftx(i, j) * rdx) * (fx(i + 1, j, k) - fx(i, j, k));
          if (ocol) t = t - open_x_term(i, j, k);
          if (orow) t = t - open_y_term(i, j, k);
        } else {
          t = -(msftx(i, j) * rdx) * (fx(i + 1, j, k) - fx(i, j, k));
          if (do_y) t -= (msftx(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
        }
        const double rmsft = 1.0 / msfty(i, j);
        if (k < nz) {
          t -= c.rdn[k] * (fz(i, j, k) - fz(i, j, k - 1));
          const double cqw_ = mc_cqw(i, j, k), cq1 = 1.0 / (1.0 + cqw_), cq2 = cqw_ * cq1;
          t += rmsft * kG * (cq1 * c.rdn[k] * (pp(i, j, k) - pp(i, j, k - 1)) - c.c1f[k] * mu2(i, j) -
                             cq2 * (c.c1f[k] * mub(i, j) + c.c2f[k]));
          if (wd_dt != 0.0) {
            const double muf = c.c1f[k] * mut(i, j) + c.c2f[k];
            double cfl = ww(i, j, k) / muf * c.rdnw[k] * wd_dt;
            cfl = cfl < 0.0 ? -cfl : cfl;
            if (cfl > 1.0) t = t - (w2(i, j, k) < 0.0 ? -1.0 : 1.0) * 0.3 * (cfl - w_crit) * muf;
          }
          const double ruw =
              H * (c.fnm[k] * (ru(i, j, k) + ru(i + 1, j, k)) + c.fnp[k] * (ru(i, j, k - 1) + ru(i + 1, j, k - 1)));
          const double rvw =
              H * (c.fnm[k] * (rv(i, j, k) + rv(i, j + 1, k)) + c.fnp[k] * (rv(i, j, k - 1) + rv(i, j + 1, k - 1)));
          const double rmt = msftx(i, j) / msfty(i, j);
          if (cor) t = t + ecor(i, j) * (cosa(i, j) * ruw - rmt * sina(i, j) * rvw);
          const double uw =
              H * (c.fnm[k] * (u2(i, j, k) + u2(i + 1, j, k)) + c.fnp[k] * (u2(i, j, k - 1) + u2(i + 1, j, k - 1)));
          const double vw =
              H * (c.fnm[k] * (v2(i, j, k) + v2(i, j + 1, k)) + c.fnp[k] * (v2(i, j, k - 1) + v2(i, j + 1, k - 1)));
          t = t + kReradius * (ruw * uw + rmt * rvw * vw);
        } else {
          t += 2.0 * c.rdn[nz - 1] * fz(i, j, nz - 1);
          const double cqw_ = mc_cqw(i, j, nz - 1), cq1 = 1.0 / (1.0 + cqw_), cq2 = cqw_ * cq1;
          t += rmsft * kG * (cq1 * 2.0 * c.rdnw[nz - 1] * (-pp(i, j, nz - 1)) - c.c1f[nz] * mu2(i, j) -
                             cq2 * (c.c1f[nz] * mub(i, j) + c.c2f[nz]));
        }
        if (use_tendf) t += rw_tendf(i, j, k) / msfty(i, j);
        rw_tend(i, j, k) = t;
      });
    }
    // KTendT (large_step.cpp:502-565)
    {
      auto fx = [&](int i, int j, int k) {
        double vel = ru(i, j, k);
        const int o = ord_face(i, nxo, flx);
        if (o == 2) return H * vel * (t2(i, j, k) + t2(i - 1, j, k));
        if (o == 3) return vel * flux3(t2(i - 2, j, k), t2(i - 1, j, k), t2(i, j, k), t2(i + 1, j, k), vel, hup);
        return vel * flux5(t2(i - 3, j, k), t2(i - 2, j, k), t2(i - 1, j, k), t2(i, j, k), t2(i + 1, j, k), t2(i + 2, j, k),
                           vel, hup);
      };
      auto fy = [&](int i, int j, int k) {
        double vel = rv(i, j, k);
        const int o = ord_face(j, nyo, fly);
        if (o == 2) return H * vel * (t2(i, j, k) + t2(i, j - 1, k));
        if (o == 3) return vel * flux3(t2(i, j - 2, k), t2(i, j - 1, k), t2(i, j, k), t2(i, j + 1, k), vel, hup);
        return vel * flux5(t2(i, j - 3, k), t2(i, j - 2, k), t2(i, j - 1, k), t2(i, j, k), t2(i, j + 1, k), t2(i, j + 2, k),
                           vel, hup);
      };
      auto fz = [&](int i, int j, int k) {
        if (k == 0 || k == nz) return 0.0;
        double vel = ww(i, j, k);
        if (k == 1 || k == nz - 1) return vel * (c.fnm[k] * t2(i, j, k) + c.fnp[k] * t2(i, j, k - 1));
        return vel * flux3(t2(i, j, k - 2), t2(i, j, k - 1), t2(i, j, k), t2(i, j, k + 1), -vel, vup);
      };
      pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
        double r;
        const bool ocol = ox && (i == 0 || i == nx - 1), orow = oy && (j == 0 || j == ny - 1);
        if (ocol || orow) {  // open boundary: WRF order y flux, x flux, open x terms, open y terms, vertical
          r = 0.0;
          if (do_y && !orow) r = r - (msftx(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
          if (!ocol) r = r - (msftx(i, j) * rdx) * (fx(i + 1, j, k) - fx(i, j, k));
          if (ocol) r = r - open_x_scalar(t2, t2, ru, rdx, i, j, k, nx);
          if (orow) r = r - open_y_scalar(t2, t2, rv, rdy, i, j, k, ny);
        } else {
          r = -(msftx(i, j) * rdx) * (fx(i + 1, j, k) - fx(i, j, k));
          if (do_y) r -= (msftx(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
        }
        r -= c.rdnw[k] * (fz(i, j, k + 1) - fz(i, j, k));
        if (want_thf) {  // set_tend(RTHFTEN): the advection times msfty; specified edges skip their edge flux
          const bool scol = spec && (i == 0 || i == nx - 1), srow = spec && (j == 0 || j == ny - 1);
          if (!scol && !srow) {
            thf(i, j, k) = r * msfty(i, j);
          } else {
            double q = 0.0;
            if (!scol) q = -(msftx(i, j) * rdx) * (fx(i + 1, j, k) - fx(i, j, k));
            if (do_y && !srow) q -= (msftx(i, j) * rdy) * (fy(i, j + 1, k) - fy(i, j, k));
            q -= c.rdnw[k] * (fz(i, j, k + 1) - fz(i, j, k));
            thf(i, j, k) = q * msfty(i, j);
          }
        }
        if (use_tendf) r += t_tendf(i, j, k) / msfty(i, j);
        // rk_addtend_dry's coupled heating (large_step.cpp:557), with moisture
        if (moist()) r += (c.c1h[k] * mut(i, j) + c.c2h[k]) * hdia(i, j, k) / msfty(i, j);
        t_tend(i, j, k) = r;
      });
    }
    // KTendPh rhs_ph (large_step.cpp:570-668), interior (no specified/open edges)
    {
      auto wdwn = [&](int i, int j, int m) {
        return H * (ww(i, j, m + 1) + ww(i, j, m)) * c.rdnw[m] *
               (ph2(i, j, m + 1) - ph2(i, j, m) + phb(i, j, m + 1) - phb(i, j, m));
      };
      auto grad6x = [&](int i, int j, int k) {
        return 45.0 * (ph2(i + 1, j, k) - ph2(i - 1, j, k)) - 9.0 * (ph2(i + 2, j, k) - ph2(i - 2, j, k)) +
               (ph2(i + 3, j, k) - ph2(i - 3, j, k)) + 45.0 * (phb(i + 1, j, k) - phb(i - 1, j, k)) -
               9.0 * (phb(i + 2, j, k) - phb(i - 2, j, k)) + (phb(i + 3, j, k) - phb(i - 3, j, k));
      };
      auto grad6y = [&](int i, int j, int k) {
        return 45.0 * (ph2(i, j + 1, k) - ph2(i, j - 1, k)) - 9.0 * (ph2(i, j + 2, k) - ph2(i, j - 2, k)) +
               (ph2(i, j + 3, k) - ph2(i, j - 3, k)) + 45.0 * (phb(i, j + 1, k) - phb(i, j - 1, k)) -
               9.0 * (phb(i, j + 2, k) - phb(i, j - 2, k)) + (phb(i, j + 3, k) - phb(i, j - 3, k));
      };
      auto grad4x = [&](int i, int j, int k) {
        return 8.0 * (ph2(i + 1, j, k) - ph2(i - 1, j, k)) - (ph2(i + 2, j, k) - ph2(i - 2, j, k)) +
               8.0 * (phb(i + 1, j, k) - phb(i - 1, j, k)) - (phb(i + 2, j, k) - phb(i - 2, j, k));
      };
      auto grad4y = [&](int i, int j, int k) {
        return 8.0 * (ph2(i, j + 1, k) - ph2(i, j - 1, k)) - (ph2(i, j + 2, k) - ph2(i, j - 2, k)) +
               8.0 * (phb(i, j + 1, k) - phb(i, j - 1, k)) - (phb(i, j + 2, k) - phb(i, j - 2, k));
      };
      auto dfwdx = [&](int i, int j, int k) { return phb(i + 1, j, k) - phb(i, j, k) + ph2(i + 1, j, k) - ph2(i, j, k); };
      auto dfwdy = [&](int i, int j, int k) { return phb(i, j + 1, k) - phb(i, j, k) + ph2(i, j + 1, k) - ph2(i, j, k); };
      // rhs_ph next to specified edges (large_step.cpp:596-605): none on the outer row, 2nd order on the
      // next, then 4th order in y (none in x), 6th from the fourth row in
      auto horder = [&](int i, int n, bool on, bool is_x) {
        if (!on) return 6;
        if (i <= 0 || i >= n - 1) return 0;
        if (i == 1 || i == n - 2) return 2;
        if (i == 2 || i == n - 3) return (is_x && !ox) ? 0 : 4;  // open x: 4th order (open_xs/xe only)
        return 6;
      };
      pfor(0, nx, 0, ny, 1, nz + 1, [&](int i, int j, int k) {
        double t = 0;
        if (k < nz) t = t - (c.fnm[k] * wdwn(i, j, k) + c.fnp[k] * wdwn(i, j, k - 1));
        t = t + (c.c1f[k] * mut(i, j) + c.c2f[k]) * kG * w2(i, j, k) / msfty(i, j);
        const double r60 = 1.0 / 60.0, r12 = 1.0 / 12.0;
        if (do_y) {
          const double mvp = c.c1f[k] * muv(i, j + 1) + c.c2f[k], mvm = c.c1f[k] * muv(i, j) + c.c2f[k];
          const double vp =
              k < nz ? v2(i, j + 1, k) + v2(i, j + 1, k - 1) : c.cfn * v2(i, j + 1, nz - 1) + c.cfn1 * v2(i, j + 1, nz - 2);
          const double vm = k < nz ? v2(i, j, k) + v2(i, j, k - 1) : c.cfn * v2(i, j, nz - 1) + c.cfn1 * v2(i, j, nz - 2);
          const double cy = (k < nz ? 0.25 : H) * rdy / msfty(i, j);
          const int o = horder(j, ny, fly, false);
          if (o == 6)
            t = t - cy * ((mvp * vp * msfvy(i, j + 1) + mvm * vm * msfvy(i, j)) * r60 * grad6y(i, j, k));
          else if (o == 4)
            t = t - cy * ((mvp * vp * msfvy(i, j + 1) + mvm * vm * msfvy(i, j)) * r12 * grad4y(i, j, k));
          else if (o == 2)
            t = t - cy * (mvp * vp * msfvy(i, j + 1) * dfwdy(i, j, k) + mvm * vm * msfvy(i, j) * dfwdy(i, j - 1, k));
        }
        const double mup = c.c1f[k] * muu(i + 1, j) + c.c2f[k], mum = c.c1f[k] * muu(i, j) + c.c2f[k];
        const double up =
            k < nz ? u2(i + 1, j, k) + u2(i + 1, j, k - 1) : c.cfn * u2(i + 1, j, nz - 1) + c.cfn1 * u2(i + 1, j, nz - 2);
        const double um = k < nz ? u2(i, j, k) + u2(i, j, k - 1) : c.cfn * u2(i, j, nz - 1) + c.cfn1 * u2(i, j, nz - 2);
        const double cx = (k < nz ? 0.25 : H) * rdx / msfty(i, j);
        const int o = horder(i, nx, flx, true);
        if (o == 6)
          t = t - cx * ((mup * up * msfux(i + 1, j) + mum * um * msfux(i, j)) * r60 * grad6x(i, j, k));
        else if (o == 2)
          t = t - cx * (mup * up * msfux(i + 1, j) * dfwdx(i, j, k) + mum * um * msfux(i, j) * dfwdx(i - 1, j, k));
        else if (o == 4)
          t = t - cx * ((mup * up * msfux(i + 1, j) + mum * um * msfux(i, j)) * r12 * grad4x(i, j, k));
        // open boundaries (large_step.cpp:636-661), south/north then west/east: inflow advection of ph by the
        // normal wind at the w level
        if (oy && (j == 0 || j == ny - 1)) {
          const int kz = k < nz - 1 ? k : nz - 1;
          const int a = j == 0 ? 0 : ny - 1, b = j == 0 ? 1 : ny;
          const double vb = H * (c.fnm[kz] * (v2(i, b, kz) + v2(i, a, kz)) + c.fnp[kz] * (v2(i, b, kz - 1) + v2(i, a, kz - 1)));
          const double muf = c.c1f[k] * mut(i, j) + c.c2f[k];
          if (j == 0) {
            const double vl = vb < 0.0 ? vb : 0.0;
            t = t - rdy * muf * (vl * (ph2(i, 1, k) - ph2(i, 0, k)));
          } else {
            const double vr = vb > 0.0 ? vb : 0.0;
            t = t - rdy * muf * (vr * (ph2(i, j, k) - ph2(i, j - 1, k)));
          }
        }
        if (ox && (i == 0 || i == nx - 1) && !(i == 0 && k == nz)) {
          const int kz = k < nz ? k : nz - 1;
          const int a = i == 0 ? 0 : nx - 1, b = i == 0 ? 1 : nx;
          const double ub = H * (c.fnm[kz] * (u2(b, j, kz) + u2(a, j, kz)) + c.fnp[kz] * (u2(b, j, kz - 1) + u2(a, j, kz - 1)));
          const double muf = c.c1f[k] * mut(i, j) + c.c2f[k];
          if (i == 0) {
            const double ul = ub < 0.0 ? ub : 0.0;
            t = t - (msftx(i, j) / msfty(i, j)) * rdx * muf * (ul * (ph2(1, j, k) - ph2(0, j, k)));
          } else {
            const double ur = ub > 0.0 ? ub : 0.0;
            t = t - (msftx(i, j) / msfty(i, j)) * rdx * muf * (ur * (ph2(i, j, k) - ph2(i - 1, j, k)));
          }
        }
        if (spec) t = t + ph_tendf(i, j, k) / msfty(i, j);  // rk_addtend_dry with a LateralBC (large_step.cpp:663)
        ph_tend(i, j, k) = t;
      });
    }
    if (spec) spec_dry();  // spec_bdy_dry (large_step.cpp:859)
  }

  // ------------------------------------------------------------------ SmallStep::run_stage (small_step.cpp:490-607)
  // p_rho_point (small_step.cpp:97-114)
  void p_rho_point(int i, int j, int k, double smdiv, bool damp) {
    const double m = c.c1h[k] * muts(i, j) + c.c2h[k];
    const double a = -((1. / m) * (alt(i, j, k) * (c.c1h[k] * mu2(i, j)) + c.rdnw[k] * (ph2(i, j, k + 1) - ph2(i, j, k))));
    al(i, j, k) = a;
    double p_ = c2a(i, j, k) *
                (alt(i, j, k) * (t2(i, j, k) - c.c1h[k] * mu2(i, j) * t_save(i, j, k)) / (m * (kT0 + t_save(i, j, k))) - a);
    if (damp) {
      const double ptmp = p_;
      p_ = p_ + smdiv * (p_ - pm1(i, j, k));
      pm1(i, j, k) = ptmp;
    } else {
      pm1(i, j, k) = p_;
    }
    pp(i, j, k) = p_;
  }

  void small_steps(int rk) {
    const int nx = g.nx, ny = g.ny, nz = g.nz, nyv_ = nyv();
    const bool do_y = !g.two_d();
    const double dt = g.dt, dts_full = dt / n_sound;
    const double dt_rk = rk == 1 ? dt / 3.0 : (rk == 2 ? 0.5 * dt : dt);
    const double dts = rk == 1 ? dt_rk : dts_full;
    const int ns = rk == 1 ? 1 : (rk == 2 ? n_sound / 2 : n_sound);
    const double epssm = p.epssm;
    if (rk == 1) std::fill(mudf.a.begin(), mudf.a.end(), 0.0);
    // KPrep small_step_prep (small_step.cpp:39-93)
    const bool first = rk == 1;
    pfor(0, nx + 1, 0, nyv_, 0, nz + 1, [&](int i, int j, int k) {
      const double mts = mub(i, j) + mu1(i, j);
      const double mus_u = first ? muu(i, j) : H * (mts + mub(i - 1, j) + mu1(i - 1, j));
      const double mus_v = first ? muv(i, j) : H * (mts + mub(i, j - 1) + mu1(i, j - 1));
      if (k == 0) {
        if (i < nx && j < ny) { muts(i, j) = mts; mu_save(i, j) = mu2(i, j); mu2(i, j) = mu1(i, j) - mu2(i, j); }
        if (j < ny) muus(i, j) = mus_u;
        if (i < nx && j < nyv_) muvs(i, j) = mus_v;
      }
      if (j < ny && k < nz) {
        const double ui = u2(i, j, k);
        u_save(i, j, k) = ui;
        u2(i, j, k) = ((c.c1h[k] * mus_u + c.c2h[k]) * u1(i, j, k) - (c.c1h[k] * muu(i, j) + c.c2h[k]) * ui) / msfuy(i, j);
      }
      if (i < nx && j < nyv_ && k < nz) {
        const double vi = v2(i, j, k);
        v_save(i, j, k) = vi;
        v2(i, j, k) =
            ((c.c1h[k] * mus_v + c.c2h[k]) * v1(i, j, k) - (c.c1h[k] * muv(i, j) + c.c2h[k]) * vi) * msfvx_inv(i, j);
      }
      if (i < nx && j < ny) {
        if (k < nz) {
          const double ti = t2(i, j, k);
          t_save(i, j, k) = ti;
          t2(i, j, k) = (c.c1h[k] * mts + c.c2h[k]) * t1(i, j, k) - (c.c1h[k] * mut(i, j) + c.c2h[k]) * ti;
          c2a(i, j, k) = kCpovcv * (pb(i, j, k) + pp(i, j, k)) / alt(i, j, k);
        }
        const double wi = w2(i, j, k), phi = ph2(i, j, k);
        w_save(i, j, k) = wi;
        w2(i, j, k) = ((c.c1f[k] * mts + c.c2f[k]) * w1(i, j, k) - (c.c1f[k] * mut(i, j) + c.c2f[k]) * wi) / msfty(i, j);
        ph_save(i, j, k) = phi;
        ph2(i, j, k) = ph1(i, j, k) - phi;
        ww1(i, j, k) = (k == 0 || k == nz) ? 0.0 : ww(i, j, k);
      }
    });
    for (F* f : {&u2, &v2, &ph2, &t_save, &u_save, &v_save}) fill(*f, g);
    for (F* f : {&mu2, &muts}) fill(*f, g);
    // KPRho calc_p_rho (small_step.cpp:116-121)
    pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) { p_rho_point(i, j, k, 0.0, false); });
    for (F* f : {&pp, &al}) fill(*f, g);
    // KCoefW calc_coef_w (small_step.cpp:124-148), dry (cq1 = 1)
    const double cof_h = 0.5 * dts * kG * (1.0 + epssm);
    const double cof = cof_h * cof_h;
    const double lid = p.top_lid ? 0.0 : 1.0;
    cfor(nx, ny, [&](int i, int j) {
      auto mh = [&](int k, double m) { return c.c1h[k] * m + c.c2h[k]; };
      auto mf = [&](int k, double m) { return c.c1f[k] * m + c.c2f[k]; };
      const double m = mut(i, j);
      ca(i, j, 0) = 0; calpha(i, j, 0) = 0; cgamma(i, j, 0) = 0;
      ca(i, j, 1) = 0;
      ca(i, j, nz) = -2.0 * cof * (c.rdnw[nz - 1] * c.rdnw[nz - 1]) * c2a(i, j, nz - 1) * lid / (mh(nz - 1, m) * mf(nz - 1, m));
      for (int k = 2; k < nz; ++k)  // c1f at the level below, as the reference (small_step.cpp:132-133)
        ca(i, j, k) = -mc_cq1(i, j, k) * cof * c.rdn[k] * c.rdnw[k - 1] * c2a(i, j, k - 1) / (mh(k - 1, m) * mf(k - 1, m));
      for (int k = 1; k < nz; ++k) {
        const double cq1 = mc_cq1(i, j, k);
        const double b = 1.0 + cq1 * cof * c.rdn[k] *
                                   (c.rdnw[k] * c2a(i, j, k) / (mh(k, m) * mf(k, m)) +
                                    c.rdnw[k - 1] * c2a(i, j, k - 1) / (mh(k - 1, m) * mf(k, m)));
        const double cc = -cq1 * cof * c.rdn[k] * c.rdnw[k] * c2a(i, j, k) / (mh(k, m) * mf(k + 1, m));
        calpha(i, j, k) = 1.0 / (b - ca(i, j, k) * cgamma(i, j, k - 1));
        cgamma(i, j, k) = cc * calpha(i, j, k);
      }
      const double b = 1.0 + 2.0 * cof * (c.rdnw[nz - 1] * c.rdnw[nz - 1]) * c2a(i, j, nz - 1) / (mh(nz - 1, m) * mf(nz, m));
      calpha(i, j, nz) = 1.0 / (b - ca(i, j, nz) * cgamma(i, j, nz - 1));
      cgamma(i, j, nz) = 0;
    });

    const double emdx = p.emdiv * (1. / c.rdx), emdy = do_y ? p.emdiv * (1. / c.rdy) : 0.0;
    const int ui0 = g.u_i0(), ui1 = nx, vj0 = g.v_j0(), vj1 = ny;
    for (int it = 1; it <= ns; ++it) {
      // KUV advance_uv (small_step.cpp:151-190)
      auto dpn = [&](int di, int dj, int i, int j, int k) {
        auto ps = [&](int kk) { return pp(i, j, kk) + pp(i - di, j - dj, kk); };
        if (k == 0) return H * (c.cf1 * ps(0) + c.cf2 * ps(1) + c.cf3 * ps(2));
        if (k == nz) return p.top_lid ? H * (c.cf1 * ps(nz - 1) + c.cf2 * ps(nz - 2) + c.cf3 * ps(nz - 3)) : 0.0;
        return H * (c.fnm[k] * ps(k) + c.fnp[k] * ps(k - 1));
      };
      auto dp = [&](int di, int dj, int i, int j, int k, double mface, double rd, double rm) {
        double d = rm * H * rd * (c.c1h[k] * mface + c.c2h[k]) *
                   (((ph2(i, j, k + 1) - ph2(i - di, j - dj, k + 1)) + (ph2(i, j, k) - ph2(i - di, j - dj, k))) +
                    (alt(i, j, k) + alt(i - di, j - dj, k)) * (pp(i, j, k) - pp(i - di, j - dj, k)) +
                    (al(i, j, k) + al(i - di, j - dj, k)) * (pb(i, j, k) - pb(i - di, j - dj, k)));
        d += rm * rd * (php(i, j, k) - php(i - di, j - dj, k)) *
             (c.rdnw[k] * (dpn(di, dj, i, j, k + 1) - dpn(di, dj, i, j, k)) -
              H * (c.c1h[k] * mu2(i - di, j - dj) + c.c1h[k] * mu2(i, j)));
        return d;
      };
      // Edge (small_step.cpp:39-45): the spec zone's faces are advanced by spec_bdyupdate instead
      auto uface = [&](int i, int j) { return spec && (i <= 0 || i >= nx || j == 0 || j == ny - 1); };
      auto vface = [&](int i, int j) { return spec && (i == 0 || i == nx - 1 || j <= 0 || j >= ny); };
      // open boundaries (small_step.cpp:175-184, 531): the boundary faces take the large-step tendency only
      pfor(0, nx + (ox ? 1 : 0), 0, ny + (oy ? 1 : 0), 0, nz, [&](int i, int j, int k) {
        if (ox && (i == 0 || i == nx) && j < ny) {
          u2(i, j, k) = u2(i, j, k) + dts * ru_tend(i, j, k);
        } else if (i >= ui0 && i < ui1 && !uface(i, j) && j < ny) {
          const double mudf_x = -emdx * (mudf(i, j) - mudf(i - 1, j)) / msfuy(i, j);
          u2(i, j, k) = u2(i, j, k) + dts * ru_tend(i, j, k) -
                        dts * mc_cqu(i, j, k) * dp(1, 0, i, j, k, muu(i, j), c.rdx, msfux(i, j) / msfuy(i, j)) + c.c1h[k] * mudf_x;
        }
        if (do_y && oy && (j == 0 || j == ny) && i < nx) {
          v2(i, j, k) = v2(i, j, k) + dts * rv_tend(i, j, k);
        } else if (do_y && j >= vj0 && j < vj1 && !vface(i, j) && i < nx) {
          const double mudf_y = -emdy * (mudf(i, j) - mudf(i, j - 1)) * msfvx_inv(i, j);
          v2(i, j, k) = v2(i, j, k) + dts * rv_tend(i, j, k) -
                        dts * mc_cqv(i, j, k) * dp(0, 1, i, j, k, muv(i, j), c.rdy, msfvy(i, j) / msfvx(i, j)) + c.c1h[k] * mudf_y;
        }
      });
      if (spec) {  // LateralBC::update_uv (lateral_bc.cpp:356-364)
        spec_update('u', nz, u2, ru_tend, dts);
        spec_update('v', nz, v2, rv_tend, dts);
      }
      if (do_y) { fill(u2, g); fill(v2, g); }
      else fill(u2, g);
      // KMuTDiv (small_step.cpp:194-207)
      pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
        const double fx = (u2(i + 1, j, k) + (c.c1h[k] * muu(i + 1, j) + c.c2h[k]) * u_save(i + 1, j, k) / msfuy(i + 1, j)) -
                          (u2(i, j, k) + (c.c1h[k] * muu(i, j) + c.c2h[k]) * u_save(i, j, k) / msfuy(i, j));
        double d = c.rdx * fx;
        if (do_y)
          d += c.rdy * ((v2(i, j + 1, k) + (c.c1h[k] * muv(i, j + 1) + c.c2h[k]) * v_save(i, j + 1, k) * msfvx_inv(i, j + 1)) -
                        (v2(i, j, k) + (c.c1h[k] * muv(i, j) + c.c2h[k]) * v_save(i, j, k) * msfvx_inv(i, j)));
        dvdxi(i, j, k) = msftx(i, j) * msfty(i, j) * d;
      });
      // KMuTCol (small_step.cpp:209-248)
      const bool firstit = it == 1, last = it == ns;
      const double rn = double(ns);
      auto ring0 = [&](int i, int j) { return spec && (i == 0 || i == nx - 1 || j == 0 || j == ny - 1); };
      cfor(nx, ny, [&](int i, int j) {
        if (ring0(i, j)) {  // spec zone: no advance_mu_t; sumflux still accumulates ww (small_step.cpp:214-220)
          for (int k = 0; k <= nz; ++k) {
            const double sw = (firstit ? 0.0 : ww_m(i, j, k)) + ww(i, j, k);
            ww_m(i, j, k) = last ? sw / rn + ww1(i, j, k) : sw;
          }
          return;
        }
        double dmdt = 0;
        for (int k = 0; k < nz; ++k) dmdt += c.dnw[k] * dvdxi(i, j, k);
        const double mtend = mu_tend(i, j);
        const double mu_old = mu2(i, j);
        const double mu_new = mu_old + dts * (dmdt + mtend);
        mu2(i, j) = mu_new;
        mudf(i, j) = dmdt + mtend;
        muts(i, j) = mut(i, j) + mu_new;
        muave(i, j) = H * ((1.0 + epssm) * mu_new + (1.0 - epssm) * mu_old);
        auto acc_ww = [&](int k, double wpp) {
          const double sw = (firstit ? 0.0 : ww_m(i, j, k)) + wpp;
          ww_m(i, j, k) = last ? sw / rn + ww1(i, j, k) : sw;
        };
        double ww_full = 0;
        ww(i, j, 0) = 0;
        acc_ww(0, 0.0);
        for (int k = 0; k < nz; ++k) {
          double wpp_hi = 0;
          if (k + 1 < nz) {
            ww_full = ww_full - c.dnw[k] * (c.c1h[k] * dmdt + dvdxi(i, j, k) + c.c1h[k] * mtend) / msfty(i, j);
            wpp_hi = ww_full - ww1(i, j, k + 1);
            ww(i, j, k + 1) = wpp_hi;
          } else {
            ww(i, j, nz) = 0;
          }
          acc_ww(k + 1, wpp_hi);
        }
      });
      // KMuTPoint (small_step.cpp:250-276)
      pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
        auto wdtn = [&](int kk) {
          if (kk == 0 || kk == nz) return 0.0;
          return ww(i, j, kk) * (c.fnm[kk] * t_save(i, j, kk) + c.fnp[kk] * t_save(i, j, kk - 1));
        };
        const double su = (firstit ? 0.0 : ru_m(i, j, k)) + u2(i, j, k);
        ru_m(i, j, k) = last ? su / rn + (c.c1h[k] * muu(i, j) + c.c2h[k]) * u_save(i, j, k) / msfuy(i, j) : su;
        if (do_y) {
          const double sv = (firstit ? 0.0 : rv_m(i, j, k)) + v2(i, j, k);
          rv_m(i, j, k) = last ? sv / rn + (c.c1h[k] * muv(i, j) + c.c2h[k]) * v_save(i, j, k) * msfvx_inv(i, j) : sv;
        }
        const double tk = t2(i, j, k);
        t_2save(i, j, k) = tk;
        if (ring0(i, j)) return;  // spec zone: theta'' by spec_bdyupdate
        double hflux = H * c.rdx *
                       (u2(i + 1, j, k) * (t_save(i + 1, j, k) + t_save(i, j, k)) -
                        u2(i, j, k) * (t_save(i, j, k) + t_save(i - 1, j, k)));
        if (do_y)
          hflux += H * c.rdy *
                   (v2(i, j + 1, k) * (t_save(i, j + 1, k) + t_save(i, j, k)) -
                    v2(i, j, k) * (t_save(i, j, k) + t_save(i, j - 1, k)));
        const double tn = tk + msfty(i, j) * dts * t_tend(i, j, k);
        t2(i, j, k) = tn - dts * msfty(i, j) * (msftx(i, j) * hflux + c.rdnw[k] * (wdtn(k + 1) - wdtn(k)));
      });
      if (spec || ox || oy) {
        // KSumEdge (small_step.cpp:279-290, 560-569): the east and north domain faces are real faces
        for (int k = 0; k < nz; ++k) {
          if (spec || ox)
          for (int j = 0; j < ny; ++j) {
            const double su = (firstit ? 0.0 : ru_m(nx, j, k)) + u2(nx, j, k);
            const double cpl = (c.c1h[k] * muu(nx, j) + c.c2h[k]) * u_save(nx, j, k);
            ru_m(nx, j, k) = last ? su / rn + cpl / msfuy(nx, j) : su;
          }
          if (spec || oy)
          for (int i = 0; i < nx; ++i) {
            const double sv = (firstit ? 0.0 : rv_m(i, ny, k)) + v2(i, ny, k);
            const double cpl = (c.c1h[k] * muv(i, ny) + c.c2h[k]) * v_save(i, ny, k);
            rv_m(i, ny, k) = last ? sv / rn + cpl * msfvx_inv(i, ny) : sv;
          }
        }
      }
      if (spec) {
        // LateralBC::update_mu_t (lateral_bc.cpp:366-372)
        spec_update('t', nz, t2, t_tend, dts);
        spec_update('t', 1, mu2, mu_tend, dts);
        spec_update('t', 1, muts, mu_tend, dts);
      }
      for (F* f : {&mu2, &mudf, &muts}) fill(*f, g);
      // advance_w: KWRhs, KWExpl, KWSolve, KWPRho (small_step.cpp:292-416)
      auto mh = [&](int k, double m) { return c.c1h[k] * m + c.c2h[k]; };
      auto mf = [&](int k, double m) { return c.c1f[k] * m + c.c2f[k]; };
      pfor(0, nx, 0, ny, 0, nz + 1, [&](int i, int j, int k) {
        if (ring0(i, j)) return;  // KWRhs: not in the spec zone
        // KWCommon::rhs (small_step.cpp:312-319)
        double R;
        if (k == 0) R = 0;
        else if (p.top_lid && k == nz) R = 0;
        else {
          auto wdwn = [&](int m) {
            return H * (ww(i, j, m) + ww(i, j, m - 1)) * c.rdnw[m - 1] *
                   (ph_save(i, j, m) - ph_save(i, j, m - 1) + phb(i, j, m) - phb(i, j, m - 1));
          };
          double r = dts * (ph_tend(i, j, k) + H * kG * (1.0 - epssm) * w2(i, j, k));
          if (k < nz) r -= dts * (c.fnm[k] * wdwn(k + 1) + c.fnp[k] * wdwn(k));
          R = ph2(i, j, k) + msfty(i, j) * r / mf(k, mut(i, j));
        }
        w_rhs(i, j, k) = R;
        // KWCommon::t2ave (small_step.cpp:302-305)
        if (k < nz) {
          const double tav = H * ((1.0 + epssm) * t2(i, j, k) + (1.0 - epssm) * t_2save(i, j, k));
          t2a(i, j, k) = (tav + c.c1h[k] * muave(i, j) * kT0) / (mh(k, muts(i, j)) * (kT0 + t_save(i, j, k)));
        }
      });
      pfor(0, nx, 0, ny, 0, nz + 1, [&](int i, int j, int k) {  // KWExpl (small_step.cpp:331-378)
        if (ring0(i, j)) return;  // spec zone: zero_grad_bdy
        if (k == 0) {  // the surface w'' from the small-step u'', v'' and the terrain slope (small_step.cpp:339-353)
          double ws = 0;
          if (terrain) {
            ws = msftx(i, j) * H * c.rdx *
                 ((ht(i + 1, j) - ht(i, j)) * (c.cf1 * u2(i + 1, j, 0) + c.cf2 * u2(i + 1, j, 1) + c.cf3 * u2(i + 1, j, 2)) +
                  (ht(i, j) - ht(i - 1, j)) * (c.cf1 * u2(i, j, 0) + c.cf2 * u2(i, j, 1) + c.cf3 * u2(i, j, 2)));
            if (do_y)
              ws = msfty(i, j) * H * c.rdy *
                       ((ht(i, j + 1) - ht(i, j)) * (c.cf1 * v2(i, j + 1, 0) + c.cf2 * v2(i, j + 1, 1) + c.cf3 * v2(i, j + 1, 2)) +
                        (ht(i, j) - ht(i, j - 1)) * (c.cf1 * v2(i, j, 0) + c.cf2 * v2(i, j, 1) + c.cf3 * v2(i, j, 2))) +
                   ws;
          }
          w2(i, j, 0) = ws;
          return;
        }
        const double m = mut(i, j), mav = muave(i, j);
        const double ope = 1.0 + epssm, ome = 1.0 - epssm;
        const double msft_inv = 1.0 / msfty(i, j);
        const double r_k = w_rhs(i, j, k), r_km1 = w_rhs(i, j, k - 1), ph_k = ph2(i, j, k), ph_km1 = ph2(i, j, k - 1);
        if (k < nz) {
          const double r_kp1 = w_rhs(i, j, k + 1), ph_kp1 = ph2(i, j, k + 1);
          w2(i, j, k) = w2(i, j, k) + dts * rw_tend(i, j, k) +
                        msft_inv * mc_cq1(i, j, k) *
                            (H * dts * kG * c.rdn[k] *
                             (c2a(i, j, k) * c.rdnw[k] / mh(k, m) * (ope * (r_kp1 - r_k) + ome * (ph_kp1 - ph_k)) -
                              c2a(i, j, k - 1) * c.rdnw[k - 1] / mh(k - 1, m) * (ope * (r_k - r_km1) + ome * (ph_k - ph_km1)))) +
                        dts * kG * msft_inv *
                            (c.rdn[k] * (c2a(i, j, k) * alt(i, j, k) * t2a(i, j, k) - c2a(i, j, k - 1) * alt(i, j, k - 1) * t2a(i, j, k - 1)) -
                             c.c1f[k] * mav);
        } else {
          const int km = nz - 1;
          double wn = w2(i, j, k) + dts * rw_tend(i, j, k) +
                      msft_inv * (-H * dts * kG / mh(km, m) * (c.rdnw[km] * c.rdnw[km]) * 2.0 * c2a(i, j, km) *
                                      (ope * (r_k - r_km1) + ome * (ph_k - ph_km1)) -
                                  dts * kG * (2.0 * c.rdnw[km] * c2a(i, j, km) * alt(i, j, km) * t2a(i, j, km) + c.c1f[k] * mav));
          if (p.top_lid) wn = 0;
          w2(i, j, k) = wn;
        }
      });
      const double dampmag = p.damp_opt == 3 ? dts * p.dampcoef : 0.0, hdepth = p.zdamp;
      cfor(nx, ny, [&](int i, int j) {  // KWSolve (small_step.cpp:381-407)
        if (ring0(i, j)) return;  // spec zone: spec_bdyupdate_ph
        const double m = mut(i, j), mts = muts(i, j), ope = 1.0 + epssm;
        for (int k = 1; k <= nz; ++k) w2(i, j, k) = (w2(i, j, k) - ca(i, j, k) * w2(i, j, k - 1)) * calpha(i, j, k);
        for (int k = nz - 1; k >= 1; --k) w2(i, j, k) = w2(i, j, k) - cgamma(i, j, k) * w2(i, j, k + 1);
        if (dampmag != 0.0) {
          const double pi = 4.0 * std::atan(1.0);
          const double htop = (ph_save(i, j, nz) + phb(i, j, nz)) / kG, hbot = htop - hdepth;
          for (int k = nz; k >= 1; --k) {
            const double hk = (ph_save(i, j, k) + phb(i, j, k)) / kG;
            double dampwt = 0;
            if (hk >= hbot) {
              const double sn = std::sin(H * pi * (hk - hbot) / hdepth);
              dampwt = dampmag * sn * sn;
            }
            w2(i, j, k) = (w2(i, j, k) - dampwt * mf(k, m) * w_save(i, j, k)) / (1.0 + dampwt);
          }
        }
        for (int k = nz; k >= 1; --k)
          ph2(i, j, k) = w_rhs(i, j, k) + msfty(i, j) * H * dts * kG * ope * w2(i, j, k) / mf(k, mts);
      });
      if (spec) update_ph_w(dts);  // spec_bdyupdate_ph, zero_grad_bdy(w) (small_step.cpp:588)
      pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) { p_rho_point(i, j, k, p.smdiv, true); });  // KWPRho
      for (F* f : {&ph2, &pp, &al}) fill(*f, g);
    }
    for (F* f : {&muts, &ru_m, &rv_m, &ww_m}) fill(*f, g);
    // KFinish small_step_finish (small_step.cpp:419-446); hd_dt removes the microphysics heating at
    // the last stage with moisture (small_step.cpp:602)
    const double hd_dt = (rk == 3 && moist()) ? dts * double(ns) : 0.0;
    pfor(0, nx + 1, 0, nyv_, 0, nz + 1, [&](int i, int j, int k) {
      if (j < ny && k < nz) {
        const double mus_u = H * (muts(i, j) + muts(i - 1, j));
        u2(i, j, k) = (msfuy(i, j) * u2(i, j, k) + u_save(i, j, k) * (c.c1h[k] * muu(i, j) + c.c2h[k])) / (c.c1h[k] * mus_u + c.c2h[k]);
      }
      if (i < nx && j < nyv_ && k < nz) {
        const double mus_v = H * (muts(i, j) + muts(i, j - 1));
        v2(i, j, k) = (msfvx(i, j) * v2(i, j, k) + v_save(i, j, k) * (c.c1h[k] * muv(i, j) + c.c2h[k])) / (c.c1h[k] * mus_v + c.c2h[k]);
      }
      if (i < nx && j < ny) {
        const double m = mut(i, j), mts = muts(i, j);
        w2(i, j, k) = (msfty(i, j) * w2(i, j, k) + w_save(i, j, k) * (c.c1f[k] * m + c.c2f[k])) / (c.c1f[k] * mts + c.c2f[k]);
        ph2(i, j, k) = ph2(i, j, k) + ph_save(i, j, k);
        ww(i, j, k) = ww(i, j, k) + ww1(i, j, k);
        if (k < nz)
          t2(i, j, k) = (t2(i, j, k) - (hd_dt != 0.0 ? hd_dt * (c.c1h[k] * m + c.c2h[k]) * hdia(i, j, k) : 0.0) +
                         t_save(i, j, k) * (c.c1h[k] * m + c.c2h[k])) /
                        (c.c1h[k] * mts + c.c2h[k]);
      }
    });
    // KFinishMu (small_step.cpp:448-454)
    pfor(0, nx + 1, 0, nyv_, 0, 1, [&](int i, int j, int) {
      if (i < nx && j < ny) mu2(i, j) = mu2(i, j) + mu_save(i, j);
      if (j < ny) muus(i, j) = H * (muts(i, j) + muts(i - 1, j));
      if (i < nx && j < nyv_) muvs(i, j) = H * (muts(i, j) + muts(i, j - 1));
    });
    for (F* f : {&u2, &v2, &w2, &t2, &ph2, &ww, &mu2}) fill(*f, g);
  }


  // ------------------------------------------------------------------ ScalarTransport (scalars.cpp:295-461)
  // Face fluxes of one species (Fluxes, scalars.cpp:42-140), periodic / symmetric edges: q = q_2 (stage),
  // qo = q_1, the acoustic-step mass fluxes ru_m, rv_m, ww_m and the stage-end column mass muts.
  struct Fl {
    const Dom& d;
    const F &q, &qo;
    double dt, hup, vup;
    bool pd;
    double f5(double a, double b, double cc, double e, double f, double h, double v) const {
      return pd ? flux5_pd(a, b, cc, e, f, h, v, hup) : flux5(a, b, cc, e, f, h, v, hup);
    }
    double f3(double a, double b, double cc, double e, double v, double up) const {
      return pd ? flux3_pd(a, b, cc, e, v, up) : flux3(a, b, cc, e, v, up);
    }
    double fx(int i, int j, int k) const {
      const double vel = d.ru_m(i, j, k);
      const int o = ord_face(i, d.g.nx, d.flx);
      if (o == 2) return H * vel * (q(i, j, k) + q(i - 1, j, k));
      if (o == 3) return vel * f3(q(i - 2, j, k), q(i - 1, j, k), q(i, j, k), q(i + 1, j, k), vel, hup);
      return vel * f5(q(i - 3, j, k), q(i - 2, j, k), q(i - 1, j, k), q(i, j, k), q(i + 1, j, k), q(i + 2, j, k), vel);
    }
    double fy(int i, int j, int k) const {
      const double vel = d.rv_m(i, j, k);
      const int o = ord_face(j, d.g.ny, d.fly);
      if (o == 2) return H * vel * (q(i, j, k) + q(i, j - 1, k));
      if (o == 3) return vel * f3(q(i, j - 2, k), q(i, j - 1, k), q(i, j, k), q(i, j + 1, k), vel, hup);
      return vel * f5(q(i, j - 3, k), q(i, j - 2, k), q(i, j - 1, k), q(i, j, k), q(i, j + 1, k), q(i, j + 2, k), vel);
    }
    double fz(int i, int j, int k) const {
      const int nz = d.g.nz;
      if (k == 0 || k == nz) return 0;
      const double vel = d.ww_m(i, j, k);
      if (k == 1 || k == nz - 1) return vel * (d.c.fnm[k] * q(i, j, k) + d.c.fnp[k] * q(i, j, k - 1));
      return vel * f3(q(i, j, k - 2), q(i, j, k - 1), q(i, j, k), q(i, j, k + 1), -vel, vup);
    }
    double mh(int i, int j, int k) const { return d.c.c1h[k] * d.muts(i, j) + d.c.c2h[k]; }
    double fxl(int i, int j, int k) const {
      const double dx = 2.0 / (d.msfty(i, j) + d.msfty(i - 1, j)) / d.c.rdx, mu = H * (mh(i, j, k) + mh(i - 1, j, k)),
                   vel = d.ru_m(i, j, k);
      // advect_scalar_pd, degrade_xs, i == ids+1 only: vel = ru/mu, cr = vel*dt/dx (scalars.cpp:88-89)
      const double cr = (d.flx && i == 1) ? (vel / mu) * dt / dx : vel * dt / dx / mu;
      return mu * (dx / dt) * flux_upwind(qo(i - 1, j, k), qo(i, j, k), cr);
    }
    double fyl(int i, int j, int k) const {
      const double dy = 2.0 / (d.msftx(i, j) + d.msftx(i, j - 1)) / d.c.rdy, mu = H * (mh(i, j, k) + mh(i, j - 1, k)),
                   vel = d.rv_m(i, j, k);
      const double cr = vel * dt / dy / mu;
      return mu * (dy / dt) * flux_upwind(qo(i, j - 1, k), qo(i, j, k), cr);
    }
    double fzl(int i, int j, int k) const {
      if (k == 0 || k == d.g.nz) return 0;
      const double dz = 2.0 / (d.c.rdnw[k] + d.c.rdnw[k - 1]), mu = mh(i, j, k), vel = d.ww_m(i, j, k);
      const double cr = vel * dt / dz / mu;
      return mu * (dz / dt) * flux_upwind(qo(i, j, k - 1), qo(i, j, k), cr);
    }
    // the spec zone (scalars.cpp:100-103): not advected
    bool spec_cell(int i, int j) const {
      return d.spec && (i == 0 || i == d.g.nx - 1 || j == 0 || j == d.g.ny - 1);
    }
    // the outermost row of a specified or open boundary: advect_scalar_pd does not limit it (scalars.cpp:104-107)
    bool edge_cell(int i, int j) const {
      return (d.flx && (i == 0 || i == d.g.nx - 1)) || (d.fly && (j == 0 || j == d.g.ny - 1));
    }
    bool open_col(int i) const { return d.ox && (i == 0 || i == d.g.nx - 1); }
    // advect_scalar in a spec-zone cell, for RQVFTEN (scalars.cpp:119-125): x flux off in the outer columns, y
    // off in the outer rows
    double adv_spec(int i, int j, int k) const {
      const bool scol = i == 0 || i == d.g.nx - 1, srow = j == 0 || j == d.g.ny - 1;
      double r = 0.0;
      if (!scol) r = -(d.msftx(i, j) * d.c.rdx) * (fx(i + 1, j, k) - fx(i, j, k));
      if (!d.g.two_d() && !srow) r -= (d.msftx(i, j) * d.c.rdy) * (fy(i, j + 1, k) - fy(i, j, k));
      return r - d.c.rdnw[k] * (fz(i, j, k + 1) - fz(i, j, k));
    }
    bool open_row(int j) const { return d.oy && (j == 0 || j == d.g.ny - 1); }
    // the open terms, x then y (scalars.cpp:110-116); fo: q itself (advect_scalar) or q_1 (advect_scalar_pd)
    double open_terms(double r, int i, int j, int k, const F& fo) const {
      if (open_col(i)) r = r - open_x_scalar(q, fo, d.ru_m, d.c.rdx, i, j, k, d.g.nx);
      if (open_row(j)) r = r - open_y_scalar(q, fo, d.rv_m, d.c.rdy, i, j, k, d.g.ny);
      return r;
    }
    double adv(int i, int j, int k) const {  // advect_scalar, horizontal terms times msftx (scalars.cpp:126-139)
      if (spec_cell(i, j)) return 0;
      double r;
      if (open_col(i) || open_row(j)) {  // WRF order: y flux, x flux, open x, open y, vertical
        r = 0.0;
        if (!d.g.two_d() && !open_row(j)) r = r - (d.msftx(i, j) * d.c.rdy) * (fy(i, j + 1, k) - fy(i, j, k));
        if (!open_col(i)) r = r - (d.msftx(i, j) * d.c.rdx) * (fx(i + 1, j, k) - fx(i, j, k));
        r = open_terms(r, i, j, k, q);
      } else {
        r = -(d.msftx(i, j) * d.c.rdx) * (fx(i + 1, j, k) - fx(i, j, k));
        if (!d.g.two_d()) r -= (d.msftx(i, j) * d.c.rdy) * (fy(i, j + 1, k) - fy(i, j, k));
      }
      return r - d.c.rdnw[k] * (fz(i, j, k + 1) - fz(i, j, k));
    }
  };
  double m_full(int k, double mub_, double mu_) const { return c.c1h[k] * (mub_ + mu_) + c.c2h[k]; }  // scalars.cpp:142

  // rk_update_scalar (KUpdate, scalars.cpp:145-158) of species n into qnew[n]
  void update_plain(int n, double dt_rk) {
    const double hup = p.h_adv_order == 5 ? 1.0 : 0.0, vup = p.v_adv_order == 3 ? 1.0 : 0.0;
    const Fl f{*this, q2[n], q1[n], dt_rk, hup, vup, false};
    const bool ften = want_qvf && n == 0;
    pfor(0, g.nx, 0, g.ny, 0, g.nz, [&](int i, int j, int k) {
      double tend = f.adv(i, j, k) * msfty(i, j);
      if (ften) qvf(i, j, k) = f.spec_cell(i, j) ? f.adv_spec(i, j, k) * msfty(i, j) : tend;  // scalars.cpp:150
      tend += (use_tendf ? qtw : qtf)[n](i, j, k);
      qnew[n](i, j, k) = (m_full(k, mub(i, j), mu1(i, j)) * q1[n](i, j, k) + dt_rk * tend) / m_full(k, mub(i, j), mu2(i, j));
    });
  }
  // rk_update_scalar_pd (KPreUpdatePD, KFacePD*, KScalePD, KUpdatePD; scalars.cpp:174-291, 380-434)
  void update_pd(int n, double dt_rk) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool do_y = !g.two_d();
    const double hup = p.h_adv_order == 5 ? 1.0 : 0.0, vup = p.v_adv_order == 3 ? 1.0 : 0.0;
    F& qo = q1[n];
    pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
      const double m = m_full(k, mub(i, j), mu1(i, j));
      qo(i, j, k) = (m * qo(i, j, k) + dt_rk * (use_tendf ? qtw : qtf)[n](i, j, k)) / m;
    });
    fill(qo, g);
    const Fl f{*this, q2[n], qo, dt_rk, hup, vup, true};
    pfor(0, nx + 1, 0, ny, 0, nz, [&](int i, int j, int k) {
      const double l = f.fxl(i, j, k);
      dhx(i, j, k) = f.fx(i, j, k) - l;
      lox(i, j, k) = l;
    });
    if (do_y)
      pfor(0, nx, 0, ny + 1, 0, nz, [&](int i, int j, int k) {
        const double l = f.fyl(i, j, k);
        dhy(i, j, k) = f.fy(i, j, k) - l;
        loy(i, j, k) = l;
      });
    pfor(0, nx, 0, ny, 0, nz + 1, [&](int i, int j, int k) {
      const double l = f.fzl(i, j, k);
      dhz(i, j, k) = f.fz(i, j, k) - l;
      loz(i, j, k) = l;
    });
    pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {  // KScalePD
      if (f.edge_cell(i, j)) { sc_scale(i, j, k) = 1; return; }  // WRF limits interior cells only
      const double mxy = msftx(i, j) * msfty(i, j);
      double hlow = c.rdx * (lox(i + 1, j, k) - lox(i, j, k));
      if (do_y) hlow += c.rdy * (loy(i, j + 1, k) - loy(i, j, k));
      const double ph_low = ((c.c1h[k] * mub(i, j) + c.c2h[k]) + c.c1h[k] * mu1(i, j)) * qo(i, j, k) -
                            dt_rk * (mxy * hlow + msfty(i, j) * c.rdnw[k] * (loz(i, j, k + 1) - loz(i, j, k)));
      const double fxp = dhx(i + 1, j, k), fxm = dhx(i, j, k);
      double hout = c.rdx * ((fxp > 0 ? fxp : 0.0) - (fxm < 0 ? fxm : 0.0));
      if (do_y) {
        const double fyp = dhy(i, j + 1, k), fym = dhy(i, j, k);
        hout += c.rdy * ((fyp > 0 ? fyp : 0.0) - (fym < 0 ? fym : 0.0));
      }
      const double fzp = dhz(i, j, k + 1), fzm = dhz(i, j, k);
      const double out = mxy * hout + msfty(i, j) * c.rdnw[k] * ((fzp < 0 ? fzp : 0.0) - (fzm > 0 ? fzm : 0.0));
      const double flux_out = dt_rk * out;
      double sc = 1;
      if (flux_out > ph_low) {
        sc = ph_low / (flux_out + 1e-20);
        if (sc < 0) sc = 0;
      }
      sc_scale(i, j, k) = sc;
    });
    fill(sc_scale, g);
    auto ddx = [&](int i, int j, int k) {
      const double dd = dhx(i, j, k);
      return dd * (dd > 0 ? sc_scale(i - 1, j, k) : (dd < 0 ? sc_scale(i, j, k) : 1.0));
    };
    auto ddy = [&](int i, int j, int k) {
      const double dd = dhy(i, j, k);
      return dd * (dd > 0 ? sc_scale(i, j - 1, k) : (dd < 0 ? sc_scale(i, j, k) : 1.0));
    };
    auto ddz = [&](int i, int j, int k) {
      if (k == 0 || k == nz) return 0.0;
      const double dd = dhz(i, j, k);
      return dd * (dd > 0 ? sc_scale(i, j, k) : (dd < 0 ? sc_scale(i, j, k - 1) : 1.0));
    };
    const bool ften = want_qvf && n == 0;
    pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {  // KUpdatePD
      if (f.spec_cell(i, j)) {
        if (ften) {  // WRF still forms the tendency here, without the edge's own flux (scalars.cpp:262-269)
          const bool scol = i == 0 || i == nx - 1, srow = j == 0 || j == ny - 1;
          double t = -c.rdnw[k] * (ddz(i, j, k + 1) - ddz(i, j, k) + loz(i, j, k + 1) - loz(i, j, k));
          if (!scol) t -= msftx(i, j) * (c.rdx * (ddx(i + 1, j, k) - ddx(i, j, k) + lox(i + 1, j, k) - lox(i, j, k)));
          if (do_y && !srow)
            t -= msftx(i, j) * (c.rdy * (ddy(i, j + 1, k) - ddy(i, j, k) + loy(i, j + 1, k) - loy(i, j, k)));
          qvf(i, j, k) = t * msfty(i, j);
        }
        qnew[n](i, j, k) = m_full(k, mub(i, j), mu1(i, j)) * qo(i, j, k) / m_full(k, mub(i, j), mu2(i, j));
        return;
      }
      double tend;
      if (f.open_col(i) || f.open_row(j)) {  // open: open terms (with q_1), vertical, x (not outer columns), y
        tend = f.open_terms(0.0, i, j, k, qo);
        tend = tend - c.rdnw[k] * (ddz(i, j, k + 1) - ddz(i, j, k) + loz(i, j, k + 1) - loz(i, j, k));
        if (!f.open_col(i))
          tend -= msftx(i, j) * (c.rdx * (ddx(i + 1, j, k) - ddx(i, j, k) + lox(i + 1, j, k) - lox(i, j, k)));
        if (do_y && !f.open_row(j))
          tend -= msftx(i, j) * (c.rdy * (ddy(i, j + 1, k) - ddy(i, j, k) + loy(i, j + 1, k) - loy(i, j, k)));
      } else {
        tend = -c.rdnw[k] * (ddz(i, j, k + 1) - ddz(i, j, k) + loz(i, j, k + 1) - loz(i, j, k));
        tend -= msftx(i, j) * (c.rdx * (ddx(i + 1, j, k) - ddx(i, j, k) + lox(i + 1, j, k) - lox(i, j, k)));
        if (do_y) tend -= msftx(i, j) * (c.rdy * (ddy(i, j + 1, k) - ddy(i, j, k) + loy(i, j + 1, k) - loy(i, j, k)));
      }
      if (ften) qvf(i, j, k) = tend * msfty(i, j);
      qnew[n](i, j, k) = (m_full(k, mub(i, j), mu1(i, j)) * qo(i, j, k) + dt_rk * (tend * msfty(i, j))) /
                         m_full(k, mub(i, j), mu2(i, j));
    });
  }
  // ScalarTransport::run_stage (scalars.cpp:436-461): plain species first, then the PD ones; each new
  // value replaces q_2 (swap), then the halos of all scalars
  void scalars(int rk) {
    if (n_sc == 0) return;
    const double dt_rk = rk == 1 ? g.dt / 3.0 : (rk == 2 ? 0.5 * g.dt : g.dt);
    if (rk == 1 && use_tendf) {  // zero_tendf + the caller's tendency, then DiffusionDriver::scalar (mut = muts)
      for (int n = 0; n < n_sc; ++n) {
        F& w = qtw[size_t(n)];
        pfor(0, g.nx, 0, g.ny, 0, g.nz, [&](int i, int j, int k) { w(i, j, k) = 0.0 + qtf[size_t(n)](i, j, k); });
        const F& qq = q2[size_t(n)];
        auto fq = [&](int i, int j, int k) { return qq(i, j, k); };
        auto mu_s = [&](int i, int j) { return muts(i, j); };
        if (p.diff_opt == 1) {  // DiffusionDriver::scalar (diffusion_driver.cpp:297-327), mut there is muts
          hdiff('m', fq, w, mu_s, xkhh, c.c1h, c.c2h);
          if (n == 0 && n_moist > 0) {  // qv: vertical_diffusion_mp with qv_base
            const std::function<double(int, int, int)> qb = [&](int, int, int k) { return qv_base[size_t(k)]; };
            vdiff('s', fq, w, kvdq, &qb);
          } else {
            vdiff('s', fq, w, kvdq, nullptr);
          }
        }
        if (p.diff_6th_opt != 0) sixth('m', fq, w, mu_s, dt_rk, c.c1h, c.c2h);
        // relax_bdy_scalar (lateral_bc.cpp:389-404): accumulated straight into the tendency
        if (spec && qbdy[size_t(n)].have)
          relax('q', g.nz, qbdy[size_t(n)], [&](int i, int j, int k) { return qq(i, j, k) * (c.c1h[k] * mut(i, j) + c.c2h[k]); },
                [&](int i, int j, int k, int b, double f0, double sm) { w(i, j, k) = w(i, j, k) + fcx[b] * f0 - gcx[b] * sm; });
      }
    }
    if (spec)  // spec_bdy_scalar, every stage (lateral_bc.cpp:406-418)
      for (int n = 0; n < n_sc; ++n)
        if (qbdy[size_t(n)].have) spec_tend('q', g.nz, qbdy[size_t(n)], qtw[size_t(n)]);
    std::vector<int> pd_n, plain_n;
    for (int n = 0; n < n_sc; ++n) {
      const bool pd = rk == 3 && (n < n_moist ? p.moist_adv_opt : p.scalar_adv_opt) == 1;
      (pd ? pd_n : plain_n).push_back(n);
    }
    // flow_dep_bdy of the species without boundary data after each update (scalars.cpp:420-429)
    for (int n : plain_n) {
      update_plain(n, dt_rk);
      std::swap(q2[n].a, qnew[n].a);
      if (spec && !qbdy[size_t(n)].have) flow_dep(q2[n]);
    }
    for (int n : pd_n) {
      update_pd(n, dt_rk);
      std::swap(q2[n].a, qnew[n].a);
      if (spec && !qbdy[size_t(n)].have) flow_dep(q2[n]);
    }
    for (int n = 0; n < n_sc; ++n) fill(q2[n], g);
  }

  // ------------------------------------------------------------------ calc_p_rho_phi (small_step.cpp:456-477,609-617)
  void calc_p_rho_phi() { calc_p_rho_phi_range(0, g.nx, 0, g.ny); }
  // KPRhoPhi over [i0, i1) x [j0, j1): the libm lanes count from i0 over i1 - i0 trips (small_step.cpp:619-633)
  void calc_p_rho_phi_range(int i0, int i1, int j0, int j1) {
    const int nx = i1 - i0;
    pfor(i0, i1, j0, j1, 0, g.nz, [&](int i, int j, int k) {
      const double a = -1.0 / (c.c1h[k] * muts(i, j) + c.c2h[k]) *
                       (alb(i, j, k) * (c.c1h[k] * mu2(i, j)) + c.rdnw[k] * (ph2(i, j, k + 1) - ph2(i, j, k)));
      al(i, j, k) = a;
      const double base = kRd * (kT0 + t2(i, j, k)) / (kP1000 * (a + alb(i, j, k)));
      double pw;
      // vm::pow lanes (backend.hpp:144,169-175): libmvec for pairs of the nx-trip loop, scalar for an odd tail
      if (libmvec && nx > 1 && !((nx & 1) && i - i0 == nx - 1))
        pw = _mm_cvtsd_f64(_ZGVbN2vv_pow(_mm_set1_pd(base), _mm_set1_pd(kCpovcv)));
      else
        pw = std::pow(base, kCpovcv);
      pp(i, j, k) = kP1000 * pw - pb(i, j, k);
    });
    fill(pp, g);
    fill(al, g);
  }

  // LargeStep::cfl_max (large_step.cpp:862-876, FVertCfl / FHorizCfl :672-687)
  void cfl_max(double& vert, double& horiz) {
    double v = 0, h = 0;
    const double dt = g.dt;
    for (int k = 1; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          double x = ww(i, j, k) / (c.c1f[k] * mut(i, j) + c.c2f[k]) * c.rdnw[k] * dt;
          x = x < 0.0 ? -x : x;
          if (x > v) v = x;  // a NaN never takes the maximum
          double a = u2(i, j, k) * c.rdx * msfux(i, j) * dt, b = v2(i, j, k) * c.rdy * msfvy(i, j) * dt;
          a = a < 0.0 ? -a : a;
          b = b < 0.0 ? -b : b;
          const double m = b > a ? b : a;
          if (m > h) h = m;
        }
    vert = v;
    horiz = h;
  }

  // RK3::step's copies (rk3.cpp:94-98)
  void step_begin() {
    if (spec) dtbc += g.dt;  // LateralBC::advance_time (rk3.cpp:93)
    u1.a = u2.a; v1.a = v2.a; w1.a = w2.a; t1.a = t2.a; ph1.a = ph2.a; mu1.a = mu2.a;
    fill(mu1, g);
    for (int n = 0; n < n_sc; ++n) q1[n].a = q2[n].a;  // rk3.cpp:99
  }
  // set_w_surface (rk3.cpp:173-200, fill = false): the kinematic surface w from the new u, v
  void set_w_surface() {
    fill(u2, g);  // set_w_surface refreshes the u, v halos first (rk3.cpp:194)
    fill(v2, g);
    cfor(g.nx, g.ny, [&](int i, int j) {
      const double ws =
          msfty(i, j) * 0.5 * c.rdy *
              ((ht(i, j + 1) - ht(i, j)) * (c.cf1 * v2(i, j + 1, 0) + c.cf2 * v2(i, j + 1, 1) + c.cf3 * v2(i, j + 1, 2)) +
               (ht(i, j) - ht(i, j - 1)) * (c.cf1 * v2(i, j, 0) + c.cf2 * v2(i, j, 1) + c.cf3 * v2(i, j, 2))) +
          msftx(i, j) * 0.5 * c.rdx *
              ((ht(i + 1, j) - ht(i, j)) * (c.cf1 * u2(i + 1, j, 0) + c.cf2 * u2(i + 1, j, 1) + c.cf3 * u2(i + 1, j, 2)) +
               (ht(i, j) - ht(i - 1, j)) * (c.cf1 * u2(i, j, 0) + c.cf2 * u2(i, j, 1) + c.cf3 * u2(i, j, 2)));
      w2(i, j, 0) = ws;
    });
    fill(w2, g);
  }
  // the end of RK3::step (rk3.cpp:159-167): spec_bdy_final, calc_p_rho_phi on the interior, set_w_surface.
  // With scalar pow the interior recomputation gives back the same bits (its inputs are unchanged
  // there), so it runs only with the libmvec lanes, whose split it changes.
  // first_rk_step_part1 (rk3.cpp:106-111): init_zero_tendency of every held tendency, then part1
  void physics_part1() {
    if (!has_phys) return;
    for (F* f : {&ru_tendf, &rv_tendf, &rw_tendf, &t_tendf, &ph_tendf}) std::fill(f->a.begin(), f->a.end(), 0.0);
    for (F& f : qtf) std::fill(f.a.begin(), f.a.end(), 0.0);
    if (phys.part1) phys.part1(phys.user);
  }
  // ------------------------------------------------------------------ microphysics (physics_coupling.cpp:262-455)
  // libm of WRF's vectorised loops (backend.hpp:144-175): libmvec's 2-lane function for the pairs of an n-trip loop,
  // the scalar one for an odd last trip; scalar everywhere without the lanes
  static bool lane(int i, int n) { return n > 1 && !((n & 1) && i == n - 1); }
  double vsin(double a, int i, int n) const {
    if (libmvec && lane(i, n)) return _mm_cvtsd_f64(_ZGVbN2v_sin(_mm_set1_pd(a)));
    return std::sin(a);
  }
  double vcos(double a, int i, int n) const {
    if (libmvec && lane(i, n)) return _mm_cvtsd_f64(_ZGVbN2v_cos(_mm_set1_pd(a)));
    return std::cos(a);
  }
  double vpow(double a, double b, int i, int n) const {
    if (libmvec && lane(i, n)) return _mm_cvtsd_f64(_ZGVbN2vv_pow(_mm_set1_pd(a), _mm_set1_pd(b)));
    return std::pow(a, b);
  }
  double vexp(double a, int i, int n) const {
    if (libmvec && lane(i, n)) return _mm_cvtsd_f64(_ZGVbN2v_exp(_mm_set1_pd(a)));
    return std::exp(a);
  }
  // moist_physics_prep_em, Kessler (kessler.hpp:16-151), moist_physics_finish_em, then the halos
  void microphysics() {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double dt = g.dt;
    F& qv = q2[0];
    F& qc = q2[1];
    F& qr = q2[2];
    // KPrep over every mass point; the lanes of WRF's i loop start one column in from specified sides
    const int pz = spec ? 1 : 0, pi0 = g.bcx == MI_BC_PERIODIC ? 0 : pz, pni = nx - 2 * pi0;
    pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
      const double zw0 = (ph2(i, j, k) + phb(i, j, k)) / kG;
      const double zw1 = (ph2(i, j, k + 1) + phb(i, j, k + 1)) / kG;
      dz8w(i, j, k) = zw1 - zw0;
      zmid(i, j, k) = 0.5 * (zw0 + zw1);
      const double q = qv(i, j, k);
      const double th = (t2(i, j, k) + kT0) / (1.0 + kRvovrd * q);
      th_phy(i, j, k) = th;
      hdia(i, j, k) = th;
      qvd(i, j, k) = q;
      mrho(i, j, k) = 1.0 / (al(i, j, k) + alb(i, j, k));
      p_phy(i, j, k) = pp(i, j, k) + pb(i, j, k);
      pii(i, j, k) = vpow((pp(i, j, k) + pb(i, j, k)) / kP1000, kRcp, i - pi0, pni);
      qcd(i, j, k) = qc(i, j, k);  // KSaveQc
    });
    // Kessler over the columns outside the spec zone; its lanes count from the first one
    const int sx = spec ? 1 : 0, sy = spec ? 1 : 0, i0 = sx, ni = nx - 2 * sx;
    const double xlv = 2.5e6, cp = kCp, ep2 = kRd / 461.6, svp1 = 0.6112, svp2 = 17.67, svp3 = 29.65,
                 svpt0 = 273.15, rhowater = 1000.0;
    cfor(nx, ny, [&](int i, int j) {
      if (i < sx || i >= nx - sx || j < sy || j >= ny - sy) return;
      const size_t nzs = static_cast<size_t>(nz);
      std::vector<double> vt(nzs), prodk(nzs), vtden(nzs), rdzk(nzs), rhok(nzs), factor(nzs), rdzw(nzs), prod(nzs);
      const double f5 = svp2 * (svpt0 - svp3) * xlv / cp;
      const double c1 = 0.001, c2 = 0.001, c3 = 2.2, c4 = 0.875, max_cr_sedimentation = 0.75;
      double crmax = 0;
      for (int k = 0; k < nz; ++k) {  // the first fall-speed loop stays scalar in WRF's object (kessler.hpp:42-44)
        prodk[k] = qr(i, j, k);
        rhok[k] = mrho(i, j, k);
        const double qrr = std::max(0.0, qr(i, j, k) * 0.001 * rhok[k]);
        vtden[k] = std::sqrt(rhok[0] / rhok[k]);
        vt[k] = 36.34 * std::pow(qrr, 0.1364) * vtden[k];
        rdzw[k] = 1.0 / dz8w(i, j, k);
        crmax = std::max(vt[k] * dt * rdzw[k], crmax);
      }
      for (int k = 0; k < nz - 1; ++k) rdzk[k] = 1.0 / (zmid(i, j, k + 1) - zmid(i, j, k));
      rdzk[nz - 1] = 1.0 / (zmid(i, j, nz - 1) - zmid(i, j, nz - 2));
      int nfall = std::max(1, int(std::round(0.5 + crmax / max_cr_sedimentation)));
      double dtfall = dt / double(nfall);
      double time_sediment = dt;
      while (nfall > 0) {
        time_sediment = time_sediment - dtfall;
        for (int k = 0; k < nz - 1; ++k) factor[k] = dtfall * rdzk[k] / rhok[k];
        factor[nz - 1] = dtfall * rdzk[nz - 1];
        const double ppt = rhok[0] * prodk[0] * vt[0] * dtfall / rhowater;
        rainncv(i, j) = ppt * 1000.0;
        rainnc(i, j) = rainnc(i, j) + ppt * 1000.0;
        for (int k = 0; k < nz - 1; ++k)
          prodk[k] = prodk[k] - factor[k] * (rhok[k] * prodk[k] * vt[k] - rhok[k + 1] * prodk[k + 1] * vt[k + 1]);
        prodk[nz - 1] = prodk[nz - 1] - factor[nz - 1] * prodk[nz - 1] * vt[nz - 1];
        if (nfall > 1) {
          nfall = nfall - 1;
          crmax = 0;
          for (int k = 0; k < nz; ++k) {  // vectorised over k in WRF's object
            const double qrr = std::max(0.0, prodk[k] * 0.001 * rhok[k]);
            vt[k] = 36.34 * vpow(qrr, 0.1364, k, nz) * vtden[k];
            crmax = std::max(vt[k] * time_sediment * rdzw[k], crmax);
          }
          const int nfall_new = std::max(1, int(std::round(0.5 + crmax / max_cr_sedimentation)));
          if (nfall_new != nfall) {
            nfall = nfall_new;
            dtfall = time_sediment / double(nfall);
          }
        } else {
          for (int k = 0; k < nz; ++k) prod[k] = prodk[k];
          nfall = 0;
        }
      }
      const int li = i - i0;
      for (int k = 0; k < nz; ++k) {
        const double factorn = 1.0 / (1.0 + c3 * dt * vpow(std::max(0.0, qr(i, j, k)), c4, li, ni));
        const double qrprod = qc(i, j, k) * (1.0 - factorn) + factorn * c1 * dt * std::max(qc(i, j, k) - c2, 0.0);
        const double rcgs = 0.001 * mrho(i, j, k);
        qc(i, j, k) = std::max(qc(i, j, k) - qrprod, 0.0);
        qr(i, j, k) = (qr(i, j, k) + prod[k] - qr(i, j, k));
        qr(i, j, k) = std::max(qr(i, j, k) + qrprod, 0.0);
        const double temp = pii(i, j, k) * th_phy(i, j, k);
        const double pressure = 1.000e+05 * vpow(pii(i, j, k), 1004.0 / 287.0, li, ni);
        const double gam = 2.5e+06 / (1004.0 * pii(i, j, k));
        const double es = 1000.0 * svp1 * vexp(svp2 * (temp - svpt0) / (temp - svp3), li, ni);
        const double qvs = ep2 * es / (pressure - es);
        prod[k] = (qv(i, j, k) - qvs) / (1.0 + pressure / (pressure - es) * qvs * f5 / ((temp - svp3) * (temp - svp3)));
        const double dim_qvs_qv = std::max(0.0, qvs - qv(i, j, k));
        const double ern_num = (1.6 + 124.9 * vpow(rcgs * qr(i, j, k), 0.2046, li, ni)) * vpow(rcgs * qr(i, j, k), 0.525, li, ni);
        const double ern_den = 2.55e8 / (pressure * qvs) + 5.4e5;
        const double ern_term1 = dt * (ern_num / ern_den) * (dim_qvs_qv / (rcgs * qvs));
        const double ern_term2 = std::max(-prod[k] - qc(i, j, k), 0.0);
        const double ern = std::min(ern_term1, std::min(ern_term2, qr(i, j, k)));
        const double product = std::max(prod[k], -qc(i, j, k));
        th_phy(i, j, k) = th_phy(i, j, k) + gam * (product - ern);
        qv(i, j, k) = std::max(qv(i, j, k) - product + ern, 0.0);
        qc(i, j, k) = qc(i, j, k) + product;
        qr(i, j, k) = qr(i, j, k) - ern;
      }
    });
    // KFinish over every mass point (use_theta_m = 1)
    const double lim = mp_lim;
    pfor(0, nx, 0, ny, 0, nz, [&](int i, int j, int k) {
      const double th = th_phy(i, j, k), th_old = hdia(i, j, k), q = qv(i, j, k), q_old = qvd(i, j, k);
      double mpten = th - th_old;
      mpten = mpten < lim * dt ? mpten : lim * dt;
      mpten = mpten > -lim * dt ? mpten : -lim * dt;
      const double qvten = q - q_old;
      t2(i, j, k) = th_old * (1.0 + kRvovrd * q_old) + mpten * (1.0 + kRvovrd * q) + kRvovrd * qvten * th - kT0;
      hdia(i, j, k) = (mpten * (1.0 + kRvovrd * q) + kRvovrd * qvten * th) / dt;
      qvd(i, j, k) = qvten / dt;
      qcd(i, j, k) = (qc(i, j, k) - qcd(i, j, k)) / dt;
    });
    fill(t2, g);
    fill(hdia, g);
    for (int n = 0; n < n_moist; ++n) fill(q2[size_t(n)], g);
    // solve_em after the microphysics: calc_p_rho_phi, without the spec zone with specified boundaries
    // (rk3.cpp:151-158)
    if (spec) calc_p_rho_phi_range(kSpecZone, nx - kSpecZone, kSpecZone, ny - kSpecZone);
    else calc_p_rho_phi();
  }
  void step_end() {
    if (has_phys && phys.after_rk) phys.after_rk(phys.user);  // phy_prep_part2 (rk3.cpp:150)
    if (mp) microphysics();
    step_end_dyn();
    if (has_phys && phys.after_all_rk_steps) phys.after_all_rk_steps(phys.user);  // rk3.cpp:166
  }
  void step_end_dyn() {
    if (spec) {
      spec_final();
      if (libmvec) {
        calc_p_rho_phi_range(kSpecZone, g.nx - kSpecZone, kSpecZone, g.ny - kSpecZone);
        fill(pp, g);
        fill(al, g);
      }
    }
    if (terrain) set_w_surface();
  }
  void step() {  // RK3::step (rk3.cpp:86-170), no microphysics
    step_begin();
    for (int rk = 1; rk <= 3; ++rk) {
      rk_prep();
      if (rk == 1) physics_part1();
      rk_tendency(rk);
      small_steps(rk);
      scalars(rk);
      calc_p_rho_phi();
    }
    step_end();
  }

  F* field(int id) {
    switch (id) {
      case MI_U: return &u2; case MI_V: return &v2; case MI_W: return &w2; case MI_T: return &t2;
      case MI_PH: return &ph2; case MI_MU: return &mu2; case MI_P: return &pp; case MI_AL: return &al;
      case MI_WW: return &ww; case MI_RU_M: return &ru_m; case MI_RV_M: return &rv_m; case MI_WW_M: return &ww_m;
      case MI_ALT: return &alt; case MI_PHP: return &php; case MI_MUTS: return &muts;
      case MI_RU_TEND: return &ru_tend; case MI_RV_TEND: return &rv_tend; case MI_RW_TEND: return &rw_tend;
      case MI_T_TEND: return &t_tend; case MI_PH_TEND: return &ph_tend;
      case MI_MUB: return &mub; case MI_PB: return &pb; case MI_ALB: return &alb; case MI_PHB: return &phb;
      case MI_T_INIT: return &t_init; case MI_MSFTX: return &msftx; case MI_MSFTY: return &msfty;
      case MI_MSFUX: return &msfux; case MI_MSFUY: return &msfuy; case MI_MSFVX: return &msfvx;
      case MI_MSFVY: return &msfvy; case MI_MSFVX_INV: return &msfvx_inv;
      case MI_HT: return &ht;
      case MI_F: return &fcor; case MI_E: return &ecor; case MI_SINA: return &sina; case MI_COSA: return &cosa;
      case MI_CQU: return &cqu; case MI_CQV: return &cqv; case MI_CQW: return &cqw; case MI_H_DIABATIC: return &hdia;
      case MI_RAINNC: return mp ? &rainnc : nullptr;
      case MI_RAINNCV: return mp ? &rainncv : nullptr;
      case MI_XKMH: return p.diff_opt == 1 ? &xkmh : nullptr;
      case MI_XKHH: return p.diff_opt == 1 ? &xkhh : nullptr;
      case MI_RU_TENDF: return use_tendf ? &ru_tendf : nullptr;
      case MI_RV_TENDF: return use_tendf ? &rv_tendf : nullptr;
      case MI_RW_TENDF: return use_tendf ? &rw_tendf : nullptr;
      case MI_T_TENDF: return use_tendf ? &t_tendf : nullptr;
      case MI_TH_FORCING: return want_thf ? &thf : nullptr;
      case MI_QV_FORCING: return want_qvf ? &qvf : nullptr;
      default:
        if (id >= MI_Q && id < MI_Q + n_sc) return &q2[size_t(id - MI_Q)];
        if (id >= MI_Q_TENDF && id < MI_Q_TENDF + n_sc) return &qtf[size_t(id - MI_Q_TENDF)];
        return nullptr;
    }
  }
};

bool is_base(int id) { return id >= MI_MUB && id <= MI_HT; }

int fail(const std::string& m, int code = MI_EINVAL) {
  g_err = m;
  return code;
}

}  // namespace

struct orc_domain {
  std::unique_ptr<Dom> d;
};

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int mi_auto_sound_steps_oracle(double dt, double dx, double dy) { return Dom::mi_auto_sound_steps_(dt, dx, dy); }

int orc_domain_create(const mi_grid* g, const mi_params* p, const mi_levels* lv, orc_domain** out) {
  try {
    if (!g || !p || !lv || !out) return fail("null argument");
    if (g->nx < 1 || g->ny < 1 || g->nz < 3) return fail("bad grid size");
    if (g->bc_x > MI_BC_OPEN || g->bc_y > MI_BC_OPEN || g->bc_x < 0 || g->bc_y < 0)
      return fail("oracle: periodic, symmetric, specified or open boundaries");
    if ((g->bc_x == MI_BC_SPECIFIED) != (g->bc_y == MI_BC_SPECIFIED))
      return fail("specified boundaries on both x and y or on neither");
    if (g->bc_x == MI_BC_SPECIFIED && (g->ny == 1 || g->nx < 10 || g->ny < 10))
      return fail("specified boundaries need a 3-D grid of at least 10 x 10 columns");
    if ((g->bc_x == MI_BC_OPEN && g->nx < 7) || (g->bc_y == MI_BC_OPEN && g->ny > 1 && g->ny < 7))
      return fail("an open direction needs at least 7 columns");
    if (p->h_adv_order != 5 && p->h_adv_order != 6) return fail("h_adv_order must be 5 or 6");
    if (p->v_adv_order != 3 && p->v_adv_order != 4) return fail("v_adv_order must be 3 or 4");
    if (p->hypsometric_opt != 1) return fail("oracle: hypsometric_opt 1 only");
    if (p->damp_opt < 0 || p->damp_opt > 3) return fail("oracle: damp_opt 0..3");
    if ((p->damp_opt == 1 || p->damp_opt == 2) && g->nz >= kMaxLev) return fail("oracle: damp_opt 1, 2 need nz < 256");
    if (p->n_moist < 0 || p->n_moist > MI_MAX_MOIST || p->n_scalars < p->n_moist || p->n_scalars > MI_MAX_SCALARS)
      return fail("bad n_moist / n_scalars");
    if (p->diff_6th_opt < 0 || p->diff_6th_opt > 2) return fail("diff_6th_opt must be 0, 1 or 2");
    if (p->diff_opt < 0 || p->diff_opt > 1) return fail("oracle: diff_opt 0 or 1 only");
    if (p->diff_opt == 1 && p->km_opt != 1 && p->km_opt != 4) return fail("oracle: km_opt 1 or 4 only");
    if (p->diff_opt == 1 && p->km_opt == 4 && g->nz + 1 >= kMaxLev) return fail("oracle: km_opt 4 needs nz < 255");
    if (p->mp_physics != 0 && p->mp_physics != 1) return fail("oracle: mp_physics 0 or 1 (Kessler) only");
    if (p->mp_physics == 1 && (p->n_moist < 3 || g->nz > 128)) return fail("Kessler needs qv, qc, qr and nz <= 128");
    auto* o = new orc_domain;
    o->d = std::make_unique<Dom>(*g, *p, *lv);
    if (o->d->n_sound < 2 || o->d->n_sound % 2) {
      delete o;
      return fail("time_step_sound must be even and >= 2");
    }
    *out = o;
    return MI_OK;
  } catch (const std::exception& e) {
    return fail(e.what(), MI_ENOMEM);
  }
}

int orc_domain_destroy(orc_domain* d) {
  delete d;
  return MI_OK;
}
int orc_domain_n_sound(const orc_domain* d) { return d->d->n_sound; }
int orc_set_libmvec(orc_domain* d, int on) {
  d->d->libmvec = on != 0;
  return MI_OK;
}
int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return MI_OK;
}

// the diffusion's 1-D base profiles (nz values)
static std::vector<double>* base_profile(Dom& d, int id) {
  return id == MI_U_BASE ? &d.u_base : id == MI_V_BASE ? &d.v_base : id == MI_QV_BASE ? &d.qv_base
         : id == MI_T_BASE ? &d.t_base : id == MI_Z_BASE ? &d.z_base : nullptr;
}

int orc_set_field(orc_domain* od, int id, const double* h, int64_t n) {
  Dom& d = *od->d;
  if (std::vector<double>* bp = base_profile(d, id)) {
    if (n != int64_t(bp->size())) return fail("base profiles take nz values");
    bp->assign(h, h + n);
    return MI_OK;
  }
  F* f = d.field(id);
  if (!f) return fail("unknown field id " + std::to_string(id));
  const int64_t full = int64_t(f->nxs) * f->nys * f->nzs;
  if (is_base(id)) {
    // base fields: a profile (nzs values; 2-D: one) repeated over the allocation halos included, or
    // the logical points with the halo filled by the field's boundary rule
    if (n == f->nzs) {
      for (int k = 0; k < f->nzs; ++k)
        for (int j = -f->hy; j < f->nys + f->hy; ++j)
          for (int i = -3; i < f->nxs + 3; ++i) (*f)(i, j, k) = h[k];
      return MI_OK;
    }
    if (n != full) return fail("base field size mismatch");
  } else if (n != full) {
    return fail("field size mismatch");
  }
  for (int k = 0; k < f->nzs; ++k)
    for (int j = 0; j < f->nys; ++j)
      for (int i = 0; i < f->nxs; ++i) (*f)(i, j, k) = h[(size_t(k) * f->nys + j) * f->nxs + i];
  fill(*f, d.g);
  return MI_OK;
}

int orc_get_field(orc_domain* od, int id, double* h, int64_t n) {
  Dom& d = *od->d;
  F* f = d.field(id);
  if (!f) return fail("unknown field id " + std::to_string(id));
  if (n != int64_t(f->nxs) * f->nys * f->nzs) return fail("field size mismatch");
  for (int k = 0; k < f->nzs; ++k)
    for (int j = 0; j < f->nys; ++j)
      for (int i = 0; i < f->nxs; ++i) h[(size_t(k) * f->nys + j) * f->nxs + i] = (*f)(i, j, k);
  return MI_OK;
}

int orc_step(orc_domain* d, int nsteps) {
  for (int s = 0; s < nsteps; ++s) d->d->step();
  return MI_OK;
}
int orc_step_begin(orc_domain* d) {
  d->d->step_begin();
  return MI_OK;
}
int orc_rk_prep(orc_domain* d, int rk) {
  if (rk < 1 || rk > 3) return fail("rk must be 1, 2 or 3");
  d->d->rk_prep();
  if (rk == 1) d->d->physics_part1();
  return MI_OK;
}
int orc_set_diffusion(orc_domain* od, int opt, double factor) {
  Dom& d = *od->d;
  if (opt < 0 || opt > 2) return fail("diff_6th_opt must be 0, 1 or 2");
  d.p.diff_6th_opt = opt;
  d.p.diff_6th_factor = factor;
  d.held_base = opt != 0 || d.p.diff_opt != 0 || d.p.damp_opt == 2 || d.spec;
  d.use_tendf = d.held_base || d.has_phys;
  return MI_OK;
}
int orc_set_column_physics(orc_domain* od, const orc_column_physics* cp) {
  Dom& d = *od->d;
  if (cp && cp->qv_forcing && d.n_moist == 0) return fail("qv_forcing needs moisture (n_moist > 0)");
  d.has_phys = cp != nullptr;
  d.phys = cp ? *cp : orc_column_physics{};
  d.use_tendf = d.held_base || d.has_phys;
  d.want_thf = cp && cp->th_forcing;
  d.want_qvf = cp && cp->qv_forcing;
  return MI_OK;
}
int orc_rk_tendency(orc_domain* d, int rk) {
  if (rk < 1 || rk > 3) return fail("rk must be 1, 2 or 3");
  d->d->rk_tendency(rk);
  return MI_OK;
}
int orc_small_steps(orc_domain* d, int rk) {
  if (rk < 1 || rk > 3) return fail("rk must be 1, 2 or 3");
  d->d->small_steps(rk);
  return MI_OK;
}
int orc_scalar_transport(orc_domain* d, int rk) {
  if (rk < 1 || rk > 3) return fail("rk must be 1, 2 or 3");
  d->d->scalars(rk);
  return MI_OK;
}
int orc_set_dt(orc_domain* d, double dt, int n_sound) {
  if (!(dt > 0)) return fail("dt must be > 0");
  if (n_sound < 2 || n_sound % 2) return fail("time_step_sound must be even and >= 2");
  d->d->set_dt(dt, n_sound);
  return MI_OK;
}
int orc_cfl_max(orc_domain* d, double* vert, double* horiz) {
  d->d->cfl_max(*vert, *horiz);
  return MI_OK;
}
// the flux operators on their own, for the closed-form checks: which = 0 flux5, 1 flux3, 2 flux5_pd, 3 flux3_pd
// (q = qm3..qp2 or qm2..qp1), 4 flux_upwind (q = qm1, q0; ua is the Courant number)
int orc_flux_eval(int which, const double* q, double ua, double up, double* out) {
  if (!q || !out) return fail("null argument");
  switch (which) {
    case 0: *out = flux5(q[0], q[1], q[2], q[3], q[4], q[5], ua, up); return MI_OK;
    case 1: *out = flux3(q[0], q[1], q[2], q[3], ua, up); return MI_OK;
    case 2: *out = flux5_pd(q[0], q[1], q[2], q[3], q[4], q[5], ua, up); return MI_OK;
    case 3: *out = flux3_pd(q[0], q[1], q[2], q[3], ua, up); return MI_OK;
    case 4: *out = flux_upwind(q[0], q[1], ua); return MI_OK;
    default: return fail("which must be 0..4");
  }
}
int orc_calc_p_rho_phi_interior(orc_domain* od, int sz) {  // SmallStep::calc_p_rho_phi_interior (small_step.cpp:619-633)
  Dom& d = *od->d;
  if (sz < 0) return fail("spec_zone must be >= 0");
  const bool px = d.g.bcx == MI_BC_PERIODIC, py = d.g.bcy == MI_BC_PERIODIC || d.g.two_d();
  const int i0 = px ? 0 : sz, i1 = px ? d.g.nx : d.g.nx - sz, j0 = py ? 0 : sz, j1 = py ? d.g.ny : d.g.ny - sz;
  if (i1 > i0 && j1 > j0) d.calc_p_rho_phi_range(i0, i1, j0, j1);
  fill(d.pp, d.g);
  fill(d.al, d.g);
  return MI_OK;
}
int orc_calc_p_rho_phi(orc_domain* d, int) {
  d->d->calc_p_rho_phi();
  return MI_OK;
}
int orc_step_end(orc_domain* d) {
  d->d->step_end();
  return MI_OK;
}
int orc_bdy_set(orc_domain* od, int var, int side, const double* values, const double* tends, int64_t n) {
  Dom& d = *od->d;
  if (!d.spec) return fail("the domain has no specified boundaries");
  if (side < 0 || side > 3 || !values || !tends) return fail("bad side or null data");
  Dom::Bdy* b = nullptr;
  if (var >= MI_BDY_U && var <= MI_BDY_MU) b = &d.bdy[var];
  else if (var >= MI_BDY_Q && var < MI_BDY_Q + d.n_sc) b = &d.qbdy[size_t(var - MI_BDY_Q)];
  if (!b) return fail("bad boundary variable");
  const size_t want = size_t(b->npos[side]) * size_t(b->nk) * Dom::kBw;
  if (n < 0 || size_t(n) != want) return fail("wrong boundary slab size");
  b->v[side].assign(values, values + n);
  b->t[side].assign(tends, tends + n);
  b->have = true;
  return MI_OK;
}
int orc_bdy_set_time(orc_domain* od, double dtbc) {
  od->d->dtbc = dtbc;
  return MI_OK;
}
int orc_bdy_get_time(orc_domain* od, double* dtbc) {
  *dtbc = od->d->dtbc;
  return MI_OK;
}

}  // extern "C"
