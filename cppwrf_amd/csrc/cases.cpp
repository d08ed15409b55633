// Synthetic initial states (host only): the reference's idealized-case recipe and this library's 3-D
// warm-bubble workload.
//
// The recipe follows the reference case assembly step by step, because its golden files are made from
// it: the hydrostatic Exner sounding integrated with Simpson's rule on a 0.5 m table
// (src/cases/cases.cpp:17-33), p(z) and its bisection inverse (cases.cpp:41-73), eta levels equally
// spaced in height (cases.cpp:194-201), BaseState::set_levels (src/dyn/state.cpp:86-116) and the dry
// hydrostatic base and perturbation state (cases.cpp:224-287), with p_top the sounding's pressure at
// ztop (src/cases/model.cpp:93-96).  Arithmetic is the host's double with no contraction, so the
// fields equal the reference's bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cppwrf_amd.h"
#include "errors.hpp"

namespace {

constexpr double kG = 9.81, kRd = 287.0, kCp = 7.0 * kRd / 2.0, kCv = kCp - kRd;
constexpr double kCvpm = -kCv / kCp, kRcp = kRd / kCp, kP1000 = 100000.0, kT0 = 300.0;
constexpr double kPi = 3.141592653589793;  // std::numbers::pi

constexpr double kRvovrd = 461.6 / kRd;  // R_v / R_d (reference constants.hpp)
constexpr double kEp2 = kRd / 461.6;      // R_d / R_v

// Exner function of height for a given theta(z), hydrostatic, on a table of step dz (reference Sounding,
// src/cases/cases.cpp:17-87).  With a vapour profile qv(z) the integration uses the density potential
// temperature theta (1 + R_v/R_d qv) / (1 + qv), and the vapour load (the weight of the vapour above z,
// trapezoid rule from the top) gives the dry pressure pd = p - load.
class Sounding {
 public:
  Sounding(std::function<double(double)> th, double p_surf, double zmax, double dz = 0.5,
           std::function<double(double)> qv = nullptr)
      : th_(std::move(th)), qv_(std::move(qv)), dz_(dz) {
    const int n = int(std::ceil(zmax / dz)) + 2;
    pi_.resize(size_t(n));
    pi_[0] = std::pow(p_surf / kP1000, kRcp);
    for (int m = 1; m < n; ++m) {
      const double z0 = (m - 1) * dz, z1 = m * dz;
      const double f = (1.0 / theta_rho(z0) + 4.0 / theta_rho(0.5 * (z0 + z1)) + 1.0 / theta_rho(z1)) / 6.0;
      pi_[size_t(m)] = pi_[size_t(m - 1)] - kG / kCp * f * dz;
    }
    if (qv_) {
      auto rq = [&](int m) {  // vapour density times g: rho_d qv with rho_d = p / (R_d theta_m pi)
        const double z = m * dz, pi = pi_[size_t(m)], pp = kP1000 * std::pow(pi, 1.0 / kRcp);
        return pp / (kRd * th_(z) * (1.0 + kRvovrd * qv_(z)) * pi) * qv_(z);
      };
      w_.assign(size_t(n), 0.0);
      for (int m = n - 2; m >= 0; --m) w_[size_t(m)] = w_[size_t(m + 1)] + 0.5 * kG * dz * (rq(m) + rq(m + 1));
    }
  }
  double theta(double z) const { return th_(z); }
  double qv(double z) const { return qv_ ? qv_(z) : 0.0; }
  double theta_rho(double z) const {
    if (!qv_) return th_(z);
    const double q = qv_(z);
    return th_(z) * (1.0 + kRvovrd * q) / (1.0 + q);
  }
  double p(double z) const {
    const double s = z / dz_;
    const int m = int(std::floor(s));
    if (m < 0 || m + 1 >= int(pi_.size())) throw std::runtime_error("sounding: height out of range");
    const double z0 = m * dz_, h = z - z0;
    const double f = (1.0 / theta_rho(z0) + 4.0 / theta_rho(z0 + 0.5 * h) + 1.0 / theta_rho(z)) / 6.0;
    const double pi = pi_[size_t(m)] - kG / kCp * f * h;
    return kP1000 * std::pow(pi, 1.0 / kRcp);
  }
  double vapour_load(double z) const {
    if (!qv_) return 0.0;
    const double s = z / dz_;
    const int m = int(std::floor(s));
    if (m < 0 || m + 1 >= int(w_.size())) throw std::runtime_error("sounding: height out of range");
    const double a = s - m;
    return (1.0 - a) * w_[size_t(m)] + a * w_[size_t(m + 1)];
  }
  double pd(double z) const { return p(z) - vapour_load(z); }
  double z_of_p(double pp) const { return bisect(pp, false); }
  double z_of_pd(double pp) const { return bisect(pp, qv_ != nullptr); }

 private:
  // the height where p (dry: pd) falls to pp, by bisection to 1e-9 m
  double bisect(double pp, bool dry) const {
    auto f = [&](double z) { return dry ? pd(z) : p(z); };
    double lo = 0, hi = double(pi_.size() - 2) * dz_;
    if (pp > f(lo) || pp < f(hi)) throw std::runtime_error("sounding: pressure out of range");
    for (int it = 0; it < 200 && hi - lo > 1e-9; ++it) {
      const double mid = 0.5 * (lo + hi);
      (f(mid) > pp ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
  }
  std::function<double(double)> th_, qv_;
  double dz_;
  std::vector<double> pi_, w_;
};

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

}  // namespace

struct mi_case {
  mi_grid g{};
  int kind = 0;
  int nz = 0;
  std::vector<double> dnw, rdnw, dn, rdn, fnm, fnp, c1h, c2h, c3h, c4h, c1f, c2f, c3f, c4f, znw, znu;
  double cf1 = 0, cf2 = 0, cf3 = 0, cfn = 0, cfn1 = 0, rdx = 0, rdy = 0, p_top = 0;
  double mub = 0;
  std::vector<double> pb, alb, phb, t_init;             // profiles
  std::vector<double> u_base, v_base, qv_base;          // the diffusion's base profiles (model.cpp:149-177)
  std::vector<double> t_base, z_base;                   // and those of the Rayleigh damping (damp_opt 2)
  std::vector<double> u, v, w, t, ph, mu, p, al;        // dense state
  // hills: the base state in full, terrain, map factors, Coriolis, moisture (dense arrays)
  bool full = false;
  std::vector<double> mub2, pb3, alb3, phb3, t_init3, ht, fcor, ecor, qv;
  std::vector<double> msftx, msfty, msfux, msfuy, msfvx, msfvy, msfvx_inv;
};

namespace {


// BaseState::set_levels, hybrid_opt = 0 (state.cpp:86-116), from the eta levels e
void set_levels(mi_case& c, const std::vector<double>& e, const mi_grid& G) {
  const int nz = G.nz, ny = G.ny;
  c.znw = e;
  c.znu.assign(size_t(nz), 0.0);
  c.dnw.assign(size_t(nz), 0.0);
  c.rdnw.assign(size_t(nz), 0.0);
  c.dn.assign(size_t(nz) + 1, 0.0);
  c.rdn.assign(size_t(nz) + 1, 0.0);
  c.fnm.assign(size_t(nz) + 1, 0.0);
  c.fnp.assign(size_t(nz) + 1, 0.0);
  for (int k = 0; k < nz; ++k) {
    c.dnw[k] = c.znw[k + 1] - c.znw[k];
    c.rdnw[k] = 1.0 / c.dnw[k];
    c.znu[k] = 0.5 * (c.znw[k + 1] + c.znw[k]);
  }
  for (int k = 1; k < nz; ++k) {
    c.dn[k] = 0.5 * (c.dnw[k] + c.dnw[k - 1]);
    c.rdn[k] = 1.0 / c.dn[k];
    c.fnp[k] = 0.5 * c.dnw[k] / c.dn[k];
    c.fnm[k] = 0.5 * c.dnw[k - 1] / c.dn[k];
  }
  const double cof1 = (2. * c.dn[1] + c.dn[2]) / (c.dn[1] + c.dn[2]) * c.dnw[0] / c.dn[1];
  const double cof2 = c.dn[1] / (c.dn[1] + c.dn[2]) * c.dnw[0] / c.dn[2];
  c.cf1 = c.fnp[1] + cof1;
  c.cf2 = c.fnm[1] - cof1 - cof2;
  c.cf3 = cof2;
  c.cfn = (0.5 * c.dnw[nz - 1] + c.dn[nz - 1]) / c.dn[nz - 1];
  c.cfn1 = -0.5 * c.dnw[nz - 1] / c.dn[nz - 1];
  c.c1h.assign(size_t(nz), 1.0);
  c.c2h.assign(size_t(nz), 0.0);
  c.c3h = c.znu;
  c.c4h.assign(size_t(nz), 0.0);
  c.c1f.assign(size_t(nz) + 1, 1.0);
  c.c2f.assign(size_t(nz) + 1, 0.0);
  c.c3f = c.znw;
  c.c4f.assign(size_t(nz) + 1, 0.0);
  c.rdx = 1.0 / G.dx;
  c.rdy = ny == 1 ? 0.0 : 1.0 / G.dy;
}

void build(mi_case& c, const mi_case_spec& sp) {
  const mi_grid& G = sp.grid;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  if (nx < 1 || ny < 1 || nz < 3) throw std::invalid_argument("case: bad grid size");
  if (!(sp.ztop > 0)) throw std::invalid_argument("case: ztop must be > 0");
  c.g = G;
  c.nz = nz;
  // case definition (reference src/cases/cases.cpp:94-191)
  std::function<double(double)> theta_env;
  std::function<double(double)> u_env = [](double) { return 0.0; };
  std::function<double(int, int, int, double, double, double)> pert = [](int, int, int, double, double, double) {
    return 0.0;
  };
  const double p_surf = 1.0e5;
  const double xcen = 0.5 * nx * G.dx;
  std::vector<double> noise;
  std::vector<double> exn(static_cast<size_t>(nz));  // the base state's Exner function (set below)
  std::function<double(double)> qv_env;              // a vapour profile: the full moist state below
  if (sp.qv0 > 0) {  // the optional unsaturated profile of any case (reference case.qv0, cases.cpp:184-188)
    const double hq = sp.qv_hscale > 0 ? sp.qv_hscale : 2500.0, qv0 = sp.qv0;
    qv_env = [=](double z) { return qv0 * std::exp(-z / hq); };
  }
  switch (sp.kind) {
    case MI_CASE_REST_STATE: {
      const double th0 = 300.0, N = 0.01, u0 = sp.u0;
      theta_env = [=](double z) { return th0 * std::exp(N * N * z / kG); };
      u_env = [=](double) { return u0; };
      break;
    }
    case MI_CASE_WARM_BUBBLE: {
      const double xc = xcen, zc = 2000.0, r = 2000.0, dth = 2.0;
      theta_env = [](double) { return 300.0; };
      pert = [=](int, int, int, double x, double, double z) {
        const double qx = (x - xc) / r, qz = (z - zc) / r;
        const double L = std::sqrt(qx * qx + qz * qz);
        if (!(L <= 1.0)) return 0.0;
        const double cs = std::cos(0.5 * kPi * L);
        return dth * (cs * cs);
      };
      break;
    }
    case MI_CASE_STRAKA: {
      // neutral 300 K with a cold bubble: a -15 K temperature perturbation, cos-shaped within the ellipse of
      // radii 4 km x 2 km centred at 3 km (reference straka, cases.cpp:103-113; Straka et al. 1993); the
      // squares by the library pow, as the reference
      const double xc = xcen, zc = 3000.0, xr = 4000.0, zr = 2000.0, dT = -15.0;
      theta_env = [](double) { return 300.0; };
      pert = [=, &exn](int, int, int k, double x, double, double z) {
        volatile double two = 2.0;  // glibc pow, not a multiply folded by the compiler
        const double L = std::sqrt(std::pow((x - xc) / xr, two) + std::pow((z - zc) / zr, two));
        return L <= 1.0 ? dT * 0.5 * (std::cos(kPi * L) + 1.0) / exn[size_t(k)] : 0.0;
      };
      break;
    }
    case MI_CASE_SQUALL2D: {
      // Weisman & Klemp (1982) analytic sounding with a 2 K warm bubble of radii 10 km x 1.4 km at 1.4 km
      // (reference squall2d, cases.cpp:135-179): theta = theta0 + (theta_tr - theta0) (z/z_tr)^(5/4) below z_tr,
      // isothermal T_tr above; RH = 1 - 0.75 (z/z_tr)^(5/4) (0.25 above); qv from RH, capped at 14 g/kg, by a
      // fixed-point iteration with the moist sounding on a 50 m table (qv needs p(z), which needs qv); no wind
      const double th0 = 300.0, thtr = 343.0, ztr = 12000.0, ttr = 213.0, qvmax = 0.014;
      const double xc = xcen, zc = 1400.0, xr = 10000.0, zr = 1400.0, dth = 2.0;
      theta_env = [=](double z) {
        return z <= ztr ? th0 + (thtr - th0) * std::pow(z / ztr, 1.25) : thtr * std::exp(kG / (kCp * ttr) * (z - ztr));
      };
      auto rh = [=](double z) { return z <= ztr ? 1.0 - 0.75 * std::pow(z / ztr, 1.25) : 0.25; };
      const double zmax = sp.ztop * 1.2 + 1000.0, dzt = 50.0;
      const int nt = int(zmax / dzt) + 2;
      auto qtab = std::make_shared<std::vector<double>>(size_t(nt), 0.0);
      auto qv_of = [qtab, dzt, nt](double z) {
        const double f = std::clamp(z / dzt, 0.0, double(nt - 1) - 1e-9);
        const int n = int(f);
        return (*qtab)[size_t(n)] + (f - n) * ((*qtab)[size_t(n) + 1] - (*qtab)[size_t(n)]);
      };
      for (int it = 0; it < 8; ++it) {
        const Sounding snd(theta_env, p_surf, zmax, 0.5,
                           it == 0 ? std::function<double(double)>{} : std::function<double(double)>(qv_of));
        std::vector<double> q(static_cast<size_t>(nt));
        for (int n = 0; n < nt; ++n) {
          const double z = std::min(n * dzt, zmax), pz = snd.p(z), T = theta_env(z) * std::pow(pz / kP1000, kRcp);
          const double es = 611.2 * std::exp(17.67 * (T - 273.15) / (T - 29.65));  // Pa
          q[size_t(n)] = std::min(qvmax, rh(z) * kEp2 * es / (pz - es));
        }
        *qtab = q;
      }
      qv_env = qv_of;
      pert = [=](int, int, int, double x, double, double z) {
        volatile double two = 2.0;  // glibc pow, not a multiply folded by the compiler
        const double L = std::sqrt(std::pow((x - xc) / xr, two) + std::pow((z - zc) / zr, two));
        return L <= 1.0 ? dth * std::pow(std::cos(0.5 * kPi * L), two) : 0.0;
      };
      break;
    }
    case MI_CASE_IGW: {
      const double th0 = 300.0, N = 0.01, U = 20.0, dth = 0.01, a = 5000.0, xc = 100000.0, Ht = sp.ztop;
      theta_env = [=](double z) { return th0 * std::exp(N * N * z / kG); };
      u_env = [=](double) { return U; };
      pert = [=](int, int, int, double x, double, double z) {
        const double q = (x - xc) / a;
        return dth * std::sin(kPi * z / Ht) / (1.0 + q * q);
      };
      break;
    }
    case MI_CASE_BUBBLE3D: {
      // N = 0.01 atmosphere at rest with a cos^2 warm bubble and seeded uniform noise (this library's
      // 64x64x40 workload); the noise is drawn in k, j, i order from splitmix64(seed)
      const double th0 = 300.0, N = 0.01;
      const double xc = xcen, yc = 0.5 * ny * G.dy, zc = sp.bubble_zc, r = sp.bubble_radius, dth = sp.bubble_dtheta;
      if (!(r > 0)) throw std::invalid_argument("case: bubble_radius must be > 0");
      theta_env = [=](double z) { return th0 * std::exp(N * N * z / kG); };
      noise.assign(size_t(nx) * ny * nz, 0.0);
      uint64_t s = sp.seed;
      for (size_t n = 0; n < noise.size(); ++n) {
        const double u01 = double(splitmix64(s) >> 11) * 0x1p-53;
        noise[n] = sp.noise_amp * (2.0 * u01 - 1.0);
      }
      const bool two_d = ny == 1;
      pert = [=, &noise](int i, int j, int k, double x, double y, double z) {
        const double qx = (x - xc) / r, qy = two_d ? 0.0 : (y - yc) / r, qz = (z - zc) / r;
        const double L = std::sqrt(qx * qx + qy * qy + qz * qz);
        double b = 0.0;
        if (L <= 1.0) {
          const double cs = std::cos(0.5 * kPi * L);
          b = dth * (cs * cs);
        }
        return b + noise[(size_t(k) * ny + j) * nx + i];
      };
      break;
    }
    default:
      throw std::invalid_argument("case: unknown kind " + std::to_string(sp.kind));
  }
  Sounding snd(theta_env, p_surf, sp.ztop * 1.2 + 1000.0);
  c.p_top = snd.p(sp.ztop);
  // eta levels equally spaced in height (cases.cpp:197-201)
  std::vector<double> e(size_t(nz) + 1);
  {
    const double ps = snd.p(0.0), pt = snd.p(sp.ztop);
    for (int k = 0; k <= nz; ++k) e[size_t(k)] = (snd.p(sp.ztop * k / nz) - pt) / (ps - pt);
    e[0] = 1.0;
    e[size_t(nz)] = 0.0;
  }
  set_levels(c, e, G);
  // dry hydrostatic base state and perturbation state (cases.cpp:224-287)
  const double mub = p_surf - c.p_top;
  c.mub = mub;
  std::vector<double> tb(static_cast<size_t>(nz)), zm(static_cast<size_t>(nz));
  c.pb.assign(size_t(nz), 0.0);
  c.alb.assign(size_t(nz), 0.0);
  c.phb.assign(size_t(nz) + 1, 0.0);
  c.t_init.assign(size_t(nz), 0.0);
  for (int k = 0; k < nz; ++k) {
    c.pb[k] = c.c3h[k] * mub + c.c4h[k] + c.p_top;
    tb[k] = snd.theta(snd.z_of_p(c.pb[k]));
    c.alb[k] = (kRd / kP1000) * tb[k] * std::pow(c.pb[k] / kP1000, kCvpm);
    exn[k] = std::pow(c.pb[k] / kP1000, kRcp);
  }
  c.phb[0] = 0.0;
  for (int k = 0; k < nz; ++k) c.phb[k + 1] = c.phb[k] - c.dnw[k] * (c.c1h[k] * mub + c.c2h[k]) * c.alb[k];
  for (int k = 0; k < nz; ++k) zm[k] = 0.5 * (c.phb[k] + c.phb[k + 1]) / kG;
  for (int k = 0; k < nz; ++k) c.t_init[k] = tb[k] - kT0;
  const size_t nm = size_t(nx) * ny * nz, nw = size_t(nx) * ny * (nz + 1);
  const int nxu = nx + 1, nyv = ny > 1 ? ny + 1 : 1;
  c.t.assign(nm, 0.0);
  c.al.assign(nm, 0.0);
  c.p.assign(nm, 0.0);
  c.ph.assign(nw, 0.0);
  c.w.assign(nw, 0.0);
  c.mu.assign(size_t(nx) * ny, 0.0);
  c.u.assign(size_t(nxu) * ny * nz, 0.0);
  c.v.assign(size_t(nx) * nyv * nz, 0.0);
  auto I = [&](int i, int j, int k) { return (size_t(k) * ny + j) * nx + i; };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double x = (i + 0.5) * G.dx, y = (j + 0.5) * G.dy;
        const double th = tb[k] + pert(i, j, k, x, y, zm[k]);
        const double alt = (kRd / kP1000) * th * std::pow(c.pb[k] / kP1000, kCvpm);
        c.al[I(i, j, k)] = alt - c.alb[k];
        c.t[I(i, j, k)] = th - kT0;
      }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double ph = 0.0;
      c.ph[(size_t(0) * ny + j) * nx + i] = 0.0;
      for (int k = 0; k < nz; ++k) {
        ph -= c.dnw[k] * ((c.c1h[k] * mub + c.c2h[k]) * c.al[I(i, j, k)]);
        c.ph[(size_t(k + 1) * ny + j) * nx + i] = ph;
      }
    }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nxu; ++i) {
        double uu = u_env(zm[k]);
        if (G.bc_x == MI_BC_SYMMETRIC && (i == 0 || i == nx)) uu = 0.0;
        c.u[(size_t(k) * ny + j) * nxu + i] = uu;
      }
  if (qv_env) {
    // the full moist state (reference initialize_moist, cases.cpp:290-352; WRF module_initialize_ideal with
    // use_theta_m = 1): qv(z) from a moist sounding, mu' from the vapour load, p' in hydrostatic balance with the
    // moisture loading of pg_buoy_w from the top down, theta_m, al and ph
    const Sounding msnd(theta_env, p_surf, sp.ztop * 1.2 + 1000.0, 0.5, qv_env);
    const double pd_surf = p_surf - msnd.vapour_load(0.0);
    const double mu1 = pd_surf - c.p_top - mub;
    std::vector<double> qv(static_cast<size_t>(nz)), thd(static_cast<size_t>(nz));
    for (int k = 0; k < nz; ++k) {
      const double p_level = c.c3h[k] * (pd_surf - c.p_top) + c.c4h[k] + c.p_top;
      const double z = msnd.z_of_pd(p_level);
      qv[k] = msnd.qv(z);
      thd[k] = msnd.theta(z);
    }
    c.qv.assign(nm, 0.0);
    std::vector<double> th(static_cast<size_t>(nz));
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double x = (i + 0.5) * G.dx, y = (j + 0.5) * G.dy;
        for (int k = 0; k < nz; ++k) th[k] = thd[k] + pert(i, j, k, x, y, zm[k]);
        auto eos = [&](int k, double pp) {
          const double qvf = 1.0 + kRvovrd * qv[k];
          const double alt = (kRd / kP1000) * th[k] * qvf * std::pow((pp + c.pb[k]) / kP1000, kCvpm);
          c.al[I(i, j, k)] = alt - c.alb[k];
          c.p[I(i, j, k)] = pp;
        };
        int k = nz - 1;
        double qvf1 = 0.5 * (qv[k] + qv[k]);
        double qvf2 = 1.0 / (1.0 + qvf1);
        qvf1 = qvf1 * qvf2;
        double pp = -0.5 * ((c.c1f[nz] * mu1) + qvf1 * (c.c1f[nz] * mub + c.c2f[nz])) / c.rdnw[k] / qvf2;
        eos(k, pp);
        for (k = nz - 2; k >= 0; --k) {
          qvf1 = 0.5 * (qv[k] + qv[k + 1]);
          qvf2 = 1.0 / (1.0 + qvf1);
          qvf1 = qvf1 * qvf2;
          pp = pp - ((c.c1f[k + 1] * mu1) + qvf1 * (c.c1f[k + 1] * mub + c.c2f[k + 1])) / qvf2 / c.rdn[k + 1];
          eos(k, pp);
        }
        double ph = 0.0;
        c.ph[(size_t(0) * ny + j) * nx + i] = 0.0;
        for (k = 0; k < nz; ++k) {
          ph = ph - c.dnw[k] * (((c.c1h[k] * mub + c.c2h[k]) + c.c1h[k] * mu1) * c.al[I(i, j, k)] + c.c1h[k] * mu1 * c.alb[k]);
          c.ph[(size_t(k + 1) * ny + j) * nx + i] = ph;
        }
        for (k = 0; k < nz; ++k) {
          c.t[I(i, j, k)] = th[k] * (1.0 + kRvovrd * qv[k]) - kT0;
          c.qv[I(i, j, k)] = qv[k];
        }
        c.mu[size_t(j) * nx + i] = mu1;
      }
  }
}


// US standard atmosphere (1976) temperature, K, to 47 km
double us_std_t(double z) {
  if (z < 11000.0) return 288.15 - 0.0065 * z;
  if (z < 20000.0) return 216.65;
  if (z < 32000.0) return 216.65 + 0.001 * (z - 20000.0);
  return 228.65 + 0.0028 * (z - 32000.0);
}
// its pressure from the layer-wise barometric formula (g0 = 9.80665, R = 287.053), Pa
double us_std_p(double z) {
  const double g0 = 9.80665, R = 287.053;
  struct L { double z0, t0, lapse; };
  const L ls[] = {{0.0, 288.15, -0.0065}, {11000.0, 216.65, 0.0}, {20000.0, 216.65, 0.001}, {32000.0, 228.65, 0.0028}};
  double p = 101325.0;
  for (int n = 0; n < 4; ++n) {
    const double z1 = n < 3 ? ls[n + 1].z0 : 1e9, top = std::min(z, z1), dz = top - ls[n].z0;
    if (dz <= 0) break;
    if (ls[n].lapse == 0.0) p *= std::exp(-g0 * dz / (R * ls[n].t0));
    else p *= std::pow((ls[n].t0 + ls[n].lapse * dz) / ls[n].t0, -g0 / (R * ls[n].lapse));
    if (z <= z1) break;
  }
  return p;
}

// A real-data-shaped domain: seeded Gaussian hills (summed over periodic images, capped smoothly), the US standard
// atmosphere, a hybrid coordinate (WRF hybrid_opt 2 with eta_c), seeded u and v jets with +-1 m/s
// noise, map factors 1 +- map_amp, Coriolis from 30 N to 50 N, optional qv from a relative humidity of
// 80 % at the ground falling to 10 % at 10 km.  The base state is dry and hydrostatic column by column
// (each column's surface pressure is the sounding's at its terrain height), as the reference's cases
// build it (src/cases/cases.cpp:224-287) with a horizontally varying mub; theta of the initial state is
// theta_m with moisture, its al and ph the hydrostatic ones at the base pressure.
void build_hills(mi_case& c, const mi_case_spec& sp) {
  const mi_grid& G = sp.grid;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const bool two_d = ny == 1;
  c.full = true;
  uint64_t seed = sp.seed;
  auto u01 = [&]() { return double(splitmix64(seed) >> 11) * 0x1p-53; };
  // sounding: theta of the US standard atmosphere, integrated hydrostatically as the other cases
  auto theta_std = [](double z) { return us_std_t(z) * std::pow(kP1000 / us_std_p(z), kRcp); };
  const double p_surf = 101325.0;
  const double hmax = sp.hill_height;
  Sounding snd(theta_std, p_surf, 40000.0);
  if (!(sp.p_top > 0) || sp.p_top >= p_surf) throw std::invalid_argument("case: p_top must be in (0, p_surf)");
  c.p_top = sp.p_top;
  const double ztop = snd.z_of_p(sp.p_top);
  // eta levels equally spaced in height over flat ground
  std::vector<double> e(size_t(nz) + 1);
  for (int k = 0; k <= nz; ++k) e[size_t(k)] = (snd.p(ztop * k / nz) - c.p_top) / (p_surf - c.p_top);
  e[0] = 1.0;
  e[size_t(nz)] = 0.0;
  set_levels(c, e, G);
  // hybrid coefficients (WRF hybrid_opt 2): B(eta) a cubic from 0 at eta_c to 1 at the ground
  const double ec = sp.etac, pd = kP1000 - c.p_top;
  auto Bf = [&](double eta) {
    if (ec <= 0.0) return eta;
    if (eta < ec) return 0.0;
    const double b1 = 2. * ec * ec * (1. - ec), b2 = -ec * (4. - 3. * ec - ec * ec * ec), b3 = 2. * (1. - ec * ec * ec),
                 b4 = -(1. - ec * ec), b5 = (1. - ec) * (1. - ec) * (1. - ec) * (1. - ec);
    return (b1 + b2 * eta + b3 * eta * eta + b4 * eta * eta * eta) / b5;
  };
  for (int k = 0; k <= nz; ++k) {
    c.c3f[k] = Bf(c.znw[k]);
    c.c4f[k] = (c.znw[k] - c.c3f[k]) * pd;
  }
  for (int k = 0; k < nz; ++k) {
    c.c3h[k] = 0.5 * (c.c3f[k + 1] + c.c3f[k]);
    c.c4h[k] = (c.znu[k] - c.c3h[k]) * pd;
    c.c1h[k] = (c.c3f[k + 1] - c.c3f[k]) / (c.znw[k + 1] - c.znw[k]);
    c.c2h[k] = (1.0 - c.c1h[k]) * pd;
  }
  c.c1f[0] = 1.0;
  for (int k = 1; k < nz; ++k) c.c1f[k] = (c.c3h[k] - c.c3h[k - 1]) / (c.znu[k] - c.znu[k - 1]);
  c.c1f[nz] = 0.0;
  for (int k = 0; k <= nz; ++k) c.c2f[k] = (1.0 - c.c1f[k]) * pd;
  // terrain: Gaussian hills, periodic images
  const double Lx = nx * G.dx, Ly = ny * G.dy;
  struct Hill { double x, y, h, w; };
  std::vector<Hill> hills;
  for (int n = 0; n < sp.n_hills; ++n) {
    Hill h;
    h.x = u01() * Lx;
    h.y = two_d ? 0.0 : u01() * Ly;
    h.h = u01() * hmax;
    h.w = sp.hill_wmin + u01() * (sp.hill_wmax - sp.hill_wmin);
    hills.push_back(h);
  }
  const size_t n2 = size_t(nx) * ny;
  c.ht.assign(n2, 0.0);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double x = (i + 0.5) * G.dx, y = (j + 0.5) * G.dy;
      double h = 0;
      for (const Hill& hl : hills)
        for (int a = -1; a <= 1; ++a)
          for (int b = (two_d ? 0 : -1); b <= (two_d ? 0 : 1); ++b) {
            const double dx = (x - hl.x - a * Lx) / hl.w, dy = two_d ? 0.0 : (y - hl.y - b * Ly) / hl.w;
            h += hl.h * std::exp(-(dx * dx + dy * dy));
          }
      // overlapping hills (and their periodic images on small domains) are capped smoothly at 1.2 times the
      // tallest single hill, as real terrain stays within a few km
      const double cap = 1.2 * hmax;
      c.ht[size_t(j) * nx + i] = cap > 0 ? cap * std::tanh(h / cap) : 0.0;
    }
  // map factors at each staggering, Coriolis
  const int nxu = nx + 1, nyv = two_d ? 1 : ny + 1;
  auto mf = [&](double xi, double yj) {
    return 1.0 + sp.map_amp * std::sin(kPi * xi / nx) * (two_d ? 1.0 : std::cos(kPi * yj / ny));
  };
  c.msftx.assign(n2, 0.0);
  c.msfux.assign(size_t(nxu) * ny, 0.0);
  c.msfvx.assign(size_t(nx) * nyv, 0.0);
  c.fcor.assign(n2, 0.0);
  c.ecor.assign(n2, 0.0);
  const double omega2 = 2.0 * 7.2921e-5;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      c.msftx[size_t(j) * nx + i] = mf(i + 0.5, j + 0.5);
      const double lat = (30.0 + 20.0 * (j + 0.5) / ny) * kPi / 180.0;
      c.fcor[size_t(j) * nx + i] = omega2 * std::sin(lat);
      c.ecor[size_t(j) * nx + i] = omega2 * std::cos(lat);
    }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nxu; ++i) c.msfux[size_t(j) * nxu + i] = mf(i, j + 0.5);
  for (int j = 0; j < nyv; ++j)
    for (int i = 0; i < nx; ++i) c.msfvx[size_t(j) * nx + i] = mf(i + 0.5, j);
  c.msfty = c.msftx;
  c.msfuy = c.msfux;
  c.msfvy = c.msfvx;
  c.msfvx_inv = c.msfvx;
  for (double& m : c.msfvx_inv) m = 1.0 / m;
  // base state, column by column
  const size_t nm = n2 * size_t(nz), nw = n2 * size_t(nz + 1);
  c.mub2.assign(n2, 0.0);
  c.pb3.assign(nm, 0.0);
  c.alb3.assign(nm, 0.0);
  c.t_init3.assign(nm, 0.0);
  c.phb3.assign(nw, 0.0);
  c.t.assign(nm, 0.0);
  c.al.assign(nm, 0.0);
  c.p.assign(nm, 0.0);
  c.ph.assign(nw, 0.0);
  c.w.assign(nw, 0.0);
  c.mu.assign(n2, 0.0);
  if (sp.moist) c.qv.assign(nm, 0.0);
  std::vector<double> zm(nm);
  auto I = [&](int i, int j, int k) { return (size_t(k) * ny + j) * nx + i; };
  const double rvrd = 461.6 / kRd;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t q = size_t(j) * nx + i;
      const double mub = snd.p(c.ht[q]) - c.p_top;
      c.mub2[q] = mub;
      double phb = kG * c.ht[q], ph = 0.0;
      c.phb3[I(i, j, 0)] = phb;
      for (int k = 0; k < nz; ++k) {
        const double pb = c.c3h[k] * mub + c.c4h[k] + c.p_top;
        const double z = snd.z_of_p(pb);
        const double tb = snd.theta(z);
        const double alb = (kRd / kP1000) * tb * std::pow(pb / kP1000, kCvpm);
        c.pb3[I(i, j, k)] = pb;
        c.alb3[I(i, j, k)] = alb;
        c.t_init3[I(i, j, k)] = tb - kT0;
        double th = tb;
        if (sp.moist) {
          const double tk = tb * std::pow(pb / kP1000, kRcp);
          const double rh = z < 10000.0 ? 0.8 - 0.7 * z / 10000.0 : 0.1;
          const double es = 611.2 * std::exp(17.67 * (tk - 273.15) / (tk - 29.65));  // Bolton (1980)
          const double qv = rh * 0.622 * es / (pb - es);
          c.qv[I(i, j, k)] = qv;
          th = tb * (1.0 + rvrd * qv);  // theta_m
        }
        c.t[I(i, j, k)] = th - kT0;
        const double alt = (kRd / kP1000) * th * std::pow(pb / kP1000, kCvpm);
        c.al[I(i, j, k)] = alt - alb;
        phb -= c.dnw[k] * (c.c1h[k] * mub + c.c2h[k]) * alb;
        ph -= c.dnw[k] * ((c.c1h[k] * mub + c.c2h[k]) * c.al[I(i, j, k)]);
        c.phb3[I(i, j, k + 1)] = phb;
        c.ph[I(i, j, k + 1)] = ph;
        zm[I(i, j, k)] = z;
      }
    }
  // jets: u(z), v(z) Gaussian in height around 10 km, maxima seeded, plus +-1 m/s noise
  const double uj = sp.jet_min + u01() * (sp.jet_max - sp.jet_min), vj = sp.jet_min + u01() * (sp.jet_max - sp.jet_min);
  auto jet = [](double amp, double z) { const double s = (z - 10000.0) / 4000.0; return amp * std::exp(-s * s); };
  c.u.assign(size_t(nxu) * ny * nz, 0.0);
  c.v.assign(size_t(nx) * nyv * nz, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nxu; ++i) {
        const double z = zm[I(i < nx ? i : 0, j, k)];
        double uu = jet(uj, z) + (2.0 * u01() - 1.0);
        if (G.bc_x == MI_BC_SYMMETRIC && (i == 0 || i == nx)) uu = 0.0;
        c.u[(size_t(k) * ny + j) * nxu + i] = uu;
      }
  if (!two_d)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < nyv; ++j)
        for (int i = 0; i < nx; ++i) {
          const double z = zm[I(i, j < ny ? j : 0, k)];
          double vv = jet(vj, z) + (2.0 * u01() - 1.0);
          if (G.bc_y == MI_BC_SYMMETRIC && (j == 0 || j == ny)) vv = 0.0;
          c.v[(size_t(k) * nyv + j) * nx + i] = vv;
        }
  // periodic u at i = nx, v at j = ny: the values of i = 0, j = 0
  if (G.bc_x == MI_BC_PERIODIC)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) c.u[(size_t(k) * ny + j) * nxu + nx] = c.u[(size_t(k) * ny + j) * nxu + 0];
  if (!two_d && G.bc_y == MI_BC_PERIODIC)
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i) c.v[(size_t(k) * nyv + ny) * nx + i] = c.v[(size_t(k) * nyv + 0) * nx + i];
  (void)ztop;
}

}  // namespace

extern "C" {

int mi_case_create(const mi_case_spec* spec, mi_case** out) {
  return mi_detail::guard([&] {
    if (!spec || !out) throw std::invalid_argument("null argument");
    auto c = std::make_unique<mi_case>();
    c->kind = spec->kind;
    if (spec->kind == MI_CASE_HILLS) {
      const mi_grid& G = spec->grid;
      if (G.nx < 1 || G.ny < 1 || G.nz < 3) throw std::invalid_argument("case: bad grid size");
      c->g = G;
      c->nz = G.nz;
      build_hills(*c, *spec);
    } else {
      build(*c, *spec);
    }
    // the 1-D base profiles of column (0, 0) (module_initialize_ideal as the reference, model.cpp:149-177): u, v,
    // qv; theta of t_init; the height of the mass levels from phb
    const int nx = c->g.nx, ny = c->g.ny, nz = c->nz, nxu = nx + 1;
    for (auto* v : {&c->u_base, &c->v_base, &c->qv_base, &c->t_base, &c->z_base}) v->assign(size_t(nz), 0.0);
    const size_t col = size_t(ny) * size_t(nx);  // the level stride of a dense mass or w field
    auto phb0 = [&](int k) { return c->full ? c->phb3[size_t(k) * col] : c->phb[size_t(k)]; };
    for (int k = 0; k < nz; ++k) {
      c->u_base[size_t(k)] = c->u[size_t(k) * ny * nxu];
      c->v_base[size_t(k)] = c->v[size_t(k) * (ny > 1 ? ny + 1 : 1) * nx];
      if (!c->qv.empty()) c->qv_base[size_t(k)] = c->qv[size_t(k) * ny * nx];
      c->t_base[size_t(k)] = c->full ? c->t_init3[size_t(k) * col] : c->t_init[size_t(k)];
      c->z_base[size_t(k)] = 0.5 * (phb0(k) + phb0(k + 1)) / 9.81;
    }
    *out = c.release();
  });
}

int mi_case_destroy(mi_case* c) {
  delete c;
  return MI_OK;
}

int mi_case_levels(const mi_case* c, mi_levels* o) {
  return mi_detail::guard([&] {
    if (!c || !o) throw std::invalid_argument("null argument");
    o->dnw = c->dnw.data(); o->rdnw = c->rdnw.data();
    o->dn = c->dn.data(); o->rdn = c->rdn.data(); o->fnm = c->fnm.data(); o->fnp = c->fnp.data();
    o->c1h = c->c1h.data(); o->c2h = c->c2h.data(); o->c3h = c->c3h.data(); o->c4h = c->c4h.data();
    o->c1f = c->c1f.data(); o->c2f = c->c2f.data(); o->c3f = c->c3f.data(); o->c4f = c->c4f.data();
    o->cf1 = c->cf1; o->cf2 = c->cf2; o->cf3 = c->cf3; o->cfn = c->cfn; o->cfn1 = c->cfn1;
    o->rdx = c->rdx; o->rdy = c->rdy; o->p_top = c->p_top;
  });
}

int mi_case_get_field(const mi_case* c, int id, double* h, int64_t n) {
  return mi_detail::guard([&] {
    if (!c || !h) throw std::invalid_argument("null argument");
    const mi_grid& g = c->g;
    const int nz = c->nz;
    auto put = [&](const std::vector<double>& src) {
      if (n != int64_t(src.size())) throw std::invalid_argument("case field size mismatch");
      std::memcpy(h, src.data(), src.size() * sizeof(double));
    };
    // horizontally uniform base profile expanded to nxs*nys*nzs points
    auto put_profile = [&](const std::vector<double>& prof) {
      const int64_t plane = int64_t(g.nx) * g.ny;
      if (n != plane * int64_t(prof.size())) throw std::invalid_argument("case field size mismatch");
      for (size_t k = 0; k < prof.size(); ++k)
        for (int64_t q = 0; q < plane; ++q) h[int64_t(k) * plane + q] = prof[k];
    };
    if (c->full) {  // hills: the base state in full and the fields of a real-data-shaped domain
      switch (id) {
        case MI_MUB: put(c->mub2); return;
        case MI_PB: put(c->pb3); return;
        case MI_ALB: put(c->alb3); return;
        case MI_PHB: put(c->phb3); return;
        case MI_T_INIT: put(c->t_init3); return;
        case MI_HT: put(c->ht); return;
        case MI_MSFTX: put(c->msftx); return;
        case MI_MSFTY: put(c->msfty); return;
        case MI_MSFUX: put(c->msfux); return;
        case MI_MSFUY: put(c->msfuy); return;
        case MI_MSFVX: put(c->msfvx); return;
        case MI_MSFVY: put(c->msfvy); return;
        case MI_MSFVX_INV: put(c->msfvx_inv); return;
        case MI_F: put(c->fcor); return;
        case MI_E: put(c->ecor); return;
        case MI_Q:
          if (c->qv.empty()) throw std::invalid_argument("case: dry, no qv");
          put(c->qv);
          return;
        default: break;
      }
    }
    switch (id) {
      case MI_U: put(c->u); break;
      case MI_V: put(c->v); break;
      case MI_W: put(c->w); break;
      case MI_T: put(c->t); break;
      case MI_PH: put(c->ph); break;
      case MI_MU: put(c->mu); break;
      case MI_P: put(c->p); break;
      case MI_AL: put(c->al); break;
      case MI_MUB: put_profile(std::vector<double>{c->mub}); break;
      case MI_PB: put_profile(c->pb); break;
      case MI_ALB: put_profile(c->alb); break;
      case MI_PHB: put_profile(c->phb); break;
      case MI_T_INIT: put_profile(c->t_init); break;
      case MI_U_BASE: put(c->u_base); break;
      case MI_V_BASE: put(c->v_base); break;
      case MI_QV_BASE: put(c->qv_base); break;
      case MI_T_BASE: put(c->t_base); break;
      case MI_Z_BASE: put(c->z_base); break;
      case MI_Q:
        if (c->qv.empty()) throw std::invalid_argument("case: dry, no qv");
        put(c->qv);
        break;
      default: throw std::invalid_argument("case: no field " + std::to_string(id));
    }
    (void)nz;
  });
}

int mi_case_load(const mi_case* c, mi_domain* d) {
  return mi_detail::guard([&] {
    if (!c || !d) throw std::invalid_argument("null argument");
    auto chk = [](int rc) {
      if (rc != MI_OK) throw mi_detail::Error(rc, mi_last_error());
    };
    auto set = [&](int id, const std::vector<double>& v) { chk(mi_domain_set_field(d, id, v.data(), int64_t(v.size()))); };
    if (c->full) {
      set(MI_MUB, c->mub2);
      set(MI_PB, c->pb3);
      set(MI_ALB, c->alb3);
      set(MI_PHB, c->phb3);
      set(MI_T_INIT, c->t_init3);
      set(MI_HT, c->ht);
      set(MI_MSFTX, c->msftx);
      set(MI_MSFTY, c->msfty);
      set(MI_MSFUX, c->msfux);
      set(MI_MSFUY, c->msfuy);
      set(MI_MSFVX, c->msfvx);
      set(MI_MSFVY, c->msfvy);
      set(MI_MSFVX_INV, c->msfvx_inv);
      set(MI_F, c->fcor);
      set(MI_E, c->ecor);
      if (!c->qv.empty()) set(MI_Q, c->qv);
    } else {
      chk(mi_domain_set_field(d, MI_MUB, &c->mub, 1));
      chk(mi_domain_set_field(d, MI_PB, c->pb.data(), int64_t(c->pb.size())));
      chk(mi_domain_set_field(d, MI_ALB, c->alb.data(), int64_t(c->alb.size())));
      chk(mi_domain_set_field(d, MI_PHB, c->phb.data(), int64_t(c->phb.size())));
      chk(mi_domain_set_field(d, MI_T_INIT, c->t_init.data(), int64_t(c->t_init.size())));
      if (!c->qv.empty()) set(MI_Q, c->qv);
    }
    chk(mi_domain_set_field(d, MI_U, c->u.data(), int64_t(c->u.size())));
    chk(mi_domain_set_field(d, MI_V, c->v.data(), int64_t(c->v.size())));
    chk(mi_domain_set_field(d, MI_W, c->w.data(), int64_t(c->w.size())));
    chk(mi_domain_set_field(d, MI_T, c->t.data(), int64_t(c->t.size())));
    chk(mi_domain_set_field(d, MI_PH, c->ph.data(), int64_t(c->ph.size())));
    chk(mi_domain_set_field(d, MI_MU, c->mu.data(), int64_t(c->mu.size())));
    chk(mi_domain_set_field(d, MI_P, c->p.data(), int64_t(c->p.size())));
    chk(mi_domain_set_field(d, MI_AL, c->al.data(), int64_t(c->al.size())));
    set(MI_U_BASE, c->u_base);
    set(MI_V_BASE, c->v_base);
    set(MI_QV_BASE, c->qv_base);
    set(MI_T_BASE, c->t_base);
    set(MI_Z_BASE, c->z_base);
  });
}

}  // extern "C"
