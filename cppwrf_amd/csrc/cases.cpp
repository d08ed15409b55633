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

// Exner function of height for a given theta(z), hydrostatic, on a table of step dz.
class Sounding {
 public:
  Sounding(std::function<double(double)> th, double p_surf, double zmax, double dz = 0.5) : th_(std::move(th)), dz_(dz) {
    const int n = int(std::ceil(zmax / dz)) + 2;
    pi_.resize(size_t(n));
    pi_[0] = std::pow(p_surf / kP1000, kRcp);
    for (int m = 1; m < n; ++m) {
      const double z0 = (m - 1) * dz, z1 = m * dz;
      const double f = (1.0 / th_(z0) + 4.0 / th_(0.5 * (z0 + z1)) + 1.0 / th_(z1)) / 6.0;
      pi_[size_t(m)] = pi_[size_t(m - 1)] - kG / kCp * f * dz;
    }
  }
  double theta(double z) const { return th_(z); }
  double p(double z) const {
    const double s = z / dz_;
    const int m = int(std::floor(s));
    if (m < 0 || m + 1 >= int(pi_.size())) throw std::runtime_error("sounding: height out of range");
    const double z0 = m * dz_, h = z - z0;
    const double f = (1.0 / th_(z0) + 4.0 / th_(z0 + 0.5 * h) + 1.0 / th_(z)) / 6.0;
    const double pi = pi_[size_t(m)] - kG / kCp * f * h;
    return kP1000 * std::pow(pi, 1.0 / kRcp);
  }
  double z_of_p(double pp) const {
    double lo = 0, hi = double(pi_.size() - 2) * dz_;
    if (pp > p(lo) || pp < p(hi)) throw std::runtime_error("sounding: pressure out of range");
    for (int it = 0; it < 200 && hi - lo > 1e-9; ++it) {
      const double mid = 0.5 * (lo + hi);
      (p(mid) > pp ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
  }

 private:
  std::function<double(double)> th_;
  double dz_;
  std::vector<double> pi_;
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
  int nz = 0;
  std::vector<double> dnw, rdnw, dn, rdn, fnm, fnp, c1h, c2h, c3h, c4h, c1f, c2f, c3f, c4f, znw, znu;
  double cf1 = 0, cf2 = 0, cf3 = 0, cfn = 0, cfn1 = 0, rdx = 0, rdy = 0, p_top = 0;
  double mub = 0;
  std::vector<double> pb, alb, phb, t_init;             // profiles
  std::vector<double> u, v, w, t, ph, mu, p, al;        // dense state
};

namespace {


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
  // BaseState::set_levels, hybrid_opt = 0 (state.cpp:86-116)
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
  // dry hydrostatic base state and perturbation state (cases.cpp:224-287)
  const double mub = p_surf - c.p_top;
  c.mub = mub;
  std::vector<double> tb(static_cast<size_t>(nz)), zm(static_cast<size_t>(nz)), exn(static_cast<size_t>(nz));
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
}

}  // namespace

extern "C" {

int mi_case_create(const mi_case_spec* spec, mi_case** out) {
  return mi_detail::guard([&] {
    if (!spec || !out) throw std::invalid_argument("null argument");
    auto c = std::make_unique<mi_case>();
    build(*c, *spec);
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
    chk(mi_domain_set_field(d, MI_MUB, &c->mub, 1));
    chk(mi_domain_set_field(d, MI_PB, c->pb.data(), int64_t(c->pb.size())));
    chk(mi_domain_set_field(d, MI_ALB, c->alb.data(), int64_t(c->alb.size())));
    chk(mi_domain_set_field(d, MI_PHB, c->phb.data(), int64_t(c->phb.size())));
    chk(mi_domain_set_field(d, MI_T_INIT, c->t_init.data(), int64_t(c->t_init.size())));
    chk(mi_domain_set_field(d, MI_U, c->u.data(), int64_t(c->u.size())));
    chk(mi_domain_set_field(d, MI_V, c->v.data(), int64_t(c->v.size())));
    chk(mi_domain_set_field(d, MI_W, c->w.data(), int64_t(c->w.size())));
    chk(mi_domain_set_field(d, MI_T, c->t.data(), int64_t(c->t.size())));
    chk(mi_domain_set_field(d, MI_PH, c->ph.data(), int64_t(c->ph.size())));
    chk(mi_domain_set_field(d, MI_MU, c->mu.data(), int64_t(c->mu.size())));
    chk(mi_domain_set_field(d, MI_P, c->p.data(), int64_t(c->p.size())));
    chk(mi_domain_set_field(d, MI_AL, c->al.data(), int64_t(c->al.size())));
  });
}

}  // extern "C"
