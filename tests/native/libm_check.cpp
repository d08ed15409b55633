// Compares the glibc pow/exp/sin ports of cppwrf_amd/csrc/libm_port.hpp (host build of the same source the
// device compiles) with the host glibc on seeded samples over the ranges the dynamics path uses:
// pow(x, cp/cv) for calc_p_rho_phi, pow(x, y) on a wide range, sin on [0, 2.4262] for the Rayleigh layer, cos on
// [0, pi/2] for the damping of the eddy viscosities, and
// the Kessler microphysics' exp (saturation vapour pressure) and pow (fall speed, rates, pressure from Exner),
// including pow(0, y) = 0.
// Exit status 0 when every result is bit-identical.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../cppwrf_amd/csrc/libm_port.hpp"

static uint64_t state = 0x5eed0123456789ULL;
static uint64_t splitmix64() {
  uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static double u01() { return double(splitmix64() >> 11) * 0x1p-53; }
static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  const double cpovcv = 1004.5 / (1004.5 - 287.0);
  long bad[6] = {0, 0, 0, 0, 0, 0};
  const double kexp[5] = {0.1364, 0.875, 1004.0 / 287.0, 0.2046, 0.525};
  for (long t = 0; t < n; ++t) {
    double x = 0.2 + 3.0 * u01();
    double a = std::pow(x, cpovcv), b = mi_libm::pow(x, cpovcv);
    if (!same(a, b) && bad[0]++ < 3) std::printf("pow(%a, cp/cv): glibc %a port %a\n", x, a, b);
    x = 0.01 + 10.0 * u01();
    const double y = -5.0 + 10.0 * u01();
    a = std::pow(x, y);
    b = mi_libm::pow(x, y);
    if (!same(a, b) && bad[1]++ < 3) std::printf("pow(%a, %a): glibc %a port %a\n", x, y, a, b);
    const double z = 2.4262 * u01();
    a = std::sin(z);
    b = mi_libm::sin(z);
    if (!same(a, b) && bad[2]++ < 3) std::printf("sin(%a): glibc %a port %a\n", z, a, b);
    // cos over [0, pi/2] (the damping layer of the eddy viscosities), with tiny arguments now and then
    const double zc = (t % 53 == 0) ? std::ldexp(u01(), -int(20 + splitmix64() % 20)) : 1.5707963267948966 * u01();
    a = std::cos(zc);
    b = mi_libm::cos(zc);
    if (!same(a, b) && bad[5]++ < 3) std::printf("cos(%a): glibc %a port %a\n", zc, a, b);
    const double e = -60.0 + 70.0 * u01();
    a = std::exp(e);
    b = mi_libm::exp(e);
    if (!same(a, b) && bad[3]++ < 3) std::printf("exp(%a): glibc %a port %a\n", e, a, b);
    const double ye = kexp[t % 5];
    // mixing ratios and densities over many decades, Exner values, and exact zeros
    double xk = (t % 5 == 2) ? 0.3 + 0.8 * u01() : std::ldexp(u01(), -int(splitmix64() % 40));
    if (t % 97 == 0) xk = 0.0;
    a = std::pow(xk, ye);
    b = mi_libm::pow(xk, ye);
    if (!same(a, b) && bad[4]++ < 3) std::printf("pow(%a, %a): glibc %a port %a\n", xk, ye, a, b);
  }
  std::printf("samples %ld, mismatches: pow(x, cp/cv) %ld, pow(x, y) %ld, sin %ld, exp %ld, Kessler pow %ld, cos %ld\n",
              n, bad[0], bad[1], bad[2], bad[3], bad[4], bad[5]);
  return bad[0] + bad[1] + bad[2] + bad[3] + bad[4] + bad[5] != 0;
}
