// CPU check of the interior-tile FP64 helpers (axisymmetricjet_amd/csrc/ajet_fastmath.h): the
// host build of the same functions against the C library and the unfused stencil sums.  Prints
// "ok" or the first failure.
#include <cmath>
#include <cstdio>
#include <random>

#include "ajet_fastmath.h"

int main() {
  std::mt19937_64 g(7);
  double worst = 0.0, worst_x = 0.0;
  auto check_pow = [&](double x) {
    const double want = std::pow(x, 0.666), got = ajet::pow0666(x);
    const double rel = std::fabs(got - want) / want;
    if (rel > worst) { worst = rel; worst_x = x; }
  };
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (int k = 0; k < 2000000; ++k) check_pow(0.05 + 20.0 * u(g));        // temperatures of the flows
  for (int k = 0; k < 200000; ++k) check_pow(std::exp2(-8.0 + 16.0 * u(g)));  // the whole fast range
  for (double x : {1.0, 0.5, 2.0, 0.7071067811865476, 1.4142135623730951, 0x1.000001p-8, 0x1.fffffp7})
    check_pow(x);
  if (worst > 1e-15) {
    std::printf("pow0666: relative error %.3e at x = %.17g\n", worst, worst_x);
    return 1;
  }
  // outside the fast range: the library's pow itself
  for (double x : {0x1p-8, 0x1p8, 1e-30, 3e5, 1e300, 0.0})
    if (ajet::pow0666(x) != std::pow(x, 0.666)) {
      std::printf("pow0666: fallback at %.17g\n", x);
      return 1;
    }
  if (!std::isnan(ajet::pow0666(NAN)) || ajet::pow0666(INFINITY) != INFINITY || !std::isnan(ajet::pow0666(-1.0))) {
    std::printf("pow0666: special values\n");
    return 1;
  }
  // rcp within an ulp
  for (int k = 0; k < 1000000; ++k) {
    const double x = std::exp2(-60.0 + 120.0 * u(g)) * (u(g) < 0.5 ? -1.0 : 1.0);
    const double want = 1.0 / x, got = ajet::rcp(x);
    if (std::fabs(got - want) > 1.0001 * std::fabs(std::nextafter(want, 2.0 * want) - want)) {
      std::printf("rcp: %.17g -> %.17g, want %.17g\n", x, got, want);
      return 1;
    }
  }
  // stencils against the unfused sums (a few ulp of the inputs) and exact on polynomials
  for (int k = 0; k < 100000; ++k) {
    double c[5];
    for (double& v : c) v = 1.0 + 0.1 * u(g);
    const double d1 = ((1.0 / 12 * c[0] + -(2.0 / 3) * c[1]) + 2.0 / 3 * c[3]) + -(1.0 / 12) * c[4];
    const double d2 = ((((-(1.0 / 12)) * c[0] + 4.0 / 3 * c[1]) + -(5.0 / 2) * c[2]) + 4.0 / 3 * c[3]) + -(1.0 / 12) * c[4];
    if (std::fabs(ajet::d1_int(c[0], c[1], c[3], c[4]) - d1) > 1e-15 ||
        std::fabs(ajet::d2_int(c[0], c[1], c[2], c[3], c[4]) - d2) > 2e-15) {
      std::printf("stencil mismatch\n");
      return 1;
    }
  }
  for (int p = 0; p <= 4; ++p) {
    // D1 is exact for degree <= 4, D2 for degree <= 5: unit spacing, x = -2..2 around 0.3
    auto f = [&](double x) { return std::pow(0.3 + x, p); };
    const double want1 = p == 0 ? 0.0 : p * std::pow(0.3, p - 1);
    const double want2 = p < 2 ? 0.0 : p * (p - 1) * std::pow(0.3, p - 2);
    if (std::fabs(ajet::d1_int(f(-2), f(-1), f(1), f(2)) - want1) > 1e-13 ||
        std::fabs(ajet::d2_int(f(-2), f(-1), f(0), f(1), f(2)) - want2) > 1e-13) {
      std::printf("stencil exactness, degree %d\n", p);
      return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
