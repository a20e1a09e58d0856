// Host harness of the library's host-only code (dfk_model.cpp: disk models,
// equilibrium, Toomre-Q solve, mass, growth extractor, random streams), built
// with AddressSanitizer and UndefinedBehaviorSanitizer (Makefile target
// bin/model_check_asan) and run by tests/test_host_sanitize.py.  It calls
// every entry point over small and odd grids, unreachable Q targets, and
// extractor inputs that are short, flat, non-finite or never grow, and
// checks the documented return codes; the sanitizers check the memory
// accesses and arithmetic.  Prints "model_check: ok" on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "diskfft_model.h"

namespace {

int fails = 0;

void expect(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "model_check: FAILED %s\n", what);
    ++fails;
  }
}

void states(int kind, int Nu, int Nphi, bool with_g, double qmin) {
  dfk_model m;
  dfk_model_defaults(kind, &m);
  if (qmin > 0) m.Qmin = qmin;
  const size_t N = (size_t)Nu * Nphi;
  std::vector<double> sig0(Nu), g(Nu), S(N), U(N), W(N), Q(Nu);
  expect(dfk_model_sigma0(&m, Nu, Nphi, sig0.data()) == 0, "sigma0");
  for (int i = 0; i < Nu; ++i) g[i] = -0.1 * sig0[i];
  const int rc = dfk_model_state(&m, Nu, Nphi, with_g ? g.data() : nullptr, S.data(), U.data(),
                                 W.data(), Q.data());
  expect(rc == 0, "state");
  bool finite = true;
  for (size_t k = 0; k < N; ++k) finite &= std::isfinite(S[k]) && std::isfinite(U[k]) && std::isfinite(W[k]);
  expect(finite, "finite state");
  const double mass = dfk_model_mass(&m, Nu, Nphi, S.data());
  expect(std::isfinite(mass) && mass > 0, "mass");
  if (kind == DFK_MODEL_POWERLAW) expect(m.K > 0 && std::isfinite(m.Qmin_got), "K solve");
}

void extract(std::vector<double> t, std::vector<double> c0, std::vector<double> c2,
             std::vector<double> ph, int want_rc, const char *what) {
  dfk_growth g;
  const int n = (int)t.size();
  const int r1 = dfk_extract_growth(t.data(), c0.data(), c2.data(), ph.data(), n, 3.0, 0.3, 10.0, &g);
  const int r2 = dfk_extract_growth_ex(t.data(), c0.data(), c2.data(), ph.data(), n, 10.0, 100.0, 0.3,
                                       10.0, 0.25, 2.0, &g);
  expect(r1 == want_rc && r2 == want_rc, what);
}

}  // namespace

int main() {
  const int grids[][2] = {{3, 4}, {17, 46}, {35, 40}, {64, 64}, {128, 96}};
  for (const auto &gr : grids) {
    states(DFK_MODEL_LKA98, gr[0], gr[1], true, 0);
    states(DFK_MODEL_LKA98, gr[0], gr[1], false, 0);
    states(DFK_MODEL_POWERLAW, gr[0], gr[1], true, 0);
    states(DFK_MODEL_POWERLAW, gr[0], gr[1], true, 1e9);   // unreachable: best K
  }
  // extractor: a growing, rotating mode; a flat series; short; non-finite
  const int n = 400;
  std::vector<double> t(n), c0(n, 1.0), c2(n), ph(n);
  for (int k = 0; k < n; ++k) {
    t[k] = 0.05 * k;
    c2[k] = 1e-3 * std::exp(0.7 * t[k]) / (1.0 + 1e-3 * std::exp(0.7 * t[k]) / 0.2);
    ph[k] = std::remainder(-2.0 * 1.3 * t[k], 2.0 * M_PI);
  }
  extract(t, c0, c2, ph, 0, "growing series");
  extract(t, c0, std::vector<double>(n, 1e-3), ph, 0, "flat series");
  extract({0.0, 0.1, 0.2}, {1, 1, 1}, {1e-3, 2e-3, 4e-3}, {0, 0, 0}, 1, "three samples");
  std::vector<double> bad = c2;
  bad[10] = NAN;
  bad[200] = INFINITY;
  {
    dfk_growth g;
    dfk_extract_growth(t.data(), c0.data(), bad.data(), ph.data(), n, 3.0, 0.3, 10.0, &g);
    dfk_extract_growth_ex(t.data(), c0.data(), bad.data(), ph.data(), n, 10.0, 100.0, 0.3, 10.0, 0.25,
                          2.0, &g);
  }
  {
    dfk_growth g;
    expect(dfk_extract_growth_ex(t.data(), c0.data(), c2.data(), ph.data(), n, 10.0, 100.0, 0.3, 10.0,
                                 0.0, 2.0, &g) != 0, "sat_frac 0 refused");
  }
  unsigned long long st = 42;
  double s = 0, s2 = 0;
  for (int k = 0; k < 100000; ++k) {
    const double z = dfk_normal(&st);
    s += z;
    s2 += z * z;
  }
  expect(std::fabs(s / 1e5) < 0.02 && std::fabs(s2 / 1e5 - 1.0) < 0.02, "normal moments");
  if (fails) return 1;
  std::printf("model_check: ok\n");
  return 0;
}
