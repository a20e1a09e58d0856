// Disk models and equilibrium initial conditions (host code, see
// include/diskfft_model.h).  The LKA98 ring follows diskmain.cpp:28-63 line
// for line; the power-law model is this project's own (the reference has no
// counterpart) and is pinned by its own tests (mass, Q_min, noise statistics,
// determinism).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/diskfft_model.h"

namespace {

struct Grid {
  std::vector<double> R, Rf;
  double du, dphi;
};

Grid make_grid(const dfk_model *m, int Nu, int Nphi) {
  Grid g;
  const double umin = std::log(m->Rin), umax = std::log(m->Rout);   // hydro.hpp:31-38
  g.du = (umax - umin) / Nu;
  g.dphi = 2.0 * M_PI / Nphi;
  g.R.resize(Nu);
  g.Rf.resize(Nu + 1);
  for (int i = 0; i <= Nu; ++i) g.Rf[i] = std::exp(umin + i * g.du);
  for (int i = 0; i < Nu; ++i) g.R[i] = std::exp(umin + (i + 0.5) * g.du);
  return g;
}

// LKA98 ring, diskmain.cpp:28,33
const double kR0 = 0.45, kW2 = 0.05, kS0 = 0.372;
double lka_sigma0(double R) { return kS0 * std::exp(-(R - kR0) * (R - kR0) / kW2); }

// power law x tanh tapers, unnormalised shape and its derivative
double pl_shape(const dfk_model *m, double R) {
  const double t1 = 0.5 * (1.0 + std::tanh((R - m->taper_in) / m->taper_w));
  const double t2 = 0.5 * (1.0 - std::tanh((R - m->taper_out) / m->taper_w));
  return std::pow(R, m->slope) * t1 * t2;
}
double pl_dshape(const dfk_model *m, double R) {
  const double x1 = (R - m->taper_in) / m->taper_w, x2 = (R - m->taper_out) / m->taper_w;
  const double t1 = 0.5 * (1.0 + std::tanh(x1)), t2 = 0.5 * (1.0 - std::tanh(x2));
  const double c1 = 1.0 / std::cosh(x1), c2 = 1.0 / std::cosh(x2);
  const double d1 = 0.5 * c1 * c1 / m->taper_w, d2 = -0.5 * c2 * c2 / m->taper_w;
  const double p = std::pow(R, m->slope);
  return m->slope * std::pow(R, m->slope - 1.0) * t1 * t2 + p * d1 * t2 + p * t1 * d2;
}

double pl_norm(const dfk_model *m, const Grid &g) {
  double s = 0.0;
  for (size_t i = 0; i < g.R.size(); ++i)
    s += pl_shape(m, g.R[i]) * g.R[i] * (g.Rf[i + 1] - g.Rf[i]) * 2.0 * M_PI;
  return m->mass_ratio * (m->gmstar / m->G) / s;
}

// Toomre Q profile for a given K; kappa^2 = R^-3 d(R^2 v_phi^2)/dR on the grid.
void q_profile(const dfk_model *m, const Grid &g, const std::vector<double> &sig0,
               const std::vector<double> &dsig0, const double *gRdisk, double K,
               std::vector<double> &Q, std::vector<double> &v2) {
  const int Nu = (int)g.R.size();
  v2.resize(Nu);
  Q.resize(Nu);
  std::vector<double> L(Nu);
  for (int i = 0; i < Nu; ++i) {
    const double R = g.R[i], s = sig0[i];
    const double dPdR = K * m->gamma * std::pow(s, m->gamma - 1.0) * dsig0[i];
    v2[i] = m->gmstar / R + (s > 0 ? (R / s) * dPdR : 0.0) - R * (gRdisk ? gRdisk[i] : 0.0);
    L[i] = R * R * v2[i];
  }
  for (int i = 0; i < Nu; ++i) {
    const int a = (i > 0) ? i - 1 : i, b = (i < Nu - 1) ? i + 1 : i;
    const double dLdR = (L[b] - L[a]) / (g.R[b] - g.R[a]);
    const double kap2 = dLdR / (g.R[i] * g.R[i] * g.R[i]);
    const double s = sig0[i];
    const double cs2 = m->gamma * K * std::pow(s, m->gamma - 1.0);
    Q[i] = (kap2 > 0 && s > 0) ? std::sqrt(cs2 * kap2) / (M_PI * m->G * s) : 0.0;
  }
}

double min_q(const dfk_model *m, const Grid &g, const std::vector<double> &Q) {
  double q = INFINITY;
  for (size_t i = 0; i < g.R.size(); ++i)
    if (g.R[i] >= m->q_in && g.R[i] <= m->q_out && Q[i] < q) q = Q[i];
  return q;
}

}  // namespace

extern "C" {

unsigned long long dfk_splitmix64(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double dfk_normal(unsigned long long *s) {
  // u1 in (0,1], u2 in [0,1), 53-bit uniforms
  const double u1 = ((dfk_splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (dfk_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

void dfk_model_defaults(int kind, dfk_model *m) {
  std::memset(m, 0, sizeof *m);
  m->kind = kind;
  m->G = 1.0;
  m->gamma = 2.0;
  if (kind == DFK_MODEL_LKA98) {
    m->Rin = 0.25; m->Rout = 1.0; m->gmstar = 0.6; m->K = 0.25; m->amp = 1e-3;   // diskmain.cpp:28-29
  } else {
    m->Rin = 0.1; m->Rout = 1.0; m->gmstar = 1.0; m->K = 0.0;
    m->mass_ratio = 1.0; m->Qmin = 1.3; m->slope = -1.5;
    m->taper_in = 0.15; m->taper_out = 0.9; m->taper_w = 0.02;
    m->q_in = 0.2; m->q_out = 0.8;
    m->noise_rms = 1e-3; m->mmin = 1; m->mmax = 8; m->seed = 42;
  }
}

int dfk_model_sigma0(const dfk_model *m, int Nu, int Nphi, double *sig0) {
  if (!m || Nu < 1 || Nphi < 1 || !sig0) return 1;
  Grid g = make_grid(m, Nu, Nphi);
  if (m->kind == DFK_MODEL_LKA98) {
    for (int i = 0; i < Nu; ++i) sig0[i] = lka_sigma0(g.R[i]);
  } else {
    const double A = pl_norm(m, g);
    for (int i = 0; i < Nu; ++i) sig0[i] = A * pl_shape(m, g.R[i]);
  }
  return 0;
}

int dfk_model_state(dfk_model *m, int Nu, int Nphi, const double *gRdisk, double *Sig, double *vR,
                    double *vp, double *Qout) {
  if (!m || Nu < 3 || Nphi < 1 || !Sig || !vR || !vp) return 1;
  Grid g = make_grid(m, Nu, Nphi);
  if (m->kind == DFK_MODEL_LKA98) {
    // diskmain.cpp:50-63
    for (int i = 0; i < Nu; ++i) {
      const double R = g.R[i];
      const double dSig = lka_sigma0(R) * (-2.0 * (R - kR0) / kW2);
      const double dPdR = m->K * m->gamma * std::pow(lka_sigma0(R), m->gamma - 1.0) * dSig;
      const double v2 = m->gmstar / R + (R / lka_sigma0(R)) * dPdR - R * (gRdisk ? gRdisk[i] : 0.0);
      const double vphi = (v2 > 0) ? std::sqrt(v2) : 0.0;
      for (int j = 0; j < Nphi; ++j) {
        const double phi = (j + 0.5) * g.dphi;
        Sig[i * Nphi + j] = lka_sigma0(R) * (1.0 + m->amp * std::cos(2.0 * phi));
        vR[i * Nphi + j] = 0.0;
        vp[i * Nphi + j] = vphi;
      }
    }
    if (Qout) {
      std::vector<double> s0(Nu), ds0(Nu), Q, v2;
      for (int i = 0; i < Nu; ++i) {
        s0[i] = lka_sigma0(g.R[i]);
        ds0[i] = s0[i] * (-2.0 * (g.R[i] - kR0) / kW2);
      }
      q_profile(m, g, s0, ds0, gRdisk, m->K, Q, v2);
      for (int i = 0; i < Nu; ++i) Qout[i] = Q[i];
    }
    return 0;
  }
  // power-law model
  const double A = pl_norm(m, g);
  std::vector<double> s0(Nu), ds0(Nu), Q, v2;
  for (int i = 0; i < Nu; ++i) {
    s0[i] = A * pl_shape(m, g.R[i]);
    ds0[i] = A * pl_dshape(m, g.R[i]);
  }
  // min Q rises ~ sqrt(K) from 0, then falls once pressure makes kappa^2 < 0
  // somewhere: bracket the first crossing of Qmin on a fine geometric scan
  double lo = 0.0, hi = -1.0, bestK = 0.0, bestQ = -1.0;
  for (double K = 1e-8; K < 1e3; K *= 1.05) {
    q_profile(m, g, s0, ds0, gRdisk, K, Q, v2);
    const double mq = min_q(m, g, Q);
    if (mq > bestQ) { bestQ = mq; bestK = K; }
    if (mq >= m->Qmin) { hi = K; break; }
    lo = K;
  }
  if (hi > 0) {
    for (int b = 0; b < 100; ++b) {
      const double mid = 0.5 * (lo + hi);
      q_profile(m, g, s0, ds0, gRdisk, mid, Q, v2);
      if (min_q(m, g, Q) >= m->Qmin) hi = mid; else lo = mid;
    }
  } else {
    hi = bestK;
  }
  m->K = hi;
  q_profile(m, g, s0, ds0, gRdisk, m->K, Q, v2);
  m->Qmin_got = min_q(m, g, Q);
  if (Qout)
    for (int i = 0; i < Nu; ++i) Qout[i] = Q[i];
  unsigned long long rng = m->seed;
  const int nm = m->mmax - m->mmin + 1;
  const double an = (nm > 0) ? m->noise_rms / std::sqrt((double)nm) : 0.0;
  std::vector<double> a(nm > 0 ? nm : 1), b(nm > 0 ? nm : 1);
  for (int i = 0; i < Nu; ++i) {
    for (int q = 0; q < nm; ++q) {
      a[q] = dfk_normal(&rng);
      b[q] = dfk_normal(&rng);
    }
    const double vphi = (v2[i] > 0) ? std::sqrt(v2[i]) : 0.0;
    for (int j = 0; j < Nphi; ++j) {
      const double phi = (j + 0.5) * g.dphi;
      double d = 0.0;
      for (int q = 0; q < nm; ++q) {
        const int mm = m->mmin + q;
        d += a[q] * std::cos(mm * phi) + b[q] * std::sin(mm * phi);
      }
      Sig[i * Nphi + j] = s0[i] * (1.0 + an * d);
      vR[i * Nphi + j] = 0.0;
      vp[i * Nphi + j] = vphi;
    }
  }
  return 0;
}

double dfk_model_mass(const dfk_model *m, int Nu, int Nphi, const double *Sig) {
  Grid g = make_grid(m, Nu, Nphi);
  double mD = 0.0;   // diskmain.cpp:65-70
  for (int i = 0; i < Nu; ++i) {
    const double dA = g.R[i] * (g.Rf[i + 1] - g.Rf[i]) * g.dphi;
    for (int j = 0; j < Nphi; ++j) mD += Sig[i * Nphi + j] * dA;
  }
  return mD;
}

int dfk_extract_growth(const double *t, const double *c0, const double *c2, const double *ph2,
                       int n, double lo_factor, double hi_frac, double grow_factor,
                       dfk_growth *out) {
  return dfk_extract_growth_ex(t, c0, c2, ph2, n, lo_factor, HUGE_VAL, hi_frac, grow_factor, 0.25,
                               HUGE_VAL, out);
}

int dfk_extract_growth_ex(const double *t, const double *c0, const double *c2, const double *ph2,
                          int n, double lo_factor, double hi_factor, double hi_frac,
                          double grow_factor, double sat_frac, double peak_window,
                          dfk_growth *out) {
  if (!t || !c0 || !c2 || !ph2 || !out || n < 4 || !(sat_frac > 0.0 && sat_frac <= 1.0)) return 1;
  std::vector<double> a(n), ph(n);
  double amax = 0.0;
  for (int k = 0; k < n; ++k) {
    a[k] = c2[k] / c0[k];
    if (a[k] > amax) amax = a[k];
  }
  // unwrap the phase
  ph[0] = ph2[0];
  for (int k = 1; k < n; ++k) {
    double d = ph2[k] - ph2[k - 1];
    d -= 2.0 * M_PI * std::floor((d + M_PI) / (2.0 * M_PI));
    ph[k] = ph[k - 1] + d;
  }
  int lo = 0, hi = n - 1, grew = 0;
  if (amax >= grow_factor * a[0]) {
    int l = -1, h = -1;
    for (int k = 0; k < n; ++k)
      if (l < 0 && a[k] >= lo_factor * a[0]) l = k;
    const double hi_level = std::min(hi_frac * amax, hi_factor * a[0]);
    for (int k = 0; k < n; ++k)
      if (h < 0 && a[k] >= hi_level) h = k;
    if (l >= 0 && h >= 0 && h - l + 1 >= 4) {
      lo = l;
      hi = h;
      grew = 1;
    }
  }
  // least squares over [lo, hi]
  const int m = hi - lo + 1;
  double st = 0, sa = 0, sp = 0;
  for (int k = lo; k <= hi; ++k) { st += t[k]; sa += std::log(a[k]); sp += ph[k]; }
  const double tm = st / m, am = sa / m, pm = sp / m;
  double stt = 0, sta = 0, stp = 0;
  for (int k = lo; k <= hi; ++k) {
    const double dt = t[k] - tm;
    stt += dt * dt;
    sta += dt * (std::log(a[k]) - am);
    stp += dt * (ph[k] - pm);
  }
  out->gamma = (stt > 0) ? sta / stt : 0.0;
  out->omega_p = (stt > 0) ? -0.5 * stp / stt : 0.0;
  const int q0 = n - std::max(1, (int)(sat_frac * n));
  double ss = 0;
  for (int k = q0; k < n; ++k) ss += a[k];
  out->sat = ss / (n - q0);
  out->t_lo = t[lo];
  out->t_hi = t[hi];
  out->n_fit = m;
  out->grew = grew;
  double pk = 0.0;
  for (int k = 0; k < n; ++k)
    if (!grew || (t[k] >= t[lo] && t[k] <= t[lo] + peak_window)) pk = std::max(pk, a[k]);
  out->sat_peak = pk;
  return 0;
}

}  // extern "C"
