/* diskfft on MI355X -- disk models and initial conditions (host code, no GPU).
 *
 * Two models:
 *   DFK_MODEL_LKA98     the reference driver's Gaussian ring (diskmain.cpp:28-63):
 *                       Rin 0.25, Rout 1, Sigma0 = 0.372 exp(-(R-0.45)^2/0.05),
 *                       K 0.25, gamma 2, GM* 0.6, seed Sigma0 (1 + amp cos 2phi).
 *   DFK_MODEL_POWERLAW  a power-law disk Sigma0 ~ R^slope with tanh tapers,
 *                       normalised to a disk/star mass ratio, K set by a
 *                       bracket + bisection so that min Toomre Q over the disk
 *                       body [q_in, q_out] equals Qmin (the smallest such K;
 *                       if no K reaches Qmin, the K maximising min Q, with
 *                       Qmin_got < Qmin), and seeded 1e-3 rms white noise in
 *                       azimuthal modes m = 1..8 (SplitMix64 + Box-Muller).
 *                       No reference counterpart (the reference only runs LKA98).
 *
 * The equilibrium needs the azimuthally averaged self-gravity g_R of Sigma0
 * (diskmain.cpp:36-48), which the caller computes with the solver of its
 * choice (dfk_potential/dfk_forces on the GPU) and passes in:
 *   1. dfk_model_sigma0     -> Sigma0(R_i)
 *   2. caller: gRdisk_i = phi-average of g_R of the axisymmetric Sigma0
 *   3. dfk_model_state      -> K (and Q profile), Sigma, vR, vphi
 */
#ifndef DISKFFT_MODEL_H
#define DISKFFT_MODEL_H

#ifdef __cplusplus
extern "C" {
#endif

#define DFK_MODEL_LKA98 0
#define DFK_MODEL_POWERLAW 1

typedef struct {
  int kind;
  double Rin, Rout, gmstar, G, gamma;
  double K;                  /* LKA98: fixed; POWERLAW: output of the Q solve */
  double amp;                /* LKA98 cos(2 phi) seed amplitude */
  /* POWERLAW */
  double mass_ratio;         /* M_disk / M_star */
  double Qmin;               /* target min Toomre Q over [q_in, q_out] */
  double slope;              /* Sigma0 ~ R^slope */
  double taper_in, taper_out, taper_w;
  double q_in, q_out;        /* radial window of the Q minimum (the disk body) */
  double Qmin_got;           /* output: min Q actually reached (< Qmin if unreachable) */
  double noise_rms;          /* delta Sigma / Sigma rms */
  int mmin, mmax;            /* noise modes */
  unsigned long long seed;
} dfk_model;

/* Reference LKA98 defaults (diskmain.cpp:28-29) and the power-law defaults
 * (M_d/M* 1, Qmin 1.3 over [0.2, 0.8], slope -1.5, tapers 0.15/0.9 width 0.02,
 * noise 1e-3 in m = 1..8, seed 42, r in [0.1, 1], GM* = 1). */
void dfk_model_defaults(int kind, dfk_model *m);

/* Sigma0 on the cell centres R (length Nu); log-grid geometry as hydro.hpp. */
int dfk_model_sigma0(const dfk_model *m, int Nu, int Nphi, double *sig0);

/* Equilibrium state.  gRdisk (length Nu, may be NULL = no self-gravity).  For
 * POWERLAW, solves K and writes it to m->K; Q (length Nu, may be NULL)
 * receives the Toomre Q profile.  Sig/vR/vp are Nu*Nphi row-major. */
int dfk_model_state(dfk_model *m, int Nu, int Nphi, const double *gRdisk, double *Sig,
                    double *vR, double *vp, double *Q);

/* Disk mass sum Sigma R dR dphi (diskmain.cpp:65-70). */
double dfk_model_mass(const dfk_model *m, int Nu, int Nphi, const double *Sig);

/* Growth-rate / pattern-speed / saturation extractor for the m = 2 mode from
 * the driver's output series (t, |C0|, |C2|, arg C2) of diskmain.cpp:93-98.
 * The reference only writes the series; this definition is the build's own:
 *   a(t) = |C2|/|C0|;  growth window = from the first sample with
 *   a >= lo_factor * a(0) to the first sample with a >= hi_frac * max a;
 *   gamma  = least-squares slope of ln a over the window;
 *   omega_p = -1/2 least-squares slope of unwrap(arg C2) over the window
 *            (C_m ~ exp(-i m Omega_p t) for a prograde pattern);
 *   sat    = mean of a over the last quarter of the samples.
 * If the mode never grows past grow_factor * a(0) (or the window has fewer
 * than 4 samples) the window is the whole series and grew = 0. */
typedef struct {
  double gamma, omega_p, sat;
  double t_lo, t_hi;         /* fit window */
  int n_fit;                 /* samples in the window */
  int grew;
  double sat_peak;           /* dfk_extract_growth_ex only: max a over [t_lo, t_lo + peak_window] */
} dfk_growth;
int dfk_extract_growth(const double *t, const double *c0, const double *c2, const double *ph2,
                       int n, double lo_factor, double hi_frac, double grow_factor,
                       dfk_growth *out);

/* The same with more knobs: the window also ends no later than the first
 * sample with a >= hi_factor * a(0) (so a later, larger peak cannot stretch
 * the window of the first linear growth across a plateau); sat is the mean of
 * a over the last sat_frac of the samples (0 < sat_frac <= 1); sat_peak is
 * the largest a in [t_lo, t_lo + peak_window], the amplitude at which that
 * first growth saturates (whole series if the mode never grew).
 * dfk_extract_growth is hi_factor = +inf, sat_frac = 0.25. */
int dfk_extract_growth_ex(const double *t, const double *c0, const double *c2, const double *ph2,
                          int n, double lo_factor, double hi_factor, double hi_frac,
                          double grow_factor, double sat_frac, double peak_window,
                          dfk_growth *out);

/* SplitMix64 stream: next 64-bit value (state updated), and a standard normal
 * by Box-Muller (one value per call; the pair's sine half is discarded). */
unsigned long long dfk_splitmix64(unsigned long long *state);
double dfk_normal(unsigned long long *state);

#ifdef __cplusplus
}
#endif
#endif
