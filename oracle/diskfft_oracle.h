/* diskfft oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C (+OpenMP) restatement of the reference diskfft time-step loop,
 * used as the parity checker for the HIP product path and as the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates
 * (paths relative to the oklo/heavy-disks checkout).
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_hydro_*            bit-exact against the reference hydro.hpp compiled
 *                          from its own source (oracle/_ref, ref_harness.cpp),
 *                          plus closed forms (mass conservation, uniform-state
 *                          invariance, source kick, Courant-1 shift).
 *   orc_fft_*              direct O(n^2) DFT and numpy.fft on random input.
 *   orc_kal_potential      orc_kal_potential_direct (restates kalnajs.hpp:83-99)
 *                          at <= 1e-10 relative, the test_kalnajs.cpp:10-42
 *                          known-answer check; K-hat realness.
 *   orc_kal_forces         finite-difference exactness on linear / cosine Phi.
 *   orc_cm                 closed form |C2|/|C0| = amp/2 on Sigma0(1+amp cos 2phi).
 *   orc_run                loop semantics (output schedule, last-step clip)
 *                          against hand-traced short runs.
 * FFTW itself (the reference's third-party FFT, version unpinned) is absent
 * from this image; its role is taken by the FFT below, pinned as stated.
 */
#ifndef DISKFFT_ORACLE_H
#define DISKFFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- DiskHydro (hydro.hpp:20-46) ---- */
typedef struct {
  int Nu, Nphi;
  double umin, umax, du, dphi, Kpoly, gamma_gas, gmstar;
  double *R, *Rf, *Vol;                 /* Nu, Nu+1, Nu */
  double *Sig, *vR, *vp;                /* Nu*Nphi      */
  double *wA, *wB, *mR, *mp;            /* Nu*Nphi      */
  double *Fm, *FmR, *FmP;               /* (Nu+1)*Nphi  */
} orc_hydro;

orc_hydro *orc_hydro_new(int Nu, int Nphi, double Rin, double Rout,
                         double K, double gamma, double gmstar);
void orc_hydro_free(orc_hydro *h);
double orc_hydro_timestep(const orc_hydro *h, double cfl);
void orc_hydro_source(orc_hydro *h, double dt, const double *gR, const double *gp);
void orc_hydro_transportR(orc_hydro *h, double dt);
void orc_hydro_transportPhi(orc_hydro *h, double dt);
void orc_hydro_floor(orc_hydro *h);
void orc_hydro_step(orc_hydro *h, double dt, const double *gR, const double *gp);
/* accessors for ctypes */
double *orc_hydro_ptr(orc_hydro *h, int which); /* 0 Sig 1 vR 2 vp 3 R 4 Rf 5 Vol */

/* ---- mixed-radix complex FFT (replaces FFTW, kalnajs.hpp:15,39-40) ---- */
typedef struct {
  int n, nf;
  int fac[40];
  double *twr, *twi;   /* exp(-2 pi i t / n), t in [0,n) */
} orc_fftplan;
int orc_fft_plan_init(orc_fftplan *p, int n);
void orc_fft_plan_free(orc_fftplan *p);
/* in-place (re,im) transform; sign -1 forward, +1 inverse, unnormalised.
 * wr/wi: scratch of length n. */
void orc_fft_exec(const orc_fftplan *p, double *re, double *im,
                  double *wr, double *wi, int sign);
/* convenience for tests: allocate scratch internally */
int orc_fft(int n, double *re, double *im, int sign);

/* ---- KalnajsSolver (kalnajs.hpp:19-117) ---- */
typedef struct {
  int Nu, Nphi, M, Nc;
  double umin, umax, du, dphi, eps, G;
  double *R, *sqrtR, *R32;              /* Nu */
  double *Khr, *Khi;                    /* M*Nc, complex K-hat */
  double *cr, *ci;                      /* M*Nc work spectrum */
  orc_fftplan prow, pcol;
} orc_kalnajs;

orc_kalnajs *orc_kal_new(int Nu, int Nphi, double Rin, double Rout,
                         double eps, double G);
void orc_kal_free(orc_kalnajs *k);
void orc_kal_potential(orc_kalnajs *k, const double *Sig, double *Phi);
void orc_kal_forces(const orc_kalnajs *k, const double *Phi, double *gR, double *gp);
void orc_kal_potential_direct(const orc_kalnajs *k, const double *Sig, double *Phi);
/* 2-D r2c / c2r of an M x Nphi real array, row-major, FFTW layout
 * (M x Nc complex, unnormalised) -- exposed for the numpy.fft pin. */
void orc_kal_r2c2d(orc_kalnajs *k, const double *in, double *outr, double *outi);
void orc_kal_c2r2d(orc_kalnajs *k, const double *inr, const double *ini, double *out);
const double *orc_kal_khat(const orc_kalnajs *k, int imag);

/* ---- driver pieces (diskmain.cpp) ---- */
/* C_m diagnostic, diskmain.cpp:72-84 (mag is |C_m|/mD) */
void orc_cm(const orc_hydro *h, int m, double mD, double *mag, double *phase);
double orc_disk_mass(const orc_hydro *h);                     /* diskmain.cpp:65-70 */
/* LKA98 ring IC + equilibrium, diskmain.cpp:28-63 (grav may be NULL if !selfgrav) */
void orc_ic_lka98(orc_hydro *h, orc_kalnajs *grav, int selfgrav, double amp);

typedef struct { double t, c0, c2, ph2; } orc_sample;
/* The time loop of diskmain.cpp:86-113.  Writes up to cap samples (one per
 * output time) and the final (t, |C2|, steps).  max_steps<=0 means unbounded.
 * Returns number of samples. */
int orc_run(orc_hydro *h, orc_kalnajs *grav, int selfgrav, double tlim,
            double dtout, double cfl, long max_steps, double mD,
            orc_sample *samples, int cap, double *t_final, double *c2_final,
            long *steps_out);

/* The same loop continued from (*t_io, *tnext_io, *step_io), updated in place. */
int orc_run_from(orc_hydro *h, orc_kalnajs *grav, int selfgrav, double tlim,
                 double dtout, double cfl, long max_steps, double mD,
                 orc_sample *samples, int cap, double *t_io, double *tnext_io,
                 long *step_io, double *c2_final);

void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
