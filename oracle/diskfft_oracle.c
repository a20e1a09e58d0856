/* diskfft oracle -- TEST INFRASTRUCTURE ONLY (see diskfft_oracle.h).
 *
 * Plain C + OpenMP restatement of diskfft/src/{hydro.hpp,kalnajs.hpp,
 * diskmain.cpp}.  Arithmetic follows the reference expression by expression
 * (same operand order, std::max semantics kept as ternaries so NaN behaves
 * the same), so that a build with the same compiler flags reproduces the
 * reference hydro bit for bit.  The FFT is this file's own mixed-radix
 * Stockham transform (FFTW is not available in this image).
 */
#define _GNU_SOURCE
#include "diskfft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* std::max(a,b) == (a < b) ? b : a   -- NaN in a propagates, as in the reference */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ===================================================================== */
/* DiskHydro                                                             */
/* ===================================================================== */

/* hydro.hpp:16-18 */
static inline double vanleer(double a, double b) {
  return (a * b > 0.0) ? 2.0 * a * b / (a + b) : 0.0;
}

/* hydro.hpp:28-43 */
orc_hydro *orc_hydro_new(int Nu, int Nphi, double Rin, double Rout,
                         double K, double gamma, double gmstar) {
  orc_hydro *h = (orc_hydro *)calloc(1, sizeof(orc_hydro));
  size_t N = (size_t)Nu * Nphi, NF = (size_t)(Nu + 1) * Nphi;
  h->Nu = Nu; h->Nphi = Nphi; h->Kpoly = K; h->gamma_gas = gamma; h->gmstar = gmstar;
  h->umin = log(Rin); h->umax = log(Rout);
  h->du = (h->umax - h->umin) / Nu; h->dphi = 2.0 * M_PI / Nphi;
  h->R = (double *)calloc(Nu, sizeof(double));
  h->Rf = (double *)calloc(Nu + 1, sizeof(double));
  h->Vol = (double *)calloc(Nu, sizeof(double));
  for (int i = 0; i <= Nu; ++i) h->Rf[i] = exp(h->umin + i * h->du);
  for (int i = 0; i < Nu; ++i) {
    h->R[i] = exp(h->umin + (i + 0.5) * h->du);
    h->Vol[i] = h->R[i] * (h->Rf[i + 1] - h->Rf[i]) * h->dphi;
  }
  h->Sig = (double *)calloc(N, sizeof(double));
  h->vR = (double *)calloc(N, sizeof(double));
  h->vp = (double *)calloc(N, sizeof(double));
  h->wA = (double *)calloc(N, sizeof(double));
  h->wB = (double *)calloc(N, sizeof(double));
  h->mR = (double *)calloc(N, sizeof(double));
  h->mp = (double *)calloc(N, sizeof(double));
  h->Fm = (double *)calloc(NF, sizeof(double));
  h->FmR = (double *)calloc(NF, sizeof(double));
  h->FmP = (double *)calloc(NF, sizeof(double));
  return h;
}

void orc_hydro_free(orc_hydro *h) {
  if (!h) return;
  free(h->R); free(h->Rf); free(h->Vol); free(h->Sig); free(h->vR); free(h->vp);
  free(h->wA); free(h->wB); free(h->mR); free(h->mp); free(h->Fm); free(h->FmR);
  free(h->FmP); free(h);
}

double *orc_hydro_ptr(orc_hydro *h, int which) {
  switch (which) {
    case 0: return h->Sig; case 1: return h->vR; case 2: return h->vp;
    case 3: return h->R; case 4: return h->Rf; case 5: return h->Vol;
    default: return NULL;
  }
}

/* hydro.hpp:45-46 */
static inline double cs2(const orc_hydro *h, double s) {
  return h->gamma_gas * h->Kpoly * pow(s, h->gamma_gas - 1.0);
}
static inline double pres(const orc_hydro *h, double s) {
  return h->Kpoly * pow(s, h->gamma_gas);
}

/* hydro.hpp:48-61 */
double orc_hydro_timestep(const orc_hydro *h, double cfl) {
  const int Nu = h->Nu, Nphi = h->Nphi;
  double dtinv = 1e-30;
#pragma omp parallel for reduction(max : dtinv) schedule(static)
  for (int i = 0; i < Nu; ++i) {
    double dR = h->Rf[i + 1] - h->Rf[i];
    for (int j = 0; j < Nphi; ++j) {
      int c = i * Nphi + j;
      double cs = sqrt(cs2(h, smax(h->Sig[c], 1e-12)));
      double sR = (fabs(h->vR[c]) + cs) / dR;
      double sp = (fabs(h->vp[c]) + cs) / (h->R[i] * h->dphi);
      if (sR + sp > dtinv) dtinv = sR + sp;
    }
  }
  return cfl / dtinv;
}

/* hydro.hpp:73-96 (writes wA/wB, then swaps with vR/vp as at :95) */
void orc_hydro_source(orc_hydro *h, double dt, const double *gR_sg, const double *gp_sg) {
  const int Nu = h->Nu, Nphi = h->Nphi;
  double *nvR = h->wA, *nvp = h->wB;
  const double *Sig = h->Sig, *vR = h->vR, *vp = h->vp, *R = h->R;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nu; ++i) {
    int im = imax(i - 1, 0), ip = imin(i + 1, Nu - 1);
    double dRc = R[ip] - R[im];
    for (int j = 0; j < Nphi; ++j) {
      int c = i * Nphi + j;
      int jm = i * Nphi + (j - 1 + Nphi) % Nphi, jp = i * Nphi + (j + 1) % Nphi;
      double s = Sig[c];
      double dPdR = (pres(h, Sig[ip * Nphi + j]) - pres(h, Sig[im * Nphi + j])) / dRc;
      double dPdphi = (pres(h, Sig[jp]) - pres(h, Sig[jm])) / (2.0 * h->dphi);
      double gR = -h->gmstar / (R[i] * R[i]);
      double gp = 0.0;
      if (gR_sg) gR += gR_sg[c];
      if (gp_sg) gp += gp_sg[c];
      double aR = -dPdR / s + gR + vp[c] * vp[c] / R[i];
      double ap = -(1.0 / (s * R[i])) * dPdphi + gp - vR[c] * vp[c] / R[i];
      nvR[c] = vR[c] + dt * aR;
      nvp[c] = vp[c] + dt * ap;
    }
  }
  double *t;
  t = h->vR; h->vR = h->wA; h->wA = t;
  t = h->vp; h->vp = h->wB; h->wB = t;
}

/* hydro.hpp:99-140 */
void orc_hydro_transportR(orc_hydro *h, double dt) {
  const int Nu = h->Nu, Nphi = h->Nphi;
  double *Sig = h->Sig, *vR = h->vR, *vp = h->vp, *mR = h->mR, *mp = h->mp;
  double *Fm = h->Fm, *FmR = h->FmR, *FmP = h->FmP;
  const double *Rf = h->Rf, *Vol = h->Vol;
  const double dphi = h->dphi;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < Nu * Nphi; ++k) { mR[k] = Sig[k] * vR[k]; mp[k] = Sig[k] * vp[k]; }
  for (int j = 0; j < Nphi; ++j) {
    Fm[j] = FmR[j] = FmP[j] = 0.0;
    Fm[Nu * Nphi + j] = FmR[Nu * Nphi + j] = FmP[Nu * Nphi + j] = 0.0;
  }
#define SLOPER(a, i, j) \
  (((i) <= 0 || (i) >= Nu - 1) ? 0.0 \
   : vanleer((a)[((i) + 1) * Nphi + (j)] - (a)[(i) * Nphi + (j)], \
             (a)[(i) * Nphi + (j)] - (a)[((i) - 1) * Nphi + (j)]))
#pragma omp parallel for schedule(static)
  for (int f = 1; f < Nu; ++f) {
    double Af = Rf[f] * dphi;
    double dRl = Rf[f] - Rf[f - 1], dRr = Rf[f + 1] - Rf[f];
    for (int j = 0; j < Nphi; ++j) {
      double vf = 0.5 * (vR[(f - 1) * Nphi + j] + vR[f * Nphi + j]);
      double sf, vRf, vpf;
      if (vf > 0) {
        sf  = Sig[(f - 1) * Nphi + j] + 0.5 * SLOPER(Sig, f - 1, j) * (1.0 - vf * dt / dRl);
        vRf = vR[(f - 1) * Nphi + j]  + 0.5 * SLOPER(vR, f - 1, j) * (1.0 - vf * dt / dRl);
        vpf = vp[(f - 1) * Nphi + j]  + 0.5 * SLOPER(vp, f - 1, j) * (1.0 - vf * dt / dRl);
      } else {
        sf  = Sig[f * Nphi + j] - 0.5 * SLOPER(Sig, f, j) * (1.0 + vf * dt / dRr);
        vRf = vR[f * Nphi + j]  - 0.5 * SLOPER(vR, f, j) * (1.0 + vf * dt / dRr);
        vpf = vp[f * Nphi + j]  - 0.5 * SLOPER(vp, f, j) * (1.0 + vf * dt / dRr);
      }
      double fm = vf * sf * Af;
      Fm[f * Nphi + j] = fm; FmR[f * Nphi + j] = fm * vRf; FmP[f * Nphi + j] = fm * vpf;
    }
  }
#undef SLOPER
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nu; ++i)
    for (int j = 0; j < Nphi; ++j) {
      int c = i * Nphi + j;
      double ns = Sig[c] + dt * (Fm[i * Nphi + j] - Fm[(i + 1) * Nphi + j]) / Vol[i];
      double nmR = mR[c] + dt * (FmR[i * Nphi + j] - FmR[(i + 1) * Nphi + j]) / Vol[i];
      double nmP = mp[c] + dt * (FmP[i * Nphi + j] - FmP[(i + 1) * Nphi + j]) / Vol[i];
      Sig[c] = ns; vR[c] = nmR / smax(ns, 1e-10); vp[c] = nmP / smax(ns, 1e-10);
    }
}

/* hydro.hpp:143-179 */
void orc_hydro_transportPhi(orc_hydro *h, double dt) {
  const int Nu = h->Nu, Nphi = h->Nphi;
  double *Sig = h->Sig, *vR = h->vR, *vp = h->vp, *mR = h->mR, *mp = h->mp;
  double *Fm = h->Fm, *FmR = h->FmR, *FmP = h->FmP;
  const double *R = h->R, *Rf = h->Rf, *Vol = h->Vol;
  const double dphi = h->dphi;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < Nu * Nphi; ++k) { mR[k] = Sig[k] * vR[k]; mp[k] = Sig[k] * vp[k]; }
#define SLOPEP(a, i, j) \
  vanleer((a)[(i) * Nphi + ((j) + 1) % Nphi] - (a)[(i) * Nphi + (j)], \
          (a)[(i) * Nphi + (j)] - (a)[(i) * Nphi + ((j) - 1 + Nphi) % Nphi])
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nu; ++i) {
    double dl = R[i] * dphi, Aface = Rf[i + 1] - Rf[i];
    for (int j = 0; j < Nphi; ++j) {
      int jm = (j - 1 + Nphi) % Nphi;
      double vf = 0.5 * (vp[i * Nphi + jm] + vp[i * Nphi + j]);
      double sf, vRf, vpf;
      if (vf > 0) {
        sf  = Sig[i * Nphi + jm] + 0.5 * SLOPEP(Sig, i, jm) * (1.0 - vf * dt / dl);
        vRf = vR[i * Nphi + jm]  + 0.5 * SLOPEP(vR, i, jm) * (1.0 - vf * dt / dl);
        vpf = vp[i * Nphi + jm]  + 0.5 * SLOPEP(vp, i, jm) * (1.0 - vf * dt / dl);
      } else {
        sf  = Sig[i * Nphi + j] - 0.5 * SLOPEP(Sig, i, j) * (1.0 + vf * dt / dl);
        vRf = vR[i * Nphi + j]  - 0.5 * SLOPEP(vR, i, j) * (1.0 + vf * dt / dl);
        vpf = vp[i * Nphi + j]  - 0.5 * SLOPEP(vp, i, j) * (1.0 + vf * dt / dl);
      }
      double fm = vf * sf * Aface;
      Fm[i * Nphi + j] = fm; FmR[i * Nphi + j] = fm * vRf; FmP[i * Nphi + j] = fm * vpf;
    }
  }
#undef SLOPEP
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nu; ++i)
    for (int j = 0; j < Nphi; ++j) {
      int c = i * Nphi + j; int jp = i * Nphi + (j + 1) % Nphi;
      double ns = Sig[c] + dt * (Fm[c] - Fm[jp]) / Vol[i];
      double nmR = mR[c] + dt * (FmR[c] - FmR[jp]) / Vol[i];
      double nmP = mp[c] + dt * (FmP[c] - FmP[jp]) / Vol[i];
      Sig[c] = ns; vR[c] = nmR / smax(ns, 1e-10); vp[c] = nmP / smax(ns, 1e-10);
    }
}

/* hydro.hpp:68-69 */
void orc_hydro_floor(orc_hydro *h) {
  const int N = h->Nu * h->Nphi;
  double *Sig = h->Sig;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < N; ++k) Sig[k] = smax(Sig[k], 1e-10);
}

/* hydro.hpp:63-70 */
void orc_hydro_step(orc_hydro *h, double dt, const double *gR, const double *gp) {
  orc_hydro_source(h, dt, gR, gp);
  orc_hydro_transportR(h, dt);
  orc_hydro_transportPhi(h, dt);
  orc_hydro_floor(h);
}

/* ===================================================================== */
/* Mixed-radix Stockham FFT (Cooley-Tukey, self-sorting; radices 4,2,3,5, */
/* any other prime by a direct small DFT).                                */
/* ===================================================================== */

int orc_fft_plan_init(orc_fftplan *p, int n) {
  memset(p, 0, sizeof(*p));
  if (n < 1) return -1;
  p->n = n;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int q = 3; m > 1; q += 2) {
    while (m % q == 0) {
      if (p->nf >= 40) return -1;
      p->fac[p->nf++] = q; m /= q;
    }
  }
  p->twr = (double *)malloc(sizeof(double) * n);
  p->twi = (double *)malloc(sizeof(double) * n);
  for (int t = 0; t < n; ++t) {
    double a = 2.0 * M_PI * (double)t / (double)n;
    p->twr[t] = cos(a);
    p->twi[t] = -sin(a);
  }
  return 0;
}

void orc_fft_plan_free(orc_fftplan *p) {
  free(p->twr); free(p->twi);
  p->twr = p->twi = NULL;
}

void orc_fft_exec(const orc_fftplan *p, double *re, double *im,
                  double *wr, double *wi, int sign) {
  const int n = p->n;
  if (n <= 1) return;
  double *ar = re, *ai = im, *br = wr, *bi = wi;
  const double sg = (sign < 0) ? 1.0 : -1.0;   /* multiplies twi */
  int Ns = 1;
  double vrs[64], vis[64];
  double *vr = vrs, *vi = vis;
  int maxr = 0;
  for (int s = 0; s < p->nf; ++s) if (p->fac[s] > maxr) maxr = p->fac[s];
  if (maxr > 64) {
    vr = (double *)malloc(sizeof(double) * maxr);
    vi = (double *)malloc(sizeof(double) * maxr);
  }
  for (int s = 0; s < p->nf; ++s) {
    const int r = p->fac[s], nb = n / r, tstep = n / (Ns * r), rstep = n / r;
    for (int j = 0; j < nb; ++j) {
      const int k = j % Ns;
      for (int m = 0; m < r; ++m) {
        double xr = ar[j + m * nb], xi = ai[j + m * nb];
        int t = k * m * tstep;
        double c = p->twr[t], sn = sg * p->twi[t];
        vr[m] = xr * c - xi * sn;
        vi[m] = xr * sn + xi * c;
      }
      const int base = (j / Ns) * Ns * r + k;
      if (r == 2) {
        br[base] = vr[0] + vr[1];       bi[base] = vi[0] + vi[1];
        br[base + Ns] = vr[0] - vr[1];  bi[base + Ns] = vi[0] - vi[1];
      } else if (r == 4) {
        /* forward (sign -1): X1 = v0 - i v1 - v2 + i v3 */
        double ar0 = vr[0] + vr[2], ai0 = vi[0] + vi[2];
        double ar1 = vr[0] - vr[2], ai1 = vi[0] - vi[2];
        double br0 = vr[1] + vr[3], bi0 = vi[1] + vi[3];
        double br1 = vr[1] - vr[3], bi1 = vi[1] - vi[3];
        /* -i*sgn*(b1): for forward multiply (br1,bi1) by -i -> (bi1,-br1) */
        double tr = sg * bi1, ti = -sg * br1;
        br[base] = ar0 + br0;              bi[base] = ai0 + bi0;
        br[base + Ns] = ar1 + tr;          bi[base + Ns] = ai1 + ti;
        br[base + 2 * Ns] = ar0 - br0;     bi[base + 2 * Ns] = ai0 - bi0;
        br[base + 3 * Ns] = ar1 - tr;      bi[base + 3 * Ns] = ai1 - ti;
      } else {
        for (int q = 0; q < r; ++q) {
          double sr = 0.0, si = 0.0;
          for (int m = 0; m < r; ++m) {
            int t = ((m * q) % r) * rstep;
            double c = p->twr[t], sn = sg * p->twi[t];
            sr += vr[m] * c - vi[m] * sn;
            si += vr[m] * sn + vi[m] * c;
          }
          br[base + q * Ns] = sr; bi[base + q * Ns] = si;
        }
      }
    }
    double *t;
    t = ar; ar = br; br = t;
    t = ai; ai = bi; bi = t;
    Ns *= r;
  }
  if (ar != re) {
    memcpy(re, ar, sizeof(double) * n);
    memcpy(im, ai, sizeof(double) * n);
  }
  if (vr != vrs) { free(vr); free(vi); }
}

int orc_fft(int n, double *re, double *im, int sign) {
  orc_fftplan p;
  if (orc_fft_plan_init(&p, n)) return -1;
  double *wr = (double *)malloc(sizeof(double) * n), *wi = (double *)malloc(sizeof(double) * n);
  orc_fft_exec(&p, re, im, wr, wi, sign);
  free(wr); free(wi);
  orc_fft_plan_free(&p);
  return 0;
}

/* ===================================================================== */
/* KalnajsSolver                                                         */
/* ===================================================================== */

/* Row r2c of rows [0,nrows) of a (M x Nphi) real array into the (M x Nc)
 * spectrum; rows >= nrows are zero input and give zero spectrum.  Two real
 * rows are packed into one complex transform (z = a + i b). */
static void rows_r2c(orc_kalnajs *k, const double *in, int nrows, double *cr, double *ci) {
  const int Nphi = k->Nphi, Nc = k->Nc, M = k->M;
  const int npair = (nrows + 1) / 2;
#pragma omp parallel
  {
    double *zr = (double *)malloc(sizeof(double) * Nphi * 4);
    double *zi = zr + Nphi, *wr = zi + Nphi, *wi = wr + Nphi;
#pragma omp for schedule(static)
    for (int p = 0; p < npair; ++p) {
      int a = 2 * p, b = 2 * p + 1;
      for (int j = 0; j < Nphi; ++j) {
        zr[j] = in[a * Nphi + j];
        zi[j] = (b < nrows) ? in[b * Nphi + j] : 0.0;
      }
      orc_fft_exec(&k->prow, zr, zi, wr, wi, -1);
      for (int q = 0; q < Nc; ++q) {
        int qn = (Nphi - q) % Nphi;
        double Zr = zr[q], Zi = zi[q], Yr = zr[qn], Yi = -zi[qn];   /* Y = conj(Z[N-q]) */
        cr[a * Nc + q] = 0.5 * (Zr + Yr);
        ci[a * Nc + q] = 0.5 * (Zi + Yi);
        if (b < nrows) {
          /* (Z - Y) / (2i) */
          cr[b * Nc + q] = 0.5 * (Zi - Yi);
          ci[b * Nc + q] = -0.5 * (Zr - Yr);
        } else if (b < M) {
          cr[b * Nc + q] = 0.0; ci[b * Nc + q] = 0.0;
        }
      }
    }
    free(zr);
  }
  for (int r = 2 * npair; r < M; ++r)
    for (int q = 0; q < Nc; ++q) { cr[r * Nc + q] = 0.0; ci[r * Nc + q] = 0.0; }
}

/* Complex transform of every column (length M) of the (M x Nc) spectrum. */
static void cols_fft(orc_kalnajs *k, double *cr, double *ci, int sign) {
  const int M = k->M, Nc = k->Nc;
#pragma omp parallel
  {
    double *zr = (double *)malloc(sizeof(double) * M * 4);
    double *zi = zr + M, *wr = zi + M, *wi = wr + M;
#pragma omp for schedule(static)
    for (int q = 0; q < Nc; ++q) {
      for (int p = 0; p < M; ++p) { zr[p] = cr[p * Nc + q]; zi[p] = ci[p * Nc + q]; }
      orc_fft_exec(&k->pcol, zr, zi, wr, wi, sign);
      for (int p = 0; p < M; ++p) { cr[p * Nc + q] = zr[p]; ci[p * Nc + q] = zi[p]; }
    }
    free(zr);
  }
}

/* Row c2r (unnormalised, FFTW convention: the imaginary parts of the DC and,
 * for even Nphi, Nyquist bins are ignored) of rows [0,nrows) of the spectrum. */
static void rows_c2r(orc_kalnajs *k, const double *cr, const double *ci, int nrows, double *out) {
  const int Nphi = k->Nphi, Nc = k->Nc;
  const int npair = (nrows + 1) / 2;
#pragma omp parallel
  {
    double *zr = (double *)malloc(sizeof(double) * Nphi * 4);
    double *zi = zr + Nphi, *wr = zi + Nphi, *wi = wr + Nphi;
#pragma omp for schedule(static)
    for (int p = 0; p < npair; ++p) {
      int a = 2 * p, b = 2 * p + 1;
      int hasb = (b < nrows);
      for (int q = 0; q < Nphi; ++q) {
        double xar, xai, xbr = 0.0, xbi = 0.0;
        int qq = (q < Nc) ? q : Nphi - q;
        double sgn = (q < Nc) ? 1.0 : -1.0;
        xar = cr[a * Nc + qq]; xai = sgn * ci[a * Nc + qq];
        if (hasb) { xbr = cr[b * Nc + qq]; xbi = sgn * ci[b * Nc + qq]; }
        if (qq == 0 || (2 * qq == Nphi)) { xai = 0.0; xbi = 0.0; }
        zr[q] = xar - xbi;
        zi[q] = xai + xbr;
      }
      orc_fft_exec(&k->prow, zr, zi, wr, wi, +1);
      for (int j = 0; j < Nphi; ++j) {
        out[a * Nphi + j] = zr[j];
        if (hasb) out[b * Nphi + j] = zi[j];
      }
    }
    free(zr);
  }
}

void orc_kal_r2c2d(orc_kalnajs *k, const double *in, double *outr, double *outi) {
  rows_r2c(k, in, k->M, outr, outi);
  cols_fft(k, outr, outi, -1);
}

void orc_kal_c2r2d(orc_kalnajs *k, const double *inr, const double *ini, double *out) {
  const size_t S = (size_t)k->M * k->Nc;
  double *tr = (double *)malloc(sizeof(double) * S), *ti = (double *)malloc(sizeof(double) * S);
  memcpy(tr, inr, sizeof(double) * S); memcpy(ti, ini, sizeof(double) * S);
  cols_fft(k, tr, ti, +1);
  rows_c2r(k, tr, ti, k->M, out);
  free(tr); free(ti);
}

/* kalnajs.hpp:105-116 */
static void build_kernel(orc_kalnajs *k) {
  const int M = k->M, Nphi = k->Nphi;
  double *rbuf = (double *)malloc(sizeof(double) * (size_t)M * Nphi);
  for (int p = 0; p < M; ++p) {
    int ioff = (p <= M / 2) ? p : p - M;
    double Du = ioff * k->du;
    for (int q = 0; q < Nphi; ++q) {
      double Dphi = q * k->dphi;
      rbuf[p * Nphi + q] = 1.0 / sqrt(2.0 * cosh(Du) - 2.0 * cos(Dphi) + k->eps);
    }
  }
  orc_kal_r2c2d(k, rbuf, k->Khr, k->Khi);
  free(rbuf);
}

/* kalnajs.hpp:25-42 */
orc_kalnajs *orc_kal_new(int Nu, int Nphi, double Rin, double Rout, double eps, double G) {
  orc_kalnajs *k = (orc_kalnajs *)calloc(1, sizeof(orc_kalnajs));
  k->Nu = Nu; k->Nphi = Nphi; k->M = 2 * Nu; k->Nc = Nphi / 2 + 1; k->eps = eps; k->G = G;
  k->umin = log(Rin); k->umax = log(Rout);
  k->du = (k->umax - k->umin) / Nu; k->dphi = 2.0 * M_PI / Nphi;
  k->R = (double *)malloc(sizeof(double) * Nu);
  k->sqrtR = (double *)malloc(sizeof(double) * Nu);
  k->R32 = (double *)malloc(sizeof(double) * Nu);
  for (int i = 0; i < Nu; ++i) {
    k->R[i] = exp(k->umin + (i + 0.5) * k->du);
    k->sqrtR[i] = sqrt(k->R[i]); k->R32[i] = k->R[i] * k->sqrtR[i];
  }
  size_t S = (size_t)k->M * k->Nc;
  k->Khr = (double *)malloc(sizeof(double) * S); k->Khi = (double *)malloc(sizeof(double) * S);
  k->cr = (double *)malloc(sizeof(double) * S); k->ci = (double *)malloc(sizeof(double) * S);
  if (orc_fft_plan_init(&k->prow, Nphi) || orc_fft_plan_init(&k->pcol, k->M)) {
    orc_kal_free(k);
    return NULL;
  }
  build_kernel(k);
  return k;
}

void orc_kal_free(orc_kalnajs *k) {
  if (!k) return;
  free(k->R); free(k->sqrtR); free(k->R32); free(k->Khr); free(k->Khi); free(k->cr); free(k->ci);
  orc_fft_plan_free(&k->prow); orc_fft_plan_free(&k->pcol);
  free(k);
}

const double *orc_kal_khat(const orc_kalnajs *k, int imag) { return imag ? k->Khi : k->Khr; }

/* kalnajs.hpp:48-63 */
void orc_kal_potential(orc_kalnajs *k, const double *Sigma, double *Phi) {
  const int Nu = k->Nu, Nphi = k->Nphi, M = k->M, Nc = k->Nc;
  double *rbuf = (double *)malloc(sizeof(double) * (size_t)Nu * Nphi);
  for (int i = 0; i < Nu; ++i)
    for (int j = 0; j < Nphi; ++j) rbuf[i * Nphi + j] = k->R32[i] * Sigma[i * Nphi + j];
  /* rows >= Nu of the zero-padded array are zero: their row transforms are
   * zero and are not computed (identical result) */
  rows_r2c(k, rbuf, Nu, k->cr, k->ci);
  cols_fft(k, k->cr, k->ci, -1);
  for (int q = 0; q < M * Nc; ++q) {
    double a = k->cr[q], b = k->ci[q], c = k->Khr[q], d = k->Khi[q];
    k->cr[q] = a * c - b * d; k->ci[q] = a * d + b * c;
  }
  cols_fft(k, k->cr, k->ci, +1);
  rows_c2r(k, k->cr, k->ci, Nu, rbuf);   /* only rows < Nu are kept (:60) */
  double pref = -k->G * k->du * k->dphi / ((double)M * Nphi);
  for (int i = 0; i < Nu; ++i)
    for (int j = 0; j < Nphi; ++j)
      Phi[i * Nphi + j] = pref * rbuf[i * Nphi + j] / k->sqrtR[i];
  free(rbuf);
}

/* kalnajs.hpp:65-80 */
void orc_kal_forces(const orc_kalnajs *k, const double *Phi, double *gR, double *gphi) {
  const int Nu = k->Nu, Nphi = k->Nphi;
  const double du = k->du, dphi = k->dphi;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nu; ++i)
    for (int j = 0; j < Nphi; ++j) {
      int jm = (j - 1 + Nphi) % Nphi, jp = (j + 1) % Nphi;
      double dPhidphi = (Phi[i * Nphi + jp] - Phi[i * Nphi + jm]) / (2.0 * dphi);
      double dPhidu;
      if (i == 0) dPhidu = (Phi[(i + 1) * Nphi + j] - Phi[i * Nphi + j]) / du;
      else if (i == Nu - 1) dPhidu = (Phi[i * Nphi + j] - Phi[(i - 1) * Nphi + j]) / du;
      else dPhidu = (Phi[(i + 1) * Nphi + j] - Phi[(i - 1) * Nphi + j]) / (2.0 * du);
      gR[i * Nphi + j] = -dPhidu / k->R[i];
      gphi[i * Nphi + j] = -dPhidphi / k->R[i];
    }
}

/* kalnajs.hpp:83-99 */
void orc_kal_potential_direct(const orc_kalnajs *k, const double *Sigma, double *Phi) {
  const int Nu = k->Nu, Nphi = k->Nphi;
  double pref = -k->G * k->du * k->dphi;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nu; ++i)
    for (int j = 0; j < Nphi; ++j) {
      double V = 0.0;
      for (int ip = 0; ip < Nu; ++ip) {
        double Du = (i - ip) * k->du;
        for (int jp = 0; jp < Nphi; ++jp) {
          double Dphi = (j - jp) * k->dphi;
          V += k->R32[ip] * Sigma[ip * Nphi + jp]
               / sqrt(2.0 * cosh(Du) - 2.0 * cos(Dphi) + k->eps);
        }
      }
      Phi[i * Nphi + j] = pref * V / k->sqrtR[i];
    }
}

/* ===================================================================== */
/* driver (diskmain.cpp)                                                 */
/* ===================================================================== */

/* diskmain.cpp:65-70 */
double orc_disk_mass(const orc_hydro *h) {
  double mD = 0.0;
  for (int i = 0; i < h->Nu; ++i) {
    double dA = h->R[i] * (h->Rf[i + 1] - h->Rf[i]) * h->dphi;
    for (int j = 0; j < h->Nphi; ++j) mD += h->Sig[i * h->Nphi + j] * dA;
  }
  return mD;
}

/* diskmain.cpp:72-84 */
void orc_cm(const orc_hydro *h, int m, double mD, double *mag, double *phase) {
  double re = 0, im = 0;
  for (int i = 0; i < h->Nu; ++i) {
    double dA = h->R[i] * (h->Rf[i + 1] - h->Rf[i]) * h->dphi;
    for (int j = 0; j < h->Nphi; ++j) {
      double phi = (j + 0.5) * h->dphi;
      re += h->Sig[i * h->Nphi + j] * cos(m * phi) * dA;
      im -= h->Sig[i * h->Nphi + j] * sin(m * phi) * dA;
    }
  }
  *mag = sqrt(re * re + im * im) / mD;
  *phase = atan2(im, re);
}

/* diskmain.cpp:28-63 (Rin/Rout/K/gamma/GM* are taken from h) */
void orc_ic_lka98(orc_hydro *h, orc_kalnajs *grav, int selfgrav, double amp) {
  const double R0 = 0.45, w2 = 0.05, S0 = 0.372;
  const int Nu = h->Nu, Nphi = h->Nphi;
  const double Kpoly = h->Kpoly, gamma_gas = h->gamma_gas, gmstar = h->gmstar;
#define SIGMA0(R) (S0 * exp(-((R) - R0) * ((R) - R0) / w2))
  double *gRdisk = (double *)calloc(Nu, sizeof(double));
  if (selfgrav && grav) {
    size_t N = (size_t)Nu * Nphi;
    double *Sax = (double *)calloc(N, sizeof(double));   /* every element written below */
    double *Phi = (double *)malloc(sizeof(double) * N);
    double *gR = (double *)malloc(sizeof(double) * N), *gp = (double *)malloc(sizeof(double) * N);
    for (int i = 0; i < Nu; ++i)
      for (int j = 0; j < Nphi; ++j) Sax[i * Nphi + j] = SIGMA0(h->R[i]);
    orc_kal_potential(grav, Sax, Phi);
    orc_kal_forces(grav, Phi, gR, gp);
    for (int i = 0; i < Nu; ++i) {
      double s = 0; for (int j = 0; j < Nphi; ++j) s += gR[i * Nphi + j];
      gRdisk[i] = s / Nphi;
    }
    free(Sax); free(Phi); free(gR); free(gp);
  }
  for (int i = 0; i < Nu; ++i) {
    double R = h->R[i];
    double dSig = SIGMA0(R) * (-2.0 * (R - R0) / w2);
    double dPdR = Kpoly * gamma_gas * pow(SIGMA0(R), gamma_gas - 1.0) * dSig;
    double v2 = gmstar / R + (R / SIGMA0(R)) * dPdR - R * gRdisk[i];
    double vphi = (v2 > 0) ? sqrt(v2) : 0.0;
    for (int j = 0; j < Nphi; ++j) {
      double phi = (j + 0.5) * h->dphi;
      h->Sig[i * Nphi + j] = SIGMA0(R) * (1.0 + amp * cos(2.0 * phi));
      h->vR[i * Nphi + j] = 0.0;
      h->vp[i * Nphi + j] = vphi;
    }
  }
#undef SIGMA0
  free(gRdisk);
}

/* diskmain.cpp:86-113 (max_steps is the added --steps cap; <=0 = none) */
int orc_run_from(orc_hydro *h, orc_kalnajs *grav, int selfgrav, double tlim, double dtout,
                 double cfl, long max_steps, double mD, orc_sample *samples, int cap,
                 double *t_io, double *tnext_io, long *step_io, double *c2_final);
int orc_run(orc_hydro *h, orc_kalnajs *grav, int selfgrav, double tlim,
            double dtout, double cfl, long max_steps, double mD,
            orc_sample *samples, int cap, double *t_final, double *c2_final,
            long *steps_out) {
  double t = 0.0, tnext = 0.0;
  long step = 0;
  int ns = orc_run_from(h, grav, selfgrav, tlim, dtout, cfl, max_steps, mD, samples, cap,
                        &t, &tnext, &step, c2_final);
  if (t_final) *t_final = t;
  if (steps_out) *steps_out = step;
  return ns;
}

/* The same loop continued from (t, tnext, step); they are updated in place.
 * max_steps caps the total step count (<= 0: none). */
int orc_run_from(orc_hydro *h, orc_kalnajs *grav, int selfgrav, double tlim,
                 double dtout, double cfl, long max_steps, double mD,
                 orc_sample *samples, int cap, double *t_io, double *tnext_io,
                 long *step_io, double *c2_final) {
  const size_t N = (size_t)h->Nu * h->Nphi;
  double *Phi = NULL, *gR = NULL, *gp = NULL;
  if (selfgrav) {
    Phi = (double *)malloc(sizeof(double) * N);
    gR = (double *)malloc(sizeof(double) * N);
    gp = (double *)malloc(sizeof(double) * N);
  }
  double t = *t_io, tnext = *tnext_io, c0, c2, ph2, ph0;
  long step = *step_io;
  int ns = 0;
  while (t < tlim && (max_steps <= 0 || step < max_steps)) {
    if (t >= tnext - 1e-12) {
      orc_cm(h, 0, mD, &c0, &ph0); orc_cm(h, 2, mD, &c2, &ph2);
      if (ns < cap && samples) {
        samples[ns].t = t; samples[ns].c0 = c0; samples[ns].c2 = c2; samples[ns].ph2 = ph2;
      }
      ++ns;
      tnext += dtout;
    }
    double dt = orc_hydro_timestep(h, cfl);
    if (t + dt > tlim) dt = tlim - t;
    if (selfgrav) {
      orc_kal_potential(grav, h->Sig, Phi);
      orc_kal_forces(grav, Phi, gR, gp);
      orc_hydro_step(h, dt, gR, gp);
    } else {
      orc_hydro_step(h, dt, NULL, NULL);
    }
    t += dt; ++step;
  }
  orc_cm(h, 2, mD, &c2, &ph2);
  *t_io = t;
  *tnext_io = tnext;
  *step_io = step;
  if (c2_final) *c2_final = c2;
  free(Phi); free(gR); free(gp);
  return ns;
}
