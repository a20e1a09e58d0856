// oracle/_ref harness -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the reference's own diskfft/src/hydro.hpp and diskfft/src/fft.hpp,
// included from the reference checkout (include path given by the Makefile,
// nothing copied), behind a small C interface so the oracle restatement can be
// checked against the real reference hydro bit for bit.  kalnajs.hpp and
// diskmain.cpp need <fftw3.h>, which this image does not have: they are not
// built here (see the oracle header for how that part is pinned instead).
#include <cstring>
#include <vector>

#include "fft.hpp"
#include "hydro.hpp"

using diskfft::DiskHydro;

extern "C" {

void *ref_hydro_new(int Nu, int Nphi, double Rin, double Rout, double K, double g,
                    double gm) {
  return new DiskHydro(Nu, Nphi, Rin, Rout, K, g, gm);
}
void ref_hydro_free(void *p) { delete static_cast<DiskHydro *>(p); }

void ref_hydro_set(void *p, const double *Sig, const double *vR, const double *vp) {
  DiskHydro *h = static_cast<DiskHydro *>(p);
  size_t N = (size_t)h->Nu * h->Nphi;
  std::memcpy(h->Sig.data(), Sig, N * sizeof(double));
  std::memcpy(h->vR.data(), vR, N * sizeof(double));
  std::memcpy(h->vp.data(), vp, N * sizeof(double));
}
void ref_hydro_get(void *p, double *Sig, double *vR, double *vp) {
  DiskHydro *h = static_cast<DiskHydro *>(p);
  size_t N = (size_t)h->Nu * h->Nphi;
  std::memcpy(Sig, h->Sig.data(), N * sizeof(double));
  std::memcpy(vR, h->vR.data(), N * sizeof(double));
  std::memcpy(vp, h->vp.data(), N * sizeof(double));
}
void ref_hydro_geom(void *p, double *R, double *Rf, double *Vol) {
  DiskHydro *h = static_cast<DiskHydro *>(p);
  std::memcpy(R, h->R.data(), h->Nu * sizeof(double));
  std::memcpy(Rf, h->Rf.data(), (h->Nu + 1) * sizeof(double));
  std::memcpy(Vol, h->Vol.data(), h->Nu * sizeof(double));
}
double ref_hydro_timestep(void *p, double cfl) {
  return static_cast<DiskHydro *>(p)->timestep(cfl);
}
// stage: 0 = full step, 1 = source, 2 = transportR, 3 = transportPhi
void ref_hydro_stage(void *p, int stage, double dt, const double *gR, const double *gp) {
  DiskHydro *h = static_cast<DiskHydro *>(p);
  size_t N = (size_t)h->Nu * h->Nphi;
  std::vector<double> vgR, vgp;
  if (gR) vgR.assign(gR, gR + N);
  if (gp) vgp.assign(gp, gp + N);
  const std::vector<double> *pR = gR ? &vgR : nullptr, *pp = gp ? &vgp : nullptr;
  switch (stage) {
    case 0: h->step(dt, pR, pp); break;
    case 1: h->source(dt, pR, pp); break;
    case 2: h->transportR(dt); break;
    case 3: h->transportPhi(dt); break;
    default: break;
  }
}
// The reference's radix-2 complex 2-D FFT (fft.hpp:42-55); sign -1 forward.
void ref_fft2d(double *re, double *im, int n1, int n2, int sign) {
  std::vector<diskfft::cdouble> a((size_t)n1 * n2);
  for (size_t k = 0; k < a.size(); ++k) a[k] = diskfft::cdouble(re[k], im[k]);
  diskfft::fft2d(a, n1, n2, sign);
  for (size_t k = 0; k < a.size(); ++k) { re[k] = a[k].real(); im[k] = a[k].imag(); }
}
}
