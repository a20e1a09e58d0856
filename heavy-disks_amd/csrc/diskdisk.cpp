// diskdisk -- drop-in replacement of the reference driver diskfft/src/diskmain.cpp
// on MI355X.  Same command line and output:
//
//   diskdisk Nu Nphi eps tlim selfgrav amp nthreads        (defaults 128 128 0.02 8 1 1e-3 1)
//
// stdout  "t=%7.3f  |C0|=%.5f  |C2|=%.6e" at every output time (dtout 0.25),
//         "final t=%.3f  |C2|=%.6e  steps=%d" at the end;
// file    diskfft_modes.dat: "# t  |C0|  |C2|  arg(C2)" then "%.5f %.6e %.6e %.6e".
//
// `nthreads` is accepted and ignored (the loop runs on the GPU).  Options after
// the positionals (none changes the default behaviour):
//   --model lka98|powerlaw  disk model (default lka98 = the reference's ring;
//                           powerlaw = r in [0.1,1], Sigma ~ r^-1.5 with tapers,
//                           M_d/M* = 1, K from min Q = 1.3, 1e-3 noise in m=1..8)
//   --seed S  --qmin Q  --mass-ratio q  --slope s    power-law model parameters
//   --steps N               stop after N steps
//   --cfl C                 CFL number (reference: 0.3, diskmain.cpp:100)
//   --out FILE              modes file (reference: diskfft_modes.dat)
//   --device D              HIP device
//   --summary FILE          write a key=value summary (timing, K, steps)
//   --time                  print loop wall time and zone-updates/s to stderr
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/diskfft.h"
#include "../../include/diskfft_model.h"

static int die(const char *what, int rc) {
  fprintf(stderr, "diskdisk: %s failed (%d): %s\n", what, rc, dfk_last_error());
  return 1;
}

// The transform variants the autotuner picks for one disk on MI355X (the
// driver reads no tuned-shape cache): register row transforms, three-pass
// where Nphi has them (128 to 512), else two-pass, else what the library
// falls back to; the column convolution likewise (three-pass for 2Nu = 256,
// 384, 512, else two-pass, else the workgroup in-place kernel, else the
// cooperative one).  All agree with the defaults to round-off.
static int fast_transforms(dfk_handle *h) {
  int rc = dfk_set_row_transform(h, 3);
  if (rc) return rc;
  for (int mode : {4, 3, 2}) {
    int r = 0, c = 0, w = 0;
    if ((rc = dfk_set_fft_options(h, 0, 1, 1, mode))) return rc;
    if ((rc = dfk_get_fft_options(h, 0, &r, &c, &w))) return rc;
    if (w == mode) return 0;
  }
  return dfk_set_fft_options(h, 0, 1, 1, 0);
}

int main(int argc, char **argv) {
  const char *pos[7] = {0};
  int npos = 0;
  long long max_steps = 0;
  double cfl = 0.3;
  std::string outname = "diskfft_modes.dat", summary, model = "lka98";
  int device = 0, timeit = 0;
  dfk_model pl;
  dfk_model_defaults(DFK_MODEL_POWERLAW, &pl);
  for (int a = 1; a < argc; ++a) {
    const char *s = argv[a];
    const bool more = a + 1 < argc;
    if (!strcmp(s, "--steps") && more) max_steps = atoll(argv[++a]);
    else if (!strcmp(s, "--cfl") && more) cfl = atof(argv[++a]);
    else if (!strcmp(s, "--out") && more) outname = argv[++a];
    else if (!strcmp(s, "--device") && more) device = atoi(argv[++a]);
    else if (!strcmp(s, "--summary") && more) summary = argv[++a];
    else if (!strcmp(s, "--model") && more) model = argv[++a];
    else if (!strcmp(s, "--seed") && more) pl.seed = strtoull(argv[++a], nullptr, 10);
    else if (!strcmp(s, "--qmin") && more) pl.Qmin = atof(argv[++a]);
    else if (!strcmp(s, "--mass-ratio") && more) pl.mass_ratio = atof(argv[++a]);
    else if (!strcmp(s, "--slope") && more) pl.slope = atof(argv[++a]);
    else if (!strcmp(s, "--time")) timeit = 1;
    else if (npos < 7) pos[npos++] = s;
  }
  // diskmain.cpp:17-23
  int Nu = (npos > 0) ? atoi(pos[0]) : 128;
  int Nphi = (npos > 1) ? atoi(pos[1]) : 128;
  double eps = (npos > 2) ? atof(pos[2]) : 0.02;
  double tlim = (npos > 3) ? atof(pos[3]) : 8.0;
  int selfgrav = (npos > 4) ? atoi(pos[4]) : 1;
  double amp = (npos > 5) ? atof(pos[5]) : 1e-3;
  (void)((npos > 6) ? atoi(pos[6]) : 1);   // nthreads: the loop runs on the GPU

  dfk_model m;
  if (model == "lka98") {
    dfk_model_defaults(DFK_MODEL_LKA98, &m);
    m.amp = amp;
  } else if (model == "powerlaw") {
    m = pl;
  } else {
    fprintf(stderr, "diskdisk: unknown model '%s'\n", model.c_str());
    return 2;
  }

  dfk_disk_params p;
  p.Nu = Nu; p.Nphi = Nphi; p.Rin = m.Rin; p.Rout = m.Rout; p.K = m.K; p.gamma = m.gamma;
  p.gmstar = m.gmstar; p.eps = eps; p.G = m.G; p.selfgrav = selfgrav;
  dfk_handle *h = nullptr;
  int rc = dfk_create(&p, 1, device, &h);
  if (rc) return die("dfk_create", rc);
  if ((rc = fast_transforms(h))) return die("transform options", rc);

  const size_t N = (size_t)Nu * Nphi;
  std::vector<double> sig0(Nu), gRdisk(Nu, 0.0), Sig(N), vR(N), vp(N);
  dfk_model_sigma0(&m, Nu, Nphi, sig0.data());
  if (selfgrav) {
    // diskmain.cpp:36-48: phi-averaged g_R of the axisymmetric disk
    std::vector<double> Sax(N), Phi(N), gR(N), gp(N);
    for (int i = 0; i < Nu; ++i)
      for (int j = 0; j < Nphi; ++j) Sax[(size_t)i * Nphi + j] = sig0[i];
    if ((rc = dfk_potential(h, 0, Sax.data(), Phi.data()))) return die("dfk_potential", rc);
    if ((rc = dfk_forces(h, 0, Phi.data(), gR.data(), gp.data()))) return die("dfk_forces", rc);
    for (int i = 0; i < Nu; ++i) {
      double s = 0;
      for (int j = 0; j < Nphi; ++j) s += gR[(size_t)i * Nphi + j];
      gRdisk[i] = s / Nphi;
    }
  }
  if (dfk_model_state(&m, Nu, Nphi, gRdisk.data(), Sig.data(), vR.data(), vp.data(), nullptr)) {
    fprintf(stderr, "diskdisk: model set-up failed\n");
    return 1;
  }
  if (m.kind == DFK_MODEL_POWERLAW) {
    // K came out of the Q solve: rebuild the handle with it
    dfk_destroy(h);
    p.K = m.K;
    if ((rc = dfk_create(&p, 1, device, &h))) return die("dfk_create", rc);
    if ((rc = fast_transforms(h))) return die("transform options", rc);
  }
  if ((rc = dfk_set_state(h, 0, Sig.data(), vR.data(), vp.data()))) return die("dfk_set_state", rc);
  const double mD = dfk_model_mass(&m, Nu, Nphi, Sig.data());   // diskmain.cpp:65-70

  FILE *fp = fopen(outname.c_str(), "w");
  if (!fp) {
    perror(outname.c_str());
    return 1;
  }
  fprintf(fp, "# t  |C0|  |C2|  arg(C2)\n");
  dfk_run_params rp;
  rp.tlim = tlim; rp.dtout = 0.25; rp.cfl = cfl; rp.mD = mD; rp.max_steps = max_steps;
  rp.stop_on_nonfinite = 0;
  if ((rc = dfk_run_init(h, &rp, 1024))) return die("dfk_run_init", rc);
  std::vector<dfk_sample> buf(4096);
  auto t0 = std::chrono::steady_clock::now();
  int done = 0;
  while (!done) {
    if ((rc = dfk_advance(h, 1000, &done))) return die("dfk_advance", rc);
    int n = 0;
    do {
      dfk_samples(h, 0, buf.data(), (int)buf.size(), &n);
      for (int q = 0; q < n; ++q) {
        printf("t=%7.3f  |C0|=%.5f  |C2|=%.6e\n", buf[q].t, buf[q].c0, buf[q].c2);
        fprintf(fp, "%.5f %.6e %.6e %.6e\n", buf[q].t, buf[q].c0, buf[q].c2, buf[q].ph2);
      }
      fflush(fp);
    } while (n == (int)buf.size());
  }
  auto t1 = std::chrono::steady_clock::now();
  dfk_status st;
  dfk_get_status(h, 0, &st);
  printf("final t=%.3f  |C2|=%.6e  steps=%d\n", st.t, st.c2_final, (int)st.steps);
  fclose(fp);
  const double secs = std::chrono::duration<double>(t1 - t0).count();
  const double zu = (double)N * st.steps / secs;
  if (timeit) fprintf(stderr, "loop_s=%.6f zone_updates_per_s=%.6e\n", secs, zu);
  if (!summary.empty()) {
    FILE *fs = fopen(summary.c_str(), "w");
    if (fs) {
      fprintf(fs, "model=%s\nNu=%d\nNphi=%d\neps=%.17g\nK=%.17g\nsteps=%lld\nt=%.17g\n"
                  "loop_s=%.9f\nzone_updates_per_s=%.6e\nflags=%d\n",
              model.c_str(), Nu, Nphi, eps, m.K, st.steps, st.t, secs, zu, st.flags);
      fclose(fs);
    }
  }
  dfk_destroy(h);
  return 0;
}
