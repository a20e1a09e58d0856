/* oracle_diskdisk -- TEST INFRASTRUCTURE ONLY.
 * Restatement of the reference driver diskfft/src/diskmain.cpp:16-115 on the
 * oracle: same seven positional arguments, same defaults, same stdout lines
 * and the same diskfft_modes.dat file, so the product diskdisk can be diffed
 * against it.  Extra options after the positionals:
 *   --steps N   stop after N steps (the loop otherwise ends at tlim)
 *   --cfl C     CFL number (reference: 0.3 hard-coded at diskmain.cpp:100)
 *   --out FILE  modes file name (reference: diskfft_modes.dat)
 *   --time      print loop wall time and zone-updates/s to stderr
 */
#define _GNU_SOURCE
#include "diskfft_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
  const char *pos[7] = {0};
  int npos = 0;
  long max_steps = 0;
  double cfl = 0.3;
  const char *outname = "diskfft_modes.dat";
  int timeit = 0;
  for (int a = 1; a < argc; ++a) {
    if (strcmp(argv[a], "--steps") == 0 && a + 1 < argc) max_steps = atol(argv[++a]);
    else if (strcmp(argv[a], "--cfl") == 0 && a + 1 < argc) cfl = atof(argv[++a]);
    else if (strcmp(argv[a], "--out") == 0 && a + 1 < argc) outname = argv[++a];
    else if (strcmp(argv[a], "--time") == 0) timeit = 1;
    else if (npos < 7) pos[npos++] = argv[a];
  }
  /* diskmain.cpp:17-23 */
  int Nu = (npos > 0) ? atoi(pos[0]) : 128;
  int Nphi = (npos > 1) ? atoi(pos[1]) : 128;
  double eps = (npos > 2) ? atof(pos[2]) : 0.02;
  double tlim = (npos > 3) ? atof(pos[3]) : 8.0;
  int selfgrav = (npos > 4) ? atoi(pos[4]) : 1;
  double amp = (npos > 5) ? atof(pos[5]) : 1e-3;
  int nthreads = (npos > 6) ? atoi(pos[6]) : 1;
  orc_set_threads(nthreads);

  /* diskmain.cpp:28-29 */
  double Rin = 0.25, Rout = 1.0;
  double Kpoly = 0.25, gamma_gas = 2.0, gmstar = 0.6, G = 1.0;
  orc_hydro *h = orc_hydro_new(Nu, Nphi, Rin, Rout, Kpoly, gamma_gas, gmstar);
  orc_kalnajs *k = orc_kal_new(Nu, Nphi, Rin, Rout, eps, G);
  if (!k) { fprintf(stderr, "oracle: unsupported grid\n"); return 1; }
  orc_ic_lka98(h, k, selfgrav, amp);
  double mD = orc_disk_mass(h);

  FILE *fp = fopen(outname, "w");
  if (!fp) { perror(outname); return 1; }
  fprintf(fp, "# t  |C0|  |C2|  arg(C2)\n");
  /* The loop of diskmain.cpp:92-110, sample lines printed as produced. */
  int cap = 1 << 20;
  orc_sample *s = (orc_sample *)malloc(sizeof(orc_sample) * cap);
  double tf, c2f;
  long steps;
  double t0 = now_s();
  int ns = orc_run(h, k, selfgrav, tlim, 0.25, cfl, max_steps, mD, s, cap, &tf, &c2f, &steps);
  double t1 = now_s();
  for (int q = 0; q < ns && q < cap; ++q) {
    printf("t=%7.3f  |C0|=%.5f  |C2|=%.6e\n", s[q].t, s[q].c0, s[q].c2);
    fprintf(fp, "%.5f %.6e %.6e %.6e\n", s[q].t, s[q].c0, s[q].c2, s[q].ph2);
  }
  printf("final t=%.3f  |C2|=%.6e  steps=%d\n", tf, c2f, (int)steps);
  fclose(fp);
  if (timeit)
    fprintf(stderr, "loop_s=%.6f zone_updates_per_s=%.6e threads=%d\n", t1 - t0,
            (double)Nu * Nphi * steps / (t1 - t0), orc_get_threads());
  free(s);
  orc_kal_free(k);
  orc_hydro_free(h);
  return 0;
}
