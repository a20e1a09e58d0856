// Device-side data layout shared by the HIP kernels and the host runtime.
//
// One handle holds an ensemble of disks (a single disk is an ensemble of one).
// Every per-zone field is FP64, row-major [i][j] with j (azimuth) contiguous,
// pooled over disks: disk d's field starts at DiskDev::so.
//
//   sig, vr, vp, phi        state + potential pools (sum of Nu*Nphi)
//   spec                    half-spectrum pool, double2, per disk [Nc][Nu]
//                           (column-major so the column transform reads
//                           contiguous memory); holds S-hat after the row r2c
//                           and V-hat after the column convolution (in place)
//   khat                    real K-hat, per disk [Nc][Nu+1] (rows p=0..Nu of the
//                           real-even M x Nc transform; row M-p equals row p)
//   geom                    per disk GT_COUNT tables of Nu+1 doubles, then Nu+1
//                           row records of RR_STRIDE doubles
//   tw                      per transform length n: exp(-2 pi i t/n), t<n
#ifndef DFK_INTERNAL_H
#define DFK_INTERNAL_H

#include <stdint.h>

#define DFK_MAX_STAGES 16

// geometry tables (index t -> geom[go + t*(Nu+1) + i])
enum {
  GT_R = 0,      // R_i (centres)
  GT_RF,         // Rf_i (faces, Nu+1 entries)
  GT_INVVOL,     // 1/Vol_i, Vol_i = R_i (Rf_{i+1}-Rf_i) dphi
  GT_INVDRC,     // 1/(R_{i+1 clamped} - R_{i-1 clamped})   (source dP/dR)
  GT_INVR,       // 1/R_i
  GT_GSTAR,      // -GM*/(R_i R_i)
  GT_INVDR,      // 1/(Rf_{i+1}-Rf_i)
  GT_INVDL,      // 1/(R_i dphi)
  GT_AF,         // Rf_i dphi  (radial face length, Nu+1 entries)
  GT_AFACE,      // Rf_{i+1}-Rf_i (azimuthal face length)
  GT_R32,        // R_i^{3/2}
  GT_PHISC,      // pref / sqrt(R_i),  pref = -G du dphi/(M Nphi)
  GT_DA,         // R_i (Rf_{i+1}-Rf_i) dphi  (C_m area weight)
  GT_GRC,        // 1/(2 du R_i) interior, 1/(du R_i) at i=0, Nu-1 (self-gravity g_R)
  GT_GPC,        // 1/(2 dphi R_i)                                  (self-gravity g_phi)
  GT_DL,         // R_i dphi  (CFL azimuthal length, formed as hydro.hpp:56 forms it)
  GT_AFZ,        // Rf_i dphi at the interior faces 1..Nu-1, 0 at faces 0 and Nu (no flux)
  GT_HUP,        // 1 for the interior cells 1..Nu-2, 0 at cells 0 and Nu-1 (no slope): sign
  GT_HDN,        //   of the MUSCL factor of an upwind cell left (HUP) / right (HDN, -1) of its face
  GT_COUNT
};

// After a disk's GT_COUNT tables, its row records: Nu+1 records of
// RR_STRIDE doubles, record i holding the values of row (or face) i that the
// streaming update reads, so that one scalar pointer per row reaches all of
// them (values of the GT tables of the same name; record Nu holds only AFZ).
enum {
  RR_INVDRC = 0, RR_INVR, RR_GSTAR, RR_GRC, RR_GPC,   // source step of row i
  RR_INVDR, RR_HUP, RR_HDN, RR_AFZ,                   // radial face i and its cells
  RR_INVVOL, RR_INVDL, RR_AFACE, RR_DA, RR_DL,        // gather, azimuthal sweep, CFL, C_m of row i
  RR_STRIDE = 16
};
#define DFK_GEOM_DOUBLES(Nu) ((long long)(GT_COUNT + RR_STRIDE) * ((Nu) + 1))

// Byte range of the streaming update's buffer descriptors: a state buffer's
// three fields of every disk (3 ntot doubles) must lie within it.
#define DFK_STREAM_BYTES 0x7ffffff0

struct DiskDev {
  int Nu, Nphi, Nc, M;
  int gamma2;          // gamma == 2: P = K s^2, c_s^2 = 2 K s
  int selfgrav;
  double K, gamma, gmstar, G, eps;
  double du, dphi;
  double inv2dphi;     // 1/(2 dphi)
  int64_t so;          // state offset (doubles)
  int64_t spo;         // spectrum offset (double2)
  int64_t kho;         // khat offset (doubles)
  int64_t go;          // geometry offset (doubles)
  int64_t two_row;     // twiddle offset for length Nphi (double2)
  int64_t two_col;     // twiddle offset for length M (double2)
  int nst_row, nst_col;
  int rad_row[DFK_MAX_STAGES];
  int rad_col[DFK_MAX_STAGES];
  int ti, tj;          // update tile (rows x cols)
  int ntile_i, ntile_j;
  int64_t part;        // offset of this disk's tiles in the partial arrays
  int pr;              // complex rows (= 2 real rows) per workgroup in row transforms
  int cb;              // columns per workgroup in the column convolution
  int cos_off;         // offset of cos(2 phi_j)/sin(2 phi_j) tables (Nphi each) in trig pool
  int tw_lds;          // bit 0: row transforms, bit 1: column convolution stage their
                       // twiddles in LDS (else read them through the caches)
  int conv_wave;       // column convolution: 0 the workgroup cooperates on its columns
                       // (ping-pong buffers), 1 one wave per column in place, 2 the
                       // workgroup in place on one column after another, 3 the
                       // two-pass register transform, 4 the three-pass one
  int pad2_;
};

// Per-disk loop control, device resident (diskmain.cpp:86-113 state).
struct DiskCtl {
  double t, dt, tnext;
  double tlim, dtout, cfl;
  double fin_re2, fin_im2;   // C_2 sums at loop end
  long long step, max_steps;
  int cur;                   // state buffer (0/1) holding the current state
  int done;                  // loop finished (t >= tlim, step cap, or failure)
  int nsamples;              // samples recorded so far (ring buffer index)
  int flags;                 // DFK_FLAG_*
  int stop_on_nonfinite;
};

#define DFK_FLAG_NONFINITE 1

// Partials written by the update / diag kernels per tile:
//   pcfl[t]        max over the tile of (|vR|+cs)/dR + (|vphi|+cs)/(R dphi)
//   pcm[4t+0..3]   sum Sigma dA, sum Sigma cos(2phi) dA, -sum Sigma sin(2phi) dA,
//                  count of non-finite zones

// Sample record: t and the raw C_0 / C_2 sums (host finishes |C|, arg).
struct DfkSampleDev { double t, re0, re2, im2; };

#endif
