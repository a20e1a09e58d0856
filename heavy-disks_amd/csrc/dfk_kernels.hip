// HIP kernels (gfx950, FP64) for the diskfft time-step loop.
//
// Per step, for every disk of the handle (a single disk is an ensemble of one):
//   k_rows_r2c  S = R^{3/2} Sigma, row transforms of length Nphi (two real rows
//               packed per complex transform), half spectrum -> spec[Nc][Nu];
//               its trailing workgroups (one per disk) run the loop control:
//               dt = cfl / max(...) from the previous step's per-tile maxima,
//               output-time check and C_m sample, last-step clip, t += dt
//               (diskmain.cpp:92-110, hydro.hpp:48-61)
//   k_cols_conv per column: zero-padded length-2Nu transform, x real K-hat,
//               inverse, keep rows < Nu                    (kalnajs.hpp:48-57)
//   k_rows_c2r  row c2r + scale pref/sqrt(R) -> Phi       (kalnajs.hpp:57-62)
//   k_update    forces (kalnajs.hpp:65-80) + source (hydro.hpp:73-96) + radial
//               MUSCL sweep (:99-140) + azimuthal sweep (:143-179) + floor
//               (:68-69) fused on an LDS tile (k_update_stream: the same
//               fused on a wavefront streaming down a 58-column strip);
//               epilogue writes the new state's CFL maximum and C_0/C_2
//               partial sums per tile.
// k_rows_fused (an option for small disks) runs the row c2r, the update and
// the next row r2c for a block of full rows in one launch; the loop control
// then rides in the column-convolution launch.
// k_control is the loop control as a launch of its own (handles without
// self-gravity, profiling).
// All FFTs are this file's own LDS Stockham transforms (radix 8, 4, 2, 3
// unrolled; any other prime by a direct DFT stage), plus the column
// convolution's two-pass register transform (dfk_regfft.h).
#include <hip/hip_runtime.h>

#include "dfk_internal.h"
#include "dfk_regfft.h"

#define WG 256

typedef long long i64;

__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// 1/x to about 1 ulp (x finite, nonzero and normal where it is used): the
// hardware estimate r (relative error below 2^-24 on gfx950, measured over
// 16.7M operands) times 1 + e + e^2 with e = 1 - x r, the cubically
// convergent step (error ~e^3 < 2^-72 before the final rounding; it gave the
// correctly rounded 1/x for every measured operand, as two Newton steps did,
// in three FMAs instead of four).
__device__ __forceinline__ double frcp(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

// hydro.hpp:16-18.  a*b > 0 keeps a+b away from zero and from denormals.
__device__ __forceinline__ double vanleer(double a, double b) {
  return (a * b > 0.0) ? (2.0 * a * b) * frcp(a + b) : 0.0;
}
// mid + vanleer(hi - mid, mid - lo) * fac where use is set, else mid, with
// the quotient formed on every lane and then selected (no branch around the
// reciprocal; where a*b <= 0 or use is clear the quotient, perhaps inf or
// NaN, is dropped): the same operations as the branch of hydro.hpp:16-18.
__device__ __forceinline__ double slope_add(double mid, double lo, double hi, bool use,
                                            double fac) {
  const double a = hi - mid, b = mid - lo;
  const double q = (2.0 * a * b) * frcp(a + b);
  const double vl = (use && a * b > 0.0) ? q : 0.0;
  return mid + vl * fac;
}
// Half the van Leer slope of a cell, vanleer(hi - mid, mid - lo) / 2, with
// the quotient formed on every lane and then selected, so that no branch
// surrounds the reciprocal (where a*b <= 0 the quotient may be inf or NaN;
// the select drops it).  It is the reference's value halved exactly:
// RN(2 RN(ab) r) = 2 RN(RN(ab) r) with r = frcp(a + b).
__device__ __forceinline__ double half_slope(double lo, double mid, double hi) {
  const double a = hi - mid, b = mid - lo;
  const double p = a * b;
  const double q = p * frcp(a + b);
  return (p > 0.0) ? q : 0.0;
}

// Barotropic EOS (hydro.hpp:44-45).  G2: every disk of the launch has
// gamma = 2 (no pow; the general path costs the update kernels ~45 VGPRs).
template <bool G2 = false>
__device__ __forceinline__ double pres(const DiskDev &D, double s) {
  if constexpr (G2) return D.K * (s * s);
  return D.gamma2 ? D.K * (s * s) : D.K * pow(s, D.gamma);
}
template <bool G2 = false>
__device__ __forceinline__ double cs2f(const DiskDev &D, double s) {
  if constexpr (G2) return (2.0 * D.K) * s;
  return D.gamma2 ? (2.0 * D.K) * s : D.gamma * D.K * pow(s, D.gamma - 1.0);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

__device__ __forceinline__ int wrapj(int j, int n) {
  if (j < 0) j += n;
  if (j >= n) j -= n;
  if (j >= n) j -= n;
  return j;
}

__device__ __forceinline__ const double *gtab(const double *geom, const DiskDev &D, int t) {
  return geom + D.go + (i64)t * (D.Nu + 1);
}

// ------------------------------------------------------------------ FFT
// Quotient of small non-negative ints (a < 2^22, b <= 2^12) through a float
// reciprocal: (a + 0.5) / b is at least 0.5/b away from an integer.  rb comes
// from the hardware reciprocal (rcpf, within 1 ulp), so the product is within
// 1.5 * 2^-23 relative of (a + 0.5) / b, which cannot cross that distance for
// a < 2^21 (every use divides a loop index below 2^16).  Used for every
// division by a run-time value in the loop kernels: the GPU has no integer
// divide, and the emulated one (~40 instructions, scalar or per lane) was a
// large part of the FFT kernels' instruction count; so was the IEEE float
// division (~10 instructions) that 1.0f / x compiles to.
__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ int qdiv(int a, float rb) { return (int)(((float)a + 0.5f) * rb); }

__device__ __forceinline__ double2 twid(const double2 *__restrict__ tw, int t, int sign) {
  double2 w = tw[t];
  if (sign > 0) w.y = -w.y;
  return w;
}

// Inputs 1 .. R-1 of a butterfly times their twiddles tw[m kt]: every twiddle
// load is issued before the first product (one memory latency per butterfly,
// not R - 1 dependent ones), and lanes with kt = 0 multiply by tw[0] = 1,
// which is exact, instead of branching around the product.
template <int R, int SIGN>
__device__ __forceinline__ void twiddle_mul(double2 (&v)[R], const double2 *__restrict__ tw, int kt) {
  double2 w[R];
#pragma unroll
  for (int m = 1; m < R; ++m) w[m] = twid(tw, m * kt, SIGN);
#pragma unroll
  for (int m = 1; m < R; ++m) v[m] = cmul(v[m], w[m]);
}

// In-register butterflies (forward for SIGN < 0).
template <int SIGN>
__device__ __forceinline__ void bfly2(double2 &a, double2 &b) {
  const double2 t = csub(a, b);
  a = cadd(a, b);
  b = t;
}
template <int SIGN>
__device__ __forceinline__ void bfly4(double2 (&v)[4]) {
  const double sg = (SIGN < 0) ? 1.0 : -1.0;
  const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  const double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
  const double2 t = make_double2(sg * b1.y, -sg * b1.x);   // -i*b1 forward, +i*b1 inverse
  v[0] = cadd(a0, b0);
  v[1] = cadd(a1, t);
  v[2] = csub(a0, b0);
  v[3] = csub(a1, t);
}
template <int SIGN>
__device__ __forceinline__ void bfly3(double2 (&v)[3]) {
  // W = exp(SIGN 2 pi i / 3) = -1/2 + i s,  s = SIGN sqrt(3)/2
  const double s = (SIGN < 0) ? -0.86602540378443864676 : 0.86602540378443864676;
  const double2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
  const double2 c = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  const double2 is = make_double2(-s * t2.y, s * t2.x);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(c, is);
  v[2] = csub(c, is);
}

// Radix 8 as radix 2 then radix 4: a_m = v_m + v_{m+4}, b_m = (v_m - v_{m+4}) W8^m,
// X_{2k} = DFT4(a)_k, X_{2k+1} = DFT4(b)_k, W8 = exp(SIGN 2 pi i / 8).
template <int SIGN>
__device__ __forceinline__ void bfly8(double2 (&v)[8]) {
  const double h = 0.70710678118654752440;
  double2 a[4], b[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    a[m] = cadd(v[m], v[m + 4]);
    b[m] = csub(v[m], v[m + 4]);
  }
  if (SIGN < 0) {   // W8 = (1 - i)/sqrt2
    b[1] = make_double2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
    b[2] = make_double2(b[2].y, -b[2].x);
    b[3] = make_double2(-h * (b[3].x - b[3].y), -h * (b[3].x + b[3].y));
  } else {          // W8 = (1 + i)/sqrt2
    b[1] = make_double2(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y));
    b[2] = make_double2(-b[2].y, b[2].x);
    b[3] = make_double2(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y));
  }
  bfly4<SIGN>(a);
  bfly4<SIGN>(b);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = a[k];
    v[2 * k + 1] = b[k];
  }
}

// One Stockham stage of radix R in {2, 3, 4, 8} over `batch` transforms of
// length n stored contiguously in LDS (a -> b).  Ns = product of the radices
// already applied.  Twiddles tw[t] = exp(-2 pi i t / n), conjugated for SIGN > 0.
template <int R, int SIGN>
__device__ __forceinline__ void fft_stage(const double2 *__restrict__ a, double2 *__restrict__ b,
                                          int n, int batch, int Ns,
                                          const double2 *__restrict__ tw) {
  const int nb = n / R;
  const int tstep = qdiv(n, rcpf((float)(Ns * R)));   // exact quotient, no integer division
  const int total = nb * batch;
  const float rnb = rcpf((float)nb), rNs = rcpf((float)Ns);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int p = qdiv(e, rnb);
    const int j = e - p * nb;
    const int jq = qdiv(j, rNs);
    const int k = j - jq * Ns;
    const double2 *src = a + p * n + j;
    double2 *dst = b + p * n + jq * Ns * R + k;
    double2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = src[m * nb];
    if (Ns > 1) twiddle_mul<R, SIGN>(v, tw, k * tstep);
    if constexpr (R == 2) bfly2<SIGN>(v[0], v[1]);
    if constexpr (R == 3) bfly3<SIGN>(v);
    if constexpr (R == 4) bfly4<SIGN>(v);
    if constexpr (R == 8) bfly8<SIGN>(v);
#pragma unroll
    for (int q = 0; q < R; ++q) dst[q * Ns] = v[q];
  }
}

// Stage of any other prime radix R: direct R-point DFT reading its inputs
// from LDS (few registers; R^2 products per R points, so a large prime
// factor makes its transform slow but not wrong).
template <int SIGN>
__device__ __forceinline__ void fft_stage_generic(const double2 *__restrict__ a, double2 *__restrict__ b, int n,
                                  int batch, int Ns, int R, const double2 *__restrict__ tw) {
  const int nb = qdiv(n, rcpf((float)R));
  const int tstep = qdiv(n, rcpf((float)(Ns * R)));
  const int rstep = nb;
  const int total = nb * batch;
  const float rnb = rcpf((float)nb), rNs = rcpf((float)Ns);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int p = qdiv(e, rnb);
    const int j = e - p * nb;
    const int jq = qdiv(j, rNs);
    const int k = j - jq * Ns;
    const double2 *src = a + p * n + j;
    double2 *dst = b + p * n + jq * Ns * R + k;
    for (int q = 0; q < R; ++q) {
      const int base = k * tstep + q * rstep;   // exponent step per input m
      double2 s = src[0];
      int t = 0;
      for (int m = 1; m < R; ++m) {
        t += base;
        if (t >= n) t -= n;
        if (t >= n) t -= n;
        const double2 c = cmul(src[m * nb], twid(tw, t, SIGN));
        s.x += c.x;
        s.y += c.y;
      }
      dst[q * Ns] = s;
    }
  }
}

// Copy the n twiddles of a transform into LDS (read many times per stage).
// B passes of staged_for starting at element b0 + threadIdx.x.
template <int B, class T, class Ld, class St>
__device__ __forceinline__ void staged_pass(int b0, int total, Ld &ld, St &st) {
  const int nt = blockDim.x;
  T v[B];
#pragma unroll
  for (int u = 0; u < B; ++u) v[u] = ld(min(b0 + u * nt + (int)threadIdx.x, total - 1));
#pragma unroll
  for (int u = 0; u < B; ++u) {
    const int e = b0 + u * nt + threadIdx.x;
    if (e < total) st(e, v[u]);
  }
}

// st(e, ld(e)) for e = threadIdx.x, threadIdx.x + blockDim.x, ... < total,
// with the loads of up to B consecutive passes issued before the first
// store: ld gets an index clamped into [0, total) and must load without
// branching (clamped addresses, selects), so that the loads sit in one basic
// block and the workgroup waits on about one memory latency per batch
// instead of one (or more) per pass.  A batch is never longer than the
// passes left (no redundant loads when one pass covers everything).
template <int B, class T, class Ld, class St>
__device__ __forceinline__ void staged_for(int total, Ld ld, St st) {
  const int nt = blockDim.x;
  for (int b0 = 0; b0 < total;) {
    const int left = (total - b0 + nt - 1) / nt;   // passes left, the same for every thread
    if (B >= 8 && left >= 8) {
      staged_pass<8, T>(b0, total, ld, st);
      b0 += 8 * nt;
    } else if (B >= 4 && left >= 4) {
      staged_pass<4, T>(b0, total, ld, st);
      b0 += 4 * nt;
    } else if (B >= 2 && left >= 2) {
      staged_pass<2, T>(b0, total, ld, st);
      b0 += 2 * nt;
    } else {
      staged_pass<1, T>(b0, total, ld, st);
      b0 += nt;
    }
  }
}

__device__ __forceinline__ void stage_twiddles(double2 *__restrict__ dst,
                                               const double2 *__restrict__ src, int n) {
  for (int e = threadIdx.x; e < n; e += blockDim.x) dst[e] = src[e];
}

// The twiddle table a transform reads: staged in LDS at `lds` (visible after
// the next barrier; no transform's first stage uses twiddles) or, to keep
// the workgroup's LDS and so its occupancy smaller, the HBM table itself
// read through L1/L2 (a per-disk autotuner knob).
__device__ __forceinline__ const double2 *tw_table(double2 *lds, const double2 *__restrict__ src,
                                                   int n, int in_lds) {
  if (!in_lds) return src;
  stage_twiddles(lds, src, n);
  return lds;
}

// Full transform of `batch` rows of length n in LDS: returns the buffer that
// holds the result (a or b).  One barrier per stage.  Inlined (as are the
// stage helpers): as a called function the call convention alone took the
// FFT kernels from ~70 to 162 VGPRs (3 instead of 6-7 waves per SIMD).
template <int SIGN>
__device__ __forceinline__ double2 *lds_fft(double2 *a, double2 *b, int n, int batch, int nst,
                            const int *__restrict__ rad, const double2 *__restrict__ tw) {
  int Ns = 1;
  for (int s = 0; s < nst; ++s) {
    const int r = rad[s];
    if (r == 8) fft_stage<8, SIGN>(a, b, n, batch, Ns, tw);
    else if (r == 4) fft_stage<4, SIGN>(a, b, n, batch, Ns, tw);
    else if (r == 2) fft_stage<2, SIGN>(a, b, n, batch, Ns, tw);
    else if (r == 3) fft_stage<3, SIGN>(a, b, n, batch, Ns, tw);
    else fft_stage_generic<SIGN>(a, b, n, batch, Ns, r, tw);
    __syncthreads();
    double2 *t = a;
    a = b;
    b = t;
    Ns *= r;
  }
  return a;
}

// ------------------------------------------------------------------ stages with global I/O
// Stockham stages (radix 2, 3, 4 or 8) whose input or output goes straight
// to HBM instead of through an LDS staging pass, for transforms of >= 2
// stages with only those radices:
//   IO_ZP    column convolution, first forward stage (Ns = 1, no twiddles):
//            column col of the half spectrum, rows >= Nu being the zero
//            padding (for an even radix exactly the inputs m >= R/2)
//   IO_KHAT  column convolution, last forward stage: outputs x real K-hat
//            (rows M - p and p share a value)
//   IO_OUT   column convolution, last inverse stage: only rows < Nu stored
//   IO_ROUT  row c2r, last stage: batch entry p holds rows row0 + 2p (real
//            part) and row0 + 2p + 1 (imaginary part); x pref / sqrt(R) into Phi
// Butterflies and products are those of the LDS passes they replace, so the
// results are the same bits.
enum { IO_LDS = 0, IO_ZP = 1, IO_KHAT = 2, IO_OUT = 3, IO_ROUT = 4 };

struct StageIO {
  double2 *spec;          // IO_ZP / IO_OUT: the disk's half spectrum [Nc][Nu]
  const double *kh;       // IO_KHAT: the disk's real K-hat [Nc][Nu + 1]
  const double *rsc;      // IO_ROUT: pref / sqrt(R) per row
  double *P;              // IO_ROUT: Phi of the disk
  int Nu, Nc, k0, row0;   // k0: first column of the batch; row0: first row
};

template <int R, int SIGN, int IO>
__device__ __forceinline__ void io_stage(const double2 *__restrict__ a, double2 *__restrict__ b,
                                         int n, int batch, int Ns, const double2 *__restrict__ tw,
                                         const StageIO &g) {
  const int nb = n / R;
  const int tstep = qdiv(n, rcpf((float)(Ns * R)));
  const int total = nb * batch;
  const float rnb = rcpf((float)nb), rNs = rcpf((float)Ns);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int p = qdiv(e, rnb);
    const int j = e - p * nb;
    const int jq = qdiv(j, rNs);
    const int k = j - jq * Ns;
    const int col = g.k0 + p;
    const bool live = col < g.Nc;
    const int ra = g.row0 + 2 * p, rb = ra + 1;
    double2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int idx = j + m * nb;
      if constexpr (IO == IO_ZP) {
        const bool pad = (R % 2 == 0) ? (m >= R / 2) : (idx >= g.Nu);
        v[m] = (live && !pad) ? g.spec[(i64)col * g.Nu + idx] : make_double2(0.0, 0.0);
      } else {
        v[m] = a[p * n + idx];
      }
    }
    if constexpr (IO != IO_ZP) {
      if (Ns > 1) twiddle_mul<R, SIGN>(v, tw, k * tstep);
    }
    if constexpr (R == 2) bfly2<SIGN>(v[0], v[1]);
    if constexpr (R == 3) bfly3<SIGN>(v);
    if constexpr (R == 4) bfly4<SIGN>(v);
    if constexpr (R == 8) bfly8<SIGN>(v);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int o = jq * Ns * R + k + q * Ns;
      if constexpr (IO == IO_KHAT) {
        const double kv = live ? g.kh[(i64)col * (g.Nu + 1) + (o <= g.Nu ? o : n - o)] : 0.0;
        double2 z = v[q];
        z.x *= kv;
        z.y *= kv;
        b[p * n + o] = z;
      } else if constexpr (IO == IO_OUT) {
        if (live && o < g.Nu) g.spec[(i64)col * g.Nu + o] = v[q];
      } else if constexpr (IO == IO_ROUT) {
        if (ra < g.Nu) g.P[(i64)ra * n + o] = v[q].x * g.rsc[ra];
        if (rb < g.Nu) g.P[(i64)rb * n + o] = v[q].y * g.rsc[rb];
      } else {
        b[p * n + o] = v[q];
      }
    }
  }
}

template <int SIGN, int IO>
__device__ __forceinline__ void io_stage_r(int r, const double2 *a, double2 *b, int n, int batch,
                                           int Ns, const double2 *tw, const StageIO &g) {
  if (r == 8) io_stage<8, SIGN, IO>(a, b, n, batch, Ns, tw, g);
  else if (r == 4) io_stage<4, SIGN, IO>(a, b, n, batch, Ns, tw, g);
  else if (r == 2) io_stage<2, SIGN, IO>(a, b, n, batch, Ns, tw, g);
  else io_stage<3, SIGN, IO>(a, b, n, batch, Ns, tw, g);
}

// ------------------------------------------------------------------ in-place transforms
// Variants of the transforms in which a group of TG threads (one wavefront,
// or the whole workgroup) transforms a batch of transforms in place in a
// single LDS buffer (no ping-pong buffer: half the LDS of the cooperative
// kernels).  In each stage every thread loads the inputs of all its
// butterflies (at most NMAX / R / TG of them: batch x n <= NMAX), the group
// synchronises, then every thread computes and stores: no thread overwrites
// a value another one still has to read.  A wavefront synchronises without
// a workgroup barrier.  Same butterflies, twiddles and I/O variants as
// io_stage (results agree to round-off; the compiler's multiply-add
// contraction may differ).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): this wave's LDS accesses are done
  __builtin_amdgcn_wave_barrier();
}

template <int TG>
__device__ __forceinline__ void group_sync() {
  if constexpr (TG == 64) wave_lds_sync();
  else __syncthreads();
}

// Points a workgroup of TG threads transforms at once in place: 8 per thread
// (one radix-8 butterfly), at least one length-1024 column.
template <int TG>
constexpr int ip_nmax() { return 8 * TG > 1024 ? 8 * TG : 1024; }

// One stage over `batch` transforms of length n (batch x n <= NMAX) for
// thread `lane` (0 .. TG-1) of the group.  Batch entry p is column k0 + p
// (IO_ZP / IO_KHAT / IO_OUT) or rows row0 + 2p, row0 + 2p + 1 (IO_ROUT).
// BATCH = false: batch is 1 (no per-butterfly batch index; fewer registers).
template <int R, int SIGN, int IO, int TG, int NMAX, bool BATCH = true>
__device__ __forceinline__ void ip_stage(double2 *buf, int n, int batch, int Ns, int lane,
                                         const double2 *__restrict__ tw, const StageIO &g) {
  constexpr int EMAX = (NMAX / R + TG - 1) / TG;   // butterflies per thread
  const int nb = n / R, total = BATCH ? nb * batch : nb;
  const int tstep = qdiv(n, rcpf((float)(Ns * R)));
  const float rNs = rcpf((float)Ns), rnb = BATCH ? rcpf((float)nb) : 0.0f;
  double2 v[EMAX][R];
#pragma unroll
  for (int u = 0; u < EMAX; ++u) {
    const int e = lane + TG * u;
    if (e < total) {
      const int p = BATCH ? qdiv(e, rnb) : 0;
      const int j = e - p * nb;
      const int jq = qdiv(j, rNs);
      const int k = j - jq * Ns;
      const int col = g.k0 + p;
      const bool live = col < g.Nc;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int idx = j + m * nb;
        if constexpr (IO == IO_ZP) {
          const bool pad = (R % 2 == 0) ? (m >= R / 2) : (idx >= g.Nu);
          v[u][m] = (live && !pad) ? g.spec[(i64)col * g.Nu + idx] : make_double2(0.0, 0.0);
        } else {
          v[u][m] = buf[p * n + idx];
        }
      }
      if constexpr (IO != IO_ZP) {
        if (Ns > 1) twiddle_mul<R, SIGN>(v[u], tw, k * tstep);
      }
    }
  }
  group_sync<TG>();
#pragma unroll
  for (int u = 0; u < EMAX; ++u) {
    const int e = lane + TG * u;
    if (e < total) {
      const int p = BATCH ? qdiv(e, rnb) : 0;
      const int j = e - p * nb;
      const int jq = qdiv(j, rNs);
      const int k = j - jq * Ns;
      const int col = g.k0 + p;
      const bool live = col < g.Nc;
      const int ra = g.row0 + 2 * p, rb = ra + 1;
      if constexpr (R == 2) bfly2<SIGN>(v[u][0], v[u][1]);
      if constexpr (R == 3) bfly3<SIGN>(v[u]);
      if constexpr (R == 4) bfly4<SIGN>(v[u]);
      if constexpr (R == 8) bfly8<SIGN>(v[u]);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int o = jq * Ns * R + k + q * Ns;
        if constexpr (IO == IO_KHAT) {
          const double kv = live ? g.kh[(i64)col * (g.Nu + 1) + (o <= g.Nu ? o : n - o)] : 0.0;
          double2 z = v[u][q];
          z.x *= kv;
          z.y *= kv;
          buf[p * n + o] = z;
        } else if constexpr (IO == IO_OUT) {
          if (live && o < g.Nu) g.spec[(i64)col * g.Nu + o] = v[u][q];
        } else if constexpr (IO == IO_ROUT) {
          if (ra < g.Nu) g.P[(i64)ra * n + o] = v[u][q].x * g.rsc[ra];
          if (rb < g.Nu) g.P[(i64)rb * n + o] = v[u][q].y * g.rsc[rb];
        } else {
          buf[p * n + o] = v[u][q];
        }
      }
    }
  }
  group_sync<TG>();
}

template <int SIGN, int IO, int TG, int NMAX, bool BATCH = true>
__device__ __forceinline__ void ip_stage_r(int r, double2 *buf, int n, int batch, int Ns, int lane,
                                           const double2 *tw, const StageIO &g) {
  if (r == 8) ip_stage<8, SIGN, IO, TG, NMAX, BATCH>(buf, n, batch, Ns, lane, tw, g);
  else if (r == 4) ip_stage<4, SIGN, IO, TG, NMAX, BATCH>(buf, n, batch, Ns, lane, tw, g);
  else if (r == 2) ip_stage<2, SIGN, IO, TG, NMAX, BATCH>(buf, n, batch, Ns, lane, tw, g);
  else ip_stage<3, SIGN, IO, TG, NMAX, BATCH>(buf, n, batch, Ns, lane, tw, g);
}

// The whole in-place convolution of columns k0 .. k0 + batch - 1 (g.k0) in
// buf by a group of TG threads.
template <int TG, int NMAX, bool BATCH = true>
__device__ __forceinline__ void ip_conv(const DiskDev &D, double2 *buf, int batch, int lane,
                                        const double2 *TW, const StageIO &g) {
  const int M = D.M, ns = D.nst_col;
  ip_stage_r<-1, IO_ZP, TG, NMAX, BATCH>(D.rad_col[0], buf, M, batch, 1, lane, TW, g);
  int Ns = D.rad_col[0];
  for (int q = 1; q < ns - 1; ++q) {
    ip_stage_r<-1, IO_LDS, TG, NMAX, BATCH>(D.rad_col[q], buf, M, batch, Ns, lane, TW, g);
    Ns *= D.rad_col[q];
  }
  ip_stage_r<-1, IO_KHAT, TG, NMAX, BATCH>(D.rad_col[ns - 1], buf, M, batch, Ns, lane, TW, g);
  Ns = 1;
  for (int q = 0; q < ns - 1; ++q) {
    ip_stage_r<+1, IO_LDS, TG, NMAX, BATCH>(D.rad_col[q], buf, M, batch, Ns, lane, TW, g);
    Ns *= D.rad_col[q];
  }
  ip_stage_r<+1, IO_OUT, TG, NMAX, BATCH>(D.rad_col[ns - 1], buf, M, batch, Ns, lane, TW, g);
}

// Whether a transform can use the global-I/O stages.
__device__ __forceinline__ bool io_fast(int nst, const int *rad) {
  bool f = nst >= 2;
  for (int q = 0; q < nst; ++q) f = f && (rad[q] == 2 || rad[q] == 3 || rad[q] == 4 || rad[q] == 8);
  return f;
}

// Middle stages 1 .. last-1 (LDS -> LDS) of a transform whose first stage
// wrote x; returns the buffer holding the result of stage `last - 1`.
template <int SIGN>
__device__ __forceinline__ double2 *mid_stages(double2 *x, double2 *y, int n, int batch, int nst,
                               const int *rad, const double2 *tw, const StageIO &g, int *Ns) {
  for (int q = 1; q < nst - 1; ++q) {
    io_stage_r<SIGN, IO_LDS>(rad[q], x, y, n, batch, *Ns, tw, g);
    __syncthreads();
    double2 *t = x; x = y; y = t;
    *Ns *= rad[q];
  }
  return x;
}

// ------------------------------------------------------------------ reductions
// v of the lane that the DPP control CTRL names within the lane's row of 16
// (quad permutations, row mirrors), each 32-bit half moved on its own.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
// Wavefront reduction of (max, 4 sums) in a fixed order: within each row of
// 16 lanes, neighbours (quad_perm 1,0,3,2), pairs of them (quad_perm
// 2,3,0,1), quads (half-row mirror), halves (row mirror) -- every lane of a
// row then holds the row's value, the two operands of each step being the
// same whichever lane forms it -- then rows 0 + 1 + 2 + 3.  Cross-lane moves
// by DPP and readlane, not through LDS.
__device__ __forceinline__ void wave_reduce5(double &mx, double &s0, double &s1, double &s2, double &s3) {
  auto step = [&](auto ctrl) {
    constexpr int C = decltype(ctrl)::value;
    const double o = dpp_row<C>(mx);
    mx = (o > mx) ? o : mx;
    s0 += dpp_row<C>(s0);
    s1 += dpp_row<C>(s1);
    s2 += dpp_row<C>(s2);
    s3 += dpp_row<C>(s3);
  };
  step(std::integral_constant<int, 0xb1>{});    // quad_perm [1, 0, 3, 2]
  step(std::integral_constant<int, 0x4e>{});    // quad_perm [2, 3, 0, 1]
  step(std::integral_constant<int, 0x141>{});   // row_half_mirror
  step(std::integral_constant<int, 0x140>{});   // row_mirror
  auto rows = [&](double &v, bool is_max) {
    const double r0 = lane_value(v, 0), r1 = lane_value(v, 16), r2 = lane_value(v, 32),
                 r3 = lane_value(v, 48);
    if (is_max) {
      const double a = (r1 > r0) ? r1 : r0, b = (r3 > r2) ? r3 : r2;
      v = (b > a) ? b : a;
    } else {
      v = (r0 + r1) + (r2 + r3);
    }
  };
  rows(mx, true);
  rows(s0, false);
  rows(s1, false);
  rows(s2, false);
  rows(s3, false);
}

// Workgroup reduction of (max, 4 sums) in a fixed order (deterministic).
// All threads of the workgroup must call it (full wavefronts).
__device__ void wg_reduce5(double &mx, double &s0, double &s1, double &s2, double &s3,
                           double *red) {
  wave_reduce5(mx, s0, s1, s2, s3);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[w * 5 + 0] = mx; red[w * 5 + 1] = s0; red[w * 5 + 2] = s1;
    red[w * 5 + 3] = s2; red[w * 5 + 4] = s3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    mx = red[0]; s0 = red[1]; s1 = red[2]; s2 = red[3]; s3 = red[4];
    for (int q = 1; q < nw; ++q) {
      double o = red[q * 5];
      mx = (o > mx) ? o : mx;
      s0 += red[q * 5 + 1]; s1 += red[q * 5 + 2]; s2 += red[q * 5 + 3]; s3 += red[q * 5 + 4];
    }
  }
}

// Running maximum of the CFL rate (|vR| + cs)/dR + (|vphi| + cs)/(R dphi)
// of hydro.hpp:52-57, with the rate of the maximum formed exactly as the
// reference forms it: the same sqrt, two IEEE divisions and a sum (dR and
// R dphi from tables holding the reference's own products), so that dt =
// cfl / max is the reference's to the bit for the same state (the max is
// order-free).  The two divisions (about 25 instructions) are taken only
// for zones that can raise the maximum: the rate is first estimated with
// the reciprocal tables, within a few ulp of the exact one, and a zone
// whose exact rate exceeds mx has an estimate >= mx (1 - 1e-14).  NaN rates
// never win, as in the reference's `sR + sp > dtinv`.
template <bool G2 = false>
__device__ __forceinline__ void cfl_max(const DiskDev &D, double &mx, double s, double u, double w,
                                        double invdR, double invdl, double dR, double dl) {
  const double cs = sqrt(cs2f<G2>(D, smax(s, 1e-12)));
  const double au = fabs(u) + cs, aw = fabs(w) + cs;
  if (au * invdR + aw * invdl >= mx * (1.0 - 1e-14)) {
    const double r = au / dR + aw / dl;
    mx = (r > mx) ? r : mx;
  }
}

// a / b correctly rounded, given ib = 1/b correctly rounded (the host's IEEE
// 1.0 / b in the tables): q0 = a ib lies within an ulp of a / b, the
// remainder a - q0 b is exact in one FMA, and q0 + r ib rounded is a / b
// rounded (Markstein's theorem for binary64 with a fused multiply-add;
// checked against IEEE division on the GPU for 16.7M random operands, and
// the dt tests check it end to end).
__device__ __forceinline__ double div_rn(double a, double b, double ib) {
  const double q0 = a * ib;
  return fma(fma(-q0, b, a), ib, q0);
}

// The CFL rate (|vR| + cs)/dR + (|vphi| + cs)/(R dphi) of one zone formed
// exactly as hydro.hpp:52-57 forms it (the same sqrt, two correctly rounded
// quotients, one sum), with no branch.  A zone with an infinite speed gives a
// NaN remainder and rate, which the running maximum ignores (the reference's
// quotient is inf there and its dt 0); such a zone raises the non-finite
// flag.
template <bool G2 = false>
__device__ __forceinline__ double cfl_rate(const DiskDev &D, double s, double u, double w,
                                           double invdR, double invdl, double dR, double dl) {
  const double cs = sqrt(cs2f<G2>(D, smax(s, 1e-12)));
  return div_rn(fabs(u) + cs, dR, invdR) + div_rn(fabs(w) + cs, dl, invdl);
}

__device__ __forceinline__ double nonfinite(double s, double u, double w) {
  return (isfinite(s) && isfinite(u) && isfinite(w)) ? 0.0 : 1.0;
}

__device__ __forceinline__ const double *stp(const double *st, i64 ntot, int buf, int f,
                                             const DiskDev &D) {
  return st + (i64)(buf * 3 + f) * ntot + D.so;
}
__device__ __forceinline__ double *stpw(double *st, i64 ntot, int buf, int f, const DiskDev &D) {
  return st + (i64)(buf * 3 + f) * ntot + D.so;
}

// ------------------------------------------------------------------ control
// Top of one iteration of diskmain.cpp:92-101 for one disk, run by one whole
// workgroup: reduce the per-tile partials (CFL maximum, C_m sums) that the
// previous update (or the initial diag) wrote for the state in buffer `par`;
// stop at t >= tlim or at the step cap; record an output sample once t has
// reached tnext; dt = cfl / max, clipped to land on tlim; t += dt.  It
// touches only the control block and the sample ring.  c: the block as read
// at the start (its loads in flight with the partials'); C: where thread 0
// writes the changes.  A disk already done reduces and changes nothing.
__device__ void control_disk(const DiskDev &D, const DiskCtl &c, DiskCtl &C, int d,
                             const double *__restrict__ pcfl, const double *__restrict__ pcm,
                             DfkSampleDev *__restrict__ samples, int cap, int par) {
  __shared__ double red[5 * 16];
  const int nt = D.ntile_i * D.ntile_j, B = blockDim.x;
  double mx = 1e-30, s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  // thread x takes tiles x, x + B, x + 2B, ... in that order; the loads of 8
  // of them are issued at once (clamped indices, the values past the end
  // dropped), so a workgroup of few threads (64 in the register row launch)
  // does not pay one memory latency per tile
  constexpr int KF = 8;
  for (int t0 = threadIdx.x; t0 < nt; t0 += KF * B) {
    double v[KF];
    double4 c4[KF];
#pragma unroll
    for (int q = 0; q < KF; ++q) {
      const i64 t = D.part + min(t0 + q * B, nt - 1);
      v[q] = pcfl[t];
      c4[q] = *reinterpret_cast<const double4 *>(pcm + 4 * t);
    }
#pragma unroll
    for (int q = 0; q < KF; ++q) {
      if (t0 + q * B < nt) {
        mx = (v[q] > mx) ? v[q] : mx;
        s0 += c4[q].x;
        s1 += c4[q].y;
        s2 += c4[q].z;
        s3 += c4[q].w;
      }
    }
  }
  wg_reduce5(mx, s0, s1, s2, s3, red);
  if (threadIdx.x != 0 || c.done) return;
  if (s3 > 0.0) {
    C.flags = c.flags | DFK_FLAG_NONFINITE;
    if (c.stop_on_nonfinite) {
      C.done = 1;
      C.cur = par;
      C.fin_re2 = s1;
      C.fin_im2 = s2;
      return;
    }
  }
  const double t = c.t;
  if (!(t < c.tlim) || (c.max_steps > 0 && c.step >= c.max_steps)) {
    C.done = 1;
    C.cur = par;
    C.fin_re2 = s1;
    C.fin_im2 = s2;
    return;
  }
  if (t >= c.tnext - 1e-12) {
    DfkSampleDev smp;
    smp.t = t; smp.re0 = s0; smp.re2 = s1; smp.im2 = s2;
    samples[(i64)d * cap + (c.nsamples % cap)] = smp;
    C.nsamples = c.nsamples + 1;
    C.tnext = c.tnext + c.dtout;
  }
  double dt = c.cfl / mx;
  if (t + dt > c.tlim) dt = c.tlim - t;
  C.dt = dt;
  C.t = t + dt;
  C.step = c.step + 1;
  C.cur = 1 - par;   // where this iteration's update writes
}

// The loop control of disk d run by a trailing workgroup of another launch
// (row r2c, or column convolution with fused row kernels): that launch's
// other workgroups need none of its results.  Workgroups of a disk that it
// stops read `done` racily; they then only transform a spectrum that no
// later kernel of the disk reads.  mirror (null: none): a host-visible copy
// of the control blocks (coherent pinned host memory) that the control of
// a captured run's last iteration also writes, so that the host reads the
// loop's state after the run without a device-to-host copy.  Nothing later
// in that iteration writes the control block.
__device__ __forceinline__ void trailing_control(const DiskDev *__restrict__ disks, DiskCtl *__restrict__ ctl,
                                                 int d, const double *__restrict__ pcfl,
                                                 const double *__restrict__ pcm,
                                                 DfkSampleDev *__restrict__ samples, int cap, int par,
                                                 DiskCtl *__restrict__ mirror) {
  const DiskCtl c = ctl[d];
  control_disk(disks[d], c, ctl[d], d, pcfl, pcm, samples, cap, par);
  // thread 0 made every change to the block (the others left control_disk
  // after the reduction)
  if (mirror && threadIdx.x == 0) {
    mirror[d] = ctl[d];
    __threadfence_system();
  }
}

// Control of every disk as a launch of its own (one workgroup per disk).  In
// the time loop the same work rides in the row-transform launch instead
// (trailing workgroups of k_rows_r2c, which needs nothing of it).
__global__ void __launch_bounds__(WG)
k_control(const DiskDev *__restrict__ disks, DiskCtl *__restrict__ ctl,
          const double *__restrict__ pcfl, const double *__restrict__ pcm,
          DfkSampleDev *__restrict__ samples, int cap, int par, DiskCtl *__restrict__ mirror) {
  trailing_control(disks, ctl, blockIdx.x, pcfl, pcm, samples, cap, par, mirror);
}

// ------------------------------------------------------------------ diag
// Per-tile CFL maximum and C_m sums of the current state (used before the
// first step and by the timestep entry point).
__global__ void __launch_bounds__(WG)
k_diag(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
       const DiskCtl *__restrict__ ctl, const double *__restrict__ st, i64 ntot,
       const double *__restrict__ geom, const double *__restrict__ trig,
       double *__restrict__ pcfl, double *__restrict__ pcm) {
  __shared__ double red[5 * (WG / 64)];
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  const int cur = ctl[wm.x].cur;
  const int Nu = D.Nu, Nphi = D.Nphi;
  const int ti = wm.y / D.ntile_j, tj = wm.y - ti * D.ntile_j;
  const int i0 = ti * D.ti, j0 = tj * D.tj;
  const double *S = stp(st, ntot, cur, 0, D), *U = stp(st, ntot, cur, 1, D),
               *W = stp(st, ntot, cur, 2, D);
  const double *invdR = gtab(geom, D, GT_INVDR), *invdl = gtab(geom, D, GT_INVDL),
               *dR = gtab(geom, D, GT_AFACE), *dl = gtab(geom, D, GT_DL), *dA = gtab(geom, D, GT_DA);
  const double *c2 = trig + D.cos_off, *s2 = trig + D.cos_off + Nphi;
  double mx = 1e-30, a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int e = threadIdx.x; e < D.ti * D.tj; e += blockDim.x) {
    const int li = qdiv(e, rcpf((float)D.tj)), lj = e - li * D.tj;
    const int i = i0 + li, j = j0 + lj;
    if (i >= Nu || j >= Nphi) continue;
    const i64 c = (i64)i * Nphi + j;
    const double s = S[c], u = U[c], w = W[c];
    cfl_max(D, mx, s, u, w, invdR[i], invdl[i], dR[i], dl[i]);
    const double sa = s * dA[i];
    a0 += sa;
    a1 += s * c2[j] * dA[i];
    a2 -= s * s2[j] * dA[i];
    a3 += nonfinite(s, u, w);
  }
  wg_reduce5(mx, a0, a1, a2, a3, red);
  if (threadIdx.x == 0) {
    const i64 t = D.part + wm.y;
    pcfl[t] = mx;
    pcm[4 * t + 0] = a0; pcm[4 * t + 1] = a1; pcm[4 * t + 2] = a2; pcm[4 * t + 3] = a3;
  }
}

// ------------------------------------------------------------------ gravity
// Row transform I/O shared by the row kernels.  r2c_store: the spectra of
// `npair` transformed row pairs in Z (pair p = rows row0 + 2p, row0 + 2p + 1
// packed as real and imaginary parts) split into the two rows' half spectra
// and stored to spec[k][row] (rows < Nu).  c2r_load: the Hermitian rebuild of
// the same packing from spec rows row0 .. row0 + 2 npair - 1 (rows >= Nu read
// as zero; FFTW c2r convention: the imaginary parts of the DC and Nyquist
// bins are ignored).
__device__ __forceinline__ void r2c_store(const double2 *Z, double2 *__restrict__ out, int row0,
                                          int npair, int n, int Nc, int Nu) {
  const int w2 = 2 * npair;
  for (int e = threadIdx.x; e < w2 * Nc; e += blockDim.x) {
    const int k = qdiv(e, rcpf((float)w2)), rr = e - k * w2;
    const int p = rr >> 1, isb = rr & 1;
    const int row = row0 + rr;
    if (row >= Nu) continue;
    const double2 z = Z[p * n + k];
    double2 y = Z[p * n + (k == 0 ? 0 : n - k)];
    y.y = -y.y;
    double2 X;
    if (!isb) X = make_double2(0.5 * (z.x + y.x), 0.5 * (z.y + y.y));
    else X = make_double2(0.5 * (z.y - y.y), -0.5 * (z.x - y.x));
    out[(i64)k * Nu + row] = X;
  }
}

// Bin k of row pair p (rows a, b with spectra xa, xb) into the packed c2r input.
__device__ __forceinline__ void c2r_put(double2 *A, int p, int k, double2 xa, double2 xb, int n, int Nc) {
  if (k == 0 || 2 * k == n) { xa.y = 0.0; xb.y = 0.0; }
  A[p * n + k] = make_double2(xa.x - xb.y, xa.y + xb.x);
  if (k > 0 && n - k >= Nc) A[p * n + (n - k)] = make_double2(xa.x + xb.y, -xa.y + xb.x);
}

struct Bin2 { double2 a, b; };   // one bin of a row pair's two spectra

// BATCHED: the loads of several passes in flight at once (staged_for; for
// workgroups with several row pairs).  Otherwise one pass after another
// (a workgroup of one row pair has a single pass, and the shorter code
// starts faster).
template <bool BATCHED = true>
__device__ __forceinline__ void c2r_load(const double2 *__restrict__ in, double2 *A, int row0,
                                         int npair, int n, int Nc, int Nu) {
  if constexpr (!BATCHED) {
    for (int e = threadIdx.x; e < npair * Nc; e += blockDim.x) {
      const int k = qdiv(e, rcpf((float)npair)), p = e - k * npair;
      const int ra = row0 + 2 * p, rb = ra + 1;
      double2 xa = make_double2(0.0, 0.0), xb = make_double2(0.0, 0.0);
      if (ra < Nu) xa = in[(i64)k * Nu + ra];
      if (rb < Nu) xb = in[(i64)k * Nu + rb];
      c2r_put(A, p, k, xa, xb, n, Nc);
    }
    return;
  }
  const float rp = rcpf((float)npair);
  staged_for<4, Bin2>(
      npair * Nc,
      [&](int e) {
        const int k = qdiv(e, rp), p = e - k * npair;
        const int ra = row0 + 2 * p, rb = ra + 1;
        const double2 xa = in[(i64)k * Nu + min(ra, Nu - 1)], xb = in[(i64)k * Nu + min(rb, Nu - 1)];
        const double2 z = make_double2(0.0, 0.0);
        return Bin2{ra < Nu ? xa : z, rb < Nu ? xb : z};
      },
      [&](int e, Bin2 x) {
        const int k = qdiv(e, rp);
        c2r_put(A, e - k * npair, k, x.a, x.b, n, Nc);
      });
}

// Doubles of the tile part of k_rows_fused's LDS (k_update's layout for a
// TI x n tile), rounded up to whole double2; and of its staged state rows
// (Sigma on rows i0-3 .. i0+TI+2, v_R and v_phi on i0-2 .. i0+TI+1).
__host__ __device__ __forceinline__ int fused_tile_doubles(int TI, int n) {
  return (3 * (TI + 4) * (n + 4) + 3 * (TI + 1) * (n + 4) + (TI + 4) * RR_STRIDE + 2 * n + 1) & ~1;
}
__host__ __device__ __forceinline__ int fused_stage_doubles(int TI, int n) { return (3 * TI + 14) * n; }

// Row r2c: S = R^{3/2} Sigma for rows [2 pr g, 2 pr (g+1)) of the current
// state.  TG = 0: the workgroup-cooperative transform (ping-pong buffers);
// TG > 0: in place by the TG threads of the workgroup (one buffer of pr rows,
// pr n <= ip_nmax<TG>(), radices 2/3/4/8; the runtime enforces both).
template <int TG, bool BATCH = true>
__device__ __forceinline__ void rows_r2c_body(const DiskDev *__restrict__ disks, const int2 *__restrict__ map,
                                              int map_base, DiskCtl *__restrict__ ctl,
                                              const double *__restrict__ st, i64 ntot,
                                              const double *__restrict__ geom, const double2 *__restrict__ tw,
                                              double2 *__restrict__ spec, int par, int nrow_blk, int ctl0,
                                              const double *__restrict__ pcfl, const double *__restrict__ pcm,
                                              DfkSampleDev *__restrict__ samples, int cap,
                                              DiskCtl *__restrict__ mirror) {
  extern __shared__ double2 lds2[];
  if ((int)blockIdx.x >= nrow_blk) {   // the loop control of disk ctl0 + blockIdx.x - nrow_blk
    trailing_control(disks, ctl, ctl0 + blockIdx.x - nrow_blk, pcfl, pcm, samples, cap, par, mirror);
    return;
  }
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  const DiskCtl &C = ctl[wm.x];
  if (C.done) return;
  const int n = D.Nphi, Nu = D.Nu, pr = D.pr, Nc = D.Nc;
  const int row0 = 2 * pr * wm.y;
  constexpr int NBUF = TG ? 1 : 2;
  double2 *A = lds2, *B = lds2 + pr * n;
  const double2 *TW = tw_table(lds2 + NBUF * pr * n, tw + D.two_row, n, D.tw_lds & 1);
  const double *S = stp(st, ntot, par, 0, D);
  const double *R32 = gtab(geom, D, GT_R32);
  // all threads stage the rows (more loads in flight than a first stage
  // reading HBM through its few butterflies: measured faster here); with
  // several row pairs the loads of up to 8 passes at once
  const float rn = rcpf((float)n);
  if constexpr (!BATCH) {
    for (int e = threadIdx.x; e < pr * n; e += blockDim.x) {
      const int p = qdiv(e, rn), j = e - p * n;
      const int ra = row0 + 2 * p, rb = ra + 1;
      double xa = 0.0, xb = 0.0;
      if (ra < Nu) xa = R32[ra] * S[(i64)ra * n + j];
      if (rb < Nu) xb = R32[rb] * S[(i64)rb * n + j];
      A[e] = make_double2(xa, xb);
    }
  } else staged_for<8, double2>(
      pr * n,
      [&](int e) {
        const int p = qdiv(e, rn), j = e - p * n;
        const int ra = row0 + 2 * p, rb = ra + 1;
        const int ca = min(ra, Nu - 1), cb_ = min(rb, Nu - 1);
        const double xa = R32[ca] * S[(i64)ca * n + j], xb = R32[cb_] * S[(i64)cb_ * n + j];
        return make_double2(ra < Nu ? xa : 0.0, rb < Nu ? xb : 0.0);
      },
      [&](int e, double2 v) { A[e] = v; });
  __syncthreads();
  const double2 *Z = A;
  if constexpr (TG == 0) {
    Z = lds_fft<-1>(A, B, n, pr, D.nst_row, D.rad_row, TW);
  } else {
    StageIO g{};
    int Ns = 1;
    for (int q = 0; q < D.nst_row; ++q) {
      ip_stage_r<-1, IO_LDS, TG, ip_nmax<TG>(), BATCH>(D.rad_row[q], A, n, pr, Ns, threadIdx.x, TW, g);
      Ns *= D.rad_row[q];
    }
  }
  r2c_store(Z, spec + D.spo, row0, pr, n, Nc, Nu);
}

#define R2C_PARAMS                                                                              \
  const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,                \
      DiskCtl *__restrict__ ctl, const double *__restrict__ st, i64 ntot,                       \
      const double *__restrict__ geom, const double2 *__restrict__ tw,                          \
      double2 *__restrict__ spec, int par, int nrow_blk, int ctl0,                              \
      const double *__restrict__ pcfl, const double *__restrict__ pcm,                          \
      DfkSampleDev *__restrict__ samples, int cap, DiskCtl *__restrict__ mirror
#define R2C_ARGS \
  disks, map, map_base, ctl, st, ntot, geom, tw, spec, par, nrow_blk, ctl0, pcfl, pcm, samples, cap, mirror

__global__ void __launch_bounds__(WG) k_rows_r2c(R2C_PARAMS) { rows_r2c_body<0>(R2C_ARGS); }

template <int TG, bool BATCH>
__global__ void __launch_bounds__(TG) k_rows_r2c_group(R2C_PARAMS) { rows_r2c_body<TG, BATCH>(R2C_ARGS); }

// Column convolution: for columns [cb g, cb (g+1)): zero-padded forward
// transform of length M, multiply by real K-hat, inverse, keep rows < Nu.
// Loop control in trailing workgroups of a column-convolution launch (the
// time loop with fused row kernels): workgroups nconv .. run the control of
// disks ctl0, ctl0 + 1, ...
#define CONV_CTL_PARAMS                                                                         \
  int nconv, int ctl0, const double *__restrict__ pcfl, const double *__restrict__ pcm,         \
      DfkSampleDev *__restrict__ samples, int cap, int par, DiskCtl *__restrict__ mirror
#define CONV_CONTROL                                                                            \
  if ((int)blockIdx.x >= nconv) {                                                               \
    trailing_control(disks, ctl, ctl0 + blockIdx.x - nconv, pcfl, pcm, samples, cap, par,      \
                     mirror);                                                                   \
    return;                                                                                     \
  }

__global__ void __launch_bounds__(WG)
k_cols_conv(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
            DiskCtl *__restrict__ ctl, const double *__restrict__ khat,
            const double2 *__restrict__ tw, double2 *__restrict__ spec, CONV_CTL_PARAMS) {
  extern __shared__ double2 lds2[];
  CONV_CONTROL
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const int M = D.M, Nu = D.Nu, Nc = D.Nc, cb = D.cb;
  const int k0 = wm.y * cb;
  double2 *A = lds2, *B = lds2 + cb * M;
  const double2 *TW = tw_table(lds2 + 2 * cb * M, tw + D.two_col, M, D.tw_lds & 2);
  double2 *io = spec + D.spo;
  const double *kh = khat + D.kho;
  if (io_fast(D.nst_col, D.rad_col)) {
    const int ns = D.nst_col;
    StageIO g{};
    g.spec = io; g.kh = kh; g.Nu = Nu; g.Nc = Nc; g.k0 = k0;
    // forward: spectrum column -> A (first stage), ..., last stage x K-hat
    io_stage_r<-1, IO_ZP>(D.rad_col[0], nullptr, A, M, cb, 1, TW, g);
    __syncthreads();
    int Ns = D.rad_col[0];
    double2 *x = mid_stages<-1>(A, B, M, cb, ns, D.rad_col, TW, g, &Ns);
    double2 *y = (x == A) ? B : A;
    io_stage_r<-1, IO_KHAT>(D.rad_col[ns - 1], x, y, M, cb, Ns, TW, g);
    __syncthreads();
    // inverse: stages 0 .. ns-2 in LDS, the last one -> rows < Nu of the column
    x = y;
    y = (x == A) ? B : A;
    Ns = 1;
    for (int q = 0; q < ns - 1; ++q) {
      io_stage_r<+1, IO_LDS>(D.rad_col[q], x, y, M, cb, Ns, TW, g);
      __syncthreads();
      double2 *t = x; x = y; y = t;
      Ns *= D.rad_col[q];
    }
    io_stage_r<+1, IO_OUT>(D.rad_col[ns - 1], x, y, M, cb, Ns, TW, g);
    return;
  }
  for (int e = threadIdx.x; e < cb * M; e += blockDim.x) {
    const int c = qdiv(e, rcpf((float)M)), p = e - c * M;
    const int k = k0 + c;
    double2 v = make_double2(0.0, 0.0);
    if (k < Nc && p < Nu) v = io[(i64)k * Nu + p];
    A[e] = v;
  }
  __syncthreads();
  double2 *Z = lds_fft<-1>(A, B, M, cb, D.nst_col, D.rad_col, TW);
  double2 *Wb = (Z == A) ? B : A;
  for (int e = threadIdx.x; e < cb * M; e += blockDim.x) {
    const int c = qdiv(e, rcpf((float)M)), p = e - c * M;
    const int k = k0 + c;
    const int pp = (p <= Nu) ? p : M - p;
    const double kv = (k < Nc) ? kh[(i64)k * (Nu + 1) + pp] : 0.0;
    double2 z = Z[e];
    z.x *= kv;
    z.y *= kv;
    Z[e] = z;
  }
  __syncthreads();
  const double2 *Y = lds_fft<+1>(Z, Wb, M, cb, D.nst_col, D.rad_col, TW);
  for (int e = threadIdx.x; e < cb * Nu; e += blockDim.x) {
    const int c = qdiv(e, rcpf((float)Nu)), p = e - c * Nu;
    const int k = k0 + c;
    if (k < Nc) io[(i64)k * Nu + p] = Y[c * M + p];
  }
}

// Column convolution with one wavefront per column, in place (the disks of a
// handle whose column transforms the runtime marked wave-local: radices
// 2/3/4/8, >= 2 stages, M <= 1024).  LDS: cb columns of M elements, then the
// twiddle table if staged.  A kernel of its own: the wave stages hold up to
// 16 complex values per lane, which would otherwise set the register budget
// (and so the occupancy) of the workgroup-cooperative kernel too.
__global__ void __launch_bounds__(WG)
k_cols_conv_wave(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
                 DiskCtl *__restrict__ ctl, const double *__restrict__ khat,
                 const double2 *__restrict__ tw, double2 *__restrict__ spec, CONV_CTL_PARAMS) {
  extern __shared__ double2 lds2[];
  CONV_CONTROL
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const int M = D.M, Nu = D.Nu, Nc = D.Nc, cb = D.cb;
  const int k0 = wm.y * cb;
  double2 *A = lds2;
  const double2 *TW = tw_table(lds2 + cb * M, tw + D.two_col, M, D.tw_lds & 2);
  StageIO g{};
  g.spec = spec + D.spo; g.kh = khat + D.kho; g.Nu = Nu; g.Nc = Nc; g.k0 = k0;
  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
  __syncthreads();   // the twiddle table, if staged in LDS
  for (int p = w; p < cb; p += nw) {   // wave w: columns w, w + nw, ...
    StageIO gp = g;
    gp.k0 = k0 + p;
    ip_conv<64, 1024, false>(D, A + p * M, 1, threadIdx.x & 63, TW, gp);
  }
}

// Column convolution in place by the whole workgroup (TG = blockDim.x
// threads), its cb columns at once in cb LDS buffers of M elements (then the
// twiddle table if staged; cb M <= ip_nmax<TG>(), which the runtime
// enforces): half the LDS of the cooperative kernel per column, so more
// columns fit on a CU at once.  Same radix limits as the wave-local kernel.
// BATCH = false: one column per workgroup (cb = 1 for every disk of the
// launch), whose addresses stay wave-uniform; in the latency regime that
// build is ~20% faster than the batched one.
template <int TG, bool BATCH>
__global__ void __launch_bounds__(TG)
k_cols_conv_group(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
                  DiskCtl *__restrict__ ctl, const double *__restrict__ khat,
                  const double2 *__restrict__ tw, double2 *__restrict__ spec, CONV_CTL_PARAMS) {
  extern __shared__ double2 lds2[];
  CONV_CONTROL
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const int M = D.M, Nc = D.Nc, cb = D.cb;
  const int k0 = wm.y * cb;
  StageIO g{};
  g.spec = spec + D.spo; g.kh = khat + D.kho; g.Nu = D.Nu; g.Nc = Nc; g.k0 = k0;
  const int nb = BATCH ? min(cb, Nc - k0) : 1;
  // one call per twiddle location, so that each inlined transform reads its
  // twiddles through a pointer of known address space (LDS or global loads;
  // through one pointer that may be either, every twiddle read is a flat
  // load, which also waits for the workgroup's outstanding LDS accesses)
  if (D.tw_lds & 2) {
    double2 *T = lds2 + cb * M;
    stage_twiddles(T, tw + D.two_col, M);
    __syncthreads();
    ip_conv<TG, ip_nmax<TG>(), BATCH>(D, lds2, nb, threadIdx.x, T, g);
  } else {
    ip_conv<TG, ip_nmax<TG>(), BATCH>(D, lds2, nb, threadIdx.x, tw + D.two_col, g);
  }
}

// Column convolution by the two-pass register transform of dfk_regfft.h
// (the disks whose runtime marked their columns mode 3): M = N1 N2 = 2 Nu
// fixed at compile time, one 64-lane wavefront per workgroup and N2 lanes
// per column (64 / N2 columns per workgroup), so every synchronisation is a
// wavefront one.  Per column: pass 1 reads rows < Nu (for a fixed n1, the
// lanes read consecutive rows), one LDS transpose per direction in H parts,
// pass 3 writes rows < Nu.  Index arithmetic is compile-time but for the
// column.
template <int N1, int N2, int H>
__device__ __forceinline__ void conv_reg_body(const DiskDev &D, int grp, double2 *lds2,
                                              const double *__restrict__ khat,
                                              const double2 *__restrict__ tw,
                                              double2 *__restrict__ spec) {
  constexpr int Nu = N1 * N2 / 2, CPW = 64 / N2, LE = rfft::conv::lds_elems_h<N1, N2, H>();
  const int c = threadIdx.x / N2, t = threadIdx.x % N2;
  const int col = grp * CPW + c;
  const bool live = col < D.Nc;
  const int cc = live ? col : D.Nc - 1;   // columns past the spectrum read the last one, store nothing
  double2 *L = lds2 + c * LE;
  const double2 *TW = tw + D.two_col;
  double2 *io = spec + D.spo + (i64)cc * Nu;
  const double *kh = khat + D.kho + (i64)cc * (Nu + 1);
  double2 v[N1];
  rfft::conv::pass1<N1, N2>(t, [&](int r) {
    const double2 q = io[r];
    return live ? q : make_double2(0.0, 0.0);
  }, TW, v);
  double2 x[N2];
  rfft::sfor<0, H>([&](auto h_) {
    constexpr int hh = decltype(h_)::value;
    rfft::conv::fwd_write<N1, N2, H>(t, hh, v, L);
    wave_lds_sync();
    if (t < N1) rfft::conv::fwd_read_c<N1, N2, H, hh>(t, L, x);
    wave_lds_sync();
  });
  if (t < N1) {
    double kv[N2];
    rfft::conv::pass2_khat<N1, N2>(t, [&](int q) { return kh[q]; }, kv);
    rfft::conv::pass2<N1, N2>(t, x, kv, TW);
  }
  double2 z[N1];
  rfft::sfor<0, H>([&](auto h_) {
    constexpr int hh = decltype(h_)::value;
    rfft::conv::inv_write<N1, N2, H>(t, hh, x, L);
    wave_lds_sync();
    rfft::conv::inv_read_c<N1, N2, H, hh>(t, L, z);
    if (hh + 1 < H) wave_lds_sync();
  });
  rfft::conv::pass3<N1, N2>(t, z, [&](int r, double2 q) {
    if (live) io[r] = q;
  });
}

// The three-pass form (rfft::conv3: M = 64 P, P = 4, 6, 8, 12, 16), one
// column per 64-thread workgroup (two k1 groups per lane for P > 8).  The
// K-hat values are loaded with the column, before pass A, and so are the
// later passes' twiddles for P <= 8, so that no pass waits on a load after a
// transpose (256^2 disk: 3.86 -> 3.53 us); with two k1 groups the passes
// load their own (preloading measured 7.2 -> 8.0 us at 2Nu = 1024).
template <int P>
__device__ __forceinline__ void conv3_body(const DiskDev &D, int col, double2 *L,
                                           const double *__restrict__ khat,
                                           const double2 *__restrict__ tw,
                                           double2 *__restrict__ spec) {
  constexpr int Nu = 32 * P, G = rfft::conv3::groups<P>();
  const int l = threadIdx.x, r = l & 7;
  int k1[G];
#pragma unroll
  for (int g = 0; g < G; ++g) k1[g] = rfft::conv3::lane_k1<P>(l, g);
  const bool live = col < D.Nc;
  const int cc = live ? col : D.Nc - 1;   // a column past the spectrum reads the last one, stores nothing
  const double2 *TW = tw + D.two_col;
  double2 *io = spec + D.spo + (i64)cc * Nu;
  const double *kh = khat + D.kho + (i64)cc * (Nu + 1);
  double kv[G][8];
#pragma unroll
  for (int g = 0; g < G; ++g) rfft::conv3::khat_vals<P>(k1[g], r, [&](int p) { return kh[p]; }, kv[g]);
  constexpr bool PRE = G == 1;
  // passes B and C: W_64^(-+ r k); pass B' (preloaded for P <= 8)
  const rfft::conv::TwRow<8, -1> wB = PRE ? rfft::conv::TwRow<8, -1>(TW, P * r) : rfft::conv::TwRow<8, -1>();
  const rfft::conv::TwRow<8, +1> wC = PRE ? rfft::conv::TwRow<8, +1>(TW, P * r) : rfft::conv::TwRow<8, +1>();
  double2 wI[G][8];
  if constexpr (PRE) rfft::conv3::binv_tw(k1[0], r, TW, wI[0]);
  double2 v[P], x[G][8];
  rfft::conv3::passA<P>(l, [&](int row) { return io[row]; }, TW, v);
  rfft::conv3::t1_write<P>(l, v, L);
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < G; ++g) rfft::conv3::t1_read(k1[g], r, L, x[g]);
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if constexpr (PRE) rfft::conv3::passB<P>(wB, x[g]);
    else rfft::conv3::passB<P>(r, TW, x[g]);
    rfft::conv3::t2_write(k1[g], r, x[g], L);
  }
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < G; ++g) rfft::conv3::t2_read(8 * k1[g] + r, L, x[g]);
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if constexpr (PRE) rfft::conv3::passC<P>(kv[g], wC, x[g]);
    else rfft::conv3::passC<P>(r, kv[g], TW, x[g]);
    rfft::conv3::t2_write(k1[g], r, x[g], L);
  }
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < G; ++g) rfft::conv3::t2_read(8 * k1[g] + r, L, x[g]);
  wave_lds_sync();
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if constexpr (PRE) rfft::conv3::passB_inv(wI[g], x[g]);
    else rfft::conv3::passB_inv(k1[g], r, TW, x[g]);
    rfft::conv3::t4_write<P>(k1[g], r, x[g], L);
  }
  wave_lds_sync();
  rfft::conv3::t4_read<P>(l, L, v);
  rfft::conv3::passA_inv<P>(l, v, [&](int row, double2 q) {
    if (live) io[row] = q;
  });
}

// One launch for every disk of the mode, whatever its transform length: the
// workgroup branches once on its disk's M (workgroups are dispatched in map
// order, disk by disk, so a CU mostly runs one instance at a time); disks in
// column mode 4 take the three-pass form, one column per workgroup.
__global__ void __launch_bounds__(64)
k_cols_conv_reg(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
                DiskCtl *__restrict__ ctl, const double *__restrict__ khat,
                const double2 *__restrict__ tw, double2 *__restrict__ spec, CONV_CTL_PARAMS) {
  extern __shared__ double2 lds2[];
  CONV_CONTROL
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  if (D.conv_wave == 4) {
    switch (D.M) {
      case 256: conv3_body<4>(D, wm.y, lds2, khat, tw, spec); break;
      case 384: conv3_body<6>(D, wm.y, lds2, khat, tw, spec); break;
      case 512: conv3_body<8>(D, wm.y, lds2, khat, tw, spec); break;
      case 768: conv3_body<12>(D, wm.y, lds2, khat, tw, spec); break;
      case 1024: conv3_body<16>(D, wm.y, lds2, khat, tw, spec); break;
      default: break;   // the runtime marks no other length mode 4
    }
    return;
  }
  switch (D.M) {
    case 128: conv_reg_body<8, 16, 1>(D, wm.y, lds2, khat, tw, spec); break;
    case 192: conv_reg_body<12, 16, 1>(D, wm.y, lds2, khat, tw, spec); break;
    case 256: conv_reg_body<16, 16, 1>(D, wm.y, lds2, khat, tw, spec); break;
    case 384: conv_reg_body<12, 32, 2>(D, wm.y, lds2, khat, tw, spec); break;
    case 512: conv_reg_body<16, 32, 2>(D, wm.y, lds2, khat, tw, spec); break;
    case 768: conv_reg_body<24, 32, 2>(D, wm.y, lds2, khat, tw, spec); break;
    case 1024: conv_reg_body<32, 32, 2>(D, wm.y, lds2, khat, tw, spec); break;
    default: break;   // the runtime marks no other length mode 3
  }
}

// Row c2r (FFTW convention: imaginary parts of the DC and Nyquist bins are
// ignored) and Phi = pref * V / sqrt(R).  TG as for the r2c.
template <int TG, bool BATCH = true>
__device__ __forceinline__ void rows_c2r_body(const DiskDev *__restrict__ disks, const int2 *__restrict__ map,
                                              int map_base, const DiskCtl *__restrict__ ctl,
                                              const double2 *__restrict__ spec,
                                              const double *__restrict__ geom, const double2 *__restrict__ tw,
                                              double *__restrict__ phi) {
  extern __shared__ double2 lds2[];
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const int n = D.Nphi, Nu = D.Nu, pr = D.pr, Nc = D.Nc;
  const int row0 = 2 * pr * wm.y;
  constexpr int NBUF = TG ? 1 : 2;
  double2 *A = lds2, *B = lds2 + pr * n;
  const double2 *TW = tw_table(lds2 + NBUF * pr * n, tw + D.two_row, n, D.tw_lds & 1);
  c2r_load<BATCH>(spec + D.spo, A, row0, pr, n, Nc, Nu);
  __syncthreads();
  const double *sc = gtab(geom, D, GT_PHISC);
  double *P = phi + D.so;
  if constexpr (TG > 0) {
    StageIO g{};
    g.rsc = sc; g.P = P; g.Nu = Nu; g.row0 = row0;
    int Ns = 1;
    for (int q = 0; q < D.nst_row - 1; ++q) {
      ip_stage_r<+1, IO_LDS, TG, ip_nmax<TG>(), BATCH>(D.rad_row[q], A, n, pr, Ns, threadIdx.x, TW, g);
      Ns *= D.rad_row[q];
    }
    ip_stage_r<+1, IO_ROUT, TG, ip_nmax<TG>(), BATCH>(D.rad_row[D.nst_row - 1], A, n, pr, Ns, threadIdx.x, TW,
                                                      g);
    return;
  }
  if (io_fast(D.nst_row, D.rad_row)) {
    StageIO g{};
    g.rsc = sc; g.P = P; g.Nu = Nu; g.row0 = row0;
    double2 *x = A, *y = B;
    int Ns = 1;
    for (int q = 0; q < D.nst_row - 1; ++q) {
      io_stage_r<+1, IO_LDS>(D.rad_row[q], x, y, n, pr, Ns, TW, g);
      __syncthreads();
      double2 *t = x; x = y; y = t;
      Ns *= D.rad_row[q];
    }
    io_stage_r<+1, IO_ROUT>(D.rad_row[D.nst_row - 1], x, y, n, pr, Ns, TW, g);
    return;
  }
  const double2 *Z = lds_fft<+1>(A, B, n, pr, D.nst_row, D.rad_row, TW);
  for (int e = threadIdx.x; e < 2 * pr * n; e += blockDim.x) {
    const int rr = qdiv(e, rcpf((float)n)), j = e - rr * n;
    const int row = row0 + rr;
    if (row >= Nu) continue;
    const double2 z = Z[(rr >> 1) * n + j];
    const double v = (rr & 1) ? z.y : z.x;
    P[(i64)row * n + j] = v * sc[row];
  }
}

#define C2R_PARAMS                                                                              \
  const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,                \
      const DiskCtl *__restrict__ ctl, const double2 *__restrict__ spec,                        \
      const double *__restrict__ geom, const double2 *__restrict__ tw, double *__restrict__ phi
#define C2R_ARGS disks, map, map_base, ctl, spec, geom, tw, phi

__global__ void __launch_bounds__(WG) k_rows_c2r(C2R_PARAMS) { rows_c2r_body<0>(C2R_ARGS); }

template <int TG, bool BATCH>
__global__ void __launch_bounds__(TG) k_rows_c2r_group(C2R_PARAMS) { rows_c2r_body<TG, BATCH>(C2R_ARGS); }

// Row transforms by the two-pass register transform of dfk_regfft.h (the
// handles whose runtime set the rows to mode 2): n = Nphi = N1 N2 fixed at
// compile time, one 64-lane wavefront per workgroup, N2 lanes per packed row
// pair (rows 2p, 2p + 1 as real and imaginary parts; 64 / N2 pairs per
// workgroup), every synchronisation a wavefront one.  Pass 1 reads (r2c) or
// writes (c2r) the pair's rows with consecutive lanes on consecutive zones.
// Pairs past the disk's last one recompute that pair and store nothing; the
// odd row of a disk with odd Nu reads as zero.
struct RegPair {
  int ra, rb;        // the pair's rows (rb clamped to ra when it is past the disk)
  bool live, hasb;   // the pair is the disk's; its second row is
};
template <int N2>
__device__ __forceinline__ RegPair reg_pair(int grp, int Nu) {
  constexpr int CPW = 64 / N2;
  RegPair r;
  const int p = grp * CPW + (int)threadIdx.x / N2;
  r.live = 2 * p < Nu;
  r.ra = r.live ? 2 * p : 2 * ((Nu - 1) >> 1);
  r.hasb = r.ra + 1 < Nu;
  r.rb = r.hasb ? r.ra + 1 : r.ra;
  return r;
}

template <int N1, int N2>
__device__ __forceinline__ void r2c_reg_body(const DiskDev &D, int grp, double2 *lds2,
                                             const double *__restrict__ S, const double *__restrict__ R32,
                                             const double2 *__restrict__ tw, double2 *__restrict__ out) {
  constexpr int n = N1 * N2, LE = rfft::rows::lds_elems<N1, N2>();
  const int Nu = D.Nu, t = threadIdx.x % N2;
  const RegPair q = reg_pair<N2>(grp, Nu);
  double2 *L = lds2 + (threadIdx.x / N2) * LE;
  const double pa = R32[q.ra], pb = R32[q.rb];
  const double *Sa = S + (i64)q.ra * n, *Sb = S + (i64)q.rb * n;
  double2 v[N1];
  rfft::rows::pass1<N1, N2, -1>(t, [&](int j) {
    const double b = pb * Sb[j];
    return make_double2(pa * Sa[j], q.hasb ? b : 0.0);
  }, tw, v);
  rfft::conv::fwd_write<N1, N2, 1>(t, 0, v, L);
  wave_lds_sync();
  double2 x[N2];
  if (t < N1) rfft::conv::fwd_read_c<N1, N2, 1, 0>(t, L, x);
  wave_lds_sync();
  if (t < N1) {
    rfft::dft<N2, -1>(x);
    rfft::rows::spill<N1, N2>(t, x, L);
  }
  wave_lds_sync();
  if (t < N1 && q.live)
    rfft::rows::split<N1, N2>(t, x, L, [&](int k, double2 xa, double2 xb) {
      out[(i64)k * Nu + q.ra] = xa;
      if (q.hasb) out[(i64)k * Nu + q.ra + 1] = xb;
    });
}

template <int N1, int N2>
__device__ __forceinline__ void c2r_reg_body(const DiskDev &D, int grp, double2 *lds2,
                                             const double2 *__restrict__ in, const double *__restrict__ sc,
                                             const double2 *__restrict__ tw, double *__restrict__ P) {
  constexpr int n = N1 * N2, LE = rfft::rows::lds_elems<N1, N2>();
  const int Nu = D.Nu, t = threadIdx.x % N2;
  const RegPair q = reg_pair<N2>(grp, Nu);
  double2 *L = lds2 + (threadIdx.x / N2) * LE;
  double2 v[N1];
  rfft::rows::pass1<N1, N2, +1>(t, [&](int m) {
    return rfft::rows::c2r_in<N1, N2>(m, [&](int k, double2 &xa, double2 &xb) {
      xa = in[(i64)k * Nu + q.ra];
      const double2 b = in[(i64)k * Nu + q.rb];
      xb = q.hasb ? b : make_double2(0.0, 0.0);
    });
  }, tw, v);
  rfft::conv::fwd_write<N1, N2, 1>(t, 0, v, L);
  wave_lds_sync();
  double2 x[N2];
  if (t < N1) rfft::conv::fwd_read_c<N1, N2, 1, 0>(t, L, x);
  if (t < N1 && q.live) {
    rfft::dft<N2, +1>(x);
    const double sa = sc[q.ra], sb = sc[q.rb];
    double *Pa = P + (i64)q.ra * n, *Pb = P + (i64)q.rb * n;
    rfft::sfor<0, N2>([&](auto k2_) {
      constexpr int k2 = decltype(k2_)::value;
      Pa[t + N1 * k2] = x[k2].x * sa;
      if (q.hasb) Pb[t + N1 * k2] = x[k2].y * sb;
    });
  }
}

// One launch for every disk of the handle, whatever its Nphi (NMAX = 512:
// the builds for rows up to 512, whose registers do not pay for the longer
// transforms; 1024: all lengths).
template <int NMAX>
__global__ void __launch_bounds__(64) k_rows_r2c_reg(R2C_PARAMS) {
  extern __shared__ double2 lds2[];
  if ((int)blockIdx.x >= nrow_blk) {
    trailing_control(disks, ctl, ctl0 + blockIdx.x - nrow_blk, pcfl, pcm, samples, cap, par, mirror);
    return;
  }
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const double *S = stp(st, ntot, par, 0, D), *R32 = gtab(geom, D, GT_R32);
  const double2 *TW = tw + D.two_row;
  double2 *out = spec + D.spo;
  switch (D.Nphi) {
    case 128: r2c_reg_body<8, 16>(D, wm.y, lds2, S, R32, TW, out); break;
    case 192: r2c_reg_body<12, 16>(D, wm.y, lds2, S, R32, TW, out); break;
    case 256: r2c_reg_body<16, 16>(D, wm.y, lds2, S, R32, TW, out); break;
    case 384: r2c_reg_body<12, 32>(D, wm.y, lds2, S, R32, TW, out); break;
    case 512: r2c_reg_body<16, 32>(D, wm.y, lds2, S, R32, TW, out); break;
    default:
      if constexpr (NMAX > 512) {
        if (D.Nphi == 768) r2c_reg_body<24, 32>(D, wm.y, lds2, S, R32, TW, out);
        else if (D.Nphi == 1024) r2c_reg_body<32, 32>(D, wm.y, lds2, S, R32, TW, out);
      }
      break;   // the runtime sets mode 2 only when every Nphi has a case
  }
}

template <int NMAX>
__global__ void __launch_bounds__(64) k_rows_c2r_reg(C2R_PARAMS) {
  extern __shared__ double2 lds2[];
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const double *sc = gtab(geom, D, GT_PHISC);
  const double2 *TW = tw + D.two_row, *in = spec + D.spo;
  double *P = phi + D.so;
  switch (D.Nphi) {
    case 128: c2r_reg_body<8, 16>(D, wm.y, lds2, in, sc, TW, P); break;
    case 192: c2r_reg_body<12, 16>(D, wm.y, lds2, in, sc, TW, P); break;
    case 256: c2r_reg_body<16, 16>(D, wm.y, lds2, in, sc, TW, P); break;
    case 384: c2r_reg_body<12, 32>(D, wm.y, lds2, in, sc, TW, P); break;
    case 512: c2r_reg_body<16, 32>(D, wm.y, lds2, in, sc, TW, P); break;
    default:
      if constexpr (NMAX > 512) {
        if (D.Nphi == 768) c2r_reg_body<24, 32>(D, wm.y, lds2, in, sc, TW, P);
        else if (D.Nphi == 1024) c2r_reg_body<32, 32>(D, wm.y, lds2, in, sc, TW, P);
      }
      break;
  }
}

// Row transforms by the three-pass register form (rfft::rows3: n = Nphi =
// 64 P, P = 2, 3, 4, 6, 8), one row pair per 64-thread workgroup: two to
// four times the workgroups of the two-pass kernels and at most 8-point DFTs
// per lane (the latency regime).  Row pairs as in the two-pass kernels.
template <int P>
__device__ __forceinline__ void r2c3_body(const DiskDev &D, int pair, double2 *L,
                                          const double *__restrict__ S, const double *__restrict__ R32,
                                          const double2 *__restrict__ tw, double2 *__restrict__ out) {
  constexpr int n = 64 * P;
  const int Nu = D.Nu, l = threadIdx.x, r = l & 7, k1 = rfft::conv3::lane_k1<P>(l);
  const RegPair q = reg_pair<64>(pair, Nu);
  const double pa = R32[q.ra], pb = R32[q.rb];
  const double *Sa = S + (i64)q.ra * n, *Sb = S + (i64)q.rb * n;
  const rfft::conv::TwRow<8, -1> wB(tw, P * r);   // pass B's twiddles, loaded with the rows
  double2 v[P], x[8];
  rfft::rows3::passA<P, -1>(l, [&](int j) {
    const double b = pb * Sb[j];
    return make_double2(pa * Sa[j], q.hasb ? b : 0.0);
  }, tw, v);
  rfft::conv3::t1_write<P>(l, v, L);
  wave_lds_sync();
  rfft::conv3::t1_read(k1, r, L, x);
  wave_lds_sync();
  rfft::rows3::passB<P, -1>(wB, x);
  rfft::conv3::t2_write(k1, r, x, L);
  wave_lds_sync();
  rfft::conv3::t2_read(8 * k1 + r, L, x);
  wave_lds_sync();
  rfft::rows3::passC_spill<P, -1>(k1, r, x, L);
  wave_lds_sync();
  if (q.live)
    rfft::rows3::split<P>(l, L, [&](int k, double2 xa, double2 xb) {
      out[(i64)k * Nu + q.ra] = xa;
      if (q.hasb) out[(i64)k * Nu + q.ra + 1] = xb;
    });
}

template <int P>
__device__ __forceinline__ void c2r3_body(const DiskDev &D, int pair, double2 *L,
                                          const double2 *__restrict__ in, const double *__restrict__ sc,
                                          const double2 *__restrict__ tw, double *__restrict__ Pout) {
  constexpr int n = 64 * P;
  const int Nu = D.Nu, l = threadIdx.x, r = l & 7, k1 = rfft::conv3::lane_k1<P>(l);
  const RegPair q = reg_pair<64>(pair, Nu);
  const rfft::conv::TwRow<8, +1> wB(tw, P * r);   // pass B's twiddles, loaded with the bins
  double2 v[P], x[8];
  rfft::rows3::passA<P, +1>(l, [&](int m) {
    return rfft::rows::c2r_in<P, 64>(m, [&](int k, double2 &xa, double2 &xb) {
      xa = in[(i64)k * Nu + q.ra];
      const double2 b = in[(i64)k * Nu + q.rb];
      xb = q.hasb ? b : make_double2(0.0, 0.0);
    });
  }, tw, v);
  rfft::conv3::t1_write<P>(l, v, L);
  wave_lds_sync();
  rfft::conv3::t1_read(k1, r, L, x);
  wave_lds_sync();
  rfft::rows3::passB<P, +1>(wB, x);
  rfft::conv3::t2_write(k1, r, x, L);
  wave_lds_sync();
  rfft::conv3::t2_read(8 * k1 + r, L, x);
  wave_lds_sync();
  rfft::rows3::passC_spill<P, +1>(k1, r, x, L);
  wave_lds_sync();
  if (q.live) {
    const double sa = sc[q.ra], sb = sc[q.rb];
    double *Pa = Pout + (i64)q.ra * n, *Pb = Pout + (i64)q.rb * n;
    rfft::rows3::unspill<P>(l, L, [&](int j, double2 z) {
      Pa[j] = z.x * sa;
      if (q.hasb) Pb[j] = z.y * sb;
    });
  }
}

__global__ void __launch_bounds__(64) k_rows_r2c_reg3(R2C_PARAMS) {
  extern __shared__ double2 lds2[];
  if ((int)blockIdx.x >= nrow_blk) {
    trailing_control(disks, ctl, ctl0 + blockIdx.x - nrow_blk, pcfl, pcm, samples, cap, par, mirror);
    return;
  }
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const double *S = stp(st, ntot, par, 0, D), *R32 = gtab(geom, D, GT_R32);
  const double2 *TW = tw + D.two_row;
  double2 *out = spec + D.spo;
  switch (D.Nphi) {
    case 128: r2c3_body<2>(D, wm.y, lds2, S, R32, TW, out); break;
    case 192: r2c3_body<3>(D, wm.y, lds2, S, R32, TW, out); break;
    case 256: r2c3_body<4>(D, wm.y, lds2, S, R32, TW, out); break;
    case 384: r2c3_body<6>(D, wm.y, lds2, S, R32, TW, out); break;
    case 512: r2c3_body<8>(D, wm.y, lds2, S, R32, TW, out); break;
    default: break;   // the runtime sets row mode 3 only when every Nphi has a case
  }
}

__global__ void __launch_bounds__(64) k_rows_c2r_reg3(C2R_PARAMS) {
  extern __shared__ double2 lds2[];
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  if (ctl[wm.x].done) return;
  const double *sc = gtab(geom, D, GT_PHISC);
  const double2 *TW = tw + D.two_row, *in = spec + D.spo;
  double *P = phi + D.so;
  switch (D.Nphi) {
    case 128: c2r3_body<2>(D, wm.y, lds2, in, sc, TW, P); break;
    case 192: c2r3_body<3>(D, wm.y, lds2, in, sc, TW, P); break;
    case 256: c2r3_body<4>(D, wm.y, lds2, in, sc, TW, P); break;
    case 384: c2r3_body<6>(D, wm.y, lds2, in, sc, TW, P); break;
    case 512: c2r3_body<8>(D, wm.y, lds2, in, sc, TW, P); break;
    default: break;
  }
}

// ------------------------------------------------------------------ K-hat setup
// Row r2c of the M x Nphi kernel table (kalnajs.hpp:105-114), all M rows, into
// tmp[Nc][M]; then column transforms keep the real part of rows 0..Nu.
__global__ void __launch_bounds__(WG)
k_khat_rows(const DiskDev *__restrict__ D_, const double *__restrict__ ktab,
            const double2 *__restrict__ tw, double2 *__restrict__ tmp) {
  extern __shared__ double2 lds2[];
  const DiskDev &D = *D_;
  const int n = D.Nphi, M = D.M, pr = D.pr, Nc = D.Nc;
  const int row0 = 2 * pr * blockIdx.x;
  double2 *A = lds2, *B = lds2 + pr * n;
  const double2 *TW = tw_table(lds2 + 2 * pr * n, tw + D.two_row, n, D.tw_lds & 1);
  for (int e = threadIdx.x; e < pr * n; e += blockDim.x) {
    const int p = e / n, j = e - p * n;
    const int ra = row0 + 2 * p, rb = ra + 1;
    A[e] = make_double2(ra < M ? ktab[(i64)ra * n + j] : 0.0, rb < M ? ktab[(i64)rb * n + j] : 0.0);
  }
  __syncthreads();
  const double2 *Z = lds_fft<-1>(A, B, n, pr, D.nst_row, D.rad_row, TW);
  const int w2 = 2 * pr;
  for (int e = threadIdx.x; e < w2 * Nc; e += blockDim.x) {
    const int k = e / w2, rr = e - k * w2;
    const int p = rr >> 1, isb = rr & 1;
    const int row = row0 + rr;
    if (row >= M) continue;
    const double2 z = Z[p * n + k];
    double2 y = Z[p * n + (k == 0 ? 0 : n - k)];
    y.y = -y.y;
    double2 X;
    if (!isb) X = make_double2(0.5 * (z.x + y.x), 0.5 * (z.y + y.y));
    else X = make_double2(0.5 * (z.y - y.y), -0.5 * (z.x - y.x));
    tmp[(i64)k * M + row] = X;
  }
}

__global__ void __launch_bounds__(WG)
k_khat_cols(const DiskDev *__restrict__ D_, const double2 *__restrict__ tw,
            double2 *__restrict__ tmp, double *__restrict__ khat) {
  extern __shared__ double2 lds2[];
  const DiskDev &D = *D_;
  const int M = D.M, Nu = D.Nu, Nc = D.Nc, cb = D.cb;
  const int k0 = blockIdx.x * cb;
  double2 *A = lds2, *B = lds2 + cb * M;
  const double2 *TW = tw_table(lds2 + 2 * cb * M, tw + D.two_col, M, D.tw_lds & 2);
  for (int e = threadIdx.x; e < cb * M; e += blockDim.x) {
    const int c = e / M, p = e - c * M;
    const int k = k0 + c;
    A[e] = (k < Nc) ? tmp[(i64)k * M + p] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const double2 *Z = lds_fft<-1>(A, B, M, cb, D.nst_col, D.rad_col, TW);
  for (int e = threadIdx.x; e < cb * M; e += blockDim.x) {
    const int c = e / M, p = e - c * M;
    const int k = k0 + c;
    if (k >= Nc) continue;
    tmp[(i64)k * M + p] = Z[e];                       // full complex K-hat (for checks)
    if (p <= Nu) khat[D.kho + (i64)k * (Nu + 1) + p] = Z[e].x;
  }
}

// ------------------------------------------------------------------ update
// Fused step on an output tile of ti x tj zones.  LDS holds the post-source
// state on (ti+4) x (tj+4) zones, the radial face fluxes, then the post-radial
// state and the azimuthal face fluxes.  GARR: gravity from given g_R/g_phi
// arrays (unit entry point) instead of the potential.
// The body of k_update for tile `tile` of disk D (LDS from `lds`, workgroup
// reduction scratch `red`).  ld(f, i, j): field f (0 Sigma, 1 v_R, 2 v_phi)
// of the current state at zone (i, j); phi(i, j): the potential there, read
// only for self-gravity; sink(lr, j, s): called with the new Sigma of every
// zone (row i0 + lr, column j < Nphi) of the tile after it is stored, and
// with 0 for the tile's rows past the disk.
template <bool GARR, bool G2, class LdF, class PhiF, class SinkF>
__device__ __forceinline__ void update_tile_body(const DiskDev &D, double dt, int tile, double *lds,
                                                 double *red, double *__restrict__ st, i64 ntot,
                                                 const double *__restrict__ gRa,
                                                 const double *__restrict__ gpa, int has_gR, int has_gp,
                                                 const double *__restrict__ geom,
                                                 const double *__restrict__ trig,
                                                 double *__restrict__ pcfl, double *__restrict__ pcm,
                                                 int par, LdF ld, PhiF phi, SinkF sink) {
  const int Nu = D.Nu, Nphi = D.Nphi, TI = D.ti, TJ = D.tj;
  const int ti = tile / D.ntile_j, tj = tile - ti * D.ntile_j;
  const int i0 = ti * TI, j0 = tj * TJ;
  const int SW = TJ + 4, SH = TI + 4;
  double *sS = lds, *sU = sS + SH * SW, *sW = sU + SH * SW;
  double *fM = sW + SH * SW, *fU = fM + (TI + 1) * SW, *fW = fU + (TI + 1) * SW;
  // the row records of rows i0-2 .. i0+TI+1 (rt, RR_STRIDE doubles each;
  // record lr is row i0-2+lr, clamped into the records 0 .. Nu) and cos 2phi,
  // sin 2phi of the tile's columns (trg), staged with phase 1's loads so that
  // the later phases read their row tables from LDS, not from the caches
  // The first staged value of each thread is loaded here and stored after
  // phase 1's loads, so that the two memory latencies overlap; any further
  // ones (tiles with more table entries than threads) after phase 1.
  double *rt = fW + (TI + 1) * SW, *trg = rt + SH * RR_STRIDE;
  const int NST = SH * RR_STRIDE + 2 * TJ;
  const double *rec0 = gtab(geom, D, GT_COUNT);
  auto stage_ld = [&](int e) {   // entry e of rt/trg (trg = rt + SH * RR_STRIDE)
    if (e < SH * RR_STRIDE) {
      const int i = min(max(i0 - 2 + e / RR_STRIDE, 0), Nu);
      return rec0[i * RR_STRIDE + (e % RR_STRIDE)];
    }
    const int q = e - SH * RR_STRIDE, sn = q >= TJ ? 1 : 0;
    const int j = min(j0 + q - sn * TJ, Nphi - 1);
    return trig[D.cos_off + sn * Nphi + j];
  };
  const double st0 = stage_ld(min((int)threadIdx.x, NST - 1));
  auto rrow = [&](int lr, int fld) { return rt[lr * RR_STRIDE + fld]; };   // row i0-2+lr
  const double *tInvdRc = gtab(geom, D, GT_INVDRC), *tInvR = gtab(geom, D, GT_INVR),
               *tGstar = gtab(geom, D, GT_GSTAR), *tGrc = gtab(geom, D, GT_GRC),
               *tGpc = gtab(geom, D, GT_GPC);
  const bool sg = D.selfgrav != 0;

  // ---- phase 1: source (+ forces) on rows i0-2..i0+TI+1, cols j0-2..j0+TJ+1
  for (int e = threadIdx.x; e < SH * SW; e += blockDim.x) {
    const int lr = qdiv(e, rcpf((float)SW)), lc = e - lr * SW;
    const int i = i0 - 2 + lr;
    if (i < 0 || i >= Nu) continue;
    const int j = wrapj(j0 - 2 + lc, Nphi);
    const int jm = wrapj(j - 1, Nphi), jp = wrapj(j + 1, Nphi);
    const int im = (i > 0) ? i - 1 : 0, ip = (i < Nu - 1) ? i + 1 : Nu - 1;
    const i64 c = (i64)i * Nphi + j;
    const double s = ld(0, i, j), u = ld(1, i, j), w = ld(2, i, j);
    const double dPdR = (pres<G2>(D, ld(0, ip, j)) - pres<G2>(D, ld(0, im, j))) * tInvdRc[i];
    const double dPdphi = (pres<G2>(D, ld(0, i, jp)) - pres<G2>(D, ld(0, i, jm))) * D.inv2dphi;
    const double invR = tInvR[i];   // loaded ahead of the gravity branch, whose end waits
    double gR = tGstar[i];
    double gp = 0.0;
    if (GARR) {
      if (has_gR) gR += gRa[D.so + c];
      if (has_gp) gp += gpa[D.so + c];
    } else if (sg) {
      // kalnajs.hpp:71-78, one-sided at the radial edges
      const int iu = (i < Nu - 1) ? i + 1 : i, id = (i > 0) ? i - 1 : i;
      gR += -(phi(iu, j) - phi(id, j)) * tGrc[i];
      gp += -(phi(i, jp) - phi(i, jm)) * tGpc[i];
    }
    const double rs = frcp(s);
    const double aR = -dPdR * rs + gR + w * w * invR;
    const double ap = -(rs * invR) * dPdphi + gp - u * w * invR;
    sS[e] = s;
    sU[e] = u + dt * aR;
    sW[e] = w + dt * ap;
  }
  if ((int)threadIdx.x < NST) rt[threadIdx.x] = st0;
  for (int e = threadIdx.x + blockDim.x; e < NST; e += blockDim.x) rt[e] = stage_ld(e);
  __syncthreads();

  // ---- phase 2: radial faces f = i0..i0+TI (between cells f-1 and f).
  // Branch-free: both upwind choices are formed and one selected, and the
  // three slopes are formed on every lane and dropped where the reference
  // skips them (a face outside 1..Nu-1, a cell at the radial edge, a*b <= 0),
  // so the LDS reads of all three are issued together.  Every address read
  // lies inside the staged rows; values read for dropped terms are unused.
  for (int e = threadIdx.x; e < (TI + 1) * SW; e += blockDim.x) {
    const int lf = qdiv(e, rcpf((float)SW)), lc = e - lf * SW;
    const int f = i0 + lf;
    const int la = (lf + 1) * SW + lc;   // cell f-1
    const int lb = la + SW;              // cell f
    const double vf = 0.5 * (sU[la] + sU[lb]);
    const bool up = vf > 0;
    const double fa = 0.5 * (1.0 - vf * dt * rrow(lf + 1, RR_INVDR));
    const double fb = -0.5 * (1.0 + vf * dt * rrow(lf + 2, RR_INVDR));
    const int lcell = up ? la : lb, cell = up ? f - 1 : f;
    const double fac = up ? fa : fb;
    const bool lim = cell > 0 && cell < Nu - 1;
    const double ss = slope_add(sS[lcell], sS[lcell - SW], sS[lcell + SW], lim, fac);
    const double su = slope_add(sU[lcell], sU[lcell - SW], sU[lcell + SW], lim, fac);
    const double sw = slope_add(sW[lcell], sW[lcell - SW], sW[lcell + SW], lim, fac);
    const bool face = f >= 1 && f <= Nu - 1;
    const double fm = face ? vf * ss * rrow(lf + 2, RR_AFZ) : 0.0;   // the face area (f interior)
    fM[e] = fm; fU[e] = face ? fm * su : 0.0; fW[e] = face ? fm * sw : 0.0;
  }
  __syncthreads();

  // ---- phase 3: radial gather for rows i0..i0+TI-1 (in place into s*)
  for (int e = threadIdx.x; e < TI * SW; e += blockDim.x) {
    const int lr = qdiv(e, rcpf((float)SW)), lc = e - lr * SW;
    const int i = i0 + lr;
    if (i >= Nu) continue;
    const int l = (lr + 2) * SW + lc;
    const double s = sS[l];
    const double mR = s * sU[l], mP = s * sW[l];
    const double iv = dt * rrow(lr + 2, RR_INVVOL);
    const double ns = s + (fM[e] - fM[e + SW]) * iv;
    const double nmR = mR + (fU[e] - fU[e + SW]) * iv;
    const double nmP = mP + (fW[e] - fW[e + SW]) * iv;
    const double rn = frcp(smax(ns, 1e-10));
    sS[l] = ns; sU[l] = nmR * rn; sW[l] = nmP * rn;
  }
  __syncthreads();

  // ---- phase 4: azimuthal faces j = j0..j0+TJ (between j-1 and j), rows
  // i0..i0+TI-1, branch-free as phase 2
  const int FW = TJ + 1;
  for (int e = threadIdx.x; e < TI * FW; e += blockDim.x) {
    const int lr = qdiv(e, rcpf((float)FW)), lf = e - lr * FW;
    const int i = i0 + lr;
    double fm = 0.0, fu = 0.0, fw = 0.0;
    if (i < Nu) {
      const int la = (lr + 2) * SW + lf + 1;   // cell j-1
      const int lb = la + 1;                   // cell j
      const double vf = 0.5 * (sW[la] + sW[lb]);
      const bool up = vf > 0;
      const double idl = rrow(lr + 2, RR_INVDL);
      const double fa = 0.5 * (1.0 - vf * dt * idl);
      const double fb = -0.5 * (1.0 + vf * dt * idl);
      const int lcell = up ? la : lb;
      const double fac = up ? fa : fb;
      const double ss = slope_add(sS[lcell], sS[lcell - 1], sS[lcell + 1], true, fac);
      const double su = slope_add(sU[lcell], sU[lcell - 1], sU[lcell + 1], true, fac);
      const double sw = slope_add(sW[lcell], sW[lcell - 1], sW[lcell + 1], true, fac);
      fm = vf * ss * rrow(lr + 2, RR_AFACE);
      fu = fm * su;
      fw = fm * sw;
    }
    fM[lr * FW + lf] = fm; fU[lr * FW + lf] = fu; fW[lr * FW + lf] = fw;
  }
  __syncthreads();

  // ---- phase 5: azimuthal gather + floor, write, next-step CFL and C_m partials
  double *So = stpw(st, ntot, 1 - par, 0, D), *Uo = stpw(st, ntot, 1 - par, 1, D),
         *Wo = stpw(st, ntot, 1 - par, 2, D);
  double mx = 1e-30, a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int e = threadIdx.x; e < TI * TJ; e += blockDim.x) {
    const int lr = qdiv(e, rcpf((float)TJ)), lc = e - lr * TJ;
    const int i = i0 + lr, j = j0 + lc;
    if (i >= Nu || j >= Nphi) {
      if (j < Nphi) sink(lr, j, 0.0);
      continue;
    }
    const int l = (lr + 2) * SW + lc + 2;
    const int f = lr * FW + lc;
    const double s = sS[l];
    const double mR = s * sU[l], mP = s * sW[l];
    const double iv = dt * rrow(lr + 2, RR_INVVOL);
    const double ns = s + (fM[f] - fM[f + 1]) * iv;
    const double nmR = mR + (fU[f] - fU[f + 1]) * iv;
    const double nmP = mP + (fW[f] - fW[f + 1]) * iv;
    const double rn = frcp(smax(ns, 1e-10));
    const double so = smax(ns, 1e-10);
    const double uo = nmR * rn, wo = nmP * rn;
    const i64 c = (i64)i * Nphi + j;
    So[c] = so; Uo[c] = uo; Wo[c] = wo;
    sink(lr, j, so);
    const double rate = cfl_rate<G2>(D, so, uo, wo, rrow(lr + 2, RR_INVDR), rrow(lr + 2, RR_INVDL),
                                     rrow(lr + 2, RR_AFACE), rrow(lr + 2, RR_DL));
    mx = (rate > mx) ? rate : mx;
    const double dA = rrow(lr + 2, RR_DA);
    a0 += so * dA;
    a1 += so * trg[lc] * dA;
    a2 -= so * trg[TJ + lc] * dA;
    a3 += nonfinite(so, uo, wo);
  }
  wg_reduce5(mx, a0, a1, a2, a3, red);
  if (threadIdx.x == 0) {
    const i64 t = D.part + tile;
    pcfl[t] = mx;
    pcm[4 * t + 0] = a0; pcm[4 * t + 1] = a1; pcm[4 * t + 2] = a2; pcm[4 * t + 3] = a3;
  }
}

template <bool GARR, bool G2>
__global__ void __launch_bounds__(WG)
k_update(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
         const DiskCtl *__restrict__ ctl, double *__restrict__ st, i64 ntot,
         const double *__restrict__ phi, const double *__restrict__ gRa,
         const double *__restrict__ gpa, int has_gR, int has_gp,
         const double *__restrict__ geom, const double *__restrict__ trig,
         double *__restrict__ pcfl, double *__restrict__ pcm, int par) {
  extern __shared__ double lds[];
  __shared__ double red[5 * (WG / 64)];
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  const DiskCtl &C = ctl[wm.x];
  if (C.done) return;
  const double *Ph = phi + D.so;
  const double *S = stp(st, ntot, par, 0, D), *U = stp(st, ntot, par, 1, D),
               *W = stp(st, ntot, par, 2, D);
  const int Nphi = D.Nphi;
  update_tile_body<GARR, G2>(
      D, C.dt, wm.y, lds, red, st, ntot, gRa, gpa, has_gR, has_gp, geom, trig, pcfl, pcm, par,
      [&](int f, int i, int j) { return (f == 0 ? S : (f == 1 ? U : W))[(i64)i * Nphi + j]; },
      [&](int i, int j) { return Ph[(i64)i * Nphi + j]; }, [](int, int, double) {});
}

// ------------------------------------------------------------------ fused row kernel
// One loop iteration's row work for the rows [i0, i0 + TI) of a disk at full
// width (TI = D.ti, even; the layout makes the tile TI x Nphi), once the
// column convolution has left V-hat in `vin`:
//   1. row c2r of the row pairs covering i0-3 .. i0+TI+2 (the rows the
//      forces of the tile read) into LDS, as k_rows_c2r (Phi = pref V /
//      sqrt(R), applied on use);
//   2. the update of k_update on the tile, forces from that Phi;
//   3. row r2c of the tile's new Sigma, as k_rows_r2c: S-hat of the new state
//      into `sout` (a buffer other than vin: a workgroup reads V-hat rows
//      that its neighbours overwrite with S-hat), which the next iteration's
//      convolution transforms in place.
// Phi, the momenta and the row spectra never leave the workgroup.  LDS: the
// tile (as k_update), (TI/2 + 4) c2r row pairs, the staged state rows (whose
// space the TI/2 r2c row pairs reuse), then the twiddles if staged.
template <int TG, bool G2>
__global__ void __launch_bounds__(TG)
k_rows_fused(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
             const DiskCtl *__restrict__ ctl, double *__restrict__ st, i64 ntot,
             const double2 *__restrict__ vin, double2 *__restrict__ sout,
             const double *__restrict__ geom, const double2 *__restrict__ tw,
             const double *__restrict__ trig, double *__restrict__ pcfl, double *__restrict__ pcm,
             int par) {
  extern __shared__ double lds[];
  __shared__ double red[5 * (TG / 64)];
  constexpr int NMAX = ip_nmax<TG>();
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  const DiskCtl &C = ctl[wm.x];
  if (C.done) return;
  const int Nu = D.Nu, n = D.Nphi, Nc = D.Nc, TI = D.ti;
  const int i0 = wm.y * TI;
  double2 *F = (double2 *)(lds + fused_tile_doubles(TI, n));   // c2r row pairs
  double *Sst = (double *)(F + (TI / 2 + 4) * n);              // staged state rows
  double *Ust = Sst + (TI + 6) * n, *Wst = Ust + (TI + 4) * n;
  double2 *Rb = (double2 *)Sst;   // r2c row pairs (the staged rows are dead by then)
  const double2 *TW = tw_table((double2 *)(Sst + fused_stage_doubles(TI, n)), tw + D.two_row, n,
                               D.tw_lds & 1);
  const int pa = max(i0 - 4, 0) >> 1;                      // first c2r row pair
  const int np = min(i0 / 2 + TI / 2 + 2, (Nu + 1) / 2) - pa;   // c2r row pairs
  const int nrp = (min(TI, Nu - i0) + 1) >> 1;             // the tile's row pairs
  const int maxb = NMAX / n;                               // transforms per in-place pass
  // Stage everything the workgroup reads from HBM -- the V-hat bins of the
  // c2r row pairs and the state rows of the tile's stencil -- with up to
  // 2 + 8 loads in flight per thread before the first LDS store, so that the
  // kernel waits on about one memory latency instead of one per pass.
  {
    const double2 *in = vin + D.spo;
    const double *S = stp(st, ntot, par, 0, D), *U = stp(st, ntot, par, 1, D),
                 *W = stp(st, ntot, par, 2, D);
    const int nsp = np * Nc, nS = (TI + 6) * n, nst = nS + 2 * (TI + 4) * n;
    const float rnp = rcpf((float)np), rn = rcpf((float)n);
    constexpr int B1 = 2, B2 = 8;
    const int nUW = (TI + 4) * n;
    for (int base = 0; base * B1 < nsp || base * B2 < nst; base += TG) {
      // loads from clamped addresses, no branches: all of them issue before
      // the first wait
      double2 xa[B1], xb[B1];
      double v[B2];
#pragma unroll
      for (int u = 0; u < B1; ++u) {
        const int e = min(base * B1 + u * TG + (int)threadIdx.x, nsp - 1);
        const int k = qdiv(e, rnp), ra = 2 * (pa + e - k * np);
        xa[u] = in[(i64)k * Nu + min(ra, Nu - 1)];
        xb[u] = in[(i64)k * Nu + min(ra + 1, Nu - 1)];
      }
#pragma unroll
      for (int u = 0; u < B2; ++u) {
        const int e = min(base * B2 + u * TG + (int)threadIdx.x, nst - 1);
        const int f = e < nS ? 0 : (e < nS + nUW ? 1 : 2);
        const int el = f == 0 ? e : e - nS - (f - 1) * nUW;
        const int r = qdiv(el, rn);
        const int i = min(max(i0 - (f == 0 ? 3 : 2) + r, 0), Nu - 1);
        v[u] = (f == 0 ? S : (f == 1 ? U : W))[(i64)i * n + el - r * n];
      }
#pragma unroll
      for (int u = 0; u < B1; ++u) {
        const int e = base * B1 + u * TG + threadIdx.x;
        if (e < nsp) {
          const int k = qdiv(e, rnp), p = e - k * np, ra = 2 * (pa + p);
          const double2 z = make_double2(0.0, 0.0);
          c2r_put(F, p, k, ra < Nu ? xa[u] : z, ra + 1 < Nu ? xb[u] : z, n, Nc);
        }
      }
      // staged rows outside the disk are never read (the tile body clamps
      // its rows into the disk and skips the others)
#pragma unroll
      for (int u = 0; u < B2; ++u) {
        const int e = base * B2 + u * TG + threadIdx.x;
        if (e < nst) Sst[e] = v[u];
      }
    }
  }
  __syncthreads();
  for (int q0 = 0; q0 < np; q0 += maxb) {
    StageIO g{};
    int Ns = 1;
    for (int q = 0; q < D.nst_row; ++q) {
      ip_stage_r<+1, IO_LDS, TG, NMAX>(D.rad_row[q], F + q0 * n, n, min(maxb, np - q0), Ns,
                                      threadIdx.x, TW, g);
      Ns *= D.rad_row[q];
    }
  }
  const double *sc = gtab(geom, D, GT_PHISC), *R32 = gtab(geom, D, GT_R32);
  update_tile_body<false, G2>(
      D, C.dt, wm.y, lds, red, st, ntot, nullptr, nullptr, 0, 0, geom, trig, pcfl, pcm, par,
      [&](int f, int i, int j) {
        return f == 0 ? Sst[(i - i0 + 3) * n + j] : (f == 1 ? Ust : Wst)[(i - i0 + 2) * n + j];
      },
      [&](int i, int j) {
        const double2 z = F[((i >> 1) - pa) * n + j];
        return ((i & 1) ? z.y : z.x) * sc[i];
      },
      [&](int lr, int j, double so) {
        double *x = (double *)&Rb[(lr >> 1) * n + j];
        x[lr & 1] = (i0 + lr < Nu) ? R32[i0 + lr] * so : 0.0;
      });
  __syncthreads();
  for (int q0 = 0; q0 < nrp; q0 += maxb) {
    StageIO g{};
    int Ns = 1;
    for (int q = 0; q < D.nst_row; ++q) {
      ip_stage_r<-1, IO_LDS, TG, NMAX>(D.rad_row[q], Rb + q0 * n, n, min(maxb, nrp - q0), Ns,
                                      threadIdx.x, TW, g);
      Ns *= D.rad_row[q];
    }
  }
  r2c_store(Rb, sout + D.spo, i0, nrp, n, Nc, Nu);
}

// ------------------------------------------------------------------ streaming update
// Register-streaming variant of k_update for bandwidth-bound (large or
// batched) runs.  One wavefront per work unit = (disk, column group g, radial
// strip k).  Lane L owns azimuthal column c = 58 g - 3 + L (periodic); lanes
// 3..60 produce output.  The wave walks source rows i0-2 .. i1+1 of its
// strip [i0, i1) (down on even strips, up on odd ones): the radial stencil
// (Sigma/Phi windows, the last four post-source rows, the previous radial
// face) stays in registers, azimuthal neighbours -- of the source step's
// pressure and potential as well as of the azimuthal sweep -- come from
// cross-lane shifts (a 3-lane halo: the source values of lanes 1..62 are
// exact, those of the edge lanes reach no output), and the row tables are
// wave-uniform scalar loads from per-row records.  The formulas of k_update,
// arranged without branches (the results agree to round-off; dt is the
// reference's exactly in both).
#define STREAM_OUT 58
#define STREAM_HALO 3

// The streaming update reads and writes the state through buffer
// descriptors.  One descriptor covers a disk's three fields of one state
// buffer (they lie ntot doubles apart): the field and row go into the scalar
// offset, the lane's column into the vector offset, so a load or store costs
// no vector address arithmetic.  A lane with nothing to store (halo lanes,
// columns past the row's end, rows outside the strip) passes kDropOffset as
// its vector offset, and the hardware's range check discards the write: the
// descriptors span DFK_STREAM_BYTES, every real element lies inside (3 ntot
// doubles <= DFK_STREAM_BYTES, checked by dfk_set_update_kernel) and the drop
// offset outside, whether or not the range check adds the scalar offset.  So
// every store is issued by the whole wavefront outside any branch, and the
// compiler can count the stores that follow a row's loads: the wait for the
// next rows' loads no longer includes the completion of earlier stores.
static constexpr int kDropOffset = 0x7fffffff;
typedef unsigned int dfk_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t stream_rsrc(const double *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, DFK_STREAM_BYTES, 0x00020000);   // raw
}
// NT: the nontemporal cache policy (the instruction's nt bit).  The update
// reads and writes the velocities v_R, v_phi once per step (the row
// transforms do not touch them): streamed past the caches they leave the
// room to Sigma and Phi, which neighbouring strips re-read and whose new
// values the next row transform reads (ensemble loop 3.47e10 -> 3.65e10
// zone-updates/s with one chain; with the new Sigma streamed as well the
// loop measured no gain).
template <bool NT = false>
__device__ __forceinline__ void store_or_drop(__amdgpu_buffer_rsrc_t r, int voff, int soff, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(dfk_u32x2, v), r, voff, soff, NT ? 2 : 0);
}
template <bool NT = false>
__device__ __forceinline__ double load_at(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, NT ? 2 : 0));
}

// Geometry tables read through the constant address space: they are not
// written while a loop kernel runs, and a load from that space stays a
// scalar load even after the buffer stores of the same loop, which the
// compiler cannot prove are elsewhere (a global-space load after them becomes
// a vector load with a wait for it).
typedef const __attribute__((address_space(4))) double ctab_t;
__device__ __forceinline__ ctab_t *ctab(const double *geom, const DiskDev &D, int t) {
  return (ctab_t *)gtab(geom, D, t);
}

// Whole-wavefront lane shifts by one through DPP (wave_shr:1 / wave_shl:1 on
// each 32-bit half): lane L receives lane L-1 (up) or L+1 (down); the lane
// shifted in at the edge receives 0.  All 64 lanes must be active.
__device__ __forceinline__ double up1(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, true),
                          __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ double dn1(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, true),
                          __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, true));
}

template <bool GARR, bool G2>
__global__ void __launch_bounds__(64, 4)   // at most 128 VGPRs: four waves per SIMD
k_update_stream(const DiskDev *__restrict__ disks, const int2 *__restrict__ map, int map_base,
                const DiskCtl *__restrict__ ctl, double *__restrict__ st, i64 ntot,
                const double *__restrict__ phi, const double *__restrict__ gRa,
                const double *__restrict__ gpa, int has_gR, int has_gp,
                const double *__restrict__ geom, const double *__restrict__ trig,
                double *__restrict__ pcfl, double *__restrict__ pcm, int par) {
  const int2 wm = map[map_base + blockIdx.x];
  const DiskDev &D = disks[wm.x];
  const DiskCtl &C = ctl[wm.x];
  if (C.done) return;
  const double dt = C.dt;
  const int Nu = D.Nu, Nphi = D.Nphi, TS = D.ti;
  const int nrs = D.ntile_i;
  const int cg = wm.y / nrs, ks = wm.y - cg * nrs;
  const int i0 = ks * TS, i1 = min(i0 + TS, Nu);
  const int j0 = cg * STREAM_OUT;
  const int lane = threadIdx.x;
  int c = (j0 - STREAM_HALO + lane) % Nphi;
  if (c < 0) c += Nphi;
  const unsigned cb = (unsigned)c * sizeof(double);   // the lane's byte offset in a row
  const bool out_ok = lane >= STREAM_HALO && lane < STREAM_HALO + STREAM_OUT &&
                      (j0 + lane - STREAM_HALO) < Nphi;

  // the state buffer read (fields Sigma, v_R, v_phi at byte offsets 0, fb,
  // 2 fb), the one written, and the potential
  const __amdgpu_buffer_rsrc_t rIn = stream_rsrc(stp(st, ntot, par, 0, D)),
                               rOut = stream_rsrc(stpw(st, ntot, 1 - par, 0, D)),
                               rP = stream_rsrc(phi + D.so);
  const int fb = (int)(ntot * (i64)sizeof(double));
  // row records: rec(i)[RR_*] (a scalar pointer per row, immediate offsets)
  ctab_t *rec0 = ctab(geom, D, GT_COUNT);
  auto rec = [&](int i) { return rec0 + i * RR_STRIDE; };
  const double c2j = trig[D.cos_off + c], s2j = trig[D.cos_off + Nphi + c];
  // self-gravity terms are formed on every disk and kept by a select (Phi
  // of a disk without self-gravity may hold anything; the select drops it),
  // so that one build serves every handle and a disk's bits do not depend on
  // the other disks of its handle
  const bool sg = !GARR && D.selfgrav != 0;

  auto clampr = [&](int r) { return r < 0 ? 0 : (r >= Nu ? Nu - 1 : r); };
  // row r's element of this lane (r clamped into the disk)
  const int rowb = Nphi * (int)sizeof(double);
  auto at = [&](int field, int r) { return load_at(rIn, cb, field * fb + clampr(r) * rowb); };
  auto atV = [&](int field, int r) { return load_at<true>(rIn, cb, field * fb + clampr(r) * rowb); };
  auto atP = [&](int r) { return load_at(rP, cb, clampr(r) * rowb); };
  double mx = 1e-30, a0 = 0, a1 = 0, a2 = 0, a3 = 0;

  // Even strips walk down (rows i0-2 .. i1+1), odd strips walk up (i1+1 ..
  // i0-2), so two neighbouring strips reach the rows they share at the same
  // time -- both at their start or both at their end -- and the second read
  // of those rows hits L2.  The arithmetic of every face and cell is the
  // same either way; only the order of the partial sums differs.
  //
  // Per source row s: the source step (forces, pressure, geometry) on the
  // row; then the radial face f whose four cells are now post-source (f =
  // s-1 walking down, s+2 walking up); then the gather, azimuthal sweep and
  // output of row g, whose two radial faces are now known (g = f-1 down, f
  // up).  The first four rows of a strip only fill the stencil (sources,
  // then the first face); after them every row does all three.  That steady
  // part has no branches: rows past the disk's edge are read clamped and
  // give finite values that no output uses (the edge faces have zero area,
  // the edge cells a zero MUSCL factor, from the tables GT_AFZ, GT_HUP,
  // GT_HDN), and a row or lane without output stores to kDropOffset and adds
  // nothing to the CFL maximum and the sums.
  //
  // MUSCL face values (hydro.hpp:107-124, 149-163): a face takes q + d F from
  // its upwind cell, d the cell's half van Leer slope (half_slope) and F =
  // +-(1 -+ vf dt/dx) twice the reference's factor +-0.5 (1 -+ vf dt/dx), or 0
  // at a cell on the disk's edge: the same real product as the reference's
  // vanleer * factor, rounded once in the FMA.  Azimuthally every lane forms
  // its own cell's slope and takes its left neighbour's by a lane shift;
  // radially the face forms the slope of its upwind cell from the ring.
  auto strip = [&](auto down_tag) {
    constexpr bool DOWN = decltype(down_tag)::value;
    constexpr int STEP = DOWN ? 1 : -1;
    const int sfirst = DOWN ? i0 - 2 : i1 + 1;
    // Sigma and Phi windows: the row behind (b), this row (0); pressures
    // P(Sigma) are formed once per row as it arrives and shifted across lanes
    // (the same values the reference forms at each neighbour)
    double S0 = at(0, sfirst);
    double Prb = pres<G2>(D, at(0, sfirst - STEP)), Pr0 = pres<G2>(D, S0);
    double Pb = atP(sfirst - STEP), P0 = atP(sfirst);
    // post-source ring of three rows in increasing row order: after the
    // source of row s, slot k holds row s-2+k walking down (rows enter at
    // slot 2), row s+k walking up (rows enter at slot 0)
    double qs0 = 0, qs1 = 0, qs2 = 0, qu0 = 0, qu1 = 0, qu2 = 0, qw0 = 0, qw1 = 0, qw2 = 0;
    // half van Leer slopes: *n of the middle slot's cell (formed after this
    // row's source), *c of the cell formed one row before (slot 0 walking
    // down, slot 2 walking up)
    double dsn = 0, dun = 0, dwn = 0, dsc = 0, duc = 0, dwc = 0;
    double pfm = 0, pfu = 0, pfw = 0;     // the face met before this one

    // fetch(r): v_R, v_phi of row r and Sigma, Phi of row r + STEP, issued
    // one row ahead (two rows ahead measured no faster and costs 8 VGPRs);
    // past the strip's last source row (i1+1 down, i0-2 up) the fetch repeats
    // that row's loads (cache hits), so no row beyond the stencil is read
    double nU, nW, nSa, nPa;
    auto fetch = [&](int r) {
      const int rl = DOWN ? min(r, i1 + 1) : max(r, i0 - 2);
      nU = atV(1, rl);
      nW = atV(2, rl);
      nSa = at(0, rl + STEP);
      nPa = atP(rl + STEP);    // the potential pool covers every disk: loaded whether used or not
    };
    fetch(sfirst);
    int s = sfirst;

    // ---- source at row s (hydro.hpp:73-96, forces kalnajs.hpp:65-80), pushed into the ring
    auto source_row = [&]() {
      const double Sa = nSa, Pa = nPa, u = nU, w = nW;
      fetch(s + STEP);
      const double Pra = pres<G2>(D, Sa);
      ctab_t *rs_ = rec(clampr(s));
      const double Prp = DOWN ? Pra : Prb, Prm = DOWN ? Prb : Pra;
      const double dPdR = (Prp - Prm) * rs_[RR_INVDRC];
      // azimuthal neighbours from the neighbouring lanes (the edge lanes 0
      // and 63 get zeros; their results reach no output lane)
      const double dPdphi = (dn1(Pr0) - up1(Pr0)) * D.inv2dphi;
      double gR = rs_[RR_GSTAR], gp = 0.0;
      if (GARR) {
        const i64 ro = (i64)clampr(s) * Nphi + c;
        if (has_gR) gR += gRa[D.so + ro];
        if (has_gp) gp += gpa[D.so + ro];
      } else {
        const double Pp = DOWN ? Pa : Pb, Pm = DOWN ? Pb : Pa;
        const double gRs = gR + -(Pp - Pm) * rs_[RR_GRC];
        const double gps = gp + -(dn1(P0) - up1(P0)) * rs_[RR_GPC];
        gR = sg ? gRs : gR;
        gp = sg ? gps : gp;
      }
      const double rs = frcp(S0);
      const double invR = rs_[RR_INVR];
      const double aR = -dPdR * rs + gR + w * w * invR;
      const double ap = -(rs * invR) * dPdphi + gp - u * w * invR;
      const double ns_s = S0, ns_u = u + dt * aR, ns_w = w + dt * ap;
      if (DOWN) {
        qs0 = qs1; qs1 = qs2; qs2 = ns_s;
        qu0 = qu1; qu1 = qu2; qu2 = ns_u;
        qw0 = qw1; qw1 = qw2; qw2 = ns_w;
      } else {
        qs2 = qs1; qs1 = qs0; qs0 = ns_s;
        qu2 = qu1; qu1 = qu0; qu0 = ns_u;
        qw2 = qw1; qw1 = qw0; qw0 = ns_w;
      }
      Prb = Pr0; Pr0 = Pra; S0 = Sa;
      Pb = P0; P0 = Pa;
    };
    // ---- the radial half slope of the middle slot's cell (hydro.hpp:107-110)
    auto slopes = [&]() {
      dsc = dsn; duc = dun; dwc = dwn;
      dsn = half_slope(qs0, qs1, qs2);
      dun = half_slope(qu0, qu1, qu2);
      dwn = half_slope(qw0, qw1, qw2);
    };
    // ---- radial face f between cells f-1 and f (hydro.hpp:111-130): slots
    // 0 and 1 walking down (slopes c and n), 1 and 2 walking up (n and c)
    auto radial_face = [&](int f, double &fm, double &fu, double &fw) {
      const int fc = min(max(f, 0), Nu);     // faces 0 .. Nu (past them only in masked rows)
      const int fl = max(fc - 1, 0), fr = min(fc, Nu - 1);
      const double sL = DOWN ? qs0 : qs1, uL = DOWN ? qu0 : qu1, wL = DOWN ? qw0 : qw1;
      const double sR = DOWN ? qs1 : qs2, uR = DOWN ? qu1 : qu2, wR = DOWN ? qw1 : qw2;
      const double dsL = DOWN ? dsc : dsn, duL = DOWN ? duc : dun, dwL = DOWN ? dwc : dwn;
      const double dsR = DOWN ? dsn : dsc, duR = DOWN ? dun : duc, dwR = DOWN ? dwn : dwc;
      const double vf = 0.5 * (uL + uR);
      const bool up = vf > 0;
      // F = 1 - vf dt/dR_{f-1} upwind left, -(1 + vf dt/dR_f) upwind right,
      // 0 for a cell at the disk's edge (no slope there); both sides' table
      // values are loaded before the choice (scalar loads; a choice between
      // two loads would become one vector load from a per-lane address, a
      // dependent memory wait in every row)
      ctab_t *rl = rec(fl), *rr = rec(fr);
      const double iDl = rl[RR_INVDR], iDr = rr[RR_INVDR], hl = rl[RR_HUP], hr = rr[RR_HDN];
      const double iD = up ? iDl : iDr;
      const double hh = up ? hl : hr;
      const double F = hh * fma(-(fabs(vf) * dt), iD, 1.0);
      const double ss = fma(up ? dsL : dsR, F, up ? sL : sR);
      const double su = fma(up ? duL : duR, F, up ? uL : uR);
      const double sw = fma(up ? dwL : dwR, F, up ? wL : wR);
      fm = vf * ss * rec(fc)[RR_AFZ];
      fu = fm * su;
      fw = fm * sw;
    };
    // ---- radial gather of row g (faces g and g+1), azimuthal sweep, floor, output
    auto gather_row = [&](int g, double flo, double fhi, double ulo, double uhi, double wlo,
                          double whi) {
      ctab_t *rg = rec(clampr(g));
      const double s_ = DOWN ? qs0 : qs2, u_ = DOWN ? qu0 : qu2, w_ = DOWN ? qw0 : qw2;
      const double iv = dt * rg[RR_INVVOL];
      const double ns = s_ + (flo - fhi) * iv;
      const double nmR = s_ * u_ + (ulo - uhi) * iv;
      const double nmP = s_ * w_ + (wlo - whi) * iv;
      const double rn = frcp(smax(ns, 1e-10));
      const double Xs = ns, Xu = nmR * rn, Xw = nmP * rn;
      // azimuthal face at this lane's left edge (hydro.hpp:153-168): this
      // lane's cell slope, the left neighbour's value and slope by shifts
      const double sm1 = up1(Xs), um1 = up1(Xu), wm1 = up1(Xw);
      const double ds = half_slope(sm1, Xs, dn1(Xs)), du = half_slope(um1, Xu, dn1(Xu)),
                   dw = half_slope(wm1, Xw, dn1(Xw));
      const double vf = 0.5 * (wm1 + Xw);
      const bool upw = vf > 0;
      const double F = (upw ? 1.0 : -1.0) * fma(-(fabs(vf) * dt), rg[RR_INVDL], 1.0);
      const double ss = fma(upw ? up1(ds) : ds, F, upw ? sm1 : Xs);
      const double su = fma(upw ? up1(du) : du, F, upw ? um1 : Xu);
      const double sw = fma(upw ? up1(dw) : dw, F, upw ? wm1 : Xw);
      const double Fm = vf * ss * rg[RR_AFACE];
      const double FU = Fm * su, FW = Fm * sw;
      const double Fm1 = dn1(Fm), FU1 = dn1(FU), FW1 = dn1(FW);
      // azimuthal gather + floor (hydro.hpp:170-178, 68-69)
      const double ns2 = Xs + (Fm - Fm1) * iv;
      const double nmR2 = Xs * Xu + (FU - FU1) * iv;
      const double nmP2 = Xs * Xw + (FW - FW1) * iv;
      const double rn2 = frcp(smax(ns2, 1e-10));
      const double so = smax(ns2, 1e-10), uo = nmR2 * rn2, wo = nmP2 * rn2;
      // (bitwise, not short-circuit: no branch)
      const bool ok = out_ok & ((g >= i0) & (g < i1));
      const int voff = ok ? (int)cb : kDropOffset, roff = clampr(g) * rowb;
      store_or_drop(rOut, voff, roff, so);
      store_or_drop<true>(rOut, voff, fb + roff, uo);
      store_or_drop<true>(rOut, voff, 2 * fb + roff, wo);
      // the new state's CFL rate (exact, hydro.hpp:52-57) and C_0, C_2 sums
      const double rate = cfl_rate<G2>(D, so, uo, wo, rg[RR_INVDR], rg[RR_INVDL], rg[RR_AFACE], rg[RR_DL]);
      mx = (ok & (rate > mx)) ? rate : mx;
      const double m = ok ? so : 0.0;
      const double dA = rg[RR_DA];
      a0 += m * dA;
      a1 += m * c2j * dA;
      a2 -= m * s2j * dA;
      a3 += (ok & (nonfinite(so, uo, wo) != 0.0)) ? 1.0 : 0.0;
    };

    // stencil fill: sources of the first three rows and the slope of the
    // middle one, then the fourth row's source and slope and the strip's
    // first face (f = i0 down, i1 up)
#pragma unroll
    for (int k = 0; k < 3; ++k, s += STEP) source_row();
    slopes();
    source_row();
    slopes();
    radial_face(DOWN ? s - 1 : s + 2, pfm, pfu, pfw);
    s += STEP;
    // steady rows in groups of four, so that the unrolled body renames the
    // ring instead of moving it; the rows past a short strip's end compute
    // nothing that is kept
    const int ngrp = (i1 - i0 + 3) >> 2;
    for (int grp = 0; grp < ngrp; ++grp)
#pragma unroll
      for (int k = 0; k < 4; ++k, s += STEP) {
        source_row();
        slopes();
        const int f = DOWN ? s - 1 : s + 2;
        double fm, fu, fw;
        radial_face(f, fm, fu, fw);
        if (DOWN)
          gather_row(f - 1, pfm, fm, pfu, fu, pfw, fw);
        else
          gather_row(f, fm, pfm, fu, pfu, fw, pfw);
        pfm = fm; pfu = fu; pfw = fw;
      }
  };
  if (ks & 1)
    strip(std::integral_constant<bool, false>{});
  else
    strip(std::integral_constant<bool, true>{});
  // wave reduction of the partials (fixed order)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(mx, off);
    mx = (o > mx) ? o : mx;
    a0 += __shfl_xor(a0, off);
    a1 += __shfl_xor(a1, off);
    a2 += __shfl_xor(a2, off);
    a3 += __shfl_xor(a3, off);
  }
  if (lane == 0) {
    const i64 t = D.part + wm.y;
    pcfl[t] = mx;
    pcm[4 * t + 0] = a0; pcm[4 * t + 1] = a1; pcm[4 * t + 2] = a2; pcm[4 * t + 3] = a3;
  }
}

// ------------------------------------------------------------------ unit stages
// Single-stage kernels that follow the reference's loop structure one to one
// (source, radial faces/gather, azimuthal faces/gather, floor, forces).  They
// back the per-stage entry points used to localise parity differences; the
// time loop uses k_update.  One thread per zone (or face), state buffer `cur`.
__global__ void k_forces(const DiskDev *__restrict__ Dp, const double *__restrict__ phi,
                         const double *__restrict__ geom, double *__restrict__ gR,
                         double *__restrict__ gp) {
  const DiskDev &D = *Dp;
  const i64 N = (i64)D.Nu * D.Nphi;
  for (i64 c = blockIdx.x * (i64)blockDim.x + threadIdx.x; c < N; c += (i64)gridDim.x * blockDim.x) {
    const int i = (int)(c / D.Nphi), j = (int)(c - (i64)i * D.Nphi);
    const int jm = wrapj(j - 1, D.Nphi), jp = wrapj(j + 1, D.Nphi);
    const int iu = (i < D.Nu - 1) ? i + 1 : i, id = (i > 0) ? i - 1 : i;
    const double *P = phi + D.so;
    gR[D.so + c] = -(P[(i64)iu * D.Nphi + j] - P[(i64)id * D.Nphi + j]) * gtab(geom, D, GT_GRC)[i];
    gp[D.so + c] = -(P[(i64)i * D.Nphi + jp] - P[(i64)i * D.Nphi + jm]) * gtab(geom, D, GT_GPC)[i];
  }
}

// source, in place on buffer `cur` (reads only own-cell velocities)
__global__ void k_source(const DiskDev *__restrict__ Dp, double *__restrict__ st, i64 ntot, int cur,
                         double dt, const double *__restrict__ gRa, const double *__restrict__ gpa,
                         const double *__restrict__ geom) {
  const DiskDev &D = *Dp;
  const int Nu = D.Nu, Nphi = D.Nphi;
  const i64 N = (i64)Nu * Nphi;
  const double *S = stp(st, ntot, cur, 0, D);
  double *U = stpw(st, ntot, cur, 1, D), *W = stpw(st, ntot, cur, 2, D);
  for (i64 c = blockIdx.x * (i64)blockDim.x + threadIdx.x; c < N; c += (i64)gridDim.x * blockDim.x) {
    const int i = (int)(c / Nphi), j = (int)(c - (i64)i * Nphi);
    const int jm = wrapj(j - 1, Nphi), jp = wrapj(j + 1, Nphi);
    const int im = (i > 0) ? i - 1 : 0, ip = (i < Nu - 1) ? i + 1 : Nu - 1;
    const double s = S[c], u = U[c], w = W[c];
    const double dPdR = (pres(D, S[(i64)ip * Nphi + j]) - pres(D, S[(i64)im * Nphi + j])) * gtab(geom, D, GT_INVDRC)[i];
    const double dPdphi = (pres(D, S[(i64)i * Nphi + jp]) - pres(D, S[(i64)i * Nphi + jm])) * D.inv2dphi;
    double gR = gtab(geom, D, GT_GSTAR)[i], gp = 0.0;
    if (gRa) gR += gRa[D.so + c];
    if (gpa) gp += gpa[D.so + c];
    const double rs = 1.0 / s, invR = gtab(geom, D, GT_INVR)[i];
    const double aR = -dPdR * rs + gR + w * w * invR;
    const double ap = -(rs * invR) * dPdphi + gp - u * w * invR;
    U[c] = u + dt * aR;
    W[c] = w + dt * ap;
  }
}

// face fluxes; dir 0 = radial (faces (Nu+1) x Nphi), 1 = azimuthal (Nu x Nphi)
__global__ void k_faces(const DiskDev *__restrict__ Dp, const double *__restrict__ st, i64 ntot,
                        int cur, double dt, int dir, const double *__restrict__ geom,
                        double *__restrict__ F) {
  const DiskDev &D = *Dp;
  const int Nu = D.Nu, Nphi = D.Nphi;
  const double *S = stp(st, ntot, cur, 0, D), *U = stp(st, ntot, cur, 1, D), *W = stp(st, ntot, cur, 2, D);
  const i64 NF = (i64)(dir == 0 ? Nu + 1 : Nu) * Nphi;
  double *Fm = F, *FU = F + NF, *FW = F + 2 * NF;
  for (i64 c = blockIdx.x * (i64)blockDim.x + threadIdx.x; c < NF; c += (i64)gridDim.x * blockDim.x) {
    const int f = (int)(c / Nphi), j = (int)(c - (i64)f * Nphi);
    double fm = 0.0, fu = 0.0, fw = 0.0;
    if (dir == 0) {
      if (f >= 1 && f <= Nu - 1) {
        const i64 a = (i64)(f - 1) * Nphi + j, b = (i64)f * Nphi + j;
        const double vf = 0.5 * (U[a] + U[b]);
        i64 q; int cell; double fac;
        if (vf > 0) { q = a; cell = f - 1; fac = 0.5 * (1.0 - vf * dt * gtab(geom, D, GT_INVDR)[f - 1]); }
        else { q = b; cell = f; fac = -0.5 * (1.0 + vf * dt * gtab(geom, D, GT_INVDR)[f]); }
        double ss = S[q], su = U[q], sw = W[q];
        if (cell > 0 && cell < Nu - 1) {
          ss += vanleer(S[q + Nphi] - S[q], S[q] - S[q - Nphi]) * fac;
          su += vanleer(U[q + Nphi] - U[q], U[q] - U[q - Nphi]) * fac;
          sw += vanleer(W[q + Nphi] - W[q], W[q] - W[q - Nphi]) * fac;
        }
        fm = vf * ss * gtab(geom, D, GT_AF)[f];
        fu = fm * su;
        fw = fm * sw;
      }
    } else {
      const int i = f;
      const int jm = wrapj(j - 1, Nphi);
      const i64 row = (i64)i * Nphi;
      const double vf = 0.5 * (W[row + jm] + W[row + j]);
      int cj; double fac;
      if (vf > 0) { cj = jm; fac = 0.5 * (1.0 - vf * dt * gtab(geom, D, GT_INVDL)[i]); }
      else { cj = j; fac = -0.5 * (1.0 + vf * dt * gtab(geom, D, GT_INVDL)[i]); }
      const int cm = wrapj(cj - 1, Nphi), cp = wrapj(cj + 1, Nphi);
      const double ss = S[row + cj] + vanleer(S[row + cp] - S[row + cj], S[row + cj] - S[row + cm]) * fac;
      const double su = U[row + cj] + vanleer(U[row + cp] - U[row + cj], U[row + cj] - U[row + cm]) * fac;
      const double sw = W[row + cj] + vanleer(W[row + cp] - W[row + cj], W[row + cj] - W[row + cm]) * fac;
      fm = vf * ss * gtab(geom, D, GT_AFACE)[i];
      fu = fm * su;
      fw = fm * sw;
    }
    Fm[c] = fm; FU[c] = fu; FW[c] = fw;
  }
}

__global__ void k_gather(const DiskDev *__restrict__ Dp, double *__restrict__ st, i64 ntot, int cur,
                         double dt, int dir, const double *__restrict__ geom,
                         const double *__restrict__ F) {
  const DiskDev &D = *Dp;
  const int Nu = D.Nu, Nphi = D.Nphi;
  const i64 N = (i64)Nu * Nphi;
  double *S = stpw(st, ntot, cur, 0, D), *U = stpw(st, ntot, cur, 1, D), *W = stpw(st, ntot, cur, 2, D);
  const i64 NF = (i64)(dir == 0 ? Nu + 1 : Nu) * Nphi;
  const double *Fm = F, *FU = F + NF, *FW = F + 2 * NF;
  for (i64 c = blockIdx.x * (i64)blockDim.x + threadIdx.x; c < N; c += (i64)gridDim.x * blockDim.x) {
    const int i = (int)(c / Nphi), j = (int)(c - (i64)i * Nphi);
    const i64 f0 = c;
    const i64 f1 = (dir == 0) ? c + Nphi : (i64)i * Nphi + wrapj(j + 1, Nphi);
    const double s = S[c];
    const double mR = s * U[c], mP = s * W[c];
    const double iv = dt * gtab(geom, D, GT_INVVOL)[i];
    const double ns = s + (Fm[f0] - Fm[f1]) * iv;
    const double nmR = mR + (FU[f0] - FU[f1]) * iv;
    const double nmP = mP + (FW[f0] - FW[f1]) * iv;
    const double rn = 1.0 / smax(ns, 1e-10);
    S[c] = ns; U[c] = nmR * rn; W[c] = nmP * rn;
  }
}

__global__ void k_floor(const DiskDev *__restrict__ Dp, double *__restrict__ st, i64 ntot, int cur) {
  const DiskDev &D = *Dp;
  const i64 N = (i64)D.Nu * D.Nphi;
  double *S = stpw(st, ntot, cur, 0, D);
  for (i64 c = blockIdx.x * (i64)blockDim.x + threadIdx.x; c < N; c += (i64)gridDim.x * blockDim.x)
    S[c] = smax(S[c], 1e-10);
}

// ------------------------------------------------------------------ host launchers
// (plain C++ wrappers so the runtime file stays free of kernel syntax)
#include "dfk_launch.h"

// The in-place group kernel for nthr threads, batched or not.
template <class K>
K pick_tg(int nthr, int batched, K a0, K a1, K b0, K b1, K c0, K c1, K d0, K d1) {
  if (nthr == 64) return batched ? a1 : a0;
  if (nthr == 128) return batched ? b1 : b0;
  if (nthr == 192) return batched ? c1 : c0;
  return batched ? d1 : d0;
}
#define PICK_GROUP(K)                                                                   \
  pick_tg(nthr, batched, K<64, false>, K<64, true>, K<128, false>, K<128, true>,        \
          K<192, false>, K<192, true>, K<256, false>, K<256, true>)

#define CHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return e_; } while (0)

hipError_t dfk_launch_control(hipStream_t s, int ndisk, const DiskDev *disks, DiskCtl *ctl,
                              const double *pcfl, const double *pcm, DfkSampleDev *samples, int cap,
                              int par, DiskCtl *mirror) {
  if (ndisk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_control, dim3(ndisk), dim3(WG), 0, s, disks, ctl, pcfl, pcm, samples, cap,
                     par, mirror);
  return hipGetLastError();
}

hipError_t dfk_launch_diag(hipStream_t s, int nblk, const DiskDev *disks, const int2 *map, int base,
                           const DiskCtl *ctl, const double *st, i64 ntot, const double *geom,
                           const double *trig, double *pcfl, double *pcm) {
  if (nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_diag, dim3(nblk), dim3(WG), 0, s, disks, map, base, ctl, st, ntot, geom, trig,
                     pcfl, pcm);
  return hipGetLastError();
}

hipError_t dfk_launch_r2c(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                          const int2 *map, int base, DiskCtl *ctl, const double *st, i64 ntot,
                          const double *geom, const double2 *tw, double2 *spec, int par,
                          int nctl, int ctl0, const double *pcfl, const double *pcm,
                          DfkSampleDev *samples, int cap, int inplace, DiskCtl *mirror) {
  if (nblk + nctl <= 0) return hipSuccess;
  if (inplace == 5) {
    hipLaunchKernelGGL(k_rows_r2c_reg3, dim3(nblk + nctl), dim3(64), lds, s, disks, map, base, ctl, st,
                       ntot, geom, tw, spec, par, nblk, ctl0, pcfl, pcm, samples, cap, mirror);
    return hipGetLastError();
  }
  if (inplace >= 3) {
    auto k = inplace == 3 ? k_rows_r2c_reg<512> : k_rows_r2c_reg<1024>;
    hipLaunchKernelGGL(k, dim3(nblk + nctl), dim3(64), lds, s, disks, map, base, ctl, st, ntot, geom,
                       tw, spec, par, nblk, ctl0, pcfl, pcm, samples, cap, mirror);
    return hipGetLastError();
  }
  if (inplace) {
    const int batched = inplace == 2;
    auto k = PICK_GROUP(k_rows_r2c_group);
    if (nthr != 64 && nthr != 128 && nthr != 192 && nthr != 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(nblk + nctl), dim3(nthr), lds, s, disks, map, base, ctl, st,
                       ntot, geom, tw, spec, par, nblk, ctl0, pcfl, pcm, samples, cap, mirror);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_rows_r2c, dim3(nblk + nctl), dim3(nthr), lds, s, disks, map, base, ctl, st,
                     ntot, geom, tw, spec, par, nblk, ctl0, pcfl, pcm, samples, cap, mirror);
  return hipGetLastError();
}

#define CONV_CTL_ARGS nblk, c.ctl0, c.pcfl, c.pcm, c.samples, c.cap, c.par, c.mirror

hipError_t dfk_launch_conv(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                           const int2 *map, int base, DiskCtl *ctl, const double *khat,
                           const double2 *tw, double2 *spec, const DfkConvCtl &c) {
  if (nblk + c.nctl <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cols_conv, dim3(nblk + c.nctl), dim3(nthr), lds, s, disks, map, base, ctl, khat,
                     tw, spec, CONV_CTL_ARGS);
  return hipGetLastError();
}

hipError_t dfk_launch_conv_wave(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                                const int2 *map, int base, DiskCtl *ctl, const double *khat,
                                const double2 *tw, double2 *spec, const DfkConvCtl &c) {
  if (nblk + c.nctl <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cols_conv_wave, dim3(nblk + c.nctl), dim3(nthr), lds, s, disks, map, base, ctl,
                     khat, tw, spec, CONV_CTL_ARGS);
  return hipGetLastError();
}

hipError_t dfk_launch_conv_group(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                                 const int2 *map, int base, DiskCtl *ctl, const double *khat,
                                 const double2 *tw, double2 *spec, int batched, const DfkConvCtl &c) {
  if (nblk + c.nctl <= 0) return hipSuccess;
  auto k = PICK_GROUP(k_cols_conv_group);
  if (nthr != 64 && nthr != 128 && nthr != 192 && nthr != 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(nblk + c.nctl), dim3(nthr), lds, s, disks, map, base, ctl, khat, tw, spec,
                     CONV_CTL_ARGS);
  return hipGetLastError();
}

// (N1, N2, H) of the two-pass register convolution for transform length M, or 0s
static void conv_reg_split(int M, int *n1, int *n2, int *h) {
  *h = 2;
  switch (M) {
    case 128: *n1 = 8; *n2 = 16; *h = 1; return;
    case 192: *n1 = 12; *n2 = 16; *h = 1; return;
    case 256: *n1 = 16; *n2 = 16; *h = 1; return;
    case 384: *n1 = 12; *n2 = 32; return;
    case 512: *n1 = 16; *n2 = 32; return;
    case 768: *n1 = 24; *n2 = 32; return;
    case 1024: *n1 = 32; *n2 = 32; return;
    default: *n1 = *n2 = *h = 0;
  }
}

int dfk_conv_reg_cols(int M) {
  int n1, n2, h;
  conv_reg_split(M, &n1, &n2, &h);
  return n2 ? 64 / n2 : 0;
}

int dfk_conv_reg3_ok(int M) { return M == 256 || M == 384 || M == 512 || M == 768 || M == 1024; }

int dfk_rows_reg3_ok(int n) { return n == 128 || n == 192 || n == 256 || n == 384 || n == 512; }

size_t dfk_conv_reg3_lds(int M) {
  const int e = M == 1024 ? rfft::conv3::lds_elems<16>()
              : M == 768  ? rfft::conv3::lds_elems<12>()
                          : rfft::conv3::lds_elems<8>();
  return (size_t)e * sizeof(double2);
}

size_t dfk_rows_reg3_lds() { return (size_t)rfft::conv3::lds_elems<8>() * sizeof(double2); }

size_t dfk_conv_reg_lds(int M) {
  int n1, n2, h;
  conv_reg_split(M, &n1, &n2, &h);
  if (!n2) return 0;
  const int a = n1 * (n2 / h + 1), b = n2 * (n1 / h + 1);
  return (size_t)(64 / n2) * (a > b ? a : b) * sizeof(double2);
}

// Row transforms by the register kernels: the same (N1, N2) splits as the
// column transforms of those lengths.
int dfk_rows_reg_pairs(int n) {
  int n1, n2, h;
  conv_reg_split(n, &n1, &n2, &h);
  return n2 ? 64 / n2 : 0;
}

size_t dfk_rows_reg_lds(int n) {
  int n1, n2, h;
  conv_reg_split(n, &n1, &n2, &h);
  if (!n2) return 0;
  const int a = n1 * (n2 + 1), b = n1 * n2;
  return (size_t)(64 / n2) * (a > b ? a : b) * sizeof(double2);
}

hipError_t dfk_launch_conv_reg(hipStream_t s, int nblk, size_t lds, const DiskDev *disks,
                               const int2 *map, int base, DiskCtl *ctl, const double *khat,
                               const double2 *tw, double2 *spec, const DfkConvCtl &c) {
  if (nblk + c.nctl <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cols_conv_reg, dim3(nblk + c.nctl), dim3(64), lds, s, disks, map, base, ctl,
                     khat, tw, spec, CONV_CTL_ARGS);
  return hipGetLastError();
}

hipError_t dfk_launch_rows_fused(hipStream_t s, int nblk, int nthr, size_t lds, int g2,
                                 const DiskDev *disks, const int2 *map, int base, const DiskCtl *ctl,
                                 double *st, i64 ntot, const double2 *vin, double2 *sout,
                                 const double *geom, const double2 *tw, const double *trig,
                                 double *pcfl, double *pcm, int par) {
  if (nblk <= 0) return hipSuccess;
  auto k = nthr == 256 ? (g2 ? k_rows_fused<256, true> : k_rows_fused<256, false>)
         : nthr == 512 ? (g2 ? k_rows_fused<512, true> : k_rows_fused<512, false>)
                       : (g2 ? k_rows_fused<1024, true> : k_rows_fused<1024, false>);
  if (nthr != 256 && nthr != 512 && nthr != 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(nblk), dim3(nthr), lds, s, disks, map, base, ctl, st, ntot, vin, sout, geom,
                     tw, trig, pcfl, pcm, par);
  return hipGetLastError();
}

size_t dfk_fused_lds(int ti, int nphi, int tw_lds) {
  return (size_t)(fused_tile_doubles(ti, nphi) + fused_stage_doubles(ti, nphi)) * sizeof(double) +
         (size_t)(ti / 2 + 4 + (tw_lds ? 1 : 0)) * nphi * sizeof(double2);
}

int dfk_fused_max_nphi(int nthr) { return ip_nmax<256>() * (nthr / 256); }

hipError_t dfk_launch_c2r(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                          const int2 *map, int base, const DiskCtl *ctl, const double2 *spec,
                          const double *geom, const double2 *tw, double *phi, int inplace) {
  if (nblk <= 0) return hipSuccess;
  if (inplace == 5) {
    hipLaunchKernelGGL(k_rows_c2r_reg3, dim3(nblk), dim3(64), lds, s, disks, map, base, ctl, spec, geom,
                       tw, phi);
    return hipGetLastError();
  }
  if (inplace >= 3) {
    auto k = inplace == 3 ? k_rows_c2r_reg<512> : k_rows_c2r_reg<1024>;
    hipLaunchKernelGGL(k, dim3(nblk), dim3(64), lds, s, disks, map, base, ctl, spec, geom, tw, phi);
    return hipGetLastError();
  }
  if (inplace) {
    const int batched = inplace == 2;
    auto k = PICK_GROUP(k_rows_c2r_group);
    if (nthr != 64 && nthr != 128 && nthr != 192 && nthr != 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(nblk), dim3(nthr), lds, s, disks, map, base, ctl, spec, geom, tw, phi);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_rows_c2r, dim3(nblk), dim3(nthr), lds, s, disks, map, base, ctl, spec, geom, tw,
                     phi);
  return hipGetLastError();
}

hipError_t dfk_launch_update(hipStream_t s, int nblk, int nthr, size_t lds, int garr, int g2,
                             const DiskDev *disks,
                             const int2 *map, int base, const DiskCtl *ctl, double *st, i64 ntot,
                             const double *phi, const double *gR, const double *gp, int has_gR,
                             int has_gp, const double *geom, const double *trig, double *pcfl,
                             double *pcm, int par) {
  if (nblk <= 0) return hipSuccess;
  auto k = garr ? (g2 ? k_update<true, true> : k_update<true, false>)
                : (g2 ? k_update<false, true> : k_update<false, false>);
  hipLaunchKernelGGL(k, dim3(nblk), dim3(nthr), lds, s, disks, map, base, ctl, st, ntot, phi, gR, gp,
                     has_gR, has_gp, geom, trig, pcfl, pcm, par);
  return hipGetLastError();
}

hipError_t dfk_launch_update_stream(hipStream_t s, int nblk, int garr, int g2, const DiskDev *disks,
                                    const int2 *map, int base, const DiskCtl *ctl, double *st,
                                    i64 ntot, const double *phi, const double *gR, const double *gp,
                                    int has_gR, int has_gp, const double *geom, const double *trig,
                                    double *pcfl, double *pcm, int par) {
  if (nblk <= 0) return hipSuccess;
  auto k = garr ? (g2 ? k_update_stream<true, true> : k_update_stream<true, false>)
                : (g2 ? k_update_stream<false, true> : k_update_stream<false, false>);
  hipLaunchKernelGGL(k, dim3(nblk), dim3(64), 0, s, disks, map, base, ctl, st, ntot, phi, gR, gp,
                     has_gR, has_gp, geom, trig, pcfl, pcm, par);
  return hipGetLastError();
}

hipError_t dfk_launch_khat(hipStream_t s, const DiskDev *D, int nrow_blk, int ncol_blk,
                           size_t lds_rows, size_t lds_cols, const double *ktab, const double2 *tw,
                           double2 *tmp, double *khat) {
  hipLaunchKernelGGL(k_khat_rows, dim3(nrow_blk), dim3(WG), lds_rows, s, D, ktab, tw, tmp);
  CHK(hipGetLastError());
  hipLaunchKernelGGL(k_khat_cols, dim3(ncol_blk), dim3(WG), lds_cols, s, D, tw, tmp, khat);
  return hipGetLastError();
}

static inline int nblocks(i64 n) {
  i64 b = (n + WG - 1) / WG;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t dfk_launch_forces(hipStream_t s, const DiskDev *D, i64 N, const double *phi,
                             const double *geom, double *gR, double *gp) {
  hipLaunchKernelGGL(k_forces, dim3(nblocks(N)), dim3(WG), 0, s, D, phi, geom, gR, gp);
  return hipGetLastError();
}

hipError_t dfk_launch_stage(hipStream_t s, const DiskDev *D, i64 N, int stage, double *st, i64 ntot,
                            int cur, double dt, const double *gR, const double *gp,
                            const double *geom, double *F) {
  switch (stage) {
    case 1:
      hipLaunchKernelGGL(k_source, dim3(nblocks(N)), dim3(WG), 0, s, D, st, ntot, cur, dt, gR, gp, geom);
      break;
    case 2:
    case 3: {
      const int dir = stage - 2;
      hipLaunchKernelGGL(k_faces, dim3(nblocks(N + N)), dim3(WG), 0, s, D, st, ntot, cur, dt, dir, geom, F);
      CHK(hipGetLastError());
      hipLaunchKernelGGL(k_gather, dim3(nblocks(N)), dim3(WG), 0, s, D, st, ntot, cur, dt, dir, geom, F);
      break;
    }
    case 4:
      hipLaunchKernelGGL(k_floor, dim3(nblocks(N)), dim3(WG), 0, s, D, st, ntot, cur);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
