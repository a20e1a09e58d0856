// Host runtime of libdiskfft: the C ABI of include/diskfft.h over the HIP
// kernels of dfk_kernels.hip.
//
// Set-up (geometry tables, C_m trig tables, FFT twiddles, the Kalnajs kernel
// table) is computed on the host with libm using the reference's formulas,
// uploaded once, and stays resident; the K-hat transform itself runs on the
// device with the same LDS FFT as the time loop.  The time loop is device
// resident: dt, t, the output schedule and the C_m samples live in DiskCtl,
// and dfk_advance launches a captured HIP graph of whole steps (four kernels
// per step: row r2c with the loop control in trailing workgroups, column
// convolution, row c2r, update).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/diskfft.h"
#include "dfk_internal.h"
#include "dfk_launch.h"

typedef long long i64;

namespace {

thread_local std::string g_err;

void set_err(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      set_err("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      return DFK_EHIP;                                                             \
    }                                                                              \
  } while (0)

#define TRY(x)                  \
  do {                          \
    int r_ = (x);               \
    if (r_ != DFK_OK) return r_; \
  } while (0)

// Radix factorisation used by both transforms: 8s first, then 4, 2, and the
// odd primes in increasing order (3 unrolled in the kernels, every other
// prime by their direct-DFT stage).
bool factor(int n, int *rad, int *nst) {
  int m = n, k = 0;
  auto take = [&](int q) {
    while (m % q == 0) {
      if (k >= DFK_MAX_STAGES) return false;
      rad[k++] = q;
      m /= q;
    }
    return true;
  };
  if (!take(8) || !take(4) || !take(2)) return false;
  for (int q = 3; q * q <= m; q += 2)
    if (!take(q)) return false;
  if (m > 1) {   // the prime left over
    if (k >= DFK_MAX_STAGES) return false;
    rad[k++] = m;
  }
  *nst = k;
  return true;
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Whether a transform plan suits the in-place kernels: radices 2/3/4/8 only
// and at least `min_st` stages.
bool inplace_radices(const int *rad, int nst, int min_st) {
  bool ok = nst >= min_st;
  for (int q = 0; q < nst; ++q) ok = ok && (rad[q] == 2 || rad[q] == 3 || rad[q] == 4 || rad[q] == 8);
  return ok;
}

// Points the in-place workgroup kernels transform at once with t threads
// (the kernels' ip_nmax<t>()).
int ip_nmax(int t) { return std::max(8 * t, 1024); }

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  Reorder a
// work list so that XCD x receives one contiguous chunk of it: neighbouring
// tiles (which share halo rows) then meet in the same L2.
const int kXcds = 8;
// Largest grid: a transform of 4096 points and its ping-pong buffer take
// 128 KiB of LDS (the twiddles are then read through the caches).
const int kMaxNu = 2048, kMaxNphi = 4096;
// Largest max|Im K-hat| / max|Re K-hat| accepted at create (round-off of the
// transform of a real-even table is ~1e-16 x log2 of its size).
const double kKhatImagTol = 1e-13;
// Graph chains: up to 16 disk groups, on at most 4 streams (one per
// hardware queue of the process).
const int kMaxChainGroups = 16, kMaxChainStreams = 4;
const int kStreamCols = 58;   // output columns per wave of the streaming update (kernel's STREAM_OUT)
std::vector<int2> xcd_order(const std::vector<int2> &v) {
  const int L = (int)v.size();
  std::vector<int2> out(L);
  const int q = L / kXcds, r = L % kXcds;
  for (int b = 0; b < L; ++b) {
    const int x = b % kXcds;
    const int start = x * q + (x < r ? x : r);
    out[b] = v[start + b / kXcds];
  }
  return out;
}

}  // namespace

struct dfk_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  int nd = 0;
  std::vector<dfk_disk_params> P;
  std::vector<DiskDev> hD;
  std::vector<int> hcur;        // host mirror of DiskCtl::cur
  std::vector<double> mD;       // per-disk |C_m| normaliser of the current run
  DiskDev *dD = nullptr;
  i64 ntot = 0, nspec = 0, nkhat = 0, ngeom = 0, ntw = 0, ntrig = 0, maxN = 0, maxF = 0;
  double *st = nullptr;         // 2 buffers x 3 fields x ntot
  double *phi = nullptr, *gR = nullptr, *gp = nullptr, *F = nullptr;
  double2 *spec = nullptr, *tw = nullptr;
  double2 *spec2 = nullptr;     // second half spectrum of the fused row kernels (states in buffer 1)
  double *khat = nullptr, *geom = nullptr, *trig = nullptr;
  // layout (rebuilt when shapes change)
  bool layout_ok = false;
  double *pcfl = nullptr, *pcm = nullptr;
  i64 ntiles = 0;
  int2 *map_upd = nullptr, *map_r2c = nullptr, *map_conv = nullptr, *map_convw = nullptr,
       *map_convg = nullptr, *map_convr = nullptr;
  // the update's work list for the loop's launches, reordered per chain
  // group (loop_cb: the group bounds it was built for; empty: not built),
  // and the list in layout order it is built from
  int2 *map_loop = nullptr;
  std::vector<int> loop_cb;
  std::vector<int2> mu_host;
  std::vector<int> upd0, r2c0, conv0, convw0, convg0, convr0;   // per-disk first block (size nd+1)
  size_t lds_upd = 0, lds_r2c = 0, lds_conv = 0, lds_convw = 0, lds_convg = 0, lds_convr = 0;
  size_t lds_max = 64 * 1024;   // per workgroup, from the device at create (gfx950: 160 KiB)
  DiskCtl *ctl = nullptr;
  // host-visible copy of ctl (coherent pinned host memory; ctl_dev: its
  // device address) that the control of each captured run's last iteration
  // writes: dfk_advance reads the loop's state there after the run instead
  // of copying the control blocks back
  DiskCtl *ctl_host = nullptr, *ctl_dev = nullptr;
  // run
  bool run_ready = false;
  int cap = 0;
  DfkSampleDev *samples = nullptr;
  std::vector<int> drained;
  std::vector<std::vector<dfk_sample>> hsamples;
  std::map<int, hipGraphExec_t> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int chunk = 128;               // loop iterations per captured graph
  long long iter = 0;            // loop iterations launched since dfk_run_init
  int thr_rows = 256, thr_cols = 256, thr_upd = 256;   // threads per workgroup
  int upd_kind = 0;              // 0: LDS-tile update kernel, 1: register-streaming
  int nchains = 3;               // disk groups stepped as parallel chains of a graph
  int row_inplace = 0;           // requested row transforms: 0 cooperative, 1 in place
                                 // (k_rows_*_group), 2 register two-pass (k_rows_*_reg)
  int rows_ip = 0;               // in use (every self-gravitating disk's rows allow it)
  int rows_batched = 0;          // in place with pr > 1 for some disk
  int rows_big = 0;              // register rows with some Nphi > 512
  int conv_batched = 0;          // some disk's in-place workgroup convolution has cb > 1
  int fused_req = 0;             // requested: fused row kernels (row c2r + update + row r2c)
                                 // with this many threads per workgroup (0: off)
  int fused = 0;                 // in use (LDS tile update; every disk self-gravitating with
                                 // in-place row radices and a fitting tile)
  size_t lds_fused = 0;
  std::vector<hipStream_t> side; // streams of chains 1 .. nchains-1 (default priority)
  std::vector<hipEvent_t> join;  // their join events; ev_fork starts them
  hipEvent_t ev_fork = nullptr;
  std::string noalloc_err;      // why the no-allocation scratch is unavailable ("" if it is)
};

namespace {

int check_disk(dfk_handle *h, int d) {
  if (!h) { set_err("null handle"); return DFK_EINVAL; }
  if (d < 0 || d >= h->nd) { set_err("disk index %d out of range [0,%d)", d, h->nd); return DFK_EINVAL; }
  return DFK_OK;
}

void drop_graphs(dfk_handle *h) {
  for (auto &g : h->graphs) hipGraphExecDestroy(g.second);
  h->graphs.clear();
}

template <class T>
int dalloc(T **p, i64 n) {
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    set_err("hipMalloc(%lld bytes) failed: %s", (i64)(sizeof(T) * n), hipGetErrorString(e));
    return DFK_ENOMEM;
  }
  return DFK_OK;
}

// Layout of tiles, transform batches and workgroup maps for the current shapes.
int build_layout(dfk_handle *h) {
  HIPCHK(hipSetDevice(h->device));
  if (h->layout_ok) return DFK_OK;
  drop_graphs(h);
  h->run_ready = false;
  // update, rows, columns: cooperative, wave-local and workgroup in-place
  std::vector<int2> mu, mr, mc, mcw, mcg, mreg;
  h->convr0.assign(h->nd + 1, 0);
  h->upd0.assign(h->nd + 1, 0);
  h->r2c0.assign(h->nd + 1, 0);
  h->conv0.assign(h->nd + 1, 0);
  h->convw0.assign(h->nd + 1, 0);
  h->convg0.assign(h->nd + 1, 0);
  i64 part = 0;
  size_t lu = 0, lr = 0, lc = 0, lcw = 0, lcg = 0, lcr = 0;
  const size_t kMaxLds = h->lds_max;
  // the requested row transform where every self-gravitating disk allows it,
  // else the next simpler one (register -> in place -> cooperative)
  h->rows_ip = h->row_inplace;
  h->rows_big = 0;
  for (const DiskDev &D : h->hD)
    if (D.selfgrav && h->rows_ip == 3 && !dfk_rows_reg3_ok(D.Nphi)) h->rows_ip = 2;
  for (const DiskDev &D : h->hD)
    if (D.selfgrav) {
      if (h->rows_ip == 2 && dfk_rows_reg_pairs(D.Nphi) == 0) h->rows_ip = 1;
      h->rows_big |= D.Nphi > 512;
    }
  if (h->rows_ip == 1)
    for (const DiskDev &D : h->hD)
      if (D.selfgrav && (!inplace_radices(D.rad_row, D.nst_row, 1) || D.Nphi > ip_nmax(h->thr_rows)))
        h->rows_ip = 0;
  // fused row kernels: tiles of ti x Nphi with ti even (whole row pairs)
  h->fused = h->fused_req > 0 && h->upd_kind == 0;
  for (const DiskDev &D : h->hD) {
    const int tif = std::max(2, std::min(D.ti & ~1, D.Nu + (D.Nu & 1)));
    if (!D.selfgrav || !inplace_radices(D.rad_row, D.nst_row, 1) ||
        D.Nphi > dfk_fused_max_nphi(h->fused_req) || dfk_fused_lds(tif, D.Nphi, D.tw_lds & 1) > kMaxLds)
      h->fused = 0;
  }
  size_t lf = 0;
  for (int d = 0; d < h->nd; ++d) {
    DiskDev &D = h->hD[d];
    if (h->upd_kind == 1) D.tj = kStreamCols;   // one wave = 58 output columns
    if (h->fused) {
      D.tj = D.Nphi;
      D.ti = std::max(2, std::min(D.ti & ~1, D.Nu + (D.Nu & 1)));
    }
    D.tj = std::max(1, std::min(D.tj, D.Nphi));
    D.ti = std::max(1, std::min(D.ti, h->fused ? D.Nu + (D.Nu & 1) : D.Nu));
    if (h->fused) lf = std::max(lf, dfk_fused_lds(D.ti, D.Nphi, D.tw_lds & 1));
    D.ntile_i = ceil_div(D.Nu, D.ti);
    D.ntile_j = ceil_div(D.Nphi, D.tj);
    D.part = part;
    const int nt = D.ntile_i * D.ntile_j;
    part += nt;
    h->upd0[d] = (int)mu.size();
    if (h->upd_kind == 1) {
      // streaming units t = column group * strips + strip, listed strip by
      // strip: neighbouring column groups of a strip run side by side on
      // one XCD and share the lines at their common edge
      for (int k = 0; k < D.ntile_i; ++k)
        for (int g = 0; g < D.ntile_j; ++g) mu.push_back(make_int2(d, g * D.ntile_i + k));
    } else {
      for (int t = 0; t < nt; ++t) mu.push_back(make_int2(d, t));
    }
    // the tile, its face fluxes, the staged row records and cos/sin of its columns
    size_t l = (size_t)(3 * (D.ti + 4) * (D.tj + 4) + 3 * (D.ti + 1) * (D.tj + 4) +
                        (D.ti + 4) * RR_STRIDE + 2 * D.tj) * sizeof(double);
    if (h->upd_kind == 0) lu = std::max(lu, l);
    D.pr = std::max(1, std::min(D.pr, ceil_div(D.Nu, 2)));
    D.cb = std::max(1, std::min(D.cb, D.Nc));
    h->r2c0[d] = (int)mr.size();
    h->conv0[d] = (int)mc.size();
    h->convw0[d] = (int)mcw.size();
    h->convg0[d] = (int)mcg.size();
    h->convr0[d] = (int)mreg.size();
    if (D.selfgrav) {
      // batches shrink until they fit the LDS budget (a cached shape from a
      // build with smaller LDS needs stays usable)
      // ping-pong buffers (+ the twiddle table where staged in LDS)
      auto lds_rows = [&](int pr) {
        return (size_t)((h->rows_ip == 1 ? 1 : 2) * pr + (D.tw_lds & 1)) * D.Nphi * sizeof(double2);
      };
      // in-place columns need one buffer per column, cooperative ones two;
      // the in-place workgroup kernels take at most ip_nmax points at once
      auto lds_cols = [&](int cb) {
        return (size_t)((D.conv_wave ? 1 : 2) * cb + ((D.tw_lds >> 1) & 1)) * D.M * sizeof(double2);
      };
      while (h->rows_ip < 2 && D.pr > 1 &&
             (lds_rows(D.pr) > kMaxLds || (h->rows_ip == 1 && D.pr * D.Nphi > ip_nmax(h->thr_rows))))
        --D.pr;
      while (D.cb > 1 && (lds_cols(D.cb) > kMaxLds ||
                          (D.conv_wave == 2 && D.cb * D.M > ip_nmax(h->thr_cols))))
        --D.cb;
      // register rows: row pairs per workgroup fixed by Nphi, twiddles read
      // through the caches (the disk's own pr is kept for the other modes)
      const int prb = h->rows_ip == 3 ? 1 : (h->rows_ip == 2 ? dfk_rows_reg_pairs(D.Nphi) : D.pr);
      for (int g = 0; g < ceil_div(D.Nu, 2 * prb); ++g) mr.push_back(make_int2(d, g));
      lr = std::max(lr, h->rows_ip == 3 ? dfk_rows_reg3_lds()
                                        : (h->rows_ip == 2 ? dfk_rows_reg_lds(D.Nphi) : lds_rows(D.pr)));
      if (D.conv_wave >= 3) {   // columns per workgroup fixed by the transform length (one in mode 4)
        const int cpw = D.conv_wave == 4 ? 1 : dfk_conv_reg_cols(D.M);
        for (int g = 0; g < ceil_div(D.Nc, cpw); ++g) mreg.push_back(make_int2(d, g));
        lcr = std::max(lcr, D.conv_wave == 4 ? dfk_conv_reg3_lds(D.M) : dfk_conv_reg_lds(D.M));
      } else {
        std::vector<int2> &cm = D.conv_wave == 2 ? mcg : (D.conv_wave == 1 ? mcw : mc);
        for (int g = 0; g < ceil_div(D.Nc, D.cb); ++g) cm.push_back(make_int2(d, g));
        size_t &lcol = D.conv_wave == 2 ? lcg : (D.conv_wave == 1 ? lcw : lc);
        lcol = std::max(lcol, lds_cols(D.cb));
      }
    }
  }
  h->rows_batched = h->conv_batched = 0;
  for (const DiskDev &D : h->hD) {
    if (D.selfgrav && D.pr > 1 && h->rows_ip == 1) h->rows_batched = 1;
    if (D.selfgrav && D.conv_wave == 2 && D.cb > 1) h->conv_batched = 1;
  }
  h->upd0[h->nd] = (int)mu.size();
  h->r2c0[h->nd] = (int)mr.size();
  h->conv0[h->nd] = (int)mc.size();
  h->convw0[h->nd] = (int)mcw.size();
  h->convg0[h->nd] = (int)mcg.size();
  h->convr0[h->nd] = (int)mreg.size();
  if (lu > kMaxLds || lr > kMaxLds || lc > kMaxLds || lcw > kMaxLds || lcg > kMaxLds || lf > kMaxLds) {
    set_err("tile/batch shapes need %zu/%zu/%zu/%zu/%zu/%zu bytes of LDS (max %zu)", lu, lr, lc, lcw,
            lcg, lf, kMaxLds);
    return DFK_EINVAL;
  }
  h->lds_fused = lf;
  if (h->fused && !h->spec2) TRY(dalloc(&h->spec2, h->nspec));
  h->lds_upd = lu;
  h->lds_r2c = lr;
  h->lds_conv = lc;
  h->lds_convw = lcw;
  h->lds_convg = lcg;
  h->lds_convr = lcr;
  hipFree(h->pcfl); hipFree(h->pcm); hipFree(h->map_upd); hipFree(h->map_r2c); hipFree(h->map_conv);
  hipFree(h->map_convw); hipFree(h->map_convg); hipFree(h->map_convr); hipFree(h->map_loop);
  h->pcfl = h->pcm = nullptr;
  h->map_upd = h->map_r2c = h->map_conv = h->map_convw = h->map_convg = h->map_convr = nullptr;
  h->map_loop = nullptr;
  h->loop_cb.clear();
  h->mu_host = mu;
  h->ntiles = part;
  TRY(dalloc(&h->pcfl, part));
  TRY(dalloc(&h->pcm, 4 * part));
  TRY(dalloc(&h->map_upd, (i64)mu.size()));
  TRY(dalloc(&h->map_loop, (i64)mu.size()));
  TRY(dalloc(&h->map_r2c, (i64)mr.size()));
  TRY(dalloc(&h->map_conv, (i64)mc.size()));
  TRY(dalloc(&h->map_convw, (i64)mcw.size()));
  TRY(dalloc(&h->map_convg, (i64)mcg.size()));
  TRY(dalloc(&h->map_convr, (i64)mreg.size()));
  // unit calls launch one disk's range [upd0[d], upd0[d+1]) of the map, so
  // each disk's range is reordered on its own
  auto per_disk = [&](std::vector<int2> &m, const std::vector<int> &first) {
    for (int d = 0; d < h->nd; ++d) {
      std::vector<int2> part(m.begin() + first[d], m.begin() + first[d + 1]);
      part = xcd_order(part);
      std::copy(part.begin(), part.end(), m.begin() + first[d]);
    }
  };
  if (h->nd == 1) {
    mu = xcd_order(mu);
    mr = xcd_order(mr);
    mc = xcd_order(mc);
    mcw = xcd_order(mcw);
    mcg = xcd_order(mcg);
    mreg = xcd_order(mreg);
  } else {
    per_disk(mu, h->upd0);
    per_disk(mr, h->r2c0);
    per_disk(mc, h->conv0);
    per_disk(mcw, h->convw0);
    per_disk(mcg, h->convg0);
    per_disk(mreg, h->convr0);
  }
  HIPCHK(hipMemcpy(h->map_upd, mu.data(), sizeof(int2) * mu.size(), hipMemcpyHostToDevice));
  // (a valid list for any disk ranges until ensure_loop_map orders it)
  HIPCHK(hipMemcpy(h->map_loop, mu.data(), sizeof(int2) * mu.size(), hipMemcpyHostToDevice));
  if (!mr.empty()) HIPCHK(hipMemcpy(h->map_r2c, mr.data(), sizeof(int2) * mr.size(), hipMemcpyHostToDevice));
  if (!mc.empty()) HIPCHK(hipMemcpy(h->map_conv, mc.data(), sizeof(int2) * mc.size(), hipMemcpyHostToDevice));
  if (!mcw.empty())
    HIPCHK(hipMemcpy(h->map_convw, mcw.data(), sizeof(int2) * mcw.size(), hipMemcpyHostToDevice));
  if (!mcg.empty())
    HIPCHK(hipMemcpy(h->map_convg, mcg.data(), sizeof(int2) * mcg.size(), hipMemcpyHostToDevice));
  if (!mreg.empty())
    HIPCHK(hipMemcpy(h->map_convr, mreg.data(), sizeof(int2) * mreg.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->dD, h->hD.data(), sizeof(DiskDev) * h->nd, hipMemcpyHostToDevice));
  h->layout_ok = true;
  return DFK_OK;
}

// Host view of one disk's control block.
int read_ctl(dfk_handle *h, int d, DiskCtl *c) {
  HIPCHK(hipMemcpy(c, h->ctl + d, sizeof(DiskCtl), hipMemcpyDeviceToHost));
  return DFK_OK;
}
int write_ctl(dfk_handle *h, int d, const DiskCtl *c) {
  HIPCHK(hipMemcpy(h->ctl + d, c, sizeof(DiskCtl), hipMemcpyHostToDevice));
  return DFK_OK;
}

// Control block for a unit call: not done, dt as given.
int unit_ctl(dfk_handle *h, int d, double dt) {
  DiskCtl c;
  memset(&c, 0, sizeof c);
  c.cur = h->hcur[d];
  c.dt = dt;
  c.done = 0;
  h->run_ready = false;
  return write_ctl(h, d, &c);
}

double *field(dfk_handle *h, int buf, int f, int d) {
  return h->st + (i64)(buf * 3 + f) * h->ntot + h->hD[d].so;
}

// Launch variant of the row kernels (dfk_launch_r2c / dfk_launch_c2r).
int row_kernel(const dfk_handle *h) {
  if (h->rows_ip == 3) return 5;
  if (h->rows_ip == 2) return h->rows_big ? 4 : 3;
  return h->rows_ip ? 1 + h->rows_batched : 0;
}

// Row r2c of map range [base, base + nblk), with the loop control of disks
// ctl0 .. ctl0 + nctl - 1 in trailing workgroups of the same launch (nctl = 0:
// none), on stream s.
hipError_t launch_r2c(dfk_handle *h, hipStream_t s, int nblk, int base, int par, int nctl,
                      int ctl0 = 0, double2 *spec = nullptr, DiskCtl *mirror = nullptr) {
  return dfk_launch_r2c(s, nblk, h->thr_rows, h->lds_r2c, h->dD, h->map_r2c, base, h->ctl,
                        h->st, h->ntot, h->geom, h->tw, spec ? spec : h->spec, par, nctl, ctl0,
                        h->pcfl, h->pcm, h->samples, h->cap, row_kernel(h), mirror);
}

// The half spectrum that holds S-hat of the states in buffer par: with fused
// row kernels, buffer 1's states have theirs in spec2 (the fused kernel
// reads V-hat from one and writes S-hat of the new state to the other).
double2 *spec_of(dfk_handle *h, int par) { return (h->fused && par) ? h->spec2 : h->spec; }

// Row c2r of map range [base, base + nblk) on stream s.
hipError_t launch_c2r(dfk_handle *h, hipStream_t s, int nblk, int base) {
  return dfk_launch_c2r(s, nblk, h->thr_rows, h->lds_r2c, h->dD, h->map_r2c, base, h->ctl, h->spec,
                        h->geom, h->tw, h->phi, row_kernel(h));
}

// Column convolution of disks [a, b) on stream s, in place on `spec` (null:
// the handle's first spectrum): each of the three column kernels
// (cooperative, wave-local, workgroup in place) for the disks that use it.
// ctl_par >= 0: the last of these launches also runs the loop control of
// disks a .. b-1 (state buffer ctl_par) in trailing workgroups.
// ctl: the control blocks the kernels read (null: the handle's); mirror: as
// for the row r2c's control.
hipError_t launch_conv_range(dfk_handle *h, hipStream_t s, int a, int b, double2 *spec = nullptr,
                             int ctl_par = -1, DiskCtl *ctl = nullptr, DiskCtl *mirror = nullptr) {
  const int n = h->conv0[b] - h->conv0[a], nw = h->convw0[b] - h->convw0[a];
  const int ng = h->convg0[b] - h->convg0[a];
  if (!spec) spec = h->spec;
  if (!ctl) ctl = h->ctl;
  DfkConvCtl none, c;
  if (ctl_par >= 0) {
    c.nctl = b - a;
    c.ctl0 = a;
    c.pcfl = h->pcfl;
    c.pcm = h->pcm;
    c.samples = h->samples;
    c.cap = h->cap;
    c.par = ctl_par;
    c.mirror = mirror;
  }
  // the control rides in the last launch that has work (the cooperative one,
  // empty but for the control, when none has)
  const int nr = h->convr0[b] - h->convr0[a];
  const int last = nr > 0 ? 3 : (ng > 0 ? 2 : (nw > 0 ? 1 : 0));
  hipError_t e = dfk_launch_conv(s, n, h->thr_cols, h->lds_conv, h->dD, h->map_conv, h->conv0[a],
                                 ctl, h->khat, h->tw, spec, last == 0 ? c : none);
  if (e != hipSuccess) return e;
  e = dfk_launch_conv_wave(s, nw, h->thr_cols, h->lds_convw, h->dD, h->map_convw,
                           h->convw0[a], ctl, h->khat, h->tw, spec, last == 1 ? c : none);
  if (e != hipSuccess) return e;
  e = dfk_launch_conv_group(s, ng, h->thr_cols, h->lds_convg, h->dD, h->map_convg,
                            h->convg0[a], ctl, h->khat, h->tw, spec, h->conv_batched,
                            last == 2 ? c : none);
  if (e != hipSuccess) return e;
  return dfk_launch_conv_reg(s, nr, h->lds_convr, h->dD, h->map_convr, h->convr0[a], ctl, h->khat,
                             h->tw, spec, last == 3 ? c : none);
}

// Column convolution of disk d (all disks for d < 0) on the handle's stream.
hipError_t launch_conv(dfk_handle *h, int d) {
  return launch_conv_range(h, h->stream, d < 0 ? 0 : d, d < 0 ? h->nd : d + 1);
}

// Gravity pipeline (rows r2c -> column convolution -> rows c2r) for one disk,
// Sigma read from state buffer `par`.
int launch_gravity_disk(dfk_handle *h, int d, int par) {
  const int nr = h->r2c0[d + 1] - h->r2c0[d];
  HIPCHK(launch_r2c(h, h->stream, nr, h->r2c0[d], par, 0));
  HIPCHK(launch_conv(h, d));
  HIPCHK(launch_c2r(h, h->stream, nr, h->r2c0[d]));
  return DFK_OK;
}

// Update of map range [base, base + nblk) from state buffer par; the
// gamma = 2 kernels when every disk of the handle has gamma = 2.
hipError_t launch_update(dfk_handle *h, int nblk, int base, int garr, int has_gR, int has_gp, int par,
                         hipStream_t s = nullptr, const int2 *map = nullptr) {
  if (!s) s = h->stream;
  if (!map) map = h->map_upd;
  int g2 = 1;
  for (const DiskDev &D : h->hD) g2 &= (D.gamma2 != 0);
  if (h->upd_kind == 1)
    return dfk_launch_update_stream(s, nblk, garr, g2, h->dD, map, base, h->ctl, h->st,
                                    h->ntot, h->phi, h->gR, h->gp, has_gR, has_gp, h->geom,
                                    h->trig, h->pcfl, h->pcm, par);
  return dfk_launch_update(s, nblk, h->thr_upd, h->lds_upd, garr, g2, h->dD, map, base,
                           h->ctl, h->st, h->ntot, h->phi, h->gR, h->gp, has_gR, has_gp, h->geom,
                           h->trig, h->pcfl, h->pcm, par);
}

// Fused row kernels: two launches, the column convolution of S-hat in
// spec_of(par) (with the loop control in trailing workgroups) and the row
// kernel, which leaves S-hat of the new state in spec_of(1 - par).
int launch_step_fused(dfk_handle *h, hipStream_t s, int a, int b, int par, DiskCtl *mirror) {
  int g2 = 1;
  for (const DiskDev &D : h->hD) g2 &= (D.gamma2 != 0);
  HIPCHK(launch_conv_range(h, s, a, b, spec_of(h, par), par, nullptr, mirror));
  HIPCHK(dfk_launch_rows_fused(s, h->upd0[b] - h->upd0[a], h->fused_req, h->lds_fused, g2, h->dD,
                               h->map_upd, h->upd0[a], h->ctl, h->st, h->ntot, spec_of(h, par),
                               spec_of(h, 1 - par), h->geom, h->tw, h->trig, h->pcfl, h->pcm, par));
  return DFK_OK;
}

// One whole loop iteration for every disk (captured into the graph): four
// launches (two with fused row kernels).  Every running disk's state is in
// buffer `par`; the update writes buffer 1 - par.  The loop control rides in
// the r2c launch (a launch of its own when no disk has self-gravity).
// mirror: null, or the host-visible control copy that this iteration's
// control also writes (the last iteration of a captured run).
int launch_step_all(dfk_handle *h, int par, DiskCtl *mirror) {
  const int nu = h->upd0[h->nd], nr = h->r2c0[h->nd];
  if (h->fused) return launch_step_fused(h, h->stream, 0, h->nd, par, mirror);
  if (nr > 0) {
    HIPCHK(launch_r2c(h, h->stream, nr, 0, par, h->nd, 0, nullptr, mirror));
    HIPCHK(launch_conv(h, -1));
    HIPCHK(launch_c2r(h, h->stream, nr, 0));
  } else {
    HIPCHK(dfk_launch_control(h->stream, h->nd, h->dD, h->ctl, h->pcfl, h->pcm, h->samples, h->cap, par,
                              mirror));
  }
  HIPCHK(launch_update(h, nu, 0, 0, 0, 0, par, h->stream, h->map_loop));
  return DFK_OK;
}

// The same for the disks [a, b) on stream s (every one of them
// self-gravitating, so the control rides in their r2c launch).
int launch_step_range(dfk_handle *h, hipStream_t s, int a, int b, int par, DiskCtl *mirror) {
  if (h->fused) return launch_step_fused(h, s, a, b, par, mirror);
  const int nr = h->r2c0[b] - h->r2c0[a], nu = h->upd0[b] - h->upd0[a];
  HIPCHK(launch_r2c(h, s, nr, h->r2c0[a], par, b - a, a, nullptr, mirror));
  HIPCHK(launch_conv_range(h, s, a, b));
  HIPCHK(launch_c2r(h, s, nr, h->r2c0[a]));
  HIPCHK(launch_update(h, nu, h->upd0[a], 0, 0, 0, par, s, h->map_loop));
  return DFK_OK;
}

// Disk ranges of the graph chains: h->nchains contiguous groups of about
// equal zone counts (one chain when a disk has no self-gravity or there are
// fewer disks than chains).
std::vector<int> chain_bounds(dfk_handle *h) {
  int k = h->nchains;
  for (const DiskDev &D : h->hD) if (!D.selfgrav) k = 1;
  k = std::max(1, std::min(k, h->nd));
  std::vector<int> cut{0};
  i64 tot = 0, acc = 0;
  for (const DiskDev &D : h->hD) tot += (i64)D.Nu * D.Nphi;
  for (int d = 0; d < h->nd && (int)cut.size() < k; ++d) {
    acc += (i64)h->hD[d].Nu * h->hD[d].Nphi;
    // close the group once it holds its share, keeping a disk for every
    // later group, or when only one disk per later group is left
    const int left = h->nd - (d + 1), later = k - (int)cut.size();
    if (left == later || (left > later && acc * k >= tot * (i64)cut.size())) cut.push_back(d + 1);
  }
  cut.push_back(h->nd);
  return cut;
}

// The loop's update work list for the chain groups cb: each group's part of
// the layout-order list reordered on its own (one launch covers a group), so
// that XCD x runs one contiguous chunk of the group -- whole strips of
// several disks -- and neighbouring strips, which read each other's edge
// rows, mostly share an L2.  (map_upd is reordered per disk, for the calls
// on one disk; with the several disks of a loop launch that put most strip
// neighbours of a 128^2 .. 512^2 disk on different XCDs.)  A rebuild drops
// the cached graphs, which launch the old order.
int ensure_loop_map(dfk_handle *h, const std::vector<int> &cb) {
  if (h->loop_cb == cb) return DFK_OK;
  drop_graphs(h);
  std::vector<int2> m(h->mu_host.size());
  for (size_t g = 0; g + 1 < cb.size(); ++g) {
    const int a = h->upd0[cb[g]], b = h->upd0[cb[g + 1]];
    std::vector<int2> part(h->mu_host.begin() + a, h->mu_host.begin() + b);
    part = xcd_order(part);
    std::copy(part.begin(), part.end(), m.begin() + a);
  }
  if (!m.empty())
    HIPCHK(hipMemcpy(h->map_loop, m.data(), sizeof(int2) * m.size(), hipMemcpyHostToDevice));
  h->loop_cb = cb;
  return DFK_OK;
}

// Graph of `nsteps` iterations starting with state buffer par0 (cached).
int get_graph(dfk_handle *h, int nsteps, int par0, hipGraphExec_t *out) {
  const int key = 2 * nsteps + par0;
  auto it = h->graphs.find(key);
  if (it != h->graphs.end()) { *out = it->second; return DFK_OK; }
  const std::vector<int> cb = chain_bounds(h);
  TRY(ensure_loop_map(h, cb));
  const int ng = (int)cb.size() - 1;              // disk groups
  const int nc = std::min(ng, kMaxChainStreams);  // streams; group g runs on stream g % nc
  HIPCHK(hipSetDevice(h->device));
  while ((int)h->side.size() < nc - 1) {
    hipStream_t s;
    hipEvent_t ev;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    h->side.push_back(s);
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    h->join.push_back(ev);
  }
  if (nc > 1 && !h->ev_fork) HIPCHK(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = DFK_OK;
  if (nc == 1) {
    for (int s = 0; s < nsteps && rc == DFK_OK; ++s)
      rc = launch_step_all(h, (par0 + s) & 1, s == nsteps - 1 ? h->ctl_dev : nullptr);
  } else {
    // stream c steps the disk groups c, c + nc, ... (group g holds disks
    // [cb[g], cb[g+1])), each step group after group: the streams share
    // nothing but the launch, so one stream's kernels run while another's
    // drain, and a group's four launches follow each other while its
    // spectrum and potential are still in the caches
    auto chk = [&](hipError_t x) {
      if (x != hipSuccess && rc == DFK_OK) {
        set_err("graph chain: %s", hipGetErrorString(x));
        rc = DFK_EHIP;
      }
    };
    chk(hipEventRecord(h->ev_fork, h->stream));
    for (int c = 1; c < nc; ++c) chk(hipStreamWaitEvent(h->side[c - 1], h->ev_fork, 0));
    for (int c = 0; c < nc && rc == DFK_OK; ++c) {
      hipStream_t st = c == 0 ? h->stream : h->side[c - 1];
      for (int s = 0; s < nsteps && rc == DFK_OK; ++s)
        for (int g = c; g < ng && rc == DFK_OK; g += nc)
          rc = launch_step_range(h, st, cb[g], cb[g + 1], (par0 + s) & 1,
                                 s == nsteps - 1 ? h->ctl_dev : nullptr);
    }
    for (int c = 1; c < nc; ++c) {
      chk(hipEventRecord(h->join[c - 1], h->side[c - 1]));
      chk(hipStreamWaitEvent(h->stream, h->join[c - 1], 0));
    }
  }
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc != DFK_OK) {
    if (e == hipSuccess) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) { set_err("graph capture failed: %s", hipGetErrorString(e)); return DFK_EHIP; }
  hipGraphExec_t ex;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) { set_err("graph instantiate failed: %s", hipGetErrorString(e)); return DFK_EHIP; }
  h->graphs[key] = ex;
  *out = ex;
  return DFK_OK;
}

// Copies the samples produced since the last drain into the host buffers;
// c: the disks' control blocks as the run left them.
int drain(dfk_handle *h, const DiskCtl *c) {
  bool any = false;
  for (int d = 0; d < h->nd; ++d) any |= c[d].nsamples > h->drained[d];
  std::vector<DfkSampleDev> ring;
  if (any) {
    ring.resize((size_t)h->nd * h->cap);
    HIPCHK(hipMemcpy(ring.data(), h->samples, sizeof(DfkSampleDev) * ring.size(), hipMemcpyDeviceToHost));
  }
  for (int d = 0; d < h->nd; ++d) {
    h->hcur[d] = c[d].cur;
    const int n = c[d].nsamples, from = h->drained[d];
    if (n - from > h->cap) {
      set_err("sample ring overflow on disk %d (%d new, capacity %d)", d, n - from, h->cap);
      return DFK_ESTATE;
    }
    for (int q = from; q < n; ++q) {
      const DfkSampleDev s = ring[(size_t)d * h->cap + (q % h->cap)];
      dfk_sample o;
      const double mD = h->mD[d];
      o.t = s.t;
      o.c0 = std::sqrt(s.re0 * s.re0 + 0.0 * 0.0) / mD;    // diskmain.cpp:82 with im = 0 for m = 0
      o.c2 = std::sqrt(s.re2 * s.re2 + s.im2 * s.im2) / mD;
      o.ph2 = std::atan2(s.im2, s.re2);
      h->hsamples[d].push_back(o);
    }
    h->drained[d] = n;
  }
  return DFK_OK;
}

// Per-tile partials of disk d's current state, reduced on the host in tile
// order.  Read-only for a run in progress or finished: the diag kernel finds
// the state through the control block's `cur`, the only field written (the
// disk's clock, step count and output schedule stay), and the disk's tile
// partials, which the loop's next control reads, are restored afterwards.
int diag_disk(dfk_handle *h, int d, double *mx, double sums[4]) {
  TRY(build_layout(h));
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(&h->ctl[d].cur, &h->hcur[d], sizeof(int), hipMemcpyHostToDevice));
  const int nb = h->upd0[d + 1] - h->upd0[d];
  const DiskDev &D = h->hD[d];
  std::vector<double> pc(nb), pm(4 * (size_t)nb), pc0(nb), pm0(4 * (size_t)nb);
  HIPCHK(hipMemcpy(pc0.data(), h->pcfl + D.part, sizeof(double) * nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(pm0.data(), h->pcm + 4 * D.part, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost));
  HIPCHK(dfk_launch_diag(h->stream, nb, h->dD, h->map_upd, h->upd0[d], h->ctl, h->st, h->ntot,
                         h->geom, h->trig, h->pcfl, h->pcm));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(pc.data(), h->pcfl + D.part, sizeof(double) * nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(pm.data(), h->pcm + 4 * D.part, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(h->pcfl + D.part, pc0.data(), sizeof(double) * nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->pcm + 4 * D.part, pm0.data(), sizeof(double) * 4 * nb, hipMemcpyHostToDevice));
  double m = 1e-30;
  double s[4] = {0, 0, 0, 0};
  for (int t = 0; t < nb; ++t) {
    if (pc[t] > m) m = pc[t];
    for (int q = 0; q < 4; ++q) s[q] += pm[4 * t + q];
  }
  *mx = m;
  for (int q = 0; q < 4; ++q) sums[q] = s[q];
  return DFK_OK;
}

// Geometry / trig / twiddle tables and the K table on the host (libm).
void host_tables(const DiskDev &D, const dfk_disk_params &p, std::vector<double> &geom,
                 std::vector<double> &trig) {
  const int Nu = D.Nu, Nphi = D.Nphi, L = Nu + 1;
  geom.assign((size_t)DFK_GEOM_DOUBLES(Nu), 0.0);
  auto T = [&](int t, int i) -> double & { return geom[(size_t)t * L + i]; };
  const double umin = std::log(p.Rin);
  const double du = D.du, dphi = D.dphi;
  std::vector<double> R(Nu), Rf(Nu + 1);
  for (int i = 0; i <= Nu; ++i) Rf[i] = std::exp(umin + i * du);              // hydro.hpp:34
  for (int i = 0; i < Nu; ++i) R[i] = std::exp(umin + (i + 0.5) * du);        // hydro.hpp:36
  const double pref = -p.G * du * dphi / (static_cast<double>(D.M) * Nphi);   // kalnajs.hpp:58
  for (int i = 0; i <= Nu; ++i) {
    T(GT_RF, i) = Rf[i];
    T(GT_AF, i) = Rf[i] * dphi;                                                // hydro.hpp:113
    T(GT_AFZ, i) = (i >= 1 && i <= Nu - 1) ? T(GT_AF, i) : 0.0;               // hydro.hpp:111
    const bool inner = i >= 1 && i <= Nu - 2;                                  // hydro.hpp:118-123
    T(GT_HUP, i) = inner ? 1.0 : 0.0;
    T(GT_HDN, i) = inner ? -1.0 : 0.0;
  }
  for (int i = 0; i < Nu; ++i) {
    const int im = std::max(i - 1, 0), ip = std::min(i + 1, Nu - 1);
    const double vol = R[i] * (Rf[i + 1] - Rf[i]) * dphi;                       // hydro.hpp:37
    const double sq = std::sqrt(R[i]);
    T(GT_R, i) = R[i];
    T(GT_INVVOL, i) = 1.0 / vol;
    T(GT_INVDRC, i) = 1.0 / (R[ip] - R[im]);                                   // hydro.hpp:77-78
    T(GT_INVR, i) = 1.0 / R[i];
    T(GT_GSTAR, i) = -p.gmstar / (R[i] * R[i]);                                 // hydro.hpp:85
    T(GT_INVDR, i) = 1.0 / (Rf[i + 1] - Rf[i]);                                 // hydro.hpp:52
    T(GT_INVDL, i) = 1.0 / (R[i] * dphi);                                       // hydro.hpp:56,152
    T(GT_AFACE, i) = Rf[i + 1] - Rf[i];                                         // hydro.hpp:152
    T(GT_R32, i) = R[i] * sq;                                                   // kalnajs.hpp:33
    T(GT_PHISC, i) = pref / sq;                                                 // kalnajs.hpp:62
    T(GT_DA, i) = R[i] * (Rf[i + 1] - Rf[i]) * dphi;                            // diskmain.cpp:75
    T(GT_GRC, i) = (i == 0 || i == Nu - 1) ? 1.0 / (du * R[i]) : 1.0 / (2.0 * du * R[i]);  // kalnajs.hpp:74-77
    T(GT_GPC, i) = 1.0 / (2.0 * dphi * R[i]);                                   // kalnajs.hpp:72,78
    T(GT_DL, i) = R[i] * dphi;                                                  // hydro.hpp:56
  }
  // row records (the same values, gathered per row)
  for (int i = 0; i <= Nu; ++i) {
    double *r = &geom[(size_t)GT_COUNT * L + (size_t)i * RR_STRIDE];
    r[RR_AFZ] = T(GT_AFZ, i);
    if (i == Nu) continue;
    r[RR_INVDRC] = T(GT_INVDRC, i); r[RR_INVR] = T(GT_INVR, i); r[RR_GSTAR] = T(GT_GSTAR, i);
    r[RR_GRC] = T(GT_GRC, i); r[RR_GPC] = T(GT_GPC, i);
    r[RR_INVDR] = T(GT_INVDR, i); r[RR_HUP] = T(GT_HUP, i); r[RR_HDN] = T(GT_HDN, i);
    r[RR_INVVOL] = T(GT_INVVOL, i); r[RR_INVDL] = T(GT_INVDL, i); r[RR_AFACE] = T(GT_AFACE, i);
    r[RR_DA] = T(GT_DA, i); r[RR_DL] = T(GT_DL, i);
  }
  trig.assign((size_t)2 * Nphi, 0.0);
  for (int j = 0; j < Nphi; ++j) {
    const double phi = (j + 0.5) * dphi;                                        // diskmain.cpp:77
    trig[j] = std::cos(2 * phi);
    trig[Nphi + j] = std::sin(2 * phi);
  }
}

void twiddles(int n, std::vector<double2> &out) {
  out.resize(n);
  for (int t = 0; t < n; ++t) {
    const double a = 2.0 * M_PI * (double)t / (double)n;
    out[t] = make_double2(std::cos(a), -std::sin(a));
  }
}

// K(p,q) table of kalnajs.hpp:106-113, M x Nphi.
void ktable(const DiskDev &D, std::vector<double> &k) {
  const int M = D.M, Nphi = D.Nphi;
  k.resize((size_t)M * Nphi);
  for (int p = 0; p < M; ++p) {
    const int ioff = (p <= M / 2) ? p : p - M;
    const double Du = ioff * D.du;
    for (int q = 0; q < Nphi; ++q) {
      const double Dphi = q * D.dphi;
      k[(size_t)p * Nphi + q] = 1.0 / std::sqrt(2.0 * std::cosh(Du) - 2.0 * std::cos(Dphi) + D.eps);
    }
  }
}

// Runs the K-hat transform of disk d; keeps the real rows 0..Nu in h->khat and,
// if full != nullptr, returns the complex [Nc][M] transform.
int build_khat(dfk_handle *h, int d, std::vector<double2> *full) {
  DiskDev D = h->hD[d];
  // batch shapes for the M-row table: same pr / cb rules as the loop
  D.pr = std::max(1, std::min(1024 / std::max(D.Nphi, 1), ceil_div(D.M, 2)));
  D.cb = std::max(1, std::min(1024 / std::max(D.M, 1), D.Nc));
  D.tw_lds = 0;   // twiddles through the caches: the largest transforms fill the LDS
  std::vector<double> kt;
  ktable(D, kt);
  double *dk = nullptr;
  double2 *tmp = nullptr;
  DiskDev *dd = nullptr;
  TRY(dalloc(&dk, (i64)kt.size()));
  TRY(dalloc(&tmp, (i64)D.Nc * D.M));
  TRY(dalloc(&dd, 1));
  int rc = DFK_OK;
  hipError_t e = hipMemcpy(dk, kt.data(), sizeof(double) * kt.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dd, &D, sizeof D, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = dfk_launch_khat(h->stream, dd, ceil_div(D.M, 2 * D.pr), ceil_div(D.Nc, D.cb),
                        (size_t)(2 * D.pr) * D.Nphi * sizeof(double2),
                        (size_t)(2 * D.cb) * D.M * sizeof(double2), dk, h->tw, tmp, h->khat);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && full) {
    full->resize((size_t)D.Nc * D.M);
    e = hipMemcpy(full->data(), tmp, sizeof(double2) * full->size(), hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) {
    set_err("K-hat build failed: %s", hipGetErrorString(e));
    rc = DFK_EHIP;
  }
  hipFree(dk);
  hipFree(tmp);
  hipFree(dd);
  return rc;
}

}  // namespace

extern "C" {

const char *dfk_last_error(void) { return g_err.c_str(); }

int dfk_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int dfk_create(const dfk_disk_params *params, int ndisks, int device, dfk_handle **out) {
  g_err.clear();
  if (!params || ndisks <= 0 || !out) { set_err("dfk_create: bad arguments"); return DFK_EINVAL; }
  *out = nullptr;
  for (int d = 0; d < ndisks; ++d) {
    const dfk_disk_params &p = params[d];
    if (p.Nu < 3 || p.Nphi < 4) { set_err("disk %d: need Nu >= 3 and Nphi >= 4 (got %d x %d)", d, p.Nu, p.Nphi); return DFK_EINVAL; }
    if (!(p.Rin > 0) || !(p.Rout > p.Rin)) { set_err("disk %d: need 0 < Rin < Rout", d); return DFK_EINVAL; }
    if (p.Nu > kMaxNu || p.Nphi > kMaxNphi) {
      // one row (Nphi) or one padded column (2 Nu) and its ping-pong buffer
      // must fit a workgroup's LDS (160 KiB on gfx950)
      set_err("disk %d: need Nu <= %d and Nphi <= %d (got %d x %d)", d, kMaxNu, kMaxNphi, p.Nu,
              p.Nphi);
      return DFK_EINVAL;
    }
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    set_err("no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    return DFK_EHIP;
  }
  if (device < 0 || device >= ndev) { set_err("device %d out of range (%d devices)", device, ndev); return DFK_EINVAL; }
  dfk_handle *h = new dfk_handle();
  h->device = device;
  h->nd = ndisks;
  h->P.assign(params, params + ndisks);
  h->hD.resize(ndisks);
  h->hcur.assign(ndisks, 0);
  h->mD.assign(ndisks, 1.0);
  h->drained.assign(ndisks, 0);
  h->hsamples.resize(ndisks);
  auto fail = [&](int rc) { dfk_destroy(h); return rc; };
  if (hipSetDevice(device) != hipSuccess) { set_err("hipSetDevice failed"); return fail(DFK_EHIP); }
  int lmax = 0;
  if (hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
      lmax > 0)
    h->lds_max = (size_t)lmax;
  {
    // the no-allocation entry point's scratch address on this device (its
    // failure only disables that entry point, which then reports it)
    char msg[256] = "";
    if (dfk_noalloc_init(device, msg, sizeof msg)) h->noalloc_err = msg;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    set_err("stream create failed");
    return fail(DFK_EHIP);
  }
  hipEventCreate(&h->ev0);
  hipEventCreate(&h->ev1);

  // descriptors, pool offsets, twiddle dedup
  std::map<int, i64> twoff;
  std::vector<double2> twall;
  i64 so = 0, spo = 0, kho = 0, go = 0, co = 0;
  for (int d = 0; d < ndisks; ++d) {
    const dfk_disk_params &p = params[d];
    DiskDev &D = h->hD[d];
    memset(&D, 0, sizeof D);
    D.Nu = p.Nu; D.Nphi = p.Nphi; D.Nc = p.Nphi / 2 + 1; D.M = 2 * p.Nu;
    D.gamma2 = (p.gamma == 2.0);
    D.selfgrav = p.selfgrav ? 1 : 0;
    D.K = p.K; D.gamma = p.gamma; D.gmstar = p.gmstar; D.G = p.G; D.eps = p.eps;
    D.du = (std::log(p.Rout) - std::log(p.Rin)) / p.Nu;   // hydro.hpp:31-32
    D.dphi = 2.0 * M_PI / p.Nphi;
    D.inv2dphi = 1.0 / (2.0 * D.dphi);
    if (D.selfgrav && (!factor(D.Nphi, D.rad_row, &D.nst_row) || !factor(D.M, D.rad_col, &D.nst_col))) {
      set_err("disk %d: transform lengths %d and %d have more than %d factors", d, D.Nphi, D.M,
              DFK_MAX_STAGES);
      return fail(DFK_EUNSUPPORTED);
    }
    for (int n : {D.Nphi, D.M}) {
      if (!D.selfgrav) break;
      if (!twoff.count(n)) {
        std::vector<double2> t;
        twiddles(n, t);
        twoff[n] = (i64)twall.size();
        twall.insert(twall.end(), t.begin(), t.end());
      }
    }
    D.two_row = twoff[D.Nphi];
    D.two_col = twoff[D.M];
    D.so = so; so += (i64)p.Nu * p.Nphi;
    D.spo = spo; spo += (i64)p.Nu * D.Nc;
    D.kho = kho; kho += (i64)(p.Nu + 1) * D.Nc;
    D.go = go; go += DFK_GEOM_DOUBLES(p.Nu);
    D.cos_off = (int)co; co += 2 * (i64)p.Nphi;
    h->maxN = std::max(h->maxN, (i64)p.Nu * p.Nphi);
    h->maxF = std::max(h->maxF, (i64)(p.Nu + 1) * p.Nphi);
  }
  h->ntot = so; h->nspec = spo; h->nkhat = kho; h->ngeom = go; h->ntrig = co; h->ntw = (i64)twall.size();
  // Default shapes from the autotuner's sweeps on MI355X: a handle with few
  // zones is launch/latency bound and wants many small workgroups; a large
  // (batched) one is bandwidth bound and wants bigger tiles and batches.
  const bool small = h->ntot < (1LL << 19);
  for (int d = 0; d < ndisks; ++d) {
    DiskDev &D = h->hD[d];
    // twiddles of the row and column transforms staged in LDS, where a
    // transform, its ping-pong buffer and the table fit
    D.tw_lds = 0;
    if ((size_t)3 * D.Nphi * sizeof(double2) <= h->lds_max) D.tw_lds |= 1;
    if ((size_t)3 * D.M * sizeof(double2) <= h->lds_max) D.tw_lds |= 2;
    D.conv_wave = 0;
    if (small) {
      D.ti = 8; D.tj = 16; D.pr = 1; D.cb = 1;
    } else {
      D.ti = 16; D.tj = 32;
      D.pr = (D.Nphi <= 192) ? 4 : (D.Nphi <= 512 ? 2 : 1);
      D.cb = (D.M <= 384) ? 2 : 1;
    }
  }
  int rc;
  if ((rc = dalloc(&h->dD, ndisks)) || (rc = dalloc(&h->st, 6 * h->ntot)) ||
      (rc = dalloc(&h->phi, h->ntot)) || (rc = dalloc(&h->gR, h->ntot)) ||
      (rc = dalloc(&h->gp, h->ntot)) || (rc = dalloc(&h->F, 3 * h->maxF)) ||
      (rc = dalloc(&h->spec, h->nspec)) || (rc = dalloc(&h->khat, h->nkhat)) ||
      (rc = dalloc(&h->geom, h->ngeom)) || (rc = dalloc(&h->trig, h->ntrig)) ||
      (rc = dalloc(&h->tw, h->ntw)) || (rc = dalloc(&h->ctl, ndisks)))
    return fail(rc);
  if (hipHostMalloc((void **)&h->ctl_host, sizeof(DiskCtl) * ndisks, hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void **)&h->ctl_dev, h->ctl_host, 0) != hipSuccess) {
    set_err("pinned host control copy: hipHostMalloc / hipHostGetDevicePointer failed");
    return fail(DFK_ENOMEM);
  }
  if (hipMemset(h->st, 0, sizeof(double) * 6 * h->ntot) != hipSuccess ||
      hipMemset(h->phi, 0, sizeof(double) * h->ntot) != hipSuccess ||
      hipMemset(h->ctl, 0, sizeof(DiskCtl) * ndisks) != hipSuccess) {
    set_err("hipMemset failed");
    return fail(DFK_EHIP);
  }
  {
    std::vector<double> geom, trig, gall((size_t)h->ngeom), tall((size_t)h->ntrig);
    for (int d = 0; d < ndisks; ++d) {
      host_tables(h->hD[d], params[d], geom, trig);
      std::copy(geom.begin(), geom.end(), gall.begin() + h->hD[d].go);
      std::copy(trig.begin(), trig.end(), tall.begin() + h->hD[d].cos_off);
    }
    if (hipMemcpy(h->geom, gall.data(), sizeof(double) * gall.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->trig, tall.data(), sizeof(double) * tall.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->tw, twall.data(), sizeof(double2) * twall.size(), hipMemcpyHostToDevice) != hipSuccess) {
      set_err("table upload failed");
      return fail(DFK_EHIP);
    }
  }
  if ((rc = build_layout(h))) return fail(rc);
  for (int d = 0; d < ndisks; ++d) {
    if (!params[d].selfgrav) continue;
    // The loop multiplies by the real part of K-hat only.  K(du, dphi) is
    // real and even in both arguments, so its transform is real up to
    // round-off; check that on the transform just built.
    std::vector<double2> full;
    if ((rc = build_khat(h, d, &full))) return fail(rc);
    double re = 0.0, im = 0.0;
    for (const double2 &v : full) {
      re = std::max(re, std::fabs(v.x));
      im = std::max(im, std::fabs(v.y));
    }
    if (!(im <= kKhatImagTol * re)) {
      set_err("disk %d: K-hat is not real (max|Im| / max|Re| = %.3g > %.0e)", d, im / re,
              kKhatImagTol);
      return fail(DFK_EHIP);
    }
  }
  *out = h;
  return DFK_OK;
}

void dfk_destroy(dfk_handle *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  drop_graphs(h);
  hipFree(h->dD); hipFree(h->st); hipFree(h->phi); hipFree(h->gR); hipFree(h->gp); hipFree(h->F);
  hipFree(h->spec); hipFree(h->spec2); hipFree(h->khat); hipFree(h->geom); hipFree(h->trig); hipFree(h->tw);
  hipFree(h->pcfl); hipFree(h->pcm); hipFree(h->map_upd); hipFree(h->map_r2c); hipFree(h->map_conv); hipFree(h->map_convw);
  hipFree(h->map_convg); hipFree(h->map_convr); hipFree(h->map_loop);
  hipFree(h->ctl); hipFree(h->samples);
  if (h->ctl_host) hipHostFree(h->ctl_host);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  for (hipStream_t s : h->side) hipStreamDestroy(s);
  for (hipEvent_t e : h->join) hipEventDestroy(e);
  if (h->ev_fork) hipEventDestroy(h->ev_fork);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

int dfk_num_disks(const dfk_handle *h) { return h ? h->nd : 0; }

int dfk_set_eos(dfk_handle *h, int d, double K, double gamma) {
  TRY(check_disk(h, d));
  DiskDev &D = h->hD[d];
  D.K = K;
  D.gamma = gamma;
  D.gamma2 = (gamma == 2.0);
  h->P[d].K = K;
  h->P[d].gamma = gamma;
  h->layout_ok = false;   // re-uploads the descriptors
  return build_layout(h);
}

int dfk_geometry(dfk_handle *h, int d, double *R, double *Rf, double *Vol) {
  TRY(check_disk(h, d));
  const dfk_disk_params &p = h->P[d];
  const DiskDev &D = h->hD[d];
  const double umin = std::log(p.Rin);
  for (int i = 0; i <= p.Nu; ++i) {
    const double rf = std::exp(umin + i * D.du);
    if (Rf) Rf[i] = rf;
  }
  for (int i = 0; i < p.Nu; ++i) {
    const double r = std::exp(umin + (i + 0.5) * D.du);
    if (R) R[i] = r;
    if (Vol) Vol[i] = r * (std::exp(umin + (i + 1) * D.du) - std::exp(umin + i * D.du)) * D.dphi;
  }
  return DFK_OK;
}

int dfk_set_state(dfk_handle *h, int d, const double *Sig, const double *vR, const double *vp) {
  TRY(check_disk(h, d));
  if (!Sig || !vR || !vp) { set_err("null state pointer"); return DFK_EINVAL; }
  HIPCHK(hipSetDevice(h->device));
  const size_t n = (size_t)h->hD[d].Nu * h->hD[d].Nphi * sizeof(double);
  const int c = h->hcur[d];
  HIPCHK(hipMemcpy(field(h, c, 0, d), Sig, n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(field(h, c, 1, d), vR, n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(field(h, c, 2, d), vp, n, hipMemcpyHostToDevice));
  h->run_ready = false;
  return DFK_OK;
}

int dfk_get_state(dfk_handle *h, int d, double *Sig, double *vR, double *vp) {
  TRY(check_disk(h, d));
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  const size_t n = (size_t)h->hD[d].Nu * h->hD[d].Nphi * sizeof(double);
  const int c = h->hcur[d];
  if (Sig) HIPCHK(hipMemcpy(Sig, field(h, c, 0, d), n, hipMemcpyDeviceToHost));
  if (vR) HIPCHK(hipMemcpy(vR, field(h, c, 1, d), n, hipMemcpyDeviceToHost));
  if (vp) HIPCHK(hipMemcpy(vp, field(h, c, 2, d), n, hipMemcpyDeviceToHost));
  return DFK_OK;
}

int dfk_timestep(dfk_handle *h, int d, double cfl, double *dt) {
  TRY(check_disk(h, d));
  if (!dt) { set_err("null dt"); return DFK_EINVAL; }
  double mx, s[4];
  TRY(diag_disk(h, d, &mx, s));
  *dt = cfl / mx;
  return DFK_OK;
}

int dfk_modes(dfk_handle *h, int d, int m, double mD, double *mag, double *phase) {
  TRY(check_disk(h, d));
  if (m != 0 && m != 2) { set_err("dfk_modes: m must be 0 or 2 (the driver's diagnostics)"); return DFK_EINVAL; }
  double mx, s[4];
  TRY(diag_disk(h, d, &mx, s));
  const double re = (m == 0) ? s[0] : s[1], im = (m == 0) ? 0.0 : s[2];
  if (mag) *mag = std::sqrt(re * re + im * im) / mD;
  if (phase) *phase = std::atan2(im, re);
  return DFK_OK;
}

int dfk_disk_mass(dfk_handle *h, int d, double *mD) {
  TRY(check_disk(h, d));
  double mx, s[4];
  TRY(diag_disk(h, d, &mx, s));
  if (mD) *mD = s[0];
  return DFK_OK;
}

int dfk_potential(dfk_handle *h, int d, const double *Sigma, double *Phi) {
  TRY(check_disk(h, d));
  if (!Sigma || !Phi) { set_err("null pointer"); return DFK_EINVAL; }
  if (!h->hD[d].selfgrav) { set_err("disk %d was created with selfgrav = 0 (no K-hat)", d); return DFK_ESTATE; }
  TRY(build_layout(h));
  const int scratch = 1 - h->hcur[d];
  const size_t n = (size_t)h->hD[d].Nu * h->hD[d].Nphi * sizeof(double);
  HIPCHK(hipMemcpy(field(h, scratch, 0, d), Sigma, n, hipMemcpyHostToDevice));
  TRY(unit_ctl(h, d, 0.0));
  TRY(launch_gravity_disk(h, d, scratch));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(Phi, h->phi + h->hD[d].so, n, hipMemcpyDeviceToHost));
  return DFK_OK;
}

int dfk_potential_noalloc(dfk_handle *h, int d, const double *Sigma, double *Phi, void *stream) {
  TRY(check_disk(h, d));
  if (!Sigma || !Phi) { set_err("null pointer"); return DFK_EINVAL; }
  const DiskDev &D = h->hD[d];
  if (!D.selfgrav) { set_err("disk %d was created with selfgrav = 0 (no K-hat)", d); return DFK_ESTATE; }
  if (!h->layout_ok) { set_err("dfk_potential_noalloc: the handle's last shape change failed"); return DFK_ESTATE; }
  if ((i64)D.Nu * D.Nc > DFK_NOALLOC_MAX_SPECTRUM) {
    set_err("dfk_potential_noalloc: disk %d needs %lld spectrum entries, the scratch holds %d",
            d, (i64)D.Nu * D.Nc, DFK_NOALLOC_MAX_SPECTRUM);
    return DFK_EINVAL;
  }
  DiskCtl *ctl = (DiskCtl *)dfk_noalloc_ctl(h->device);
  double2 *spec = (double2 *)dfk_noalloc_spectrum(h->device);
  if (!ctl || !spec) {
    set_err("dfk_potential_noalloc: scratch unavailable (%s)", h->noalloc_err.c_str());
    return DFK_EHIP;
  }
  HIPCHK(hipSetDevice(h->device));
  hipStream_t s = (hipStream_t)stream;   // NULL: the device's null stream
  // The kernels address disk d's fields at its pool offsets (state D.so,
  // spectrum D.spo) and its control block at ctl[d]; these bases put those
  // addresses on the caller's arrays, the scratch spectrum and the scratch
  // control block (all zero: not done).  Address arithmetic only.
  auto shift = [](const void *p, i64 bytes) { return (uintptr_t)p - (uintptr_t)bytes; };
  const double *st = (const double *)shift(Sigma, D.so * (i64)sizeof(double));
  double *phi = (double *)shift(Phi, D.so * (i64)sizeof(double));
  double2 *spb = (double2 *)shift(spec, D.spo * (i64)sizeof(double2));
  DiskCtl *cb = (DiskCtl *)shift(ctl, (i64)d * (i64)sizeof(DiskCtl));
  const int nr = h->r2c0[d + 1] - h->r2c0[d];
  const int ip = row_kernel(h);
  HIPCHK(dfk_launch_r2c(s, nr, h->thr_rows, h->lds_r2c, h->dD, h->map_r2c, h->r2c0[d], cb, st, 0,
                        h->geom, h->tw, spb, 0, 0, 0, nullptr, nullptr, nullptr, 0, ip));
  HIPCHK(launch_conv_range(h, s, d, d + 1, spb, -1, cb));
  HIPCHK(dfk_launch_c2r(s, nr, h->thr_rows, h->lds_r2c, h->dD, h->map_r2c, h->r2c0[d], cb, spb,
                        h->geom, h->tw, phi, ip));
  return DFK_OK;
}

int dfk_forces(dfk_handle *h, int d, const double *Phi, double *gR, double *gp) {
  TRY(check_disk(h, d));
  if (!Phi || !gR || !gp) { set_err("null pointer"); return DFK_EINVAL; }
  HIPCHK(hipSetDevice(h->device));
  const DiskDev &D = h->hD[d];
  const size_t n = (size_t)D.Nu * D.Nphi * sizeof(double);
  HIPCHK(hipMemcpy(h->phi + D.so, Phi, n, hipMemcpyHostToDevice));
  HIPCHK(dfk_launch_forces(h->stream, h->dD + d, (i64)D.Nu * D.Nphi, h->phi, h->geom, h->gR, h->gp));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(gR, h->gR + D.so, n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(gp, h->gp + D.so, n, hipMemcpyDeviceToHost));
  return DFK_OK;
}

int dfk_step(dfk_handle *h, int d, double dt, const double *gR, const double *gp) {
  TRY(check_disk(h, d));
  TRY(build_layout(h));
  const DiskDev &D = h->hD[d];
  const size_t n = (size_t)D.Nu * D.Nphi * sizeof(double);
  if (gR) HIPCHK(hipMemcpy(h->gR + D.so, gR, n, hipMemcpyHostToDevice));
  if (gp) HIPCHK(hipMemcpy(h->gp + D.so, gp, n, hipMemcpyHostToDevice));
  const int c = h->hcur[d];
  TRY(unit_ctl(h, d, dt));
  const int nb = h->upd0[d + 1] - h->upd0[d];
  HIPCHK(launch_update(h, nb, h->upd0[d], 1, gR != nullptr, gp != nullptr, c));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->hcur[d] = 1 - c;
  TRY(unit_ctl(h, d, dt));   // records the new current buffer
  return DFK_OK;
}

int dfk_stage(dfk_handle *h, int d, int stage, double dt, const double *gR, const double *gp) {
  TRY(check_disk(h, d));
  if (stage < 1 || stage > 4) { set_err("stage must be 1..4"); return DFK_EINVAL; }
  HIPCHK(hipSetDevice(h->device));
  const DiskDev &D = h->hD[d];
  const size_t n = (size_t)D.Nu * D.Nphi * sizeof(double);
  if (stage == 1) {
    if (gR) HIPCHK(hipMemcpy(h->gR + D.so, gR, n, hipMemcpyHostToDevice));
    if (gp) HIPCHK(hipMemcpy(h->gp + D.so, gp, n, hipMemcpyHostToDevice));
  }
  h->run_ready = false;
  HIPCHK(dfk_launch_stage(h->stream, h->dD + d, (i64)D.Nu * D.Nphi, stage, h->st, h->ntot,
                          h->hcur[d], dt, (stage == 1 && gR) ? h->gR : nullptr,
                          (stage == 1 && gp) ? h->gp : nullptr, h->geom, h->F));
  HIPCHK(hipStreamSynchronize(h->stream));
  return DFK_OK;
}

int dfk_khat(dfk_handle *h, int d, double *re, double *im) {
  TRY(check_disk(h, d));
  if (!h->hD[d].selfgrav) { set_err("disk %d has no K-hat (selfgrav = 0)", d); return DFK_ESTATE; }
  HIPCHK(hipSetDevice(h->device));
  std::vector<double2> full;
  TRY(build_khat(h, d, &full));
  const DiskDev &D = h->hD[d];
  for (int p = 0; p < D.M; ++p)
    for (int k = 0; k < D.Nc; ++k) {
      const double2 v = full[(size_t)k * D.M + p];
      if (re) re[(size_t)p * D.Nc + k] = v.x;
      if (im) im[(size_t)p * D.Nc + k] = v.y;
    }
  return DFK_OK;
}

int dfk_run_init(dfk_handle *h, const dfk_run_params *rp, int sample_cap) {
  if (!h || !rp) { set_err("null argument"); return DFK_EINVAL; }
  TRY(build_layout(h));
  const int cap = std::max(sample_cap, 4);
  if (cap != h->cap || !h->samples) {
    hipFree(h->samples);
    h->samples = nullptr;
    TRY(dalloc(&h->samples, (i64)cap * h->nd));
    h->cap = cap;
    drop_graphs(h);   // graphs bake the sample buffer address and capacity
  }
  std::vector<DiskCtl> c(h->nd);
  for (int d = 0; d < h->nd; ++d) {
    memset(&c[d], 0, sizeof(DiskCtl));
    c[d].t = 0.0;
    c[d].tnext = 0.0;
    c[d].tlim = rp[d].tlim;
    c[d].dtout = rp[d].dtout;
    c[d].cfl = rp[d].cfl;
    c[d].max_steps = rp[d].max_steps;
    c[d].stop_on_nonfinite = rp[d].stop_on_nonfinite;
    // every disk starts the run with its state in buffer 0, so that iteration
    // g of the run reads buffer g & 1 for all running disks
    if (h->hcur[d] == 1) {
      const size_t nb = (size_t)h->hD[d].Nu * h->hD[d].Nphi * sizeof(double);
      for (int f = 0; f < 3; ++f)
        HIPCHK(hipMemcpyAsync(field(h, 0, f, d), field(h, 1, f, d), nb, hipMemcpyDeviceToDevice, h->stream));
      h->hcur[d] = 0;
    }
    c[d].cur = 0;
    h->mD[d] = rp[d].mD;
    h->drained[d] = 0;
    h->hsamples[d].clear();
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(h->ctl, c.data(), sizeof(DiskCtl) * h->nd, hipMemcpyHostToDevice));
  HIPCHK(dfk_launch_diag(h->stream, h->upd0[h->nd], h->dD, h->map_upd, 0, h->ctl, h->st, h->ntot,
                         h->geom, h->trig, h->pcfl, h->pcm));
  // fused row kernels: the first iteration's convolution needs S-hat of the
  // initial state (later ones get it from the previous row kernel)
  if (h->fused) HIPCHK(launch_r2c(h, h->stream, h->r2c0[h->nd], 0, 0, 0, 0, spec_of(h, 0)));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->iter = 0;
  h->run_ready = true;
  return DFK_OK;
}

int dfk_advance(dfk_handle *h, int nsteps, int *all_done) {
  if (!h) { set_err("null handle"); return DFK_EINVAL; }
  if (!h->run_ready) { set_err("dfk_advance before dfk_run_init (or after a unit call)"); return DFK_ESTATE; }
  if (nsteps < 0) { set_err("nsteps < 0"); return DFK_EINVAL; }
  HIPCHK(hipSetDevice(h->device));
  const int chunk_max = std::max(1, std::min(h->cap - 2, h->chunk));
  double total_ms = 0.0;
  int left = nsteps;
  while (left > 0) {
    const int chunk = std::min(left, chunk_max);
    hipGraphExec_t ex;
    TRY(get_graph(h, chunk, (int)(h->iter & 1), &ex));
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    HIPCHK(hipGraphLaunch(ex, h->stream));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    total_ms += ms;
    h->iter += chunk;
    // the run's last iteration wrote every disk's control block to ctl_host
    // before the event completed
    TRY(drain(h, h->ctl_host));
    left -= chunk;
  }
  h->last_ms = total_ms;
  if (all_done) {
    std::vector<DiskCtl> c;
    const DiskCtl *cc = h->ctl_host;
    if (nsteps == 0) {   // no run: the control blocks themselves
      c.resize(h->nd);
      HIPCHK(hipMemcpy(c.data(), h->ctl, sizeof(DiskCtl) * h->nd, hipMemcpyDeviceToHost));
      cc = c.data();
    }
    int ad = 1;
    for (int d = 0; d < h->nd; ++d) ad &= (cc[d].done != 0);
    *all_done = ad;
  }
  return DFK_OK;
}

int dfk_samples(dfk_handle *h, int d, dfk_sample *buf, int cap, int *n) {
  TRY(check_disk(h, d));
  std::vector<dfk_sample> &v = h->hsamples[d];
  const int k = std::min<int>(cap, (int)v.size());
  if (buf) std::copy(v.begin(), v.begin() + k, buf);
  v.erase(v.begin(), v.begin() + k);
  if (n) *n = k;
  return DFK_OK;
}

int dfk_get_status(dfk_handle *h, int d, dfk_status *st) {
  TRY(check_disk(h, d));
  if (!st) { set_err("null status"); return DFK_EINVAL; }
  HIPCHK(hipSetDevice(h->device));
  DiskCtl c;
  TRY(read_ctl(h, d, &c));
  st->t = c.t;
  st->steps = c.step;
  st->done = c.done;
  st->flags = c.flags;
  st->c2_final = std::sqrt(c.fin_re2 * c.fin_re2 + c.fin_im2 * c.fin_im2) / h->mD[d];
  st->ph2_final = std::atan2(c.fin_im2, c.fin_re2);
  return DFK_OK;
}

namespace {
// Snapshot header: the disk, its loop clock and every loop parameter of the
// saved run (so that a load continues that run, whatever dfk_run_init set).
struct SnapHeader {
  char magic[8];             // "DFKSNAP2"
  int Nu, Nphi;
  dfk_disk_params params;
  double t, tnext, tlim, dtout, cfl, mD;
  long long step, max_steps;
  int flags, stop_on_nonfinite;
};
const char kSnapMagic[8] = {'D', 'F', 'K', 'S', 'N', 'A', 'P', '2'};
}  // namespace

int dfk_snapshot_save(dfk_handle *h, int d, const char *path) {
  TRY(check_disk(h, d));
  if (!path) { set_err("null path"); return DFK_EINVAL; }
  if (!h->run_ready) { set_err("dfk_snapshot_save needs a run (dfk_run_init)"); return DFK_ESTATE; }
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  DiskCtl c;
  TRY(read_ctl(h, d, &c));
  SnapHeader hd;
  memset(&hd, 0, sizeof hd);
  memcpy(hd.magic, kSnapMagic, 8);
  hd.Nu = h->hD[d].Nu;
  hd.Nphi = h->hD[d].Nphi;
  hd.params = h->P[d];
  hd.t = c.t; hd.tnext = c.tnext; hd.tlim = c.tlim; hd.dtout = c.dtout; hd.cfl = c.cfl;
  hd.mD = h->mD[d];
  hd.step = c.step;
  hd.max_steps = c.max_steps;
  hd.flags = c.flags;
  hd.stop_on_nonfinite = c.stop_on_nonfinite;
  const size_t N = (size_t)hd.Nu * hd.Nphi;
  std::vector<double> buf(3 * N);
  h->hcur[d] = c.cur;
  TRY(dfk_get_state(h, d, buf.data(), buf.data() + N, buf.data() + 2 * N));
  FILE *f = fopen(path, "wb");
  if (!f) { set_err("cannot open %s for writing", path); return DFK_EINVAL; }
  const bool ok = fwrite(&hd, sizeof hd, 1, f) == 1 && fwrite(buf.data(), sizeof(double), 3 * N, f) == 3 * N;
  fclose(f);
  if (!ok) { set_err("short write to %s", path); return DFK_EINVAL; }
  return DFK_OK;
}

int dfk_snapshot_load(dfk_handle *h, int d, const char *path) {
  TRY(check_disk(h, d));
  if (!path) { set_err("null path"); return DFK_EINVAL; }
  if (!h->run_ready) { set_err("dfk_snapshot_load needs dfk_run_init first"); return DFK_ESTATE; }
  FILE *f = fopen(path, "rb");
  if (!f) { set_err("cannot open %s", path); return DFK_EINVAL; }
  SnapHeader hd;
  bool ok = fread(&hd, sizeof hd, 1, f) == 1 && memcmp(hd.magic, kSnapMagic, 8) == 0;
  const dfk_disk_params &p = h->P[d];
  if (ok && (hd.Nu != p.Nu || hd.Nphi != p.Nphi || hd.params.K != p.K || hd.params.eps != p.eps ||
             hd.params.gamma != p.gamma || hd.params.gmstar != p.gmstar || hd.params.Rin != p.Rin ||
             hd.params.Rout != p.Rout || hd.params.G != p.G || hd.params.selfgrav != p.selfgrav)) {
    fclose(f);
    set_err("snapshot %s is for a different disk (grid or parameters differ)", path);
    return DFK_EINVAL;
  }
  const size_t N = (size_t)p.Nu * p.Nphi;
  std::vector<double> buf(3 * N);
  ok = ok && fread(buf.data(), sizeof(double), 3 * N, f) == 3 * N;
  fclose(f);
  if (!ok) { set_err("%s is not a complete snapshot", path); return DFK_EINVAL; }
  HIPCHK(hipSetDevice(h->device));
  DiskCtl c;
  TRY(read_ctl(h, d, &c));
  const int cur = (int)(h->iter & 1);   // where running disks hold their state now
  h->hcur[d] = cur;
  c.cur = cur;
  const size_t nb = N * sizeof(double);
  HIPCHK(hipMemcpy(field(h, cur, 0, d), buf.data(), nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(field(h, cur, 1, d), buf.data() + N, nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(field(h, cur, 2, d), buf.data() + 2 * N, nb, hipMemcpyHostToDevice));
  c.t = hd.t;
  c.tnext = hd.tnext;
  c.step = hd.step;
  c.flags = hd.flags;
  c.tlim = hd.tlim;
  c.dtout = hd.dtout;
  c.cfl = hd.cfl;
  c.max_steps = hd.max_steps;
  c.stop_on_nonfinite = hd.stop_on_nonfinite;
  h->mD[d] = hd.mD;
  c.done = 0;
  // the disk's sample history restarts at the restored state (samples of the
  // state it replaces, e.g. run_init's t = 0 output, are dropped)
  c.nsamples = 0;
  h->drained[d] = 0;
  h->hsamples[d].clear();
  TRY(write_ctl(h, d, &c));
  // partials (CFL maximum, C_m sums) of the restored state
  const int nb_ = h->upd0[d + 1] - h->upd0[d];
  HIPCHK(dfk_launch_diag(h->stream, nb_, h->dD, h->map_upd, h->upd0[d], h->ctl, h->st, h->ntot,
                         h->geom, h->trig, h->pcfl, h->pcm));
  if (h->fused)   // and S-hat of the restored state, which the next convolution reads
    HIPCHK(launch_r2c(h, h->stream, h->r2c0[d + 1] - h->r2c0[d], h->r2c0[d], cur, 0, 0,
                      spec_of(h, cur)));
  HIPCHK(hipStreamSynchronize(h->stream));
  return DFK_OK;
}

int dfk_set_shapes(dfk_handle *h, int d, int ti, int tj, int pr, int cb) {
  TRY(check_disk(h, d));
  DiskDev &D = h->hD[d];
  if (ti > 0) D.ti = ti;
  if (tj > 0) D.tj = tj;
  if (pr > 0) D.pr = pr;
  if (cb > 0) D.cb = cb;
  h->layout_ok = false;
  return build_layout(h);
}

int dfk_set_fft_options(dfk_handle *h, int d, int rows_tw_lds, int cols_tw_lds, int cols_wave) {
  TRY(check_disk(h, d));
  DiskDev &D = h->hD[d];
  // twiddles in LDS only where a transform, its ping-pong buffer and the table fit
  const bool rfit = (size_t)3 * D.Nphi * sizeof(double2) <= h->lds_max;
  const bool cfit = (size_t)3 * D.M * sizeof(double2) <= h->lds_max;
  const int v = (rows_tw_lds && rfit ? 1 : 0) | (cols_tw_lds && cfit ? 2 : 0);
  // the in-place column transforms need radices 2/3/4/8, >= 2 stages, M <= 1024;
  // the two-pass register one a transform length it is built for
  const bool fast = D.M <= 1024 && inplace_radices(D.rad_col, D.nst_col, 2);
  int w = (fast && D.selfgrav && (cols_wave == 1 || cols_wave == 2)) ? cols_wave : 0;
  if (cols_wave == 3 && D.selfgrav && dfk_conv_reg_cols(D.M) > 0) w = 3;
  if (cols_wave == 4 && D.selfgrav && dfk_conv_reg3_ok(D.M)) w = 4;
  if (v != D.tw_lds || w != D.conv_wave) {
    D.tw_lds = v;
    D.conv_wave = w;
    h->layout_ok = false;
  }
  return build_layout(h);
}

int dfk_get_fft_options(dfk_handle *h, int d, int *rows_tw_lds, int *cols_tw_lds, int *cols_wave) {
  TRY(check_disk(h, d));
  if (rows_tw_lds) *rows_tw_lds = h->hD[d].tw_lds & 1;
  if (cols_tw_lds) *cols_tw_lds = (h->hD[d].tw_lds >> 1) & 1;
  if (cols_wave) *cols_wave = h->hD[d].conv_wave;
  return DFK_OK;
}

int dfk_get_shapes(dfk_handle *h, int d, int *ti, int *tj, int *pr, int *cb) {
  TRY(check_disk(h, d));
  const DiskDev &D = h->hD[d];
  if (ti) *ti = D.ti;
  if (tj) *tj = D.tj;
  if (pr) *pr = D.pr;
  if (cb) *cb = D.cb;
  return DFK_OK;
}

int dfk_last_advance_ms(dfk_handle *h, double *ms) {
  if (!h || !ms) { set_err("null argument"); return DFK_EINVAL; }
  *ms = h->last_ms;
  return DFK_OK;
}

int dfk_set_threads(dfk_handle *h, int rows, int cols, int update) {
  if (!h) { set_err("null handle"); return DFK_EINVAL; }
  for (int v : {rows, cols, update})
    if (v != 0 && (v < 64 || v > 256 || v % 64)) {
      set_err("threads per workgroup must be 64, 128, 192 or 256");
      return DFK_EINVAL;
    }
  if (rows) h->thr_rows = rows;
  if (cols) h->thr_cols = cols;
  if (update) h->thr_upd = update;
  h->layout_ok = false;   // the in-place transforms' batch limits depend on the threads
  return build_layout(h);
}

int dfk_set_row_transform(dfk_handle *h, int inplace) {
  if (!h || inplace < 0 || inplace > 3) {
    set_err("row transform must be 0 (cooperative), 1 (in place), 2 (register) or 3 (register, "
            "three passes)");
    return DFK_EINVAL;
  }
  if (inplace != h->row_inplace) {
    h->row_inplace = inplace;
    h->layout_ok = false;
    return build_layout(h);
  }
  return DFK_OK;
}

int dfk_get_row_transform(dfk_handle *h, int *inplace) {
  if (!h || !inplace) { set_err("null argument"); return DFK_EINVAL; }
  TRY(build_layout(h));
  *inplace = h->rows_ip;
  return DFK_OK;
}

int dfk_set_fused_rows(dfk_handle *h, int threads) {
  if (!h || (threads != 0 && threads != 256 && threads != 512 && threads != 1024)) {
    set_err("fused rows: threads per workgroup must be 0 (off), 256, 512 or 1024");
    return DFK_EINVAL;
  }
  if (threads != h->fused_req) {
    h->fused_req = threads;
    h->layout_ok = false;
    h->run_ready = false;   // the loop's spectra change meaning
    return build_layout(h);
  }
  return DFK_OK;
}

int dfk_get_fused_rows(dfk_handle *h, int *threads) {
  if (!h || !threads) { set_err("null argument"); return DFK_EINVAL; }
  TRY(build_layout(h));
  *threads = h->fused ? h->fused_req : 0;
  return DFK_OK;
}

int dfk_set_update_kernel(dfk_handle *h, int kind) {
  if (!h || kind < 0 || kind > 1) { set_err("update kernel must be 0 (LDS tile) or 1 (streaming)"); return DFK_EINVAL; }
  if (kind == 1 && 3 * h->ntot * (i64)sizeof(double) > DFK_STREAM_BYTES) {
    set_err("streaming update: the state of this handle (%lld zones) exceeds its 31-bit buffer offsets "
            "(at most %lld zones); use the LDS-tile update", (long long)h->ntot,
            (long long)(DFK_STREAM_BYTES / (3 * sizeof(double))));
    return DFK_EUNSUPPORTED;
  }
  if (kind != h->upd_kind) {
    h->upd_kind = kind;
    for (auto &D : h->hD) {
      D.tj = (kind == 1) ? kStreamCols : 32;
      if (kind == 0) D.ti = std::min(D.ti, 16);   // a strip height may not fit an LDS tile
    }
    h->layout_ok = false;
    return build_layout(h);
  }
  return DFK_OK;
}

int dfk_get_update_kernel(dfk_handle *h, int *kind) {
  if (!h || !kind) { set_err("null argument"); return DFK_EINVAL; }
  *kind = h->upd_kind;
  return DFK_OK;
}

int dfk_set_chains(dfk_handle *h, int k) {
  if (!h || k < 1 || k > kMaxChainGroups) {
    set_err("chains must be 1 .. %d", kMaxChainGroups);
    return DFK_EINVAL;
  }
  if (k != h->nchains) {
    h->nchains = k;
    drop_graphs(h);
  }
  return DFK_OK;
}

int dfk_get_chains(dfk_handle *h, int *k) {
  if (!h || !k) { set_err("null argument"); return DFK_EINVAL; }
  *k = (int)chain_bounds(h).size() - 1;
  return DFK_OK;
}

int dfk_get_threads(dfk_handle *h, int *rows, int *cols, int *update) {
  if (!h) { set_err("null handle"); return DFK_EINVAL; }
  if (rows) *rows = h->thr_rows;
  if (cols) *cols = h->thr_cols;
  if (update) *update = h->thr_upd;
  return DFK_OK;
}

int dfk_set_graph_steps(dfk_handle *h, int nsteps) {
  if (!h || nsteps < 1) { set_err("bad argument"); return DFK_EINVAL; }
  if (nsteps != h->chunk) drop_graphs(h);
  h->chunk = nsteps;
  return DFK_OK;
}

int dfk_profile_step(dfk_handle *h, int nrep, double *ms) {
  if (!h || !ms || nrep < 1) { set_err("bad argument"); return DFK_EINVAL; }
  if (!h->run_ready) { set_err("dfk_profile_step needs dfk_run_init"); return DFK_ESTATE; }
  HIPCHK(hipSetDevice(h->device));
  std::vector<DiskCtl> saved(h->nd);
  HIPCHK(hipMemcpy(saved.data(), h->ctl, sizeof(DiskCtl) * h->nd, hipMemcpyDeviceToHost));
  // running disks hold their state in buffer par and keep running; finished
  // disks stay skipped (their state may sit in the buffer the update writes)
  const int par = (int)(h->iter & 1);
  std::vector<DiskCtl> c = saved;
  for (int d = 0; d < h->nd; ++d) {
    if (!(c[d].dt > 0)) c[d].dt = 1e-6;
    c[d].tlim = 1e300;
    c[d].tnext = 1e300;   // no samples: the ring keeps the run's own
    c[d].max_steps = 0;
  }
  const int nu = h->upd0[h->nd], nr = h->r2c0[h->nd];
  hipEvent_t e0 = h->ev0, e1 = h->ev1;
  for (int k = 0; k < 5; ++k) {
    if (h->fused && (k == 1 || k == 3)) {   // no separate row transforms in the fused loop
      ms[k] = 0.0;
      continue;
    }
    HIPCHK(hipMemcpy(h->ctl, c.data(), sizeof(DiskCtl) * h->nd, hipMemcpyHostToDevice));
    HIPCHK(hipEventRecord(e0, h->stream));
    for (int r = 0; r < nrep; ++r) {
      switch (k) {
        case 0:
          HIPCHK(dfk_launch_control(h->stream, h->nd, h->dD, h->ctl, h->pcfl, h->pcm, h->samples, h->cap, par));
          break;
        case 1:   // as in the loop: with the control
          HIPCHK(launch_r2c(h, h->stream, nr, 0, par, h->nd));
          break;
        case 2:   // fused: with the control, on the spectrum of buffer par
          if (h->fused) HIPCHK(launch_conv_range(h, h->stream, 0, h->nd, spec_of(h, par), par));
          else HIPCHK(launch_conv(h, -1));
          break;
        case 3:
          HIPCHK(launch_c2r(h, h->stream, nr, 0));
          break;
        case 4:   // fused: the row kernel (writes the other state buffer and spectrum)
          if (h->fused) {
            int g2 = 1;
            for (const DiskDev &D : h->hD) g2 &= (D.gamma2 != 0);
            HIPCHK(dfk_launch_rows_fused(h->stream, nu, h->fused_req, h->lds_fused, g2, h->dD,
                                         h->map_upd, 0, h->ctl, h->st, h->ntot, spec_of(h, par),
                                         spec_of(h, 1 - par), h->geom, h->tw, h->trig, h->pcfl,
                                         h->pcm, par));
          } else {
            HIPCHK(launch_update(h, nu, 0, 0, 0, 0, par));
          }
          break;
      }
    }
    HIPCHK(hipEventRecord(e1, h->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    ms[k] = (double)t / nrep;
  }
  // the control kernel advanced t / nsamples: restore the loop position and
  // recompute the partials of the (unchanged) current state
  HIPCHK(hipMemcpy(h->ctl, saved.data(), sizeof(DiskCtl) * h->nd, hipMemcpyHostToDevice));
  HIPCHK(dfk_launch_diag(h->stream, nu, h->dD, h->map_upd, 0, h->ctl, h->st, h->ntot, h->geom,
                         h->trig, h->pcfl, h->pcm));
  // fused: the repeated convolutions transformed S-hat of buffer par in place
  if (h->fused) HIPCHK(launch_r2c(h, h->stream, nr, 0, par, 0, 0, spec_of(h, par)));
  HIPCHK(hipStreamSynchronize(h->stream));
  return DFK_OK;
}

}  // extern "C"
