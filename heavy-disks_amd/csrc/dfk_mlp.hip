// Surrogate MLP on MFMA: 4 -> 128 -> 128 -> 128 -> 3, GELU hidden layers,
// BF16 weights and activations with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16), an ensemble of `members` independent networks.
//
//   k_mlp_train  fused forward + loss + backward for 32 points per workgroup
//                (all activations in LDS), per-workgroup FP32 weight gradients
//   k_mlp_adam   reduces the workgroup gradients in a fixed order, Adam step on
//                the FP32 master weights, refreshes the BF16 compute copy
//   k_mlp_infer  fused forward pass for 64 points per workgroup
//
// Parameter layout per member (FP32 master, PyTorch convention W[out][in]):
//   W1[128][4] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[3][128] b4[3]
// BF16 compute copy, zero padded for 16 x 16 MFMA tiles:
//   W1p[128][16] W2[128][128] W3[128][128] W4p[16][128]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/diskfft_mlp.h"

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

#define H_ 128
#define NPAR 34051
#define OW1 0
#define OB1 512
#define OW2 640
#define OB2 17024
#define OW3 17152
#define OB3 33536
#define OW4 33664
#define OB4 34048
#define CW1 0
#define CW2 2048
#define CW3 18432
#define CW4 34816
#define CNW 36864
#define PT 32      // points per training workgroup
#define PI_ 64     // points per inference workgroup

// ------------------------------------------------------------------ device helpers
__device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 b) { return __uint_as_float(((unsigned)b) << 16); }

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// C[M x N] = A[M x K] B[K x N] on 16 x 16 tiles, one tile per wave at a time.
// la(m, k), lb(k, n) return BF16 operands; ep(m, n, value) consumes results.
// MFMA 16x16x16 bf16 lane layout: A row m = lane % 16, k = 4 (lane / 16) + j;
// B column n = lane % 16, k = 4 (lane / 16) + j; D row 4 (lane / 16) + j,
// column lane % 16.
template <class LA, class LB, class EP>
__device__ __forceinline__ void gemm16(int M, int N, int K, LA la, LB lb, EP ep) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int ntn = N / 16, ntiles = (M / 16) * ntn;
  const int r = lane & 15, q = lane >> 4;
  for (int t = wave; t < ntiles; t += nw) {
    const int mt = t / ntn, nt = t - mt * ntn;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
      v4s a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = (short)la(mt * 16 + r, k0 + 4 * q + j);
        b[j] = (short)lb(k0 + 4 * q + j, nt * 16 + r);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) ep(mt * 16 + 4 * q + j, nt * 16 + r, acc[j]);
  }
}

// ------------------------------------------------------------------ probe
// D = A B for one 16 x 16 x 16 tile (checks the lane layout above).
__global__ void k_mfma_probe(const float *A, const float *B, float *D) {
  gemm16(16, 16, 16, [&](int m, int k) { return f2bf(A[m * 16 + k]); },
         [&](int k, int n) { return f2bf(B[k * 16 + n]); },
         [&](int m, int n, float v) { D[m * 16 + n] = v; });
}

// ------------------------------------------------------------------ training
struct TrainLds {
  bf16 X[PT][16];
  bf16 Z1[PT][H_], Z2[PT][H_], Z3[PT][H_];
  bf16 Hh[PT][H_];
  bf16 G[PT][H_];
  float Dz[PT][H_];
  float Y[PT][16];
  bf16 Gy[PT][16];
  float red[4];
};

__global__ void __launch_bounds__(256)
k_mlp_train(const bf16 *__restrict__ wbf, const float *__restrict__ master,
            const float *__restrict__ X, const float *__restrict__ T, const int *__restrict__ idx,
            int n, int batch_off, int nbatch, float scale, float *__restrict__ gpart,
            float *__restrict__ lpart) {
  __shared__ TrainLds L;
  const int mem = blockIdx.y, wg = blockIdx.x, nwg = gridDim.x;
  const bf16 *W = wbf + (size_t)mem * CNW;
  const float *Mw = master + (size_t)mem * NPAR;
  float *gp = gpart + ((size_t)mem * nwg + wg) * NPAR;
  const int tid = threadIdx.x;
  // inputs (rows beyond the batch are zero and carry zero loss gradient)
  for (int e = tid; e < PT * 16; e += blockDim.x) {
    const int p = e >> 4, i = e & 15, b = wg * PT + p;
    float v = 0.f;
    if (i < 4 && b < nbatch) v = X[(size_t)idx[(size_t)mem * n + batch_off + b] * 4 + i];
    L.X[p][i] = f2bf(v);
  }
  __syncthreads();
  // ---- forward
  gemm16(PT, H_, 16, [&](int m, int k) { return L.X[m][k]; },
         [&](int k, int c) { return W[CW1 + c * 16 + k]; },
         [&](int m, int c, float v) { L.Z1[m][c] = f2bf(v + Mw[OB1 + c]); });
  __syncthreads();
  for (int e = tid; e < PT * H_; e += blockDim.x) L.Hh[e / H_][e % H_] = f2bf(gelu(bf2f(L.Z1[e / H_][e % H_])));
  __syncthreads();
  gemm16(PT, H_, H_, [&](int m, int k) { return L.Hh[m][k]; },
         [&](int k, int c) { return W[CW2 + c * H_ + k]; },
         [&](int m, int c, float v) { L.Z2[m][c] = f2bf(v + Mw[OB2 + c]); });
  __syncthreads();
  for (int e = tid; e < PT * H_; e += blockDim.x) L.Hh[e / H_][e % H_] = f2bf(gelu(bf2f(L.Z2[e / H_][e % H_])));
  __syncthreads();
  gemm16(PT, H_, H_, [&](int m, int k) { return L.Hh[m][k]; },
         [&](int k, int c) { return W[CW3 + c * H_ + k]; },
         [&](int m, int c, float v) { L.Z3[m][c] = f2bf(v + Mw[OB3 + c]); });
  __syncthreads();
  for (int e = tid; e < PT * H_; e += blockDim.x) L.Hh[e / H_][e % H_] = f2bf(gelu(bf2f(L.Z3[e / H_][e % H_])));
  __syncthreads();
  gemm16(PT, 16, H_, [&](int m, int k) { return L.Hh[m][k]; },
         [&](int k, int c) { return W[CW4 + c * H_ + k]; },
         [&](int m, int c, float v) { L.Y[m][c] = v + (c < 3 ? Mw[OB4 + c] : 0.f); });
  __syncthreads();
  // ---- loss: scale * sum (y - t)^2 over the batch; dL/dy
  if (tid < 4) L.red[tid] = 0.f;
  __syncthreads();
  for (int e = tid; e < PT * 16; e += blockDim.x) {
    const int p = e >> 4, o = e & 15, b = wg * PT + p;
    float g = 0.f;
    if (o < 3 && b < nbatch) {
      const float d = L.Y[p][o] - T[(size_t)idx[(size_t)mem * n + batch_off + b] * 3 + o];
      g = 2.f * scale * d;
      atomicAdd(&L.red[0], scale * d * d);
    }
    L.Gy[p][o] = f2bf(g);
  }
  __syncthreads();
  if (tid == 0) lpart[(size_t)mem * nwg + wg] = L.red[0];
  // ---- layer 4: dW4[o][i] = sum_p Gy[p][o] H3[p][i]; dH3 = Gy W4p
  gemm16(16, H_, PT, [&](int o, int p) { return L.Gy[p][o]; },
         [&](int p, int i) { return L.Hh[p][i]; },
         [&](int o, int i, float v) { if (o < 3) gp[OW4 + o * H_ + i] = v; });
  if (tid < 3) {
    float s = 0.f;
    for (int p = 0; p < PT; ++p) s += bf2f(L.Gy[p][tid]);
    gp[OB4 + tid] = s;
  }
  gemm16(PT, H_, 16, [&](int p, int o) { return L.Gy[p][o]; },
         [&](int o, int i) { return W[CW4 + o * H_ + i]; },
         [&](int p, int i, float v) { L.Dz[p][i] = v; });
  __syncthreads();
  for (int e = tid; e < PT * H_; e += blockDim.x) {
    const int p = e / H_, c = e % H_;
    L.G[p][c] = f2bf(L.Dz[p][c] * gelu_d(bf2f(L.Z3[p][c])));
    L.Hh[p][c] = f2bf(gelu(bf2f(L.Z2[p][c])));            // H2 for dW3
  }
  __syncthreads();
  // ---- layers 3, 2: dW = G^T H_prev, db = sum G, dH_prev = G W
  for (int layer = 3; layer >= 2; --layer) {
    const int ow = (layer == 3) ? OW3 : OW2, ob = (layer == 3) ? OB3 : OB2;
    const int cw = (layer == 3) ? CW3 : CW2;
    gemm16(H_, H_, PT, [&](int o, int p) { return L.G[p][o]; },
           [&](int p, int i) { return L.Hh[p][i]; },
           [&](int o, int i, float v) { gp[ow + o * H_ + i] = v; });
    if (tid < H_) {
      float s = 0.f;
      for (int p = 0; p < PT; ++p) s += bf2f(L.G[p][tid]);
      gp[ob + tid] = s;
    }
    gemm16(PT, H_, H_, [&](int p, int o) { return L.G[p][o]; },
           [&](int o, int i) { return W[cw + o * H_ + i]; },
           [&](int p, int i, float v) { L.Dz[p][i] = v; });
    __syncthreads();
    for (int e = tid; e < PT * H_; e += blockDim.x) {
      const int p = e / H_, c = e % H_;
      const bf16 zprev = (layer == 3) ? L.Z2[p][c] : L.Z1[p][c];
      L.G[p][c] = f2bf(L.Dz[p][c] * gelu_d(bf2f(zprev)));
      if (layer == 3) L.Hh[p][c] = f2bf(gelu(bf2f(L.Z1[p][c])));   // H1 for dW2
    }
    __syncthreads();
  }
  // ---- layer 1: dW1[o][i] = sum_p G[p][o] X[p][i] (i < 4)
  gemm16(H_, 16, PT, [&](int o, int p) { return L.G[p][o]; },
         [&](int p, int i) { return L.X[p][i]; },
         [&](int o, int i, float v) { if (i < 4) gp[OW1 + o * 4 + i] = v; });
  if (tid < H_) {
    float s = 0.f;
    for (int p = 0; p < PT; ++p) s += bf2f(L.G[p][tid]);
    gp[OB1 + tid] = s;
  }
}

// index of master parameter q in the padded BF16 copy (-1: bias, not mirrored)
__device__ __forceinline__ int compute_index(int q) {
  if (q < OB1) return CW1 + (q / 4) * 16 + (q % 4);
  if (q < OW2) return -1;
  if (q < OB2) return CW2 + (q - OW2);
  if (q < OW3) return -1;
  if (q < OB3) return CW3 + (q - OW3);
  if (q < OW4) return -1;
  if (q < OB4) return CW4 + (q - OW4);
  return -1;
}

__global__ void k_mlp_adam(float *__restrict__ master, float *__restrict__ mom,
                           float *__restrict__ vel, bf16 *__restrict__ wbf,
                           const float *__restrict__ gpart, int nwg, int members, float lr,
                           float b1, float b2, float eps, float bc1, float bc2,
                           float *__restrict__ gout) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x, mem = blockIdx.y;
  if (q >= NPAR || mem >= members) return;
  float g = 0.f;
  for (int w = 0; w < nwg; ++w) g += gpart[((size_t)mem * nwg + w) * NPAR + q];
  const size_t k = (size_t)mem * NPAR + q;
  if (gout) gout[k] = g;
  if (lr == 0.f) return;
  const float m = b1 * mom[k] + (1.f - b1) * g;
  const float v = b2 * vel[k] + (1.f - b2) * g * g;
  mom[k] = m;
  vel[k] = v;
  const float w = master[k] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  master[k] = w;
  const int ci = compute_index(q);
  if (ci >= 0) wbf[(size_t)mem * CNW + ci] = f2bf(w);
}

// ------------------------------------------------------------------ inference
struct InferLds {
  bf16 X[PI_][16];
  bf16 A[PI_][H_];
  bf16 B[PI_][H_];
};

__global__ void __launch_bounds__(256)
k_mlp_infer(const bf16 *__restrict__ wbf, const float *__restrict__ master,
            const float *__restrict__ X, int n, float *__restrict__ Y) {
  __shared__ InferLds L;
  const int mem = blockIdx.y, p0 = blockIdx.x * PI_;
  const bf16 *W = wbf + (size_t)mem * CNW;
  const float *Mw = master + (size_t)mem * NPAR;
  for (int e = threadIdx.x; e < PI_ * 16; e += blockDim.x) {
    const int p = e >> 4, i = e & 15;
    L.X[p][i] = f2bf((i < 4 && p0 + p < n) ? X[(size_t)(p0 + p) * 4 + i] : 0.f);
  }
  __syncthreads();
  gemm16(PI_, H_, 16, [&](int m, int k) { return L.X[m][k]; },
         [&](int k, int c) { return W[CW1 + c * 16 + k]; },
         [&](int m, int c, float v) { L.A[m][c] = f2bf(gelu(bf2f(f2bf(v + Mw[OB1 + c])))); });
  __syncthreads();
  gemm16(PI_, H_, H_, [&](int m, int k) { return L.A[m][k]; },
         [&](int k, int c) { return W[CW2 + c * H_ + k]; },
         [&](int m, int c, float v) { L.B[m][c] = f2bf(gelu(bf2f(f2bf(v + Mw[OB2 + c])))); });
  __syncthreads();
  gemm16(PI_, H_, H_, [&](int m, int k) { return L.B[m][k]; },
         [&](int k, int c) { return W[CW3 + c * H_ + k]; },
         [&](int m, int c, float v) { L.A[m][c] = f2bf(gelu(bf2f(f2bf(v + Mw[OB3 + c])))); });
  __syncthreads();
  gemm16(PI_, 16, H_, [&](int m, int k) { return L.A[m][k]; },
         [&](int k, int c) { return W[CW4 + c * H_ + k]; },
         [&](int m, int c, float v) {
           if (c < 3 && p0 + m < n) Y[((size_t)mem * n + p0 + m) * 3 + c] = v + Mw[OB4 + c];
         });
}

// ------------------------------------------------------------------ host runtime
namespace {
thread_local std::string g_mlp_err;
void merr(const char *fmt, ...) {
  char b[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b, sizeof b, fmt, ap);
  va_end(ap);
  g_mlp_err = b;
}
#define MCHK(x)                                                                       \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      merr("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__);   \
      return DFK_MLP_EHIP;                                                            \
    }                                                                                 \
  } while (0)
}  // namespace

struct dfk_mlp {
  int device = 0, members = 0;
  hipStream_t stream = nullptr;
  float *master = nullptr, *mom = nullptr, *vel = nullptr;
  bf16 *wbf = nullptr;
  long long step = 0;
};

namespace {
// BF16 copy (round to nearest even) of the master weights, on the host
unsigned short h_f2bf(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
void to_compute(const float *m, unsigned short *c) {
  memset(c, 0, sizeof(unsigned short) * CNW);
  for (int o = 0; o < H_; ++o)
    for (int i = 0; i < 4; ++i) c[CW1 + o * 16 + i] = h_f2bf(m[OW1 + o * 4 + i]);
  for (int q = 0; q < H_ * H_; ++q) {
    c[CW2 + q] = h_f2bf(m[OW2 + q]);
    c[CW3 + q] = h_f2bf(m[OW3 + q]);
  }
  for (int q = 0; q < 3 * H_; ++q) c[CW4 + q] = h_f2bf(m[OW4 + q]);
}
int upload(dfk_mlp *h, int mem, const float *params) {
  std::vector<unsigned short> c(CNW);
  to_compute(params, c.data());
  MCHK(hipMemcpy(h->master + (size_t)mem * NPAR, params, sizeof(float) * NPAR, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(h->wbf + (size_t)mem * CNW, c.data(), sizeof(unsigned short) * CNW, hipMemcpyHostToDevice));
  return DFK_MLP_OK;
}
}  // namespace

extern "C" {

const char *dfk_mlp_last_error(void) { return g_mlp_err.c_str(); }

int dfk_mlp_param_count(void) { return NPAR; }

int dfk_mlp_create(int members, int device, unsigned long long seed, dfk_mlp **out) {
  if (members < 1 || !out) { merr("bad arguments"); return DFK_MLP_EINVAL; }
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) { merr("no HIP device available"); return DFK_MLP_EHIP; }
  dfk_mlp *h = new dfk_mlp();
  h->device = device;
  h->members = members;
  *out = nullptr;
  const size_t np = (size_t)members * NPAR;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&h->master, np * 4) != hipSuccess || hipMalloc(&h->mom, np * 4) != hipSuccess ||
      hipMalloc(&h->vel, np * 4) != hipSuccess ||
      hipMalloc(&h->wbf, (size_t)members * CNW * 2) != hipSuccess ||
      hipMemset(h->mom, 0, np * 4) != hipSuccess || hipMemset(h->vel, 0, np * 4) != hipSuccess) {
    merr("device set-up failed");
    dfk_mlp_destroy(h);
    return DFK_MLP_EHIP;
  }
  // Glorot-uniform weights, zero biases, one seed stream per member
  std::vector<float> w(NPAR);
  for (int m = 0; m < members; ++m) {
    std::mt19937_64 rng(seed * 1000003ULL + (unsigned long long)m);
    std::fill(w.begin(), w.end(), 0.f);
    auto fill = [&](int off, int fo, int fi) {
      std::uniform_real_distribution<float> u(-std::sqrt(6.f / (fi + fo)), std::sqrt(6.f / (fi + fo)));
      for (int q = 0; q < fo * fi; ++q) w[off + q] = u(rng);
    };
    fill(OW1, H_, 4);
    fill(OW2, H_, H_);
    fill(OW3, H_, H_);
    fill(OW4, 3, H_);
    int rc = upload(h, m, w.data());
    if (rc) { dfk_mlp_destroy(h); return rc; }
  }
  *out = h;
  return DFK_MLP_OK;
}

void dfk_mlp_destroy(dfk_mlp *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  hipFree(h->master); hipFree(h->mom); hipFree(h->vel); hipFree(h->wbf);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

int dfk_mlp_get_params(dfk_mlp *h, int member, float *out) {
  if (!h || !out || member < 0 || member >= h->members) { merr("bad arguments"); return DFK_MLP_EINVAL; }
  MCHK(hipSetDevice(h->device));
  MCHK(hipStreamSynchronize(h->stream));
  MCHK(hipMemcpy(out, h->master + (size_t)member * NPAR, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
  return DFK_MLP_OK;
}

int dfk_mlp_set_params(dfk_mlp *h, int member, const float *in) {
  if (!h || !in || member < 0 || member >= h->members) { merr("bad arguments"); return DFK_MLP_EINVAL; }
  MCHK(hipSetDevice(h->device));
  return upload(h, member, in);
}

int dfk_mlp_forward(dfk_mlp *h, const float *x, int n, float *y) {
  if (!h || !x || !y || n < 1) { merr("bad arguments"); return DFK_MLP_EINVAL; }
  MCHK(hipSetDevice(h->device));
  float *dx = nullptr, *dy = nullptr;
  MCHK(hipMalloc(&dx, sizeof(float) * 4 * (size_t)n));
  if (hipMalloc(&dy, sizeof(float) * 3 * (size_t)n * h->members) != hipSuccess) { hipFree(dx); merr("alloc"); return DFK_MLP_EHIP; }
  hipError_t e = hipMemcpy(dx, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + PI_ - 1) / PI_, h->members), dim3(256), 0, h->stream,
                       h->wbf, h->master, dx, n, dy);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(y, dy, sizeof(float) * 3 * (size_t)n * h->members, hipMemcpyDeviceToHost);
  hipFree(dx);
  hipFree(dy);
  if (e != hipSuccess) { merr("forward failed: %s", hipGetErrorString(e)); return DFK_MLP_EHIP; }
  return DFK_MLP_OK;
}

namespace {
struct TrainBufs {
  float *X = nullptr, *T = nullptr, *gpart = nullptr, *lpart = nullptr, *gout = nullptr;
  int *idx = nullptr;
  ~TrainBufs() { hipFree(X); hipFree(T); hipFree(gpart); hipFree(lpart); hipFree(gout); hipFree(idx); }
};
}  // namespace

int dfk_mlp_grad(dfk_mlp *h, const float *x, const float *t, int n, float *grad, float *loss) {
  if (!h || !x || !t || n < 1) { merr("bad arguments"); return DFK_MLP_EINVAL; }
  MCHK(hipSetDevice(h->device));
  TrainBufs b;
  const int M = h->members, nwg = (n + PT - 1) / PT;
  MCHK(hipMalloc(&b.X, sizeof(float) * 4 * (size_t)n));
  MCHK(hipMalloc(&b.T, sizeof(float) * 3 * (size_t)n));
  MCHK(hipMalloc(&b.idx, sizeof(int) * (size_t)n * M));
  MCHK(hipMalloc(&b.gpart, sizeof(float) * (size_t)M * nwg * NPAR));
  MCHK(hipMalloc(&b.lpart, sizeof(float) * (size_t)M * nwg));
  MCHK(hipMalloc(&b.gout, sizeof(float) * (size_t)M * NPAR));
  std::vector<int> id((size_t)n * M);
  for (int m = 0; m < M; ++m)
    for (int i = 0; i < n; ++i) id[(size_t)m * n + i] = i;
  MCHK(hipMemcpy(b.X, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(b.T, t, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(b.idx, id.data(), sizeof(int) * id.size(), hipMemcpyHostToDevice));
  const float scale = 1.0f / (3.0f * n);
  hipLaunchKernelGGL(k_mlp_train, dim3(nwg, M), dim3(256), 0, h->stream, h->wbf, h->master, b.X, b.T,
                     b.idx, n, 0, n, scale, b.gpart, b.lpart);
  MCHK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_adam, dim3((NPAR + 255) / 256, M), dim3(256), 0, h->stream, h->master,
                     h->mom, h->vel, h->wbf, b.gpart, nwg, M, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, b.gout);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(h->stream));
  if (grad) MCHK(hipMemcpy(grad, b.gout, sizeof(float) * (size_t)M * NPAR, hipMemcpyDeviceToHost));
  if (loss) {
    std::vector<float> lp((size_t)M * nwg);
    MCHK(hipMemcpy(lp.data(), b.lpart, sizeof(float) * lp.size(), hipMemcpyDeviceToHost));
    for (int m = 0; m < M; ++m) {
      double s = 0;
      for (int w = 0; w < nwg; ++w) s += lp[(size_t)m * nwg + w];
      loss[m] = (float)s;
    }
  }
  return DFK_MLP_OK;
}

int dfk_mlp_train(dfk_mlp *h, const float *x, const float *t, int n, int epochs, int batch,
                  float lr, unsigned long long seed, float *loss_hist) {
  if (!h || !x || !t || n < 1 || epochs < 1 || batch < PT) { merr("bad arguments"); return DFK_MLP_EINVAL; }
  MCHK(hipSetDevice(h->device));
  TrainBufs b;
  const int M = h->members;
  batch = std::min(batch, n);
  const int nwg_max = (batch + PT - 1) / PT;
  MCHK(hipMalloc(&b.X, sizeof(float) * 4 * (size_t)n));
  MCHK(hipMalloc(&b.T, sizeof(float) * 3 * (size_t)n));
  MCHK(hipMalloc(&b.idx, sizeof(int) * (size_t)n * M));
  MCHK(hipMalloc(&b.gpart, sizeof(float) * (size_t)M * nwg_max * NPAR));
  MCHK(hipMalloc(&b.lpart, sizeof(float) * (size_t)M * nwg_max));
  MCHK(hipMemcpy(b.X, x, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(b.T, t, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice));
  std::vector<int> id((size_t)n * M);
  std::mt19937_64 rng(seed);
  std::vector<float> lp((size_t)M * nwg_max);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int m = 0; m < M; ++m) {
      int *row = id.data() + (size_t)m * n;
      for (int i = 0; i < n; ++i) row[i] = i;
      std::shuffle(row, row + n, rng);
    }
    MCHK(hipMemcpy(b.idx, id.data(), sizeof(int) * id.size(), hipMemcpyHostToDevice));
    double eloss = 0.0;
    for (int off = 0; off < n; off += batch) {
      const int nb = std::min(batch, n - off), nwg = (nb + PT - 1) / PT;
      const float scale = 1.0f / (3.0f * nb);
      hipLaunchKernelGGL(k_mlp_train, dim3(nwg, M), dim3(256), 0, h->stream, h->wbf, h->master, b.X,
                         b.T, b.idx, n, off, nb, scale, b.gpart, b.lpart);
      MCHK(hipGetLastError());
      ++h->step;
      const float bc1 = 1.f - std::pow(b1, (float)h->step), bc2 = 1.f - std::pow(b2, (float)h->step);
      hipLaunchKernelGGL(k_mlp_adam, dim3((NPAR + 255) / 256, M), dim3(256), 0, h->stream, h->master,
                         h->mom, h->vel, h->wbf, b.gpart, nwg, M, lr, b1, b2, eps, bc1, bc2,
                         (float *)nullptr);
      MCHK(hipGetLastError());
      if (loss_hist) {
        MCHK(hipMemcpyAsync(lp.data(), b.lpart, sizeof(float) * (size_t)M * nwg, hipMemcpyDeviceToHost, h->stream));
        MCHK(hipStreamSynchronize(h->stream));
        for (int q = 0; q < M * nwg; ++q) eloss += lp[q] * (double)nb / n;
      }
    }
    if (loss_hist) loss_hist[ep] = (float)(eloss / M);
  }
  MCHK(hipStreamSynchronize(h->stream));
  return DFK_MLP_OK;
}

int dfk_mlp_mfma_probe(const float *A, const float *B, float *D) {
  float *a = nullptr, *b = nullptr, *d = nullptr;
  MCHK(hipMalloc(&a, 256 * 4));
  MCHK(hipMalloc(&b, 256 * 4));
  MCHK(hipMalloc(&d, 256 * 4));
  MCHK(hipMemcpy(a, A, 256 * 4, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(b, B, 256 * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, a, b, d);
  MCHK(hipGetLastError());
  MCHK(hipDeviceSynchronize());
  MCHK(hipMemcpy(D, d, 256 * 4, hipMemcpyDeviceToHost));
  hipFree(a);
  hipFree(b);
  hipFree(d);
  return DFK_MLP_OK;
}

}  // extern "C"
