// In-register DFTs for the two-pass transforms of the column convolution
// (k_cols_conv_reg) and the row transforms (k_rows_r2c_reg,
// k_rows_c2r_reg): sizes 2, 3, 4, 8, 12, 16, 24, 32, forward (SIGN < 0,
// exp(-2 pi i nk/N)) or inverse (SIGN > 0), unnormalised, natural order in
// and out, fully unrolled at compile time (indices are template constants,
// so the arrays stay in registers and twiddles of the form +-1, +-i are
// exact sign flips and swaps).  Host and device code, so that the codelets
// and the kernel's phases can be checked against numpy on the CPU.
//
// Composite sizes N = P Q by one Cooley-Tukey step: with n = Q n1 + n2,
// X[k1 + P k2] = sum_n2 W_N^(n2 k1) W_Q^(n2 k2) sum_n1 x[Q n1 + n2] W_P^(n1 k1).
#ifndef DFK_REGFFT_H
#define DFK_REGFFT_H

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#define RF_HD __host__ __device__ __forceinline__

namespace rfft {

// cos / sin (2 pi k / 96), k < 96, correctly rounded (96 = lcm of the sizes)
constexpr double kCos96[96] = {
    1.0, 0.9978589232386035, 0.9914448613738104, 0.9807852804032304,
    0.9659258262890683, 0.9469301294951057, 0.9238795325112867, 0.8968727415326883,
    0.8660254037844386, 0.8314696123025452, 0.7933533402912352, 0.7518398074789774,
    0.7071067811865476, 0.6593458151000688, 0.6087614290087207, 0.5555702330196022,
    0.5, 0.4422886902190013, 0.3826834323650898, 0.3214394653031616,
    0.25881904510252074, 0.19509032201612828, 0.1305261922200516, 0.06540312923014306,
    0.0, -0.06540312923014306, -0.1305261922200516, -0.19509032201612828,
    -0.25881904510252074, -0.3214394653031616, -0.3826834323650898, -0.4422886902190013,
    -0.5, -0.5555702330196022, -0.6087614290087207, -0.6593458151000688,
    -0.7071067811865476, -0.7518398074789774, -0.7933533402912352, -0.8314696123025452,
    -0.8660254037844386, -0.8968727415326883, -0.9238795325112867, -0.9469301294951057,
    -0.9659258262890683, -0.9807852804032304, -0.9914448613738104, -0.9978589232386035,
    -1.0, -0.9978589232386035, -0.9914448613738104, -0.9807852804032304,
    -0.9659258262890683, -0.9469301294951057, -0.9238795325112867, -0.8968727415326883,
    -0.8660254037844386, -0.8314696123025452, -0.7933533402912352, -0.7518398074789774,
    -0.7071067811865476, -0.6593458151000688, -0.6087614290087207, -0.5555702330196022,
    -0.5, -0.4422886902190013, -0.3826834323650898, -0.3214394653031616,
    -0.25881904510252074, -0.19509032201612828, -0.1305261922200516, -0.06540312923014306,
    0.0, 0.06540312923014306, 0.1305261922200516, 0.19509032201612828,
    0.25881904510252074, 0.3214394653031616, 0.3826834323650898, 0.4422886902190013,
    0.5, 0.5555702330196022, 0.6087614290087207, 0.6593458151000688,
    0.7071067811865476, 0.7518398074789774, 0.7933533402912352, 0.8314696123025452,
    0.8660254037844386, 0.8968727415326883, 0.9238795325112867, 0.9469301294951057,
    0.9659258262890683, 0.9807852804032304, 0.9914448613738104, 0.9978589232386035,
};
constexpr double kSin96[96] = {
    0.0, 0.06540312923014306, 0.1305261922200516, 0.19509032201612828,
    0.25881904510252074, 0.3214394653031616, 0.3826834323650898, 0.4422886902190013,
    0.5, 0.5555702330196022, 0.6087614290087207, 0.6593458151000688,
    0.7071067811865476, 0.7518398074789774, 0.7933533402912352, 0.8314696123025452,
    0.8660254037844386, 0.8968727415326883, 0.9238795325112867, 0.9469301294951057,
    0.9659258262890683, 0.9807852804032304, 0.9914448613738104, 0.9978589232386035,
    1.0, 0.9978589232386035, 0.9914448613738104, 0.9807852804032304,
    0.9659258262890683, 0.9469301294951057, 0.9238795325112867, 0.8968727415326883,
    0.8660254037844386, 0.8314696123025452, 0.7933533402912352, 0.7518398074789774,
    0.7071067811865476, 0.6593458151000688, 0.6087614290087207, 0.5555702330196022,
    0.5, 0.4422886902190013, 0.3826834323650898, 0.3214394653031616,
    0.25881904510252074, 0.19509032201612828, 0.1305261922200516, 0.06540312923014306,
    0.0, -0.06540312923014306, -0.1305261922200516, -0.19509032201612828,
    -0.25881904510252074, -0.3214394653031616, -0.3826834323650898, -0.4422886902190013,
    -0.5, -0.5555702330196022, -0.6087614290087207, -0.6593458151000688,
    -0.7071067811865476, -0.7518398074789774, -0.7933533402912352, -0.8314696123025452,
    -0.8660254037844386, -0.8968727415326883, -0.9238795325112867, -0.9469301294951057,
    -0.9659258262890683, -0.9807852804032304, -0.9914448613738104, -0.9978589232386035,
    -1.0, -0.9978589232386035, -0.9914448613738104, -0.9807852804032304,
    -0.9659258262890683, -0.9469301294951057, -0.9238795325112867, -0.8968727415326883,
    -0.8660254037844386, -0.8314696123025452, -0.7933533402912352, -0.7518398074789774,
    -0.7071067811865476, -0.6593458151000688, -0.6087614290087207, -0.5555702330196022,
    -0.5, -0.4422886902190013, -0.3826834323650898, -0.3214394653031616,
    -0.25881904510252074, -0.19509032201612828, -0.1305261922200516, -0.06540312923014306,
};

template <int I>
using ic = std::integral_constant<int, I>;

// f(ic<B>), f(ic<B+1>), ..., f(ic<E-1>)
template <int B, int E, class F>
RF_HD void sfor(F &&f) {
  if constexpr (B < E) {
    f(ic<B>{});
    sfor<B + 1, E>(f);
  }
}

RF_HD double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
RF_HD double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
RF_HD double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// a * W_N^E with W_N = exp(SIGN 2 pi i / N): exact for E a multiple of N/4
template <int N, int E, int SIGN>
RF_HD double2 twc(double2 a) {
  constexpr int e = ((E % N) + N) % N;
  if constexpr (e == 0) {
    return a;
  } else if constexpr (4 * e == N) {            // SIGN i
    return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
  } else if constexpr (2 * e == N) {            // -1
    return make_double2(-a.x, -a.y);
  } else if constexpr (4 * e == 3 * N) {        // -SIGN i
    return SIGN < 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
  } else {
    static_assert(96 % N == 0, "twiddle table covers N | 96");
    constexpr int t = e * (96 / N);
    constexpr double c = kCos96[t], s = (SIGN < 0 ? -1.0 : 1.0) * kSin96[t];
    return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

template <int N>
constexpr int split_p() {   // first factor of N for one Cooley-Tukey step
  return (N % 4 == 0 && N != 4 && N != 8) ? (N == 12 || N == 16 ? 4 : 8) : (N % 2 == 0 ? 2 : 3);
}

template <int N, int SIGN>
RF_HD void dft(double2 (&v)[N]);

template <int N, int SIGN>
RF_HD void dft_ct(double2 (&v)[N]) {
  constexpr int P = split_p<N>(), Q = N / P;
  double2 t[N];   // t[n2 P + k1] = Y[n2][k1] W_N^(n2 k1)
  sfor<0, Q>([&](auto n2_) {
    constexpr int n2 = decltype(n2_)::value;
    double2 a[P];
    sfor<0, P>([&](auto n1_) { a[decltype(n1_)::value] = v[Q * decltype(n1_)::value + n2]; });
    dft<P, SIGN>(a);
    sfor<0, P>([&](auto k1_) {
      constexpr int k1 = decltype(k1_)::value;
      t[n2 * P + k1] = twc<N, n2 * k1, SIGN>(a[k1]);
    });
  });
  sfor<0, P>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    double2 b[Q];
    sfor<0, Q>([&](auto n2_) { b[decltype(n2_)::value] = t[decltype(n2_)::value * P + k1]; });
    dft<Q, SIGN>(b);
    sfor<0, Q>([&](auto k2_) { v[k1 + P * decltype(k2_)::value] = b[decltype(k2_)::value]; });
  });
}

template <int N, int SIGN>
RF_HD void dft(double2 (&v)[N]) {
  if constexpr (N == 1) {
  } else if constexpr (N == 2) {
    const double2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  } else if constexpr (N == 3) {
    constexpr double s = (SIGN < 0 ? -1.0 : 1.0) * 0.8660254037844386;   // sin(2 pi/3)
    const double2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const double2 c = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    const double2 is = make_double2(-s * t2.y, s * t2.x);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(c, is);
    v[2] = csub(c, is);
  } else if constexpr (N == 4) {
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const double2 b0 = cadd(v[1], v[3]), b1 = twc<4, 1, SIGN>(csub(v[1], v[3]));
    v[0] = cadd(a0, b0);
    v[1] = cadd(a1, b1);
    v[2] = csub(a0, b0);
    v[3] = csub(a1, b1);
  } else {
    dft_ct<N, SIGN>(v);
  }
}

// Only inputs n < N/2 are nonzero (v[n >= N/2] is not read): X[2j] is the
// N/2-point DFT of x, X[2j+1] that of x[n] W_N^n.
template <int N, int SIGN>
RF_HD void dft_half_in(double2 (&v)[N]) {
  static_assert(N % 2 == 0, "even size");
  constexpr int H = N / 2;
  double2 a[H], b[H];
  sfor<0, H>([&](auto n_) {
    constexpr int n = decltype(n_)::value;
    a[n] = v[n];
    b[n] = twc<N, n, SIGN>(v[n]);
  });
  dft<H, SIGN>(a);
  dft<H, SIGN>(b);
  sfor<0, H>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    v[2 * j] = a[j];
    v[2 * j + 1] = b[j];
  });
}

// Only outputs k < N/2 are formed (v[k >= N/2] is left undefined):
// X[k] = E[k] + W_N^k O[k] with E, O the N/2-point DFTs of the even and odd inputs.
template <int N, int SIGN>
RF_HD void dft_half_out(double2 (&v)[N]) {
  static_assert(N % 2 == 0, "even size");
  constexpr int H = N / 2;
  double2 e[H], o[H];
  sfor<0, H>([&](auto m_) {
    constexpr int m = decltype(m_)::value;
    e[m] = v[2 * m];
    o[m] = v[2 * m + 1];
  });
  dft<H, SIGN>(e);
  dft<H, SIGN>(o);
  sfor<0, H>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    v[k] = cadd(e[k], twc<N, k, SIGN>(o[k]));
  });
}

// ------------------------------------------------------------------ column convolution
// The zero-padded convolution of one spectrum column by two passes of
// in-register DFTs per direction and one LDS transpose between them:
// x = rows 0 .. Nu-1 of the column (rows Nu .. M-1 are the padding),
// y = rows 0 .. Nu-1 of IDFT_M(K . DFT_M(x)) (unnormalised; K real with
// K[p] = K[M - p], given for p <= Nu), M = N1 N2 = 2 Nu.
//   forward, n = N2 n1 + n2, k = k1 + N1 k2:
//     pass 1 (lane n2): Y[k1] = W_M^(n2 k1) DFT_N1(x[N2 n1 + n2])   -> LDS [k1][n2]
//     pass 2 (lane k1): X[k1 + N1 k2] = DFT_N2(Y[.][k1])
//   inverse, n = b + N2 a:
//     pass 2 (lane k1): Z[b] = W_M^(-b k1) IDFT_N2(K X)[b]          -> LDS [b][k1]
//     pass 3 (lane b):  y[b + N2 a] = IDFT_N1(Z[.][b])[a], a < N1/2
// A column uses T = N2 lanes; lanes >= N1 are idle in pass 2.
namespace conv {

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// W_M^(SIGN m k) for k < K from the table tw[t] = W_M^t (length M): exact
// table values for k < 8 and multiples of 8, one product otherwise.
template <int K, int SIGN>
struct TwRow {
  static constexpr int NL = K < 8 ? K : 8, NH = (K + 7) / 8;
  double2 lo[NL], hi[NH];
  RF_HD TwRow() {}   // unset (a row the caller does not use)
  RF_HD TwRow(const double2 *tw, int m) {
    sfor<0, NL>([&](auto j_) {
      constexpr int j = decltype(j_)::value;
      lo[j] = tw[m * j];
      if constexpr (SIGN > 0) lo[j].y = -lo[j].y;
    });
    sfor<0, NH>([&](auto i_) {
      constexpr int i = decltype(i_)::value;
      hi[i] = tw[8 * m * i];
      if constexpr (SIGN > 0) hi[i].y = -hi[i].y;
    });
  }
  template <int k>
  RF_HD double2 mul(double2 a) const {
    if constexpr (k == 0) return a;
    else if constexpr (k < 8) return cmul(a, lo[k]);
    else if constexpr (k % 8 == 0) return cmul(a, hi[k / 8]);
    else return cmul(cmul(a, lo[k % 8]), hi[k / 8]);
  }
};

// The transposes go through LDS in H parts (H = 1 or 2: half the LDS per
// column for H = 2, at the price of one more synchronisation pair):
// forward part h carries the values of lanes n2 in [h N2/H, (h+1) N2/H),
// inverse part h those of lanes k1 in [h N1/H, (h+1) N1/H).
template <int N1, int N2, int H>
constexpr int lds_elems_h() { return cmax(N1 * (N2 / H + 1), N2 * (N1 / H + 1)); }

// pass 1, lane n2 < N2: ld(row) returns row `row` (< Nu) of the column;
// v[k1] = Y[k1] W_M^(n2 k1).
template <int N1, int N2, class Ld>
RF_HD void pass1(int n2, Ld ld, const double2 *tw, double2 (&v)[N1]) {
  sfor<0, N1 / 2>([&](auto n1_) {
    constexpr int n1 = decltype(n1_)::value;
    v[n1] = ld(N2 * n1 + n2);
  });
  const TwRow<N1, -1> w(tw, n2);
  dft_half_in<N1, -1>(v);
  sfor<1, N1>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    v[k1] = w.template mul<k1>(v[k1]);
  });
}

// forward transpose, part h: writers are lanes n2 of the part (all their
// k1), readers lanes k1 < N1 (x[n2] for the part's n2)
template <int N1, int N2, int H>
RF_HD void fwd_write(int n2, int h, const double2 (&v)[N1], double2 *L) {
  constexpr int W = N2 / H;
  if (n2 / W != h) return;
  sfor<0, N1>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    L[k1 * (W + 1) + n2 - h * W] = v[k1];
  });
}
template <int N1, int N2, int H, int HH>
RF_HD void fwd_read_c(int k1, const double2 *L, double2 (&x)[N2]) {
  constexpr int W = N2 / H;
  sfor<0, W>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    x[HH * W + j] = L[k1 * (W + 1) + j];
  });
}

// K values lane k1 < N1 multiplies by: kv[k2] = K[p], p = k1 + N1 k2 (K[M - p] for p > Nu);
// kh(p) returns K[p], p <= Nu.
template <int N1, int N2, class Kh>
RF_HD void pass2_khat(int k1, Kh kh, double (&kv)[N2]) {
  constexpr int M = N1 * N2, Nu = M / 2;
  sfor<0, N2>([&](auto k2_) {
    constexpr int k2 = decltype(k2_)::value;
    const int p = k1 + N1 * k2;
    kv[k2] = kh(p <= Nu ? p : M - p);
  });
}

// pass 2, lane k1 < N1: forward N2-point DFT, times K, inverse, times
// W_M^(-b k1): x[b] = Z[b].
template <int N1, int N2>
RF_HD void pass2(int k1, double2 (&x)[N2], const double (&kv)[N2], const double2 *tw) {
  dft<N2, -1>(x);
  sfor<0, N2>([&](auto k2_) {
    constexpr int k2 = decltype(k2_)::value;
    x[k2].x *= kv[k2];
    x[k2].y *= kv[k2];
  });
  dft<N2, +1>(x);
  const TwRow<N2, +1> w(tw, k1);
  sfor<1, N2>([&](auto b_) {
    constexpr int b = decltype(b_)::value;
    x[b] = w.template mul<b>(x[b]);
  });
}

// inverse transpose, part h: writers lanes k1 of the part (< N1), readers lanes b < N2
template <int N1, int N2, int H>
RF_HD void inv_write(int k1, int h, const double2 (&x)[N2], double2 *L) {
  constexpr int W = N1 / H;
  if (k1 >= N1 || k1 / W != h) return;
  sfor<0, N2>([&](auto b_) {
    constexpr int b = decltype(b_)::value;
    L[b * (W + 1) + k1 - h * W] = x[b];
  });
}
template <int N1, int N2, int H, int HH>
RF_HD void inv_read_c(int b, const double2 *L, double2 (&z)[N1]) {
  constexpr int W = N1 / H;
  sfor<0, W>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    z[HH * W + j] = L[b * (W + 1) + j];
  });
}

// pass 3, lane b < N2: st(row, value) receives rows b + N2 a, a < N1/2 (all < Nu)
template <int N1, int N2, class St>
RF_HD void pass3(int b, double2 (&z)[N1], St st) {
  dft_half_out<N1, +1>(z);
  sfor<0, N1 / 2>([&](auto a_) {
    constexpr int a = decltype(a_)::value;
    st(b + N2 * a, z[a]);
  });
}

}  // namespace conv

// ------------------------------------------------------------------ column convolution, three passes
// The same zero-padded convolution for M = 64 P (P = 4, 6, 8, 12, 16: Nu = 32 P) by
// one column per wavefront and three passes of in-register DFTs per
// direction (P, 8 and 8 points), two LDS transposes per direction, so that
// no lane holds more than one 8-point DFT's values (latency: the two-pass
// form gives half of each wavefront an N2-point DFT per direction).
//   forward, n = 64 n1 + 8 n2 + n3, k = k1 + P k2 + 8 P k3:
//     pass A (lane m = 8 n2 + n3):  A[k1] = W_M^(m k1) DFT_P(x[64 n1 + m])[k1]       -> T1 [k1][m]
//     pass B (lane 8 k1 + n3):      B[k2] = W_64^(n3 k2) DFT_8(A[k1][8 n2 + n3])[k2] -> T2 [8 k1 + k2][n3]
//     pass C (lane 8 k1 + k2):      X[k1 + P k2 + 8 P k3] = DFT_8(B[8 k1 + k2][.])[k3]
//   inverse, n = c + 8 b + 64 a:
//     pass C (lane 8 k1 + k2):      C[c] = W_64^(-c k2) IDFT_8(K X)[c]                -> T2 [8 k1 + c][k2]
//     pass B' (lane 8 k1 + c):      E[b] = W_M^(-(8 b + c) k1) IDFT_8(C[.])[b]        -> T1 [k1][8 b + c]
//     pass A' (lane m = 8 b + c):   y[m + 64 a] = IDFT_P(E[.][m])[a], a < P/2
// (the exponents of W_M that the index splits leave are multiples of M).
// Passes B to B' run on lanes 8 k1 + r for k1 < 8 and, for P > 8, on the
// same lanes again for k1 + 8 (two groups of eight values per lane); a
// (lane, group) whose k1 reaches P takes k1 = P - 1 and repeats that lane's
// work (same reads, same values written), so every access stays inside the
// tables and the LDS area.
namespace conv3 {

constexpr int S1 = 65, S2 = 9;   // row strides of T1 ([P][64]) and T2 ([8 P][8]) in double2
template <int P>
constexpr int lds_elems() {      // T1 and T2 (at least 64 rows) share the area
  return P * S1 > (8 * P > 64 ? 8 * P : 64) * S2 ? P * S1 : (8 * P > 64 ? 8 * P : 64) * S2;
}

// k1 groups per lane in passes B .. B', and lane l's k1 in group g
template <int P>
constexpr int groups() { return (P + 7) / 8; }
template <int P>
RF_HD int lane_k1(int l, int g = 0) { return (l >> 3) + 8 * g < P ? (l >> 3) + 8 * g : P - 1; }

// pass A, lane m: ld(row) returns row `row` (< Nu) of the column
template <int P, class Ld>
RF_HD void passA(int m, Ld ld, const double2 *tw, double2 (&v)[P]) {
  const conv::TwRow<P, -1> w(tw, m);   // issued before the data loads
  sfor<0, P / 2>([&](auto n1_) {
    constexpr int n1 = decltype(n1_)::value;
    v[n1] = ld(64 * n1 + m);
  });
  dft_half_in<P, -1>(v);
  sfor<1, P>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    v[k1] = w.template mul<k1>(v[k1]);
  });
}

// T1: lane m of pass A writes its A[k1]; lane (k1, n3) reads A[k1][8 n2 + n3], n2 < 8
template <int P>
RF_HD void t1_write(int m, const double2 (&v)[P], double2 *L) {
  sfor<0, P>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    L[k1 * S1 + m] = v[k1];
  });
}
RF_HD void t1_read(int k1, int n3, const double2 *L, double2 (&x)[8]) {
  sfor<0, 8>([&](auto n2_) {
    constexpr int n2 = decltype(n2_)::value;
    x[n2] = L[k1 * S1 + 8 * n2 + n3];
  });
}

// pass B, lane (k1, n3), with the twiddle row W_M^(-P n3 k2) = W_64^(-n3 k2)
// already loaded (conv::TwRow<8, -1>(tw, P n3)), or loading it
template <int P>
RF_HD void passB(const conv::TwRow<8, -1> &w, double2 (&x)[8]) {
  dft<8, -1>(x);
  sfor<1, 8>([&](auto k2_) {
    constexpr int k2 = decltype(k2_)::value;
    x[k2] = w.template mul<k2>(x[k2]);
  });
}
template <int P>
RF_HD void passB(int n3, const double2 *tw, double2 (&x)[8]) {
  passB<P>(conv::TwRow<8, -1>(tw, P * n3), x);
}

// T2: lane (k1, r) writes its x[q] to row 8 k1 + q, column r; lane 8 k1 + q
// reads the row (forward: r = n3, q = k2; inverse: r = k2, q = c)
RF_HD void t2_write(int k1, int r, const double2 (&x)[8], double2 *L) {
  sfor<0, 8>([&](auto q_) {
    constexpr int q = decltype(q_)::value;
    L[(8 * k1 + q) * S2 + r] = x[q];
  });
}
RF_HD void t2_read(int row, const double2 *L, double2 (&y)[8]) {
  sfor<0, 8>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    y[r] = L[row * S2 + r];
  });
}

// K values lane (k1, k2) multiplies by: kv[k3] = K[p], p = k1 + P k2 + 8 P k3
// (K[M - p] for p > Nu); kh(p) returns K[p], p <= Nu
template <int P, class Kh>
RF_HD void khat_vals(int k1, int k2, Kh kh, double (&kv)[8]) {
  constexpr int M = 64 * P, Nu = 32 * P;
  sfor<0, 8>([&](auto k3_) {
    constexpr int k3 = decltype(k3_)::value;
    const int p = k1 + P * k2 + 8 * P * k3;
    kv[k3] = kh(p <= Nu ? p : M - p);
  });
}

// pass C, lane (k1, k2): forward DFT over n3, times K, inverse DFT over k3,
// times W_64^(-c k2)
template <int P>
RF_HD void passC(const double (&kv)[8], const conv::TwRow<8, +1> &w, double2 (&y)[8]) {
  dft<8, -1>(y);
  sfor<0, 8>([&](auto k3_) {
    constexpr int k3 = decltype(k3_)::value;
    y[k3].x *= kv[k3];
    y[k3].y *= kv[k3];
  });
  dft<8, +1>(y);
  sfor<1, 8>([&](auto c_) {
    constexpr int c = decltype(c_)::value;
    y[c] = w.template mul<c>(y[c]);
  });
}
template <int P>
RF_HD void passC(int k2, const double (&kv)[8], const double2 *tw, double2 (&y)[8]) {
  passC<P>(kv, conv::TwRow<8, +1>(tw, P * k2), y);
}

// pass B', lane (k1, c): inverse DFT over k2, times W_M^(-(8 b + c) k1)
// (table values, (8 b + c) k1 < M), loaded by binv_tw or by the pass
RF_HD void binv_tw(int k1, int c, const double2 *tw, double2 (&w)[8]) {
  sfor<0, 8>([&](auto b_) {
    constexpr int b = decltype(b_)::value;
    w[b] = tw[(8 * b + c) * k1];
    w[b].y = -w[b].y;
  });
}
RF_HD void passB_inv(const double2 (&w)[8], double2 (&z)[8]) {
  dft<8, +1>(z);
  sfor<0, 8>([&](auto b_) {
    constexpr int b = decltype(b_)::value;
    z[b] = cmul(z[b], w[b]);
  });
}
RF_HD void passB_inv(int k1, int c, const double2 *tw, double2 (&z)[8]) {
  double2 w[8];
  binv_tw(k1, c, tw, w);
  dft<8, +1>(z);
  sfor<0, 8>([&](auto b_) {
    constexpr int b = decltype(b_)::value;
    z[b] = cmul(z[b], w[b]);
  });
}

// T1 backwards: lane (k1, c) writes E[b] to [k1][8 b + c]; lane m reads E[.][m]
template <int P>
RF_HD void t4_write(int k1, int c, const double2 (&z)[8], double2 *L) {
  sfor<0, 8>([&](auto b_) {
    constexpr int b = decltype(b_)::value;
    L[k1 * S1 + 8 * b + c] = z[b];
  });
}
template <int P>
RF_HD void t4_read(int m, const double2 *L, double2 (&w)[P]) {
  sfor<0, P>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    w[k1] = L[k1 * S1 + m];
  });
}

// pass A', lane m: st(row, value) receives rows m + 64 a, a < P/2 (all < Nu)
template <int P, class St>
RF_HD void passA_inv(int m, double2 (&w)[P], St st) {
  dft_half_out<P, +1>(w);
  sfor<0, P / 2>([&](auto a_) {
    constexpr int a = decltype(a_)::value;
    st(m + 64 * a, w[a]);
  });
}

}  // namespace conv3

// ------------------------------------------------------------------ row transforms, three passes
// Row transforms of length n = 64 P (P = 2, 3, 4, 6, 8), one packed row pair per
// wavefront, by the index splits and transposes of conv3 (all P inputs of
// pass A nonzero): pass C leaves Z[k1 + P k2 + 8 P k3] on lane (k1, k2),
// which spills the transform to LDS in natural order.  r2c: rows a, b as
// real and imaginary parts, forward transform, split into the two half
// spectra (rows::split's formulas); c2r: the packed spectrum of rows::c2r_in,
// inverse transform, rows a, b as its real and imaginary parts.
namespace rows3 {

// pass A, lane m: ld(j) returns x[j], j = 64 n1 + m
template <int P, int SIGN, class Ld>
RF_HD void passA(int m, Ld ld, const double2 *tw, double2 (&v)[P]) {
  const conv::TwRow<P, SIGN> w(tw, m);   // issued before the data loads
  sfor<0, P>([&](auto n1_) {
    constexpr int n1 = decltype(n1_)::value;
    v[n1] = ld(64 * n1 + m);
  });
  dft<P, SIGN>(v);
  sfor<1, P>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    v[k1] = w.template mul<k1>(v[k1]);
  });
}

// pass B, lane (k1, n3): DFT over n2, times W_64^(SIGN n3 k2) (the twiddle
// row conv::TwRow<8, SIGN>(tw, P n3), loaded by the caller or by the pass)
template <int P, int SIGN>
RF_HD void passB(const conv::TwRow<8, SIGN> &w, double2 (&x)[8]) {
  dft<8, SIGN>(x);
  sfor<1, 8>([&](auto k2_) {
    constexpr int k2 = decltype(k2_)::value;
    x[k2] = w.template mul<k2>(x[k2]);
  });
}
template <int P, int SIGN>
RF_HD void passB(int n3, const double2 *tw, double2 (&x)[8]) {
  const conv::TwRow<8, SIGN> w(tw, P * n3);
  dft<8, SIGN>(x);
  sfor<1, 8>([&](auto k2_) {
    constexpr int k2 = decltype(k2_)::value;
    x[k2] = w.template mul<k2>(x[k2]);
  });
}

// pass C, lane (k1, k2): DFT over n3, the values into L[0, n) in natural order
template <int P, int SIGN>
RF_HD void passC_spill(int k1, int k2, double2 (&y)[8], double2 *L) {
  dft<8, SIGN>(y);
  sfor<0, 8>([&](auto k3_) {
    constexpr int k3 = decltype(k3_)::value;
    L[k1 + P * k2 + 8 * P * k3] = y[k3];
  });
}

// r2c split, lane l: bins k = l + 64 j <= n/2 of the two rows from the
// packed transform Z in L (rows::split's formulas), to st(k, Xa, Xb)
template <int P, class St>
RF_HD void split(int l, const double2 *L, St st) {
  constexpr int n = 64 * P, h = n / 2;
  sfor<0, (h + 64) / 64>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    const int k = l + 64 * j;
    if (k <= h) {
      const double2 z = L[k];
      double2 y = L[k == 0 ? 0 : n - k];
      y.y = -y.y;
      st(k, make_double2(0.5 * (z.x + y.x), 0.5 * (z.y + y.y)),
         make_double2(0.5 * (z.y - y.y), -0.5 * (z.x - y.x)));
    }
  });
}

// c2r output, lane m: points j = m + 64 a of the packed inverse in L, to st(j, z)
template <int P, class St>
RF_HD void unspill(int m, const double2 *L, St st) {
  sfor<0, P>([&](auto a_) {
    constexpr int a = decltype(a_)::value;
    st(m + 64 * a, L[m + 64 * a]);
  });
}

}  // namespace rows3

// Row transforms of length n = N1 N2 by the same two passes, N1 <= N2: pass 1
// on lanes t < N2 (inputs x[N2 n1 + t]), the transpose of conv::fwd_write /
// fwd_read_c with H = 1, pass 2 on lanes k1 < N1, which then hold bins
// k1 + N1 k2.
namespace rows {

// LDS elements per transform: the transpose, or the spectrum in natural order
template <int N1, int N2>
constexpr int lds_elems() { return conv::cmax(N1 * (N2 + 1), N1 * N2); }

// pass 1, lane t < N2: ld(j) returns x[j]; v[k1] = Y[k1] W_n^(SIGN t k1)
// with Y the N1-point DFT of x[N2 n1 + t] over n1.
template <int N1, int N2, int SIGN, class Ld>
RF_HD void pass1(int t, Ld ld, const double2 *tw, double2 (&v)[N1]) {
  sfor<0, N1>([&](auto n1_) {
    constexpr int n1 = decltype(n1_)::value;
    v[n1] = ld(N2 * n1 + t);
  });
  dft<N1, SIGN>(v);
  const conv::TwRow<N1, SIGN> w(tw, t);
  sfor<1, N1>([&](auto k1_) {
    constexpr int k1 = decltype(k1_)::value;
    v[k1] = w.template mul<k1>(v[k1]);
  });
}

// r2c, lane k1 < N1 holding Z[k1 + N1 k2] of a packed row pair (rows a, b as
// real and imaginary parts): the spectrum in natural order into L[0, n).
template <int N1, int N2>
RF_HD void spill(int k1, const double2 (&x)[N2], double2 *L) {
  sfor<0, N2>([&](auto k2_) {
    constexpr int k2 = decltype(k2_)::value;
    L[k1 + N1 * k2] = x[k2];
  });
}

// r2c, lane k1 < N1 after spill: the two rows' bins k = k1 + N1 k2 <= n/2,
// Xa = (Z[k] + conj Z[n-k]) / 2, Xb = (Z[k] - conj Z[n-k]) / 2i, to st(k, Xa, Xb).
template <int N1, int N2, class St>
RF_HD void split(int k1, const double2 (&x)[N2], const double2 *L, St st) {
  constexpr int n = N1 * N2;
  sfor<0, N2 / 2 + 1>([&](auto k2_) {
    constexpr int k2 = decltype(k2_)::value;
    if (k2 < N2 / 2 || k1 == 0) {   // k = n/2 (k2 = N2/2) only on lane 0
      const int k = k1 + N1 * k2;
      const double2 z = x[k2];
      double2 y = L[k == 0 ? 0 : n - k];
      y.y = -y.y;
      st(k, make_double2(0.5 * (z.x + y.x), 0.5 * (z.y + y.y)),
         make_double2(0.5 * (z.y - y.y), -0.5 * (z.x - y.x)));
    }
  });
}

// c2r input of lane t < N2, element m of the packed row pair from the two
// rows' half spectra: bin(k) returns (Xa[k], Xb[k]) for k <= n/2; the
// imaginary parts of the DC and Nyquist bins are dropped (FFTW's c2r
// convention), bins above n/2 are the conjugates of bins n - m.
template <int N1, int N2, class Bin>
RF_HD double2 c2r_in(int m, Bin bin) {
  constexpr int n = N1 * N2;
  const bool lo = 2 * m <= n;
  const int k = lo ? m : n - m;
  double2 xa, xb;
  bin(k, xa, xb);
  // a select, not a branch: a branch writing the registers of loads still in
  // flight makes the wavefront wait for them before its later loads issue
  const bool real = k == 0 || 2 * k == n;
  xa.y = real ? 0.0 : xa.y;
  xb.y = real ? 0.0 : xb.y;
  return lo ? make_double2(xa.x - xb.y, xa.y + xb.x) : make_double2(xa.x + xb.y, -xa.y + xb.x);
}

}  // namespace rows
}  // namespace rfft

#endif
