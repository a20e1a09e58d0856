// Host harness of the two-pass register transforms in dfk_regfft.h: runs the
// kernels' phases on the CPU, lane by lane, with the same (N1, N2) splits and
// the same LDS layouts, so that a test can compare them with numpy's FFT.
//
//   regfft_check r2c N    stdin: rows a, b (N reals each)
//                         stdout: Xa[k], Xb[k] for k <= N/2 ("re im" lines, Xa first)
//   regfft_check c2r N    stdin: Xa[k], Xb[k] for k <= N/2 ("re im" pairs, Xa first)
//                         stdout: rows a, b (N reals each, unnormalised inverse)
//   regfft_check conv M   stdin: column x[r], r < M/2 ("re im"), then K[p], p <= M/2
//                         stdout: rows r < M/2 of inverse(K x forward(x zero-padded to M))
//   regfft_check conv3 M  the same by the three-pass column transform (M = 256 .. 1024)
//   regfft_check r2c3 N / c2r3 N   the row transforms by the three-pass form (N = 128 .. 512)
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dfk_regfft.h"

namespace {

std::vector<double2> table(int n) {   // tw[t] = exp(-2 pi i t / n), as the runtime builds it
  std::vector<double2> tw(n);
  for (int t = 0; t < n; ++t) {
    const double a = -2.0 * M_PI * t / n;
    tw[t] = make_double2(std::cos(a), std::sin(a));
  }
  return tw;
}

double rd() {
  double v;
  if (std::scanf("%lf", &v) != 1) {
    std::fprintf(stderr, "regfft_check: short input\n");
    std::exit(2);
  }
  return v;
}

template <int N1, int N2>
void r2c() {
  constexpr int n = N1 * N2, LE = rfft::rows::lds_elems<N1, N2>();
  std::vector<double> a(n), b(n);
  for (auto &v : a) v = rd();
  for (auto &v : b) v = rd();
  const std::vector<double2> tw = table(n);
  std::vector<double2> L(LE);
  double2 v[N2][N1], x[N1][N2];
  for (int t = 0; t < N2; ++t) {
    rfft::rows::pass1<N1, N2, -1>(t, [&](int j) { return make_double2(a[j], b[j]); }, tw.data(), v[t]);
    rfft::conv::fwd_write<N1, N2, 1>(t, 0, v[t], L.data());
  }
  for (int k1 = 0; k1 < N1; ++k1) rfft::conv::fwd_read_c<N1, N2, 1, 0>(k1, L.data(), x[k1]);
  for (int k1 = 0; k1 < N1; ++k1) {
    rfft::dft<N2, -1>(x[k1]);
    rfft::rows::spill<N1, N2>(k1, x[k1], L.data());
  }
  std::vector<double2> Xa(n / 2 + 1), Xb(n / 2 + 1);
  std::vector<int> seen(n / 2 + 1, 0);
  for (int k1 = 0; k1 < N1; ++k1)
    rfft::rows::split<N1, N2>(k1, x[k1], L.data(), [&](int k, double2 p, double2 q) {
      Xa[k] = p;
      Xb[k] = q;
      seen[k] += 1;
    });
  for (int k = 0; k <= n / 2; ++k)
    if (seen[k] != 1) {
      std::fprintf(stderr, "regfft_check: bin %d written %d times\n", k, seen[k]);
      std::exit(3);
    }
  for (const double2 &z : Xa) std::printf("%.17g %.17g\n", z.x, z.y);
  for (const double2 &z : Xb) std::printf("%.17g %.17g\n", z.x, z.y);
}

template <int N1, int N2>
void c2r() {
  constexpr int n = N1 * N2, LE = rfft::rows::lds_elems<N1, N2>();
  std::vector<double2> Xa(n / 2 + 1), Xb(n / 2 + 1);
  for (auto &z : Xa) { z.x = rd(); z.y = rd(); }
  for (auto &z : Xb) { z.x = rd(); z.y = rd(); }
  const std::vector<double2> tw = table(n);
  std::vector<double2> L(LE);
  double2 v[N2][N1], x[N1][N2];
  for (int t = 0; t < N2; ++t) {
    rfft::rows::pass1<N1, N2, +1>(t, [&](int m) {
      return rfft::rows::c2r_in<N1, N2>(m, [&](int k, double2 &p, double2 &q) { p = Xa[k]; q = Xb[k]; });
    }, tw.data(), v[t]);
    rfft::conv::fwd_write<N1, N2, 1>(t, 0, v[t], L.data());
  }
  std::vector<double> a(n), b(n);
  for (int k1 = 0; k1 < N1; ++k1) {
    rfft::conv::fwd_read_c<N1, N2, 1, 0>(k1, L.data(), x[k1]);
    rfft::dft<N2, +1>(x[k1]);
    for (int k2 = 0; k2 < N2; ++k2) {
      a[k1 + N1 * k2] = x[k1][k2].x;
      b[k1 + N1 * k2] = x[k1][k2].y;
    }
  }
  for (double z : a) std::printf("%.17g\n", z);
  for (double z : b) std::printf("%.17g\n", z);
}

template <int N1, int N2, int H>
void conv() {
  constexpr int M = N1 * N2, Nu = M / 2, LE = rfft::conv::lds_elems_h<N1, N2, H>();
  std::vector<double2> col(Nu);
  for (auto &z : col) { z.x = rd(); z.y = rd(); }
  std::vector<double> K(Nu + 1);
  for (auto &k : K) k = rd();
  const std::vector<double2> tw = table(M);
  std::vector<double2> L(LE), out(Nu, make_double2(NAN, NAN));
  double2 v[N2][N1], x[N1][N2], z[N2][N1];
  for (int t = 0; t < N2; ++t)
    rfft::conv::pass1<N1, N2>(t, [&](int r) { return col[r]; }, tw.data(), v[t]);
  rfft::sfor<0, H>([&](auto h_) {
    constexpr int hh = decltype(h_)::value;
    for (int t = 0; t < N2; ++t) rfft::conv::fwd_write<N1, N2, H>(t, hh, v[t], L.data());
    for (int k1 = 0; k1 < N1; ++k1) rfft::conv::fwd_read_c<N1, N2, H, hh>(k1, L.data(), x[k1]);
  });
  for (int k1 = 0; k1 < N1; ++k1) {
    double kv[N2];
    rfft::conv::pass2_khat<N1, N2>(k1, [&](int p) { return K[p]; }, kv);
    rfft::conv::pass2<N1, N2>(k1, x[k1], kv, tw.data());
  }
  rfft::sfor<0, H>([&](auto h_) {
    constexpr int hh = decltype(h_)::value;
    for (int k1 = 0; k1 < N1; ++k1) rfft::conv::inv_write<N1, N2, H>(k1, hh, x[k1], L.data());
    for (int b = 0; b < N2; ++b) rfft::conv::inv_read_c<N1, N2, H, hh>(b, L.data(), z[b]);
  });
  for (int b = 0; b < N2; ++b) rfft::conv::pass3<N1, N2>(b, z[b], [&](int r, double2 q) { out[r] = q; });
  for (const double2 &q : out) std::printf("%.17g %.17g\n", q.x, q.y);
}

// The three-pass column convolution, lane by lane over the 64 lanes of one
// wavefront (lanes above 8 P in passes B .. B' repeat lane 8 (P-1) + l mod 8).
template <int P>
void conv3() {
  constexpr int M = 64 * P, Nu = M / 2;
  std::vector<double2> col(Nu);
  for (auto &z : col) { z.x = rd(); z.y = rd(); }
  std::vector<double> K(Nu + 1);
  for (auto &k : K) k = rd();
  const std::vector<double2> tw = table(M);
  constexpr int G = rfft::conv3::groups<P>();
  std::vector<double2> L(rfft::conv3::lds_elems<P>(), make_double2(NAN, NAN)), out(Nu, make_double2(NAN, NAN));
  std::vector<std::array<double2, P>> v(64);
  double2 x[64][G][8];
  double kv[64][G][8];
  auto k1 = [](int l, int g) { return rfft::conv3::lane_k1<P>(l, g); };
  for (int l = 0; l < 64; ++l)
    for (int g = 0; g < G; ++g)
      rfft::conv3::khat_vals<P>(k1(l, g), l & 7, [&](int p) { return K.at(p); }, kv[l][g]);
  for (int m = 0; m < 64; ++m) {
    double2(&vm)[P] = *reinterpret_cast<double2(*)[P]>(v[m].data());
    rfft::conv3::passA<P>(m, [&](int r) { return col.at(r); }, tw.data(), vm);
    rfft::conv3::t1_write<P>(m, vm, L.data());
  }
  for (int l = 0; l < 64; ++l)
    for (int g = 0; g < G; ++g) rfft::conv3::t1_read(k1(l, g), l & 7, L.data(), x[l][g]);
  for (int l = 0; l < 64; ++l)
    for (int g = 0; g < G; ++g) {
      rfft::conv3::passB<P>(l & 7, tw.data(), x[l][g]);
      rfft::conv3::t2_write(k1(l, g), l & 7, x[l][g], L.data());
    }
  for (int l = 0; l < 64; ++l)
    for (int g = 0; g < G; ++g) rfft::conv3::t2_read(8 * k1(l, g) + (l & 7), L.data(), x[l][g]);
  for (int l = 0; l < 64; ++l)
    for (int g = 0; g < G; ++g) {
      rfft::conv3::passC<P>(l & 7, kv[l][g], tw.data(), x[l][g]);
      rfft::conv3::t2_write(k1(l, g), l & 7, x[l][g], L.data());
    }
  for (int l = 0; l < 64; ++l)
    for (int g = 0; g < G; ++g) rfft::conv3::t2_read(8 * k1(l, g) + (l & 7), L.data(), x[l][g]);
  for (int l = 0; l < 64; ++l)
    for (int g = 0; g < G; ++g) {
      rfft::conv3::passB_inv(k1(l, g), l & 7, tw.data(), x[l][g]);
      rfft::conv3::t4_write<P>(k1(l, g), l & 7, x[l][g], L.data());
    }
  for (int m = 0; m < 64; ++m) {
    double2(&vm)[P] = *reinterpret_cast<double2(*)[P]>(v[m].data());
    rfft::conv3::t4_read<P>(m, L.data(), vm);
    rfft::conv3::passA_inv<P>(m, vm, [&](int r, double2 q) { out.at(r) = q; });
  }
  for (const double2 &q : out) std::printf("%.17g %.17g\n", q.x, q.y);
}

// The three-pass row transforms, lane by lane over one wavefront.
template <int P, int SIGN, class Ld>
void rows3_fft(Ld ld, const std::vector<double2> &tw, std::vector<double2> &L) {
  double2 v[64][P], x[64][8];
  for (int m = 0; m < 64; ++m) {
    rfft::rows3::passA<P, SIGN>(m, ld, tw.data(), v[m]);
    rfft::conv3::t1_write<P>(m, v[m], L.data());
  }
  for (int l = 0; l < 64; ++l) rfft::conv3::t1_read(rfft::conv3::lane_k1<P>(l), l & 7, L.data(), x[l]);
  for (int l = 0; l < 64; ++l) {
    rfft::rows3::passB<P, SIGN>(l & 7, tw.data(), x[l]);
    rfft::conv3::t2_write(rfft::conv3::lane_k1<P>(l), l & 7, x[l], L.data());
  }
  for (int l = 0; l < 64; ++l) rfft::conv3::t2_read(8 * rfft::conv3::lane_k1<P>(l) + (l & 7), L.data(), x[l]);
  for (int l = 0; l < 64; ++l) rfft::rows3::passC_spill<P, SIGN>(rfft::conv3::lane_k1<P>(l), l & 7, x[l], L.data());
}

template <int P>
void r2c3() {
  constexpr int n = 64 * P;
  std::vector<double> a(n), b(n);
  for (auto &v : a) v = rd();
  for (auto &v : b) v = rd();
  const std::vector<double2> tw = table(n);
  std::vector<double2> L(rfft::conv3::lds_elems<8>(), make_double2(NAN, NAN));
  rows3_fft<P, -1>([&](int j) { return make_double2(a.at(j), b.at(j)); }, tw, L);
  std::vector<double2> Xa(n / 2 + 1), Xb(n / 2 + 1);
  std::vector<int> seen(n / 2 + 1, 0);
  for (int l = 0; l < 64; ++l)
    rfft::rows3::split<P>(l, L.data(), [&](int k, double2 p, double2 q) {
      Xa.at(k) = p;
      Xb.at(k) = q;
      seen.at(k) += 1;
    });
  for (int k = 0; k <= n / 2; ++k)
    if (seen[k] != 1) {
      std::fprintf(stderr, "regfft_check: bin %d written %d times\n", k, seen[k]);
      std::exit(3);
    }
  for (const double2 &z : Xa) std::printf("%.17g %.17g\n", z.x, z.y);
  for (const double2 &z : Xb) std::printf("%.17g %.17g\n", z.x, z.y);
}

template <int P>
void c2r3() {
  constexpr int n = 64 * P;
  std::vector<double2> Xa(n / 2 + 1), Xb(n / 2 + 1);
  for (auto &z : Xa) { z.x = rd(); z.y = rd(); }
  for (auto &z : Xb) { z.x = rd(); z.y = rd(); }
  const std::vector<double2> tw = table(n);
  std::vector<double2> L(rfft::conv3::lds_elems<8>(), make_double2(NAN, NAN));
  rows3_fft<P, +1>([&](int m) {
    return rfft::rows::c2r_in<P, 64>(m, [&](int k, double2 &p, double2 &q) { p = Xa.at(k); q = Xb.at(k); });
  }, tw, L);
  std::vector<double> a(n, NAN), b(n, NAN);
  for (int m = 0; m < 64; ++m)
    rfft::rows3::unspill<P>(m, L.data(), [&](int j, double2 z) { a.at(j) = z.x; b.at(j) = z.y; });
  for (double z : a) std::printf("%.17g\n", z);
  for (double z : b) std::printf("%.17g\n", z);
}

bool rows3_dispatch(bool r2c, int n) {
  switch (n) {
    case 128: r2c ? r2c3<2>() : c2r3<2>(); return true;
    case 192: r2c ? r2c3<3>() : c2r3<3>(); return true;
    case 256: r2c ? r2c3<4>() : c2r3<4>(); return true;
    case 384: r2c ? r2c3<6>() : c2r3<6>(); return true;
    case 512: r2c ? r2c3<8>() : c2r3<8>(); return true;
    default: return false;
  }
}

bool conv3_dispatch(int M) {
  switch (M) {
    case 256: conv3<4>(); return true;
    case 384: conv3<6>(); return true;
    case 512: conv3<8>(); return true;
    case 768: conv3<12>(); return true;
    case 1024: conv3<16>(); return true;
    default: return false;
  }
}

// the (N1, N2) splits of the kernels (transposes in H parts for the columns)
template <template <int, int> class F>
bool rows_dispatch(int n) {
  switch (n) {
    case 128: F<8, 16>::run(); return true;
    case 192: F<12, 16>::run(); return true;
    case 256: F<16, 16>::run(); return true;
    case 384: F<12, 32>::run(); return true;
    case 512: F<16, 32>::run(); return true;
    case 768: F<24, 32>::run(); return true;
    case 1024: F<32, 32>::run(); return true;
    default: return false;
  }
}
template <int A, int B> struct R2C { static void run() { r2c<A, B>(); } };
template <int A, int B> struct C2R { static void run() { c2r<A, B>(); } };

bool conv_dispatch(int M) {
  switch (M) {
    case 128: conv<8, 16, 1>(); return true;
    case 192: conv<12, 16, 1>(); return true;
    case 256: conv<16, 16, 1>(); return true;
    case 384: conv<12, 32, 2>(); return true;
    case 512: conv<16, 32, 2>(); return true;
    case 768: conv<24, 32, 2>(); return true;
    case 1024: conv<32, 32, 2>(); return true;
    default: return false;
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: regfft_check r2c|c2r|conv|conv3|r2c3|c2r3 N\n");
    return 2;
  }
  const int n = std::atoi(argv[2]);
  bool ok = false;
  if (!std::strcmp(argv[1], "r2c")) ok = rows_dispatch<R2C>(n);
  else if (!std::strcmp(argv[1], "c2r")) ok = rows_dispatch<C2R>(n);
  else if (!std::strcmp(argv[1], "conv")) ok = conv_dispatch(n);
  else if (!std::strcmp(argv[1], "conv3")) ok = conv3_dispatch(n);
  else if (!std::strcmp(argv[1], "r2c3")) ok = rows3_dispatch(true, n);
  else if (!std::strcmp(argv[1], "c2r3")) ok = rows3_dispatch(false, n);
  if (!ok) {
    std::fprintf(stderr, "regfft_check: no %s transform of length %d\n", argv[1], n);
    return 2;
  }
  return 0;
}
