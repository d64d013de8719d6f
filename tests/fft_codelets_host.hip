// Host-side check of the in-register DFT codelets (dft<N>) and of the index maps
// of the wave-line split (WaveLine::gpos / rpos) against a direct O(N^2) DFT in
// double precision. The wave-line exchange is emulated lane by lane: every part
// t < S transforms its M elements, the twiddle (Cooley-Tukey split only) is
// applied, and every part t < M transforms the S values gathered across parts.
// Prints one line per case and exits non-zero on any failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../gradwave_amd/csrc/gw_regfft.h"

using gw::cpx;

static double g_worst = 0.0;
static int g_fail = 0;

static void naive(const std::vector<double2>& a, std::vector<double2>& out, int sign) {
  const int n = (int)a.size();
  out.assign(n, double2{0, 0});
  for (int k = 0; k < n; ++k) {
    double re = 0, im = 0;
    for (int j = 0; j < n; ++j) {
      const long e = ((long)j * k) % n;
      const double ang = sign * 2.0 * M_PI * (double)e / n;
      re += a[j].x * std::cos(ang) - a[j].y * std::sin(ang);
      im += a[j].x * std::sin(ang) + a[j].y * std::cos(ang);
    }
    out[k] = double2{re, im};
  }
}

static std::vector<double2> random_line(int n, unsigned seed) {
  srand(seed);
  std::vector<double2> a(n);
  for (auto& v : a) v = double2{rand() / (double)RAND_MAX - 0.5, rand() / (double)RAND_MAX - 0.5};
  return a;
}

static void report(const char* what, int n, double err) {
  const bool ok = err < 1e-12 * n;
  if (!ok) ++g_fail;
  if (err > g_worst) g_worst = err;
  printf("%-32s n=%-4d max_err=%.3e %s\n", what, n, err, ok ? "ok" : "FAIL");
}

template <int N> static void check_dft() {
  for (int sign : {+1, -1}) {
    std::vector<double2> a = random_line(N, 17 + N), ref;
    naive(a, ref, sign);
    double2 v[N];
    for (int i = 0; i < N; ++i) v[i] = a[i];
    if (sign > 0) gw::dft<N, +1, double>(v);
    else gw::dft<N, -1, double>(v);
    double err = 0;
    for (int i = 0; i < N; ++i) err = std::fmax(err, std::hypot(v[i].x - ref[i].x, v[i].y - ref[i].y));
    report(sign > 0 ? "dft inverse" : "dft forward", N, err);
  }
}

template <int N, int M, int S> static void check_waveline() {
  using W = gw::WaveLine<double, N, M, S>;
  std::vector<double2> tw(N);
  for (int m = 0; m < N; ++m) tw[m] = double2{std::cos(2 * M_PI * m / N), std::sin(2 * M_PI * m / N)};
  // both maps are permutations of 0..N-1
  std::vector<int> seen_g(N, 0), seen_r(N, 0);
  for (int t = 0; t < S; ++t)
    for (int q = 0; q < M; ++q) seen_g[W::gpos(t, q)]++;
  for (int t = 0; t < M; ++t)
    for (int k = 0; k < S; ++k) seen_r[W::rpos(t, k)]++;
  int bad = 0;
  for (int i = 0; i < N; ++i) bad += (seen_g[i] != 1) + (seen_r[i] != 1);
  if (bad) ++g_fail;
  printf("waveline %3d = %2d x %-2d %s maps %s\n", N, M, S, W::PFA ? "prime-factor" : "cooley-tukey",
         bad ? "FAIL" : "ok");
  // inverse: G-space line a -> r-space A = sum_j a_j e^{+2 pi i j k / N}
  std::vector<double2> a = random_line(N, 100 + N), ref;
  naive(a, ref, +1);
  std::vector<double2> E((size_t)S * M);
  for (int t = 0; t < S; ++t) {
    double2 x[M];
    for (int q = 0; q < M; ++q) x[q] = a[W::gpos(t, q)];
    gw::dft<M, +1, double>(x);
    if (!W::PFA)
      for (int k = 1; k < M; ++k) x[k] = gw::twmul<+1, double>(x[k], tw[t * k]);
    for (int k = 0; k < M; ++k) E[(size_t)t * M + k] = x[k];
  }
  std::vector<double2> A(N);
  for (int t = 0; t < M; ++t) {
    double2 y[S];
    for (int s = 0; s < S; ++s) y[s] = E[(size_t)s * M + t];
    gw::dft<S, +1, double>(y);
    for (int k = 0; k < S; ++k) A[W::rpos(t, k)] = y[k];
  }
  double err = 0;
  for (int i = 0; i < N; ++i) err = std::fmax(err, std::hypot(A[i].x - ref[i].x, A[i].y - ref[i].y));
  report("waveline inverse", N, err);
  // forward: r-space line b -> G-space B = sum_j b_j e^{-2 pi i j k / N}
  std::vector<double2> b = random_line(N, 200 + N);
  naive(b, ref, -1);
  for (int t = 0; t < M; ++t) {
    double2 y[S];
    for (int k = 0; k < S; ++k) y[k] = b[W::rpos(t, k)];
    gw::dft<S, -1, double>(y);
    if (!W::PFA)
      for (int s = 1; s < S; ++s) y[s] = gw::twmul<-1, double>(y[s], tw[t * s]);
    for (int s = 0; s < S; ++s) E[(size_t)s * M + t] = y[s];
  }
  std::vector<double2> B(N);
  for (int t = 0; t < S; ++t) {
    double2 x[M];
    for (int k = 0; k < M; ++k) x[k] = E[(size_t)t * M + k];
    gw::dft<M, -1, double>(x);
    for (int q = 0; q < M; ++q) B[W::gpos(t, q)] = x[q];
  }
  err = 0;
  for (int i = 0; i < N; ++i) err = std::fmax(err, std::hypot(B[i].x - ref[i].x, B[i].y - ref[i].y));
  report("waveline forward", N, err);
}

// Host emulation of the generic LDS line passes of gw_fft.h (line_fft_inverse /
// line_fft_forward): every task of a pass runs gw::line_task, the same code the
// device passes run, with the same (ntask, A, B, R, C, S) parameters.
template <int SIGN, int TW>
static void host_pass(int np, double2* a, const double2* tw, int ntask, int A, int B, int R, int C, int S) {
  for (int t = 0; t < ntask; ++t) {
    const int r = t % R, o = t / R;
    double2* base = a + C * o + B * r;
    switch (np) {
#define GW_HP(NN) case NN: gw::line_task<NN, SIGN, TW, double>(base, A, tw, S * r); break;
      GW_HP(1) GW_HP(2) GW_HP(3) GW_HP(4) GW_HP(5) GW_HP(6) GW_HP(7) GW_HP(8) GW_HP(9) GW_HP(10)
      GW_HP(12) GW_HP(14) GW_HP(15) GW_HP(16)
#undef GW_HP
      default: ++g_fail; break;
    }
  }
}

// natural index held at position pos after the inverse split (f2 = 1: two factors)
static int split_index(int pos, int f0, int f1, int f2) {
  const int M = f1 * f2;
  const int k0 = pos / M, k1 = (pos % M) / f2, k2 = pos % f2;
  return k0 + f0 * k1 + f0 * f1 * k2;
}

static void check_split(int f0, int f1, int f2) {
  const int n = f0 * f1 * f2, M = f1 * f2;
  std::vector<double2> tw(n);
  for (int m = 0; m < n; ++m) tw[m] = double2{std::cos(2 * M_PI * m / n), std::sin(2 * M_PI * m / n)};
  std::vector<double2> a = random_line(n, 300 + n), ref;
  naive(a, ref, +1);
  if (f2 == 1) {
    host_pass<+1, 1>(f0, a.data(), tw.data(), f1, f1, 1, f1, 0, 1);
    host_pass<+1, 0>(f1, a.data(), tw.data(), f0, 1, f1, f0, 0, 1);
  } else {
    host_pass<+1, 1>(f0, a.data(), tw.data(), M, M, 1, M, 0, 1);
    host_pass<+1, 1>(f1, a.data(), tw.data(), f0 * f2, f2, 1, f2, M, f0);
    host_pass<+1, 0>(f2, a.data(), tw.data(), f0 * f1, 1, f2, f0 * f1, 0, 1);
  }
  double err = 0;
  for (int pos = 0; pos < n; ++pos) {
    const double2 r = ref[split_index(pos, f0, f1, f2)];
    err = std::fmax(err, std::hypot(a[pos].x - r.x, a[pos].y - r.y));
  }
  char what[64];
  snprintf(what, sizeof(what), "lds split %dx%dx%d inverse", f0, f1, f2);
  report(what, n, err);
  // forward: the permuted layout of a natural line back to natural order
  std::vector<double2> bnat = random_line(n, 400 + n), b(n);
  naive(bnat, ref, -1);
  for (int pos = 0; pos < n; ++pos) b[pos] = bnat[split_index(pos, f0, f1, f2)];
  if (f2 == 1) {
    host_pass<-1, 1>(f1, b.data(), tw.data(), f0, 1, f1, f0, 0, 1);
    host_pass<-1, 0>(f0, b.data(), tw.data(), f1, f1, 1, f1, 0, 1);
  } else {
    host_pass<-1, 1>(f2, b.data(), tw.data(), f0 * f1, 1, f2, f1, M, f0);
    host_pass<-1, 0>(f1, b.data(), tw.data(), f0 * f2, f2, 1, f2, M, 0);
    host_pass<-1, 2>(f0, b.data(), tw.data(), M, M, 1, M, 0, 1);
  }
  err = 0;
  for (int k = 0; k < n; ++k) err = std::fmax(err, std::hypot(b[k].x - ref[k].x, b[k].y - ref[k].y));
  snprintf(what, sizeof(what), "lds split %dx%dx%d forward", f0, f1, f2);
  report(what, n, err);
}

int main() {
  check_dft<1>(); check_dft<2>(); check_dft<3>(); check_dft<4>(); check_dft<5>();
  check_dft<6>(); check_dft<7>(); check_dft<8>(); check_dft<9>(); check_dft<10>();
  check_dft<12>(); check_dft<14>(); check_dft<15>(); check_dft<16>();
  // the wave-line configurations of the box kernels and their prime-factor alternatives
  check_waveline<24, 4, 6>(); check_waveline<24, 3, 8>();
  check_waveline<36, 6, 6>(); check_waveline<36, 4, 9>();
  check_waveline<90, 9, 10>();
  check_waveline<120, 10, 12>(); check_waveline<120, 8, 15>();
  // the generic LDS kernels' two- and three-factor splits (three factors: every
  // 7-smooth box length without a two-codelet split, e.g. 125, 162, 200, 216, 243, 250, > 256)
  check_split(12, 10, 1); check_split(9, 10, 1); check_split(7, 7, 1);
  check_split(5, 5, 5); check_split(2, 9, 9); check_split(8, 5, 5); check_split(6, 6, 6);
  check_split(3, 9, 9); check_split(10, 5, 5); check_split(7, 7, 6); check_split(4, 3, 5);
  check_split(16, 16, 16); check_split(15, 14, 12);
  printf("worst %.3e failures %d\n", g_worst, g_fail);
  return g_fail ? 1 : 0;
}
