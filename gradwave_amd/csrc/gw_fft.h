// In-register DFT codelets and LDS line passes for the box FFTs of H|psi>.
//
// A 1-D transform of length n = N1*N2 (N1, N2 <= 16, 7-smooth) is done as two
// in-register passes over an LDS line with one barrier between them:
//
//   "inverse" form (natural order in, digit-permuted order out), sign +1:
//     pass 1: task j2<N2 : DFT_N1 over a[N2*q + j2], twiddle W_n^{+j2*k1}, store in place
//     pass 2: task k1<N1 : DFT_N2 over a[N2*k1 + q], store in place
//     -> position N2*k1 + k2 holds X[k1 + N1*k2]
//   "forward" form (digit-permuted order in, natural order out), sign -1:
//     pass 1: task a<N1  : DFT_N2 over a[N2*a + q], twiddle W_n^{-a*k2}, store in place
//     pass 2: task k2<N2 : DFT_N1 over a[N2*q + k2], store in place
//
// The forward form is the transpose of the inverse form, so a G->r transform
// followed by a pointwise multiply (with the multiplier stored in the same
// permuted order) and an r->G transform returns natural order with no
// reordering pass.  Both forms touch each element twice through registers.
#pragma once

#include <hip/hip_runtime.h>

namespace gw {

template <typename T> struct cpx_t;
template <> struct cpx_t<double> { using type = double2; };
template <> struct cpx_t<float> { using type = float2; };
template <typename T> using cpx = typename cpx_t<T>::type;

template <typename T> __host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  cpx<T> r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T> __host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  cpx<T> r; r.x = a.x + b.x; r.y = a.y + b.y; return r;
}
template <typename T> __host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  cpx<T> r; r.x = a.x - b.x; r.y = a.y - b.y; return r;
}
// multiply by the conjugate-or-not twiddle w (w = e^{+2 pi i t/n} from the table)
template <int SIGN, typename T> __host__ __device__ __forceinline__ cpx<T> twmul(cpx<T> a, cpx<T> w) {
  cpx<T> r;
  if constexpr (SIGN > 0) {
    r.x = a.x * w.x - a.y * w.y;
    r.y = a.x * w.y + a.y * w.x;
  } else {
    r.x = a.x * w.x + a.y * w.y;
    r.y = a.y * w.x - a.x * w.y;
  }
  return r;
}

// ---- compile-time twiddle constants (Taylor series on a reduced argument) ----
constexpr double cx_sin_red(double x) {  // |x| <= pi/4
  double term = x, sum = x, x2 = x * x;
  for (int i = 1; i < 14; ++i) { term *= -x2 / ((2 * i) * (2 * i + 1)); sum += term; }
  return sum;
}
constexpr double cx_cos_red(double x) {
  double term = 1.0, sum = 1.0, x2 = x * x;
  for (int i = 1; i < 14; ++i) { term *= -x2 / ((2 * i - 1) * (2 * i)); sum += term; }
  return sum;
}
// cos/sin of 2*pi*k/n, exact symmetry reduction on integers (octants)
constexpr double cx_cos2pi(long k, long n) {
  k %= n; if (k < 0) k += n;
  // angle = 2 pi k / n ; use 8*k relative to n to find the octant
  long k8 = 8 * k;
  int oct = (int)(k8 / n);  // 0..7
  long rem8 = k8 - (long)oct * n;  // in [0, n)
  constexpr double PI4 = 0.78539816339744830961566084581987572;
  double f = PI4 * (double)rem8 / (double)n;  // in [0, pi/4)
  double c = cx_cos_red(f), s = cx_sin_red(f), cc = cx_cos_red(PI4 - f), ss = cx_sin_red(PI4 - f);
  switch (oct) {
    case 0: return c;
    case 1: return ss;
    case 2: return -s;
    case 3: return -cc;
    case 4: return -c;
    case 5: return -ss;
    case 6: return s;
    default: return cc;
  }
}
constexpr double cx_sin2pi(long k, long n) { return cx_cos2pi(4 * k - n, 4 * n); }

template <int N> struct TwTab {
  double c[N > 0 ? N : 1];
  double s[N > 0 ? N : 1];
  constexpr TwTab() : c{}, s{} {
    for (int k = 0; k < N; ++k) { c[k] = cx_cos2pi(k, N); s[k] = cx_sin2pi(k, N); }
  }
};

constexpr int smallest_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

// Largest P < n with n = P*Q, gcd(P, Q) = 1, Q > 1; 0 for prime powers.
constexpr int coprime_split(int n) {
  for (int p = n - 1; p > 1; --p)
    if (n % p == 0 && gcd_c(p, n / p) == 1) return p;
  return 0;
}

// x with (a * x) % m == 1, for gcd(a, m) = 1
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return m == 1 ? 0 : -1;
}

// In-place DFT of N complex values in registers, natural order in and out:
// X[k] = sum_j a[j] exp(SIGN * 2 pi i j k / N).
template <int N, int SIGN, typename T> __host__ __device__ __forceinline__ void dft(cpx<T>* a) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2) {
    cpx<T> t = a[1];
    a[1] = csub<T>(a[0], t);
    a[0] = cadd<T>(a[0], t);
  } else if constexpr (N == 3) {
    constexpr T c1 = T(-0.5);
    constexpr T s1 = T(SIGN) * T(0.86602540378443864676372317075293618);
    cpx<T> t1 = cadd<T>(a[1], a[2]);
    cpx<T> t2 = csub<T>(a[1], a[2]);
    cpx<T> m; m.x = a[0].x + c1 * t1.x; m.y = a[0].y + c1 * t1.y;
    cpx<T> r; r.x = -s1 * t2.y; r.y = s1 * t2.x;  // i*s1*t2
    a[0] = cadd<T>(a[0], t1);
    a[1] = cadd<T>(m, r);
    a[2] = csub<T>(m, r);
  } else if constexpr (N == 4) {
    cpx<T> t0 = cadd<T>(a[0], a[2]), t1 = csub<T>(a[0], a[2]);
    cpx<T> t2 = cadd<T>(a[1], a[3]), t3 = csub<T>(a[1], a[3]);
    cpx<T> it3;  // SIGN*i*t3
    it3.x = -T(SIGN) * t3.y; it3.y = T(SIGN) * t3.x;
    a[0] = cadd<T>(t0, t2);
    a[2] = csub<T>(t0, t2);
    a[1] = cadd<T>(t1, it3);
    a[3] = csub<T>(t1, it3);
  } else if constexpr (smallest_factor(N) == N) {
    // prime N (5, 7, 11, 13): symmetric pair form, O(N^2/2)
    constexpr TwTab<N> tw{};
    constexpr int H = (N - 1) / 2;
    cpx<T> sp[H], sm[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) { sp[j - 1] = cadd<T>(a[j], a[N - j]); sm[j - 1] = csub<T>(a[j], a[N - j]); }
    cpx<T> out[N];
    cpx<T> a0 = a[0];
    out[0] = a0;
#pragma unroll
    for (int j = 0; j < H; ++j) { out[0].x += sp[j].x; out[0].y += sp[j].y; }
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      T re = a0.x, im = a0.y, ire = 0, iim = 0;
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const T c = T(tw.c[(j * k) % N]);
        const T s = T(SIGN) * T(tw.s[(j * k) % N]);
        re += c * sp[j - 1].x; im += c * sp[j - 1].y;
        ire += s * sm[j - 1].x; iim += s * sm[j - 1].y;
      }
      // X[k] = sum c*sp + i*s*sm ; X[N-k] = sum c*sp - i*s*sm
      out[k].x = re - iim; out[k].y = im + ire;
      out[N - k].x = re + iim; out[N - k].y = im - ire;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) a[k] = out[k];
  } else if constexpr (coprime_split(N) > 0) {
    // Good-Thomas prime-factor form, no twiddles: input j = (Q*j1 + P*j2) mod N,
    // output k = CRT(k1 mod P, k2 mod Q) = (k1*Q*(Q^-1 mod P) + k2*P*(P^-1 mod Q)) mod N.
    constexpr int P = coprime_split(N);
    constexpr int Q = N / P;
    constexpr int CP = Q * inv_mod(Q % P, P);
    constexpr int CQ = P * inv_mod(P % Q, Q);
    cpx<T> b[N];
#pragma unroll
    for (int j2 = 0; j2 < Q; ++j2) {
      cpx<T> t[P];
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) t[j1] = a[(Q * j1 + P * j2) % N];
      dft<P, SIGN, T>(t);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) b[Q * k1 + j2] = t[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      cpx<T> t[Q];
#pragma unroll
      for (int j2 = 0; j2 < Q; ++j2) t[j2] = b[Q * k1 + j2];
      dft<Q, SIGN, T>(t);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) a[(CP * k1 + CQ * k2) % N] = t[k2];
    }
  } else {
    constexpr int P = smallest_factor(N);
    constexpr int Q = N / P;
    constexpr TwTab<N> tw{};
    // j = Q*j1 + j2 ; k = k1 + P*k2
    cpx<T> b[N];
#pragma unroll
    for (int j2 = 0; j2 < Q; ++j2) {
      cpx<T> t[P];
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) t[j1] = a[Q * j1 + j2];
      dft<P, SIGN, T>(t);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        const int e = (j2 * k1) % N;
        cpx<T> w; w.x = T(tw.c[e]); w.y = T(SIGN) * T(tw.s[e]);
        b[Q * k1 + j2] = (e == 0) ? t[k1] : cmul<T>(t[k1], w);
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      cpx<T> t[Q];
#pragma unroll
      for (int j2 = 0; j2 < Q; ++j2) t[j2] = b[Q * k1 + j2];
      dft<Q, SIGN, T>(t);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) a[k1 + P * k2] = t[k2];
    }
  }
}

// One radix task of a line pass: NP elements at base[A * q] (q < NP) are
// DFT'd (size NP, sign SIGN) in place, with the twiddles W^{SIGN * S * r * q}
// (table tw holds e^{+2 pi i t / n}) applied after the DFT (TW = 1), before it
// (TW = 2) or not at all (TW = 0). Host and device code share it (the host
// test of the three-factor split runs it task by task).
template <int NP, int SIGN, int TW, typename T>
__host__ __device__ __forceinline__ void line_task(cpx<T>* base, int A, const cpx<T>* __restrict__ tw, int step) {
  cpx<T> v[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) v[q] = base[A * q];
  if constexpr (TW == 2) {
#pragma unroll
    for (int q = 1; q < NP; ++q) v[q] = twmul<SIGN, T>(v[q], tw[step * q]);
  }
  dft<NP, SIGN, T>(v);
  if constexpr (TW == 1) {
#pragma unroll
    for (int q = 1; q < NP; ++q) v[q] = twmul<SIGN, T>(v[q], tw[step * q]);
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) base[A * q] = v[q];
}

// One pass of a line transform on LDS data. For each (line, r) task:
//   elements at  lds[line*lstride + (A*q + B*r)*es], q in [0, NP)
//   are DFT'd (size NP, sign SIGN), optionally multiplied by W_n^{SIGN*r*q}
//   (table tw holds e^{+2 pi i t / n}), and stored back in place.
// Task order: if LINES_FAST, consecutive threads take consecutive lines
// (conflict-free when lines are adjacent columns); else consecutive r.
template <int NP, int SIGN, bool TWIDDLE, bool LINES_FAST, typename T>
__device__ __forceinline__ void line_pass(cpx<T>* __restrict__ lds, const cpx<T>* __restrict__ tw,
                                          int nlines, int lstride, int es, int ntask, int A, int B,
                                          const int* __restrict__ line_map) {
  const int total = nlines * ntask;
  for (int task = threadIdx.x; task < total; task += blockDim.x) {
    int line, r;
    if constexpr (LINES_FAST) { line = task % nlines; r = task / nlines; }
    else { r = task % ntask; line = task / ntask; }
    const int lrow = line_map ? line_map[line] : line;
    line_task<NP, SIGN, TWIDDLE ? 1 : 0, T>(lds + lrow * lstride + B * r * es, A * es, tw, r);
  }
}

// Pass with a two-level task index (the middle pass of a three-factor split):
// task t < ntask has r = t % R and o = t / R; its elements sit at
// lds[line*lstride + (C*o + B*r + A*q)*es] and its twiddle step is S * r.
template <int NP, int SIGN, int TW, bool LINES_FAST, typename T>
__device__ __forceinline__ void line_pass2(cpx<T>* __restrict__ lds, const cpx<T>* __restrict__ tw,
                                           int nlines, int lstride, int es, int ntask, int A, int B, int R, int C,
                                           int S, const int* __restrict__ line_map) {
  const int total = nlines * ntask;
  for (int task = threadIdx.x; task < total; task += blockDim.x) {
    int line, t;
    if constexpr (LINES_FAST) { line = task % nlines; t = task / nlines; }
    else { t = task % ntask; line = task / ntask; }
    const int r = t % R, o = t / R;
    const int lrow = line_map ? line_map[line] : line;
    line_task<NP, SIGN, TW, T>(lds + lrow * lstride + (C * o + B * r) * es, A * es, tw, S * r);
  }
}

// Runtime-length pass: the switch instantiates every codelet up to RMAX, and the
// largest one sets the register budget of the whole kernel, so kernels whose
// plans only split into factors <= 10 are built with RMAX = 10.
template <int SIGN, bool TWIDDLE, bool LINES_FAST, typename T, int RMAX = 16>
__device__ __forceinline__ void line_pass_rt(int np, cpx<T>* lds, const cpx<T>* tw, int nlines,
                                             int lstride, int es, int ntask, int A, int B,
                                             const int* line_map) {
#define GW_LP(NN) case NN: if constexpr (NN <= RMAX) line_pass<NN, SIGN, TWIDDLE, LINES_FAST, T>(lds, tw, nlines, lstride, es, ntask, A, B, line_map); break;
  switch (np) {
    GW_LP(1) GW_LP(2) GW_LP(3) GW_LP(4) GW_LP(5) GW_LP(6) GW_LP(7) GW_LP(8) GW_LP(9)
    GW_LP(10) GW_LP(12) GW_LP(14) GW_LP(15) GW_LP(16)
    default: break;
  }
#undef GW_LP
}

template <int SIGN, int TW, bool LINES_FAST, typename T, int RMAX = 16>
__device__ __forceinline__ void line_pass2_rt(int np, cpx<T>* lds, const cpx<T>* tw, int nlines, int lstride,
                                              int es, int ntask, int A, int B, int R, int C, int S,
                                              const int* line_map) {
#define GW_LP(NN) case NN: if constexpr (NN <= RMAX) line_pass2<NN, SIGN, TW, LINES_FAST, T>(lds, tw, nlines, lstride, es, ntask, A, B, R, C, S, line_map); break;
  switch (np) {
    GW_LP(1) GW_LP(2) GW_LP(3) GW_LP(4) GW_LP(5) GW_LP(6) GW_LP(7) GW_LP(8) GW_LP(9)
    GW_LP(10) GW_LP(12) GW_LP(14) GW_LP(15) GW_LP(16)
    default: break;
  }
#undef GW_LP
}

// Axis plan: n = n1f * n2f (* n3f). Inverse (G->r, natural -> permuted), then
// the matching forward (r->G, permuted -> natural). Barriers are the caller's
// between transforms; barriers sit between the passes here.
//
// Two factors (n3f == 1): the decimation-in-frequency split; the value of
// output index k1 + n1f*k2 lands at position n2f*k1 + k2.
// Three factors f0 f1 f2 (M = f1 f2): pass 1 is the same first stage over
// stride M (twiddle W_n^{r k0}); pass 2 splits each length-M sub-line the same
// way (stride f2, twiddle W_M^{r k1} = W_n^{f0 r k1}); pass 3 transforms the
// contiguous groups of f2. Output index k0 + f0*k1 + f0*f1*k2 lands at position
// M*k0 + f2*k1 + k2. The forward runs the adjoint passes in reverse order.
// RMAX is the largest codelet a kernel instantiates (10 or 16); kernels built
// with RMAX = kRmaxF3 take codelets up to 16 and also the three-factor split
// (the only instantiations that carry its code, so the two-factor kernels keep
// their register budget).
constexpr int kRmaxF3 = 17;

template <typename T, bool LINES_FAST, int RMAX = 16>
__device__ __forceinline__ void line_fft_inverse(cpx<T>* lds, const cpx<T>* tw, int n1f, int n2f,
                                                 int nlines, int lstride, int es,
                                                 const int* line_map, int n3f = 1) {
  if (RMAX < kRmaxF3 || n3f == 1) {
    // pass 1: task j2 < n2f: DFT_n1f over a[n2f*q + j2], twiddle W^{+j2*k1}
    line_pass_rt<+1, true, LINES_FAST, T, RMAX>(n1f, lds, tw, nlines, lstride, es, n2f, n2f, 1, line_map);
    __syncthreads();
    // pass 2: task k1 < n1f: DFT_n2f over a[n2f*k1 + q]
    line_pass_rt<+1, false, LINES_FAST, T, RMAX>(n2f, lds, tw, nlines, lstride, es, n1f, 1, n2f, line_map);
    return;
  }
  if constexpr (RMAX >= kRmaxF3) {
    const int f0 = n1f, f1 = n2f, f2 = n3f, M = f1 * f2;
    line_pass_rt<+1, true, LINES_FAST, T, RMAX>(f0, lds, tw, nlines, lstride, es, M, M, 1, line_map);
    __syncthreads();
    line_pass2_rt<+1, 1, LINES_FAST, T, RMAX>(f1, lds, tw, nlines, lstride, es, f0 * f2, f2, 1, f2, M, f0,
                                              line_map);
    __syncthreads();
    line_pass_rt<+1, false, LINES_FAST, T, RMAX>(f2, lds, tw, nlines, lstride, es, f0 * f1, 1, f2, line_map);
  }
}

template <typename T, bool LINES_FAST, int RMAX = 16>
__device__ __forceinline__ void line_fft_forward(cpx<T>* lds, const cpx<T>* tw, int n1f, int n2f,
                                                 int nlines, int lstride, int es,
                                                 const int* line_map, int n3f = 1) {
  if (RMAX < kRmaxF3 || n3f == 1) {
    // pass 1: task a < n1f: DFT_n2f over a[n2f*a + q], twiddle W^{-a*k2}
    line_pass_rt<-1, true, LINES_FAST, T, RMAX>(n2f, lds, tw, nlines, lstride, es, n1f, 1, n2f, line_map);
    __syncthreads();
    // pass 2: task k2 < n2f: DFT_n1f over a[n2f*q + k2]
    line_pass_rt<-1, false, LINES_FAST, T, RMAX>(n1f, lds, tw, nlines, lstride, es, n2f, n2f, 1, line_map);
    return;
  }
  if constexpr (RMAX >= kRmaxF3) {
    const int f0 = n1f, f1 = n2f, f2 = n3f, M = f1 * f2;
    // adjoint of pass 3 and of pass 2's twiddle: DFT_f2 over the contiguous group of
    // (k0, k1), then W_n^{-f0 k1 j} on its element j
    line_pass2_rt<-1, 1, LINES_FAST, T, RMAX>(f2, lds, tw, nlines, lstride, es, f0 * f1, 1, f2, f1, M, f0,
                                              line_map);
    __syncthreads();
    // adjoint of pass 2's DFT: DFT_f1 over stride f2 inside each sub-line
    line_pass2_rt<-1, 0, LINES_FAST, T, RMAX>(f1, lds, tw, nlines, lstride, es, f0 * f2, f2, 1, f2, M, 0,
                                              line_map);
    __syncthreads();
    // adjoint of pass 1: twiddle W_n^{-r q} first, then DFT_f0 over stride M
    line_pass2_rt<-1, 2, LINES_FAST, T, RMAX>(f0, lds, tw, nlines, lstride, es, M, M, 1, M, 0, 1, line_map);
  }
}

}  // namespace gw
