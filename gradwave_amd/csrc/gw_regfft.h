// Wave-cooperative line FFTs: a length-N line (N = M*S) is spread over S lanes
// of one wave, each lane holding at most max(M, S) <= 16 complex values.
//
// Lane layout: a wave holds LW = 64/S lines; lane = t*LW + l carries part t of
// line l (lanes >= S*LW idle). Exchange between the two passes goes through a
// wave-private LDS area E[l][s][k], so no workgroup barrier is involved.
//
//   inverse (sign +1)   in : x[q]  = a[gpos(t, q)],  q < M      (lane's share)
//                       out: y[k2] = A[rpos(t, k2)], k2 < S      (lanes t < M only)
//     DFT_M over q in registers, twiddle W_N^{t*k1} (Cooley-Tukey split only),
//     transpose via E, DFT_S over the S parts in registers.
//   forward (sign -1)   in : y[k2] = a[rpos(t, k2)]  (lanes t < M)
//                       out: x[m1] = A[gpos(t, m1)], m1 < M
//     the transpose of the inverse.
//
// Cooley-Tukey split (gcd(M, S) > 1): gpos = S*q + t, rpos = t + M*k2.
// Good-Thomas split (gcd(M, S) = 1, e.g. 90 = 9*10): gpos = (S*q + M*t) mod N,
// rpos = CRT(t mod M, k2 mod S), and the N twiddle multiplies disappear.
// Either way every lane knows the absolute positions of its elements, so
// G-space and r-space data are read and written in natural order: no
// permutation pass.
#pragma once

#include "gw_fft.h"

namespace gw {

__device__ __forceinline__ void wave_lds_sync() {
  // all LDS writes of this wave complete and visible before any later read
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, int N, int M, int S> struct WaveLine {
  static_assert(M * S == N, "N = M*S");
  static_assert(M <= S, "pass 2 must have no more tasks than lanes");
  static constexpr int LW = 64 / S;
  // padded k stride of the exchange area. A 16-byte FP64 element fills 4 of the
  // 64 four-byte banks, so 16 element slots repeat: with an even stride the
  // line stride S*EK is a multiple of 4 slots (of 16 when S*EK = 144) and the
  // LW lines of a wave collide; the smallest odd EK > M spreads them (at most
  // 2-way). FP32 elements (8 bytes, 32 slots) keep M + 1.
  static constexpr int EK = sizeof(T) == 8 ? (M % 2 == 0 ? M + 1 : M + 2) : M + 1;
  static constexpr int E_PER_WAVE = LW * S * EK;
  // gcd(M, S) = 1: Good-Thomas split, no twiddles between the passes
  static constexpr bool PFA = M > 1 && gcd_c(M, S) == 1;
  static constexpr int CM = PFA ? S * inv_mod(S % M, M) : 1;  // CRT weight of the M-digit
  static constexpr int CS = PFA ? M * inv_mod(M % S, S) : M;  // CRT weight of the S-digit

  // G-space position of lane part t's element q (q < M):
  //   S*q + t, or (S*q + M*t) mod N for the prime-factor split.
  __host__ __device__ static __forceinline__ int gpos(int t, int q) {
    if constexpr (PFA) {
      const int v = S * q + M * t;
      return v >= N ? v - N : v;
    } else {
      return S * q + t;
    }
  }
  // r-space position of lane part t's element k2 (t < M, k2 < S):
  //   t + M*k2, or (CM*t + CS*k2) mod N for the prime-factor split.
  __host__ __device__ static __forceinline__ int rpos(int t, int k2) {
    if constexpr (PFA) {
      const int v = (CM * t) % N + (CS * k2) % N;  // the k2 term folds once unrolled
      return v >= N ? v - N : v;
    } else {
      return t + M * k2;
    }
  }

  // inverse: x (M) -> y (S); tw = LDS table of e^{+2 pi i m/N} (unused for the prime-factor split).
  // Every lane reads E at an in-bounds slot (lanes t >= M the padding column), so y
  // needs no initial value; results of lanes t >= M are meaningless.
  __device__ static __forceinline__ void inverse(cpx<T> (&x)[M], cpx<T> (&y)[S], int t, int l,
                                                 bool part, cpx<T>* __restrict__ E,
                                                 const cpx<T>* __restrict__ tw) {
    dft<M, +1, T>(x);
    if (part) {
      if constexpr (!PFA) {
#pragma unroll
        for (int k = 1; k < M; ++k) x[k] = twmul<+1, T>(x[k], tw[t * k]);
      }
#pragma unroll
      for (int k = 0; k < M; ++k) E[(l * S + t) * EK + k] = x[k];
    }
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < S; ++s) y[s] = E[(l * S + s) * EK + min(t, M)];
    wave_lds_sync();
    dft<S, +1, T>(y);
  }

  // forward: y (S, lanes t < M) -> x (M); x of lanes that hold no part is meaningless
  __device__ static __forceinline__ void forward(cpx<T> (&y)[S], cpx<T> (&x)[M], int t, int l,
                                                 bool part, cpx<T>* __restrict__ E,
                                                 const cpx<T>* __restrict__ tw) {
    dft<S, -1, T>(y);
    if (t < M) {
      if constexpr (!PFA) {
#pragma unroll
        for (int s = 1; s < S; ++s) y[s] = twmul<-1, T>(y[s], tw[t * s]);
      }
#pragma unroll
      for (int s = 0; s < S; ++s) E[(l * S + s) * EK + t] = y[s];
    }
    wave_lds_sync();
    const int tr = min(t, S - 1);  // lanes past the last line read an in-bounds row
#pragma unroll
    for (int k = 0; k < M; ++k) x[k] = E[(l * S + tr) * EK + k];
    wave_lds_sync();
    dft<M, -1, T>(x);
  }
};

}  // namespace gw
