// Wave-cooperative line FFTs: a length-N line (N = M*S) is spread over S lanes
// of one wave, each lane holding at most max(M, S) <= 16 complex values.
//
// Lane layout: a wave holds LW = 64/S lines; lane = t*LW + l carries part t of
// line l (lanes >= S*LW idle). Exchange between the two passes goes through a
// wave-private LDS area E[l][s][k], so no workgroup barrier is involved.
//
//   inverse (sign +1)   in : x[q]  = a[S*q + t],  q < M        (lane's strided share)
//                       out: y[k2] = A[t + M*k2], k2 < S        (lanes t < M only)
//     DFT_M over q in registers, twiddle W_N^{t*k1}, transpose via E,
//     DFT_S over the S parts in registers.
//   forward (sign -1)   in : y[k2] = a[t + M*k2]  (lanes t < M)
//                       out: x[m1] = A[S*m1 + t], m1 < M
//     the transpose of the inverse.
//
// G-space data thus enters and leaves as the lanes' strided shares and r-space
// data is held as (t + M*k2) stripes in natural order: no permutation pass.
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
  static constexpr int EK = M + 1;           // padded k stride of the exchange area
  static constexpr int E_PER_WAVE = LW * S * EK;

  // inverse: x (M) -> y (S); tw = LDS table of e^{+2 pi i m/N}
  __device__ static __forceinline__ void inverse(cpx<T> (&x)[M], cpx<T> (&y)[S], int t, int l,
                                                 bool part, cpx<T>* __restrict__ E,
                                                 const cpx<T>* __restrict__ tw) {
    dft<M, +1, T>(x);
    if (part) {
#pragma unroll
      for (int k = 1; k < M; ++k) x[k] = twmul<+1, T>(x[k], tw[t * k]);
#pragma unroll
      for (int k = 0; k < M; ++k) E[(l * S + t) * EK + k] = x[k];
    }
    wave_lds_sync();
    if (t < M) {
#pragma unroll
      for (int s = 0; s < S; ++s) y[s] = E[(l * S + s) * EK + t];
    }
    wave_lds_sync();
    dft<S, +1, T>(y);
  }

  // forward: y (S, lanes t < M) -> x (M)
  __device__ static __forceinline__ void forward(cpx<T> (&y)[S], cpx<T> (&x)[M], int t, int l,
                                                 bool part, cpx<T>* __restrict__ E,
                                                 const cpx<T>* __restrict__ tw) {
    dft<S, -1, T>(y);
    if (t < M) {
#pragma unroll
      for (int s = 1; s < S; ++s) y[s] = twmul<-1, T>(y[s], tw[t * s]);
#pragma unroll
      for (int s = 0; s < S; ++s) E[(l * S + s) * EK + t] = y[s];
    }
    wave_lds_sync();
    if (part) {
#pragma unroll
      for (int k = 0; k < M; ++k) x[k] = E[(l * S + t) * EK + k];
    }
    wave_lds_sync();
    dft<M, -1, T>(x);
  }
};

}  // namespace gw
