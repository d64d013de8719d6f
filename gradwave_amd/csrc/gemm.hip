// Complex contractions of H|psi> and Davidson on MFMA (fp64 16x16x4, fp32 16x16x4).
//
// No complex MFMA exists, so both kernels run a real GEMM in which one operand
// is the natural interleaved (re, im) storage and the other is expanded on the
// fly, making the complex result come out interleaved:
//
//  inner   C[i][j] = sum_g conj(X[i][g]) Y[j][g]            (long K = npw)
//          D (M x 2N) = Xr (M x 2K) . B (2K x 2N),
//          B[kk][2j] = Yr_j[kk],  B[kk][2j+1] = (Yi, -Yr)_j[kk]
//          split-K over the grid; partials in the compute type, summed in fp64
//          by a second kernel (or the K-chunks as one rocBLAS batched GEMM).
//  combine Z[a][g] = beta Z + sum_j U'[j][a] Y[j][g]        (long N = npw)
//          Zr (A x 2G) = U2 (A x 2J) . E (2J x 2G),
//          U2[a][2j] = Re U'[j][a], U2[a][2j+1] = Im U'[j][a],
//          E[2j][2g+c] = (Yr, Yi)[c], E[2j+1][2g+c] = (-Yi, Yr)[c]
//
// The reduction index of an MFMA step is free to permute as long as A and B
// agree, so each lane reads V consecutive reals of its row and feeds them to V
// consecutive MFMA steps. k_gemm_inner / k_gemm_combine stream operands from
// global memory straight into VGPRs; the _lds and _tile kernels stage 32-complex
// K-chunks of both operands in padded LDS tiles (next chunk prefetched into
// registers). MFMA 16x16x4 lane layout: A[i=l%16][k=l/16], B[k=l/16][j=l%16],
// D col l%16 and row Mfma<T>::row.
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include <rocblas/rocblas.h>

#include "gw_internal.h"

namespace gw {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Output row of accumulator register r in lane group lk = lane/16 (probed on
// gfx950): f64 16x16x4 interleaves rows (lk + 4r), f32 16x16x4 blocks them (4lk + r).
template <> struct Mfma<double> {
  using acc = v4d;
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
  static __device__ __forceinline__ acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  using acc = v4f;
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
  static __device__ __forceinline__ acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Linear workgroup id -> work id so that consecutive work ids share an XCD
// (dispatch places workgroup L on XCD L % 8); -1 for padding ids.
__device__ __forceinline__ long gemm_work_id(long total) {
  const long L = blockIdx.x;
  const long per = (total + 7) / 8;
  const long w = (L % 8) * per + L / 8;
  return w < total ? w : -1;
}

constexpr int kV = 8;        // reals per lane per MFMA group (4 complex)
constexpr int kTile = 64;    // workgroup tile (rows x real cols), 4 waves of 32x32
constexpr int kStepC = 2 * kV;  // complex elements of K consumed per group (4 lanes-groups x V reals / 2)

template <typename T>
__global__ void __launch_bounds__(256)
k_gemm_inner(int M, int N, int K, const cpx<T>* __restrict__ X, long ldx, long sx,
             const cpx<T>* __restrict__ Y, long ldy, long sy, int splits, int kchunk,
             T* __restrict__ part, int ntx, int nty, long total) {
  using A4 = typename Mfma<T>::acc;
  // work id order: (k, split) slowest, then row tile, then column tile, so the
  // tiles of one K-chunk run together on one XCD and share its L2
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int bx = (int)(wid % ntx), by = (int)((wid / ntx) % nty), bz = (int)(wid / ((long)ntx * nty));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int i0 = by * kTile + (w >> 1) * 32;
  const int jj0 = bx * kTile + (w & 1) * 32;  // real column of D
  const int k = bz / splits, sp = bz - k * splits;
  const int gb = sp * kchunk, ge = min(K, gb + kchunk);
  const cpx<T>* Xk = X + (long)k * sx;
  const cpx<T>* Yk = Y + (long)k * sy;
  A4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = A4{0, 0, 0, 0};
  const cpx<T>* xrow[2];
  bool xok[2];
#pragma unroll
  for (int bm = 0; bm < 2; ++bm) {
    const int row = i0 + bm * 16 + li;
    xok[bm] = row < M;
    xrow[bm] = Xk + (long)(xok[bm] ? row : 0) * ldx;
  }
  const cpx<T>* yrow[2];
  bool yok[2], yodd[2];
#pragma unroll
  for (int bn = 0; bn < 2; ++bn) {
    const int jj = jj0 + bn * 16 + li;
    const int j = jj >> 1;
    yok[bn] = j < N;
    yodd[bn] = jj & 1;
    yrow[bn] = Yk + (long)(yok[bn] ? j : 0) * ldy;
  }
  for (int g0 = gb; g0 < ge; g0 += kStepC) {
    const int gl = g0 + lk * (kV / 2);  // this lane's first complex element
    T av[2][kV], bv[2][kV];
#pragma unroll
    for (int bm = 0; bm < 2; ++bm) {
#pragma unroll
      for (int v = 0; v < kV / 2; ++v) {
        const int g = gl + v;
        cpx<T> x = (xok[bm] && g < ge) ? xrow[bm][g] : cpx<T>{0, 0};
        av[bm][2 * v] = x.x;
        av[bm][2 * v + 1] = x.y;
      }
    }
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
#pragma unroll
      for (int v = 0; v < kV / 2; ++v) {
        const int g = gl + v;
        cpx<T> y = (yok[bn] && g < ge) ? yrow[bn][g] : cpx<T>{0, 0};
        if (yodd[bn]) { bv[bn][2 * v] = y.y; bv[bn][2 * v + 1] = -y.x; }
        else { bv[bn][2 * v] = y.x; bv[bn][2 * v + 1] = y.y; }
      }
    }
#pragma unroll
    for (int v = 0; v < kV; ++v)
#pragma unroll
      for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
  }
  const int ncol = 2 * N;
  T* pk = part + ((long)k * splits + sp) * M * ncol;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + bm * 16 + Mfma<T>::row(lk, r);
        const int col = jj0 + bn * 16 + li;
        if (row < M && col < ncol) pk[(long)row * ncol + col] = acc[bm][bn][r];
      }
}

// Sums the split-K partials of each real output element in FP64. The 4 waves of a
// workgroup take every 4th split of 64 consecutive elements (coalesced rows, 4
// independent loads in flight per lane) and combine through LDS.
template <typename T>
__device__ __forceinline__ double reduce_splits(const T* __restrict__ pk, long stride, long e, bool ok, int splits) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  if (ok) {
    int s = w;
    for (; s + 12 < splits; s += 16) {
      a0 += (double)pk[(long)s * stride + e];
      a1 += (double)pk[(long)(s + 4) * stride + e];
      a2 += (double)pk[(long)(s + 8) * stride + e];
      a3 += (double)pk[(long)(s + 12) * stride + e];
    }
    for (; s < splits; s += 4) a0 += (double)pk[(long)s * stride + e];
  }
  red[w][lane] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  return (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// block bx of k-point k: 64 consecutive real elements of C
template <typename T>
__device__ __forceinline__ void reduce_body(long bx, int k, int M, int N, int splits, const T* __restrict__ part,
                                            cpx<T>* __restrict__ C, long ldc, long sc, int conj) {
  const long nreal = (long)M * 2 * N;
  const long e = bx * 64 + (threadIdx.x & 63);  // real element i*2N + 2j + c
  const double v = reduce_splits<T>(part + (long)k * splits * nreal, nreal, e, e < nreal, splits);
  if (threadIdx.x < 64 && e < nreal) {
    const long i = e / (2 * N), r = e - i * 2 * N;
    reinterpret_cast<T*>(C + (long)k * sc + i * ldc + (r >> 1))[r & 1] = (T)((conj && (r & 1)) ? -v : v);
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_gemm_inner_reduce(int M, int N, int splits, const T* __restrict__ part, cpx<T>* __restrict__ C, long ldc,
                    long sc, int conj) {
  reduce_body<T>(blockIdx.x, blockIdx.y, M, N, splits, part, C, ldc, sc, conj);
}


// ---------------------------------------------------------------------------
// LDS-staged versions: a workgroup stages a 64-row x 32-complex K-chunk of each
// operand in padded LDS tiles (16-byte global loads, next chunk prefetched into
// registers while the MFMAs of the current one run); each lane then feeds the
// MFMAs from its contiguous 8-complex K segment of the tile.
constexpr int kKC = 32;  // complex K elements per chunk

template <typename T> struct Lds;
template <> struct Lds<double> { static constexpr int PAD = 1; };  // 16-byte row skew
template <> struct Lds<float> { static constexpr int PAD = 2; };
template <typename T> constexpr int kLd = kKC + Lds<T>::PAD;  // LDS row stride (complex)

// work item `wid` of one inner product: tile (bx, by) of k-point bz / splits, split bz % splits
template <typename T>
__device__ __forceinline__ void inner_lds_body(long wid, int M, int N, int K, const cpx<T>* __restrict__ X,
                                               long ldx, long sx, const cpx<T>* __restrict__ Y, long ldy, long sy,
                                               int splits, int kchunk, T* __restrict__ part, int ntx, int nty) {
  using A4 = typename Mfma<T>::acc;
  constexpr int LD = kLd<T>;
  __shared__ cpx<T> As[64 * LD];
  __shared__ cpx<T> Bs[32 * LD];
  const int bx = (int)(wid % ntx), by = (int)((wid / ntx) % nty), bz = (int)(wid / ((long)ntx * nty));
  const int k = bz / splits, sp = bz - k * splits;
  const int gb = sp * kchunk, ge = min(K, gb + kchunk);
  const int row0 = by * 64, col0 = bx * 32;  // X rows, Y rows (complex columns of C)
  const cpx<T>* Xk = X + (long)k * sx;
  const cpx<T>* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  // global -> register staging: 64x32 of X (8 per thread) and 32x32 of Y (4 per thread)
  cpx<T> ra[8], rb[4];
  auto gload = [&](int g0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + 256 * i, r = e >> 5, cc = e & 31;
      const int row = row0 + r, g = g0 + cc;
      ra[i] = (row < M && g < ge) ? Xk[(long)row * ldx + g] : cpx<T>{0, 0};
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = tid + 256 * i, r = e >> 5, cc = e & 31;
      const int row = col0 + r, g = g0 + cc;
      rb[i] = (row < N && g < ge) ? Yk[(long)row * ldy + g] : cpx<T>{0, 0};
    }
  };
  A4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = A4{0, 0, 0, 0};
  if (gb < ge) gload(gb);
  for (int g0 = gb; g0 < ge; g0 += kKC) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) { const int e = tid + 256 * i; As[(e >> 5) * LD + (e & 31)] = ra[i]; }
#pragma unroll
    for (int i = 0; i < 4; ++i) { const int e = tid + 256 * i; Bs[(e >> 5) * LD + (e & 31)] = rb[i]; }
    __syncthreads();
    if (g0 + kKC < ge) gload(g0 + kKC);
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // two halves of the lane's 8-complex K segment
      T av[2][8], bv[2][8];
#pragma unroll
      for (int bm = 0; bm < 2; ++bm) {
        const cpx<T>* ap = As + (wm * 32 + bm * 16 + li) * LD + lk * 8 + h * 4;
#pragma unroll
        for (int v = 0; v < 4; ++v) { const cpx<T> x = ap[v]; av[bm][2 * v] = x.x; av[bm][2 * v + 1] = x.y; }
      }
#pragma unroll
      for (int bn = 0; bn < 2; ++bn) {
        const int jj = wn * 32 + bn * 16 + li;
        const cpx<T>* bp = Bs + (jj >> 1) * LD + lk * 8 + h * 4;
        const bool odd = jj & 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const cpx<T> y = bp[v];
          bv[bn][2 * v] = odd ? y.y : y.x;
          bv[bn][2 * v + 1] = odd ? -y.x : y.y;
        }
      }
#pragma unroll
      for (int v = 0; v < 8; ++v)
#pragma unroll
        for (int bm = 0; bm < 2; ++bm)
#pragma unroll
          for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
    }
  }
  const int ncol = 2 * N;
  T* pk = part + ((long)k * splits + sp) * M * ncol;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * 32 + bm * 16 + Mfma<T>::row(lk, r);
        const int col = 2 * col0 + wn * 32 + bn * 16 + li;
        if (row < M && col < ncol) pk[(long)row * ncol + col] = acc[bm][bn][r];
      }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_gemm_inner_lds(int M, int N, int K, const cpx<T>* __restrict__ X, long ldx, long sx,
                 const cpx<T>* __restrict__ Y, long ldy, long sy, int splits, int kchunk,
                 T* __restrict__ part, int ntx, int nty, long total) {
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  inner_lds_body<T>(wid, M, N, K, X, ldx, sx, Y, ldy, sy, splits, kchunk, part, ntx, nty);
}

// work item `wid` of one back-projection: tile (by, bx) of k-point wid / (ntx nty)
template <typename T>
__device__ __forceinline__ void combine_lds_body(long wid, int A, int G, int J, const cpx<T>* __restrict__ U,
                                                 long ldu, long su, int conj_u, const cpx<T>* __restrict__ Y,
                                                 long ldy, long sy, cpx<T>* __restrict__ Z, long ldz, long sz,
                                                 T beta, int ntx, int nty) {
  using A4 = typename Mfma<T>::acc;
  constexpr int LD = kLd<T>;
  __shared__ cpx<T> Ut[64 * LD];  // [a][j]
  __shared__ cpx<T> Yt[32 * LD];  // [g][j]
  const int by = (int)(wid % nty), bx = (int)((wid / nty) % ntx), k = (int)(wid / ((long)ntx * nty));
  const int a0 = by * 64, g0c = bx * 32;
  const cpx<T>* Uk = U + (long)k * su;
  const cpx<T>* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  const T usign = conj_u ? T(-1) : T(1);
  cpx<T> ru[8], ry[4];
  auto gload = [&](int j0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {  // U rows j (32) x columns a (64), a fastest
      const int e = tid + 256 * i, jr = e >> 6, ac = e & 63;
      const int j = j0 + jr, a = a0 + ac;
      ru[i] = (j < J && a < A) ? Uk[(long)j * ldu + a] : cpx<T>{0, 0};
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // Y rows j (32) x columns g (32), g fastest
      const int e = tid + 256 * i, jr = e >> 5, gc = e & 31;
      const int j = j0 + jr, g = g0c + gc;
      ry[i] = (j < J && g < G) ? Yk[(long)j * ldy + g] : cpx<T>{0, 0};
    }
  };
  A4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = A4{0, 0, 0, 0};
  if (J > 0) gload(0);
  for (int j0 = 0; j0 < J; j0 += kKC) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) { const int e = tid + 256 * i; Ut[(e & 63) * LD + (e >> 6)] = ru[i]; }
#pragma unroll
    for (int i = 0; i < 4; ++i) { const int e = tid + 256 * i; Yt[(e & 31) * LD + (e >> 5)] = ry[i]; }
    __syncthreads();
    if (j0 + kKC < J) gload(j0 + kKC);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      T av[2][8], bv[2][8];
#pragma unroll
      for (int bm = 0; bm < 2; ++bm) {
        const cpx<T>* up = Ut + (wm * 32 + bm * 16 + li) * LD + lk * 8 + h * 4;
#pragma unroll
        for (int v = 0; v < 4; ++v) { const cpx<T> u = up[v]; av[bm][2 * v] = u.x; av[bm][2 * v + 1] = usign * u.y; }
      }
#pragma unroll
      for (int bn = 0; bn < 2; ++bn) {
        const int col = wn * 32 + bn * 16 + li;
        const cpx<T>* yp = Yt + (col >> 1) * LD + lk * 8 + h * 4;
        const bool im = col & 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const cpx<T> y = yp[v];
          bv[bn][2 * v] = im ? y.y : y.x;
          bv[bn][2 * v + 1] = im ? y.x : -y.y;
        }
      }
#pragma unroll
      for (int v = 0; v < 8; ++v)
#pragma unroll
        for (int bm = 0; bm < 2; ++bm)
#pragma unroll
          for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
    }
  }
  const int ncol = 2 * G;
  T* Zk = reinterpret_cast<T*>(Z + (long)k * sz);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = a0 + wm * 32 + bm * 16 + Mfma<T>::row(lk, r);
        const int col = 2 * g0c + wn * 32 + bn * 16 + li;
        if (row < A && col < ncol) {
          T* zp = Zk + (long)row * 2 * ldz + col;
          const T v = acc[bm][bn][r];
          *zp = (beta != T(0)) ? v + beta * (*zp) : v;
        }
      }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_gemm_combine_lds(int A, int G, int J, const cpx<T>* __restrict__ U, long ldu, long su, int conj_u,
                   const cpx<T>* __restrict__ Y, long ldy, long sy, cpx<T>* __restrict__ Z, long ldz, long sz,
                   T beta, int ntx, int nty, long total) {
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  combine_lds_body<T>(wid, A, G, J, U, ldu, su, conj_u, Y, ldy, sy, Z, ldz, sz, beta, ntx, nty);
}

// Tile-shape variants of the two LDS-staged kernels: each wave computes WM x WN
// MFMA blocks of 16 x 16 (2 x 2 waves per workgroup), so a workgroup covers
// 32 WM rows x 32 WN real columns. (2, 2) is the kernels above; the autotuner
// picks among (2, 2), (4, 2), (2, 4) and (4, 4) per workload.
template <typename T, int WM, int WN>
__global__ void __launch_bounds__(256)
k_gemm_inner_tile(int M, int N, int K, const cpx<T>* __restrict__ X, long ldx, long sx,
                  const cpx<T>* __restrict__ Y, long ldy, long sy, int splits, int kchunk,
                  T* __restrict__ part, int ntx, int nty, long total) {
  using A4 = typename Mfma<T>::acc;
  constexpr int LD = kLd<T>;
  constexpr int TM = 32 * WM, TN = 16 * WN;  // X rows, Y rows (complex columns of C)
  __shared__ cpx<T> As[TM * LD];
  __shared__ cpx<T> Bs[TN * LD];
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int bx = (int)(wid % ntx), by = (int)((wid / ntx) % nty), bz = (int)(wid / ((long)ntx * nty));
  const int k = bz / splits, sp = bz - k * splits;
  const int gb = sp * kchunk, ge = min(K, gb + kchunk);
  const int row0 = by * TM, col0 = bx * TN;
  const cpx<T>* Xk = X + (long)k * sx;
  const cpx<T>* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  constexpr int NA = TM * kKC / 256, NB = TN * kKC / 256;
  cpx<T> ra[NA], rb[NB];
  auto gload = [&](int g0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = tid + 256 * i, r = e >> 5, cc = e & 31;
      const int row = row0 + r, g = g0 + cc;
      ra[i] = (row < M && g < ge) ? Xk[(long)row * ldx + g] : cpx<T>{0, 0};
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int e = tid + 256 * i, r = e >> 5, cc = e & 31;
      const int row = col0 + r, g = g0 + cc;
      rb[i] = (row < N && g < ge) ? Yk[(long)row * ldy + g] : cpx<T>{0, 0};
    }
  };
  A4 acc[WM][WN];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[a][b] = A4{0, 0, 0, 0};
  if (gb < ge) gload(gb);
  for (int g0 = gb; g0 < ge; g0 += kKC) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NA; ++i) { const int e = tid + 256 * i; As[(e >> 5) * LD + (e & 31)] = ra[i]; }
#pragma unroll
    for (int i = 0; i < NB; ++i) { const int e = tid + 256 * i; Bs[(e >> 5) * LD + (e & 31)] = rb[i]; }
    __syncthreads();
    if (g0 + kKC < ge) gload(g0 + kKC);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      T av[WM][8], bv[WN][8];
#pragma unroll
      for (int bm = 0; bm < WM; ++bm) {
        const cpx<T>* ap = As + (wm * 16 * WM + bm * 16 + li) * LD + lk * 8 + h * 4;
#pragma unroll
        for (int v = 0; v < 4; ++v) { const cpx<T> x = ap[v]; av[bm][2 * v] = x.x; av[bm][2 * v + 1] = x.y; }
      }
#pragma unroll
      for (int bn = 0; bn < WN; ++bn) {
        const int jj = wn * 16 * WN + bn * 16 + li;
        const cpx<T>* bp = Bs + (jj >> 1) * LD + lk * 8 + h * 4;
        const bool odd = jj & 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const cpx<T> y = bp[v];
          bv[bn][2 * v] = odd ? y.y : y.x;
          bv[bn][2 * v + 1] = odd ? -y.x : y.y;
        }
      }
#pragma unroll
      for (int v = 0; v < 8; ++v)
#pragma unroll
        for (int bm = 0; bm < WM; ++bm)
#pragma unroll
          for (int bn = 0; bn < WN; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
    }
  }
  const int ncol = 2 * N;
  T* pk = part + ((long)k * splits + sp) * M * ncol;
#pragma unroll
  for (int bm = 0; bm < WM; ++bm)
#pragma unroll
    for (int bn = 0; bn < WN; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * 16 * WM + bm * 16 + Mfma<T>::row(lk, r);
        const int col = 2 * col0 + wn * 16 * WN + bn * 16 + li;
        if (row < M && col < ncol) pk[(long)row * ncol + col] = acc[bm][bn][r];
      }
}

template <typename T, int WM, int WN>
__global__ void __launch_bounds__(256)
k_gemm_combine_tile(int A, int G, int J, const cpx<T>* __restrict__ U, long ldu, long su, int conj_u,
                    const cpx<T>* __restrict__ Y, long ldy, long sy, cpx<T>* __restrict__ Z, long ldz, long sz,
                    T beta, int ntx, int nty, long total) {
  using A4 = typename Mfma<T>::acc;
  constexpr int LD = kLd<T>;
  constexpr int TA = 32 * WM, TG = 16 * WN;
  __shared__ cpx<T> Ut[TA * LD];  // [a][j]
  __shared__ cpx<T> Yt[TG * LD];  // [g][j]
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int by = (int)(wid % nty), bx = (int)((wid / nty) % ntx), k = (int)(wid / ((long)ntx * nty));
  const int a0 = by * TA, g0c = bx * TG;
  const cpx<T>* Uk = U + (long)k * su;
  const cpx<T>* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  const T usign = conj_u ? T(-1) : T(1);
  constexpr int NU = TA * kKC / 256, NY = TG * kKC / 256;
  cpx<T> ru[NU], ry[NY];
  auto gload = [&](int j0) {
#pragma unroll
    for (int i = 0; i < NU; ++i) {  // U rows j (32) x columns a (TA), a fastest
      const int e = tid + 256 * i, jr = e / TA, ac = e % TA;
      const int j = j0 + jr, a = a0 + ac;
      ru[i] = (j < J && a < A) ? Uk[(long)j * ldu + a] : cpx<T>{0, 0};
    }
#pragma unroll
    for (int i = 0; i < NY; ++i) {  // Y rows j (32) x columns g (TG), g fastest
      const int e = tid + 256 * i, jr = e / TG, gc = e % TG;
      const int j = j0 + jr, g = g0c + gc;
      ry[i] = (j < J && g < G) ? Yk[(long)j * ldy + g] : cpx<T>{0, 0};
    }
  };
  A4 acc[WM][WN];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[a][b] = A4{0, 0, 0, 0};
  if (J > 0) gload(0);
  for (int j0 = 0; j0 < J; j0 += kKC) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NU; ++i) { const int e = tid + 256 * i; Ut[(e % TA) * LD + (e / TA)] = ru[i]; }
#pragma unroll
    for (int i = 0; i < NY; ++i) { const int e = tid + 256 * i; Yt[(e % TG) * LD + (e / TG)] = ry[i]; }
    __syncthreads();
    if (j0 + kKC < J) gload(j0 + kKC);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      T av[WM][8], bv[WN][8];
#pragma unroll
      for (int bm = 0; bm < WM; ++bm) {
        const cpx<T>* up = Ut + (wm * 16 * WM + bm * 16 + li) * LD + lk * 8 + h * 4;
#pragma unroll
        for (int v = 0; v < 4; ++v) { const cpx<T> u = up[v]; av[bm][2 * v] = u.x; av[bm][2 * v + 1] = usign * u.y; }
      }
#pragma unroll
      for (int bn = 0; bn < WN; ++bn) {
        const int col = wn * 16 * WN + bn * 16 + li;
        const cpx<T>* yp = Yt + (col >> 1) * LD + lk * 8 + h * 4;
        const bool im = col & 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const cpx<T> y = yp[v];
          bv[bn][2 * v] = im ? y.y : y.x;
          bv[bn][2 * v + 1] = im ? y.x : -y.y;
        }
      }
#pragma unroll
      for (int v = 0; v < 8; ++v)
#pragma unroll
        for (int bm = 0; bm < WM; ++bm)
#pragma unroll
          for (int bn = 0; bn < WN; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
    }
  }
  const int ncol = 2 * G;
  T* Zk = reinterpret_cast<T*>(Z + (long)k * sz);
#pragma unroll
  for (int bm = 0; bm < WM; ++bm)
#pragma unroll
    for (int bn = 0; bn < WN; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = a0 + wm * 16 * WM + bm * 16 + Mfma<T>::row(lk, r);
        const int col = 2 * g0c + wn * 16 * WN + bn * 16 + li;
        if (row < A && col < ncol) {
          T* zp = Zk + (long)row * 2 * ldz + col;
          const T v = acc[bm][bn][r];
          *zp = (beta != T(0)) ? v + beta * (*zp) : v;
        }
      }
}

// ---------------------------------------------------------------------------
// Complex-accumulator kernels (tile variants 6-13). Each 16 x 16 complex block of
// the result keeps its own accumulators, and every complex k feeds the MFMAs
// straight from the (re, im) storage of both operands: nothing is expanded on
// the fly. A workgroup (2 x 2 waves of BM x BN blocks) covers 32 BM x 32 BN
// complex results; both operands are staged through two padded LDS buffers in
// chunks of KC complex along K, with one barrier per chunk and the next chunk's
// global loads in flight during the current chunk's MFMAs. With a = a' + i a''
// (a'' = -Im a for conj(a), +Im a otherwise) and b = b' + i b'':
//   four real products (G3 = false): re += a'b' - a''b'', im += a'b'' + a''b'
//   three real products (G3 = true, Gauss): t1 += a'b', t2 += a''b'',
//     t3 += (a' + a'')(b' + b''); re = t1 - t2, im = t3 - t1 - t2
// The three-product form does 3/4 of the MFMA work for two VALU adds per
// operand value; its rounding error stays within a small multiple of the
// four-product form's (both sum |a||b| terms in the accumulation type).
template <typename T> struct CTileCfg;
// FP64: 8 complex per chunk row (128 B); odd row stride -> conflict-free column reads
template <> struct CTileCfg<double> { static constexpr int KC = 8, PADK = 1, PADM = 0; };
// FP32: 16 complex per chunk row (128 B); row stride KC + 2 / M-rows padded by 16 complex
template <> struct CTileCfg<float> { static constexpr int KC = 16, PADK = 2, PADM = 16; };

template <bool G3> constexpr int kNacc = G3 ? 3 : 2;

template <typename T, int BM, int BN, bool CONJ_A, bool G3>
__device__ __forceinline__ void cmma_step(const cpx<T> (&a)[BM], const cpx<T> (&b)[BN],
                                          typename Mfma<T>::acc (&acc)[kNacc<G3>][BM][BN]) {
  if constexpr (G3) {
    T bs[BN];
#pragma unroll
    for (int j = 0; j < BN; ++j) bs[j] = b[j].x + b[j].y;
#pragma unroll
    for (int i = 0; i < BM; ++i) {
      const T a2 = CONJ_A ? -a[i].y : a[i].y, as = a[i].x + a2;
#pragma unroll
      for (int j = 0; j < BN; ++j) {
        acc[0][i][j] = Mfma<T>::op(a[i].x, b[j].x, acc[0][i][j]);
        acc[1][i][j] = Mfma<T>::op(a2, b[j].y, acc[1][i][j]);
        acc[2][i][j] = Mfma<T>::op(as, bs[j], acc[2][i][j]);
      }
    }
  } else {
    // the a' products of every block first, then the a'' ones: 2 BM BN
    // independent MFMAs separate the two updates of one accumulator
#pragma unroll
    for (int i = 0; i < BM; ++i)
#pragma unroll
      for (int j = 0; j < BN; ++j) {
        acc[0][i][j] = Mfma<T>::op(a[i].x, b[j].x, acc[0][i][j]);
        acc[1][i][j] = Mfma<T>::op(a[i].x, b[j].y, acc[1][i][j]);
      }
#pragma unroll
    for (int i = 0; i < BM; ++i) {
      const T ai = a[i].y, nai = -a[i].y;
#pragma unroll
      for (int j = 0; j < BN; ++j) {
        acc[0][i][j] = Mfma<T>::op(CONJ_A ? ai : nai, b[j].y, acc[0][i][j]);
        acc[1][i][j] = Mfma<T>::op(CONJ_A ? nai : ai, b[j].x, acc[1][i][j]);
      }
    }
  }
}

// (re, im) of register r of block (i, j) from the accumulators
template <typename T, int BM, int BN, bool G3>
__device__ __forceinline__ cpx<T> cmma_result(const typename Mfma<T>::acc (&acc)[kNacc<G3>][BM][BN], int i, int j,
                                              int r) {
  if constexpr (G3) {
    const T t1 = acc[0][i][j][r], t2 = acc[1][i][j][r];
    return cpx<T>{t1 - t2, acc[2][i][j][r] - t1 - t2};
  } else {
    return cpx<T>{acc[0][i][j][r], acc[1][i][j][r]};
  }
}

// Inner product C = conj(X) Y^T over split s of K (both operands K-contiguous,
// LDS rows [m][k]); partials [k][split][M][N] complex in the compute type.
template <typename T, int BM, int BN, bool G3>
__global__ void __launch_bounds__(256)
k_cgemm_inner(int M, int N, int K, const cpx<T>* __restrict__ X, long ldx, long sx,
              const cpx<T>* __restrict__ Y, long ldy, long sy, int splits, int kchunk,
              T* __restrict__ part, int ntx, int nty, long total) {
  using A4 = typename Mfma<T>::acc;
  constexpr int KC = CTileCfg<T>::KC, LD = KC + CTileCfg<T>::PADK;
  constexpr int TM = 32 * BM, TN = 32 * BN;
  constexpr int NA = TM * KC / 256, NB = TN * KC / 256;
  static_assert(NA * 256 == TM * KC && NB * 256 == TN * KC, "tile / chunk shape");
  __shared__ cpx<T> As[2][TM * LD];
  __shared__ cpx<T> Bs[2][TN * LD];
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int bx = (int)(wid % ntx), by = (int)((wid / ntx) % nty), bz = (int)(wid / ((long)ntx * nty));
  const int k = bz / splits, sp = bz - k * splits;
  const int gb = sp * kchunk, ge = min(K, gb + kchunk);
  const int row0 = by * TM, col0 = bx * TN;
  const cpx<T>* Xk = X + (long)k * sx;
  const cpx<T>* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  cpx<T> ra[NA], rb[NB];
  // every lane loads from a clamped (valid) address and zeroes what lies outside,
  // so the loop body has no branches
  auto gload = [&](int g0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = tid + 256 * i, g = g0 + e % KC, row = row0 + e / KC;
      const cpx<T> v = Xk[(long)min(row, M - 1) * ldx + min(g, ge - 1)];
      ra[i] = (row < M && g < ge) ? v : cpx<T>{0, 0};
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int e = tid + 256 * i, g = g0 + e % KC, row = col0 + e / KC;
      const cpx<T> v = Yk[(long)min(row, N - 1) * ldy + min(g, ge - 1)];
      rb[i] = (row < N && g < ge) ? v : cpx<T>{0, 0};
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NA; ++i) { const int e = tid + 256 * i; As[buf][(e / KC) * LD + e % KC] = ra[i]; }
#pragma unroll
    for (int i = 0; i < NB; ++i) { const int e = tid + 256 * i; Bs[buf][(e / KC) * LD + e % KC] = rb[i]; }
  };
  A4 acc[kNacc<G3>][BM][BN];
#pragma unroll
  for (int q = 0; q < kNacc<G3>; ++q)
#pragma unroll
    for (int i = 0; i < BM; ++i)
#pragma unroll
      for (int j = 0; j < BN; ++j) acc[q][i][j] = A4{0, 0, 0, 0};
  const int nch = ge > gb ? (ge - gb + KC - 1) / KC : 0;
  auto compute = [&](int buf) {
    const cpx<T>* as = As[buf] + (wm * 16 * BM + li) * LD + lk;
    const cpx<T>* bs = Bs[buf] + (wn * 16 * BN + li) * LD + lk;
#pragma unroll
    for (int s = 0; s < KC / 4; ++s) {
      cpx<T> a[BM], b[BN];
#pragma unroll
      for (int i = 0; i < BM; ++i) a[i] = as[i * 16 * LD + 4 * s];
#pragma unroll
      for (int j = 0; j < BN; ++j) b[j] = bs[j * 16 * LD + 4 * s];
      cmma_step<T, BM, BN, true, G3>(a, b, acc);
    }
  };
  if (nch > 0) {
    gload(gb);
    sstore(0);
    __syncthreads();
    // steady state without conditionals; the last chunk is peeled
    for (int c = 0; c + 1 < nch; ++c) {
      gload(gb + (c + 1) * KC);
      compute(c & 1);
      sstore((c + 1) & 1);
      __syncthreads();
    }
    compute((nch - 1) & 1);
  }
  cpx<T>* pk = reinterpret_cast<cpx<T>*>(part) + ((long)k * splits + sp) * M * N;
#pragma unroll
  for (int i = 0; i < BM; ++i)
#pragma unroll
    for (int j = 0; j < BN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * 16 * BM + i * 16 + Mfma<T>::row(lk, r);
        const int col = col0 + wn * 16 * BN + j * 16 + li;
        if (row < M && col < N) pk[(long)row * N + col] = cmma_result<T, BM, BN, G3>(acc, i, j, r);
      }
}

// Combine Z = beta Z + U'^T Y (U' = U or conj U, both operands stored [j][.],
// LDS rows [j][a] and [j][g]); the whole J per workgroup.
template <typename T, int BM, int BN, bool G3, bool CONJ>
__global__ void __launch_bounds__(256)
k_cgemm_combine(int A, int G, int J, const cpx<T>* __restrict__ U, long ldu, long su,
                const cpx<T>* __restrict__ Y, long ldy, long sy, cpx<T>* __restrict__ Z, long ldz, long sz,
                T beta, int ntx, int nty, long total) {
  using A4 = typename Mfma<T>::acc;
  constexpr int KC = CTileCfg<T>::KC;
  constexpr int TA = 32 * BM, TG = 32 * BN;
  constexpr int LDU = TA + CTileCfg<T>::PADM, LDY = TG + CTileCfg<T>::PADM;
  constexpr int NU = TA * KC / 256, NY = TG * KC / 256;
  static_assert(NU * 256 == TA * KC && NY * 256 == TG * KC, "tile / chunk shape");
  __shared__ cpx<T> Us[2][KC * LDU];
  __shared__ cpx<T> Ys[2][KC * LDY];
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int by = (int)(wid % nty), bx = (int)((wid / nty) % ntx), k = (int)(wid / ((long)ntx * nty));
  const int a0 = by * TA, g0 = bx * TG;
  const cpx<T>* Uk = U + (long)k * su;
  const cpx<T>* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  cpx<T> ru[NU], ry[NY];
  // clamped addresses, zeroed outside: no branches in the loop body
  auto gload = [&](int j0) {
#pragma unroll
    for (int i = 0; i < NU; ++i) {
      const int e = tid + 256 * i, j = j0 + e / TA, a = a0 + e % TA;
      const cpx<T> v = Uk[(long)min(j, J - 1) * ldu + min(a, A - 1)];
      ru[i] = (j < J && a < A) ? v : cpx<T>{0, 0};
    }
#pragma unroll
    for (int i = 0; i < NY; ++i) {
      const int e = tid + 256 * i, j = j0 + e / TG, g = g0 + e % TG;
      const cpx<T> v = Yk[(long)min(j, J - 1) * ldy + min(g, G - 1)];
      ry[i] = (j < J && g < G) ? v : cpx<T>{0, 0};
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NU; ++i) { const int e = tid + 256 * i; Us[buf][(e / TA) * LDU + e % TA] = ru[i]; }
#pragma unroll
    for (int i = 0; i < NY; ++i) { const int e = tid + 256 * i; Ys[buf][(e / TG) * LDY + e % TG] = ry[i]; }
  };
  A4 acc[kNacc<G3>][BM][BN];
#pragma unroll
  for (int q = 0; q < kNacc<G3>; ++q)
#pragma unroll
    for (int i = 0; i < BM; ++i)
#pragma unroll
      for (int j = 0; j < BN; ++j) acc[q][i][j] = A4{0, 0, 0, 0};
  const int nch = (J + KC - 1) / KC;
  auto compute = [&](int buf) {
    const cpx<T>* us = Us[buf] + lk * LDU + wm * 16 * BM + li;
    const cpx<T>* ys = Ys[buf] + lk * LDY + wn * 16 * BN + li;
#pragma unroll
    for (int s = 0; s < KC / 4; ++s) {
      cpx<T> a[BM], b[BN];
#pragma unroll
      for (int i = 0; i < BM; ++i) a[i] = us[4 * s * LDU + i * 16];
#pragma unroll
      for (int j = 0; j < BN; ++j) b[j] = ys[4 * s * LDY + j * 16];
      cmma_step<T, BM, BN, CONJ, G3>(a, b, acc);
    }
  };
  if (nch > 0) {
    gload(0);
    sstore(0);
    __syncthreads();
    for (int c = 0; c + 1 < nch; ++c) {  // the last chunk is peeled
      gload((c + 1) * KC);
      compute(c & 1);
      sstore((c + 1) & 1);
      __syncthreads();
    }
    compute((nch - 1) & 1);
  }
  cpx<T>* Zk = Z + (long)k * sz;
#pragma unroll
  for (int i = 0; i < BM; ++i)
#pragma unroll
    for (int j = 0; j < BN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = a0 + wm * 16 * BM + i * 16 + Mfma<T>::row(lk, r);
        const int col = g0 + wn * 16 * BN + j * 16 + li;
        if (row < A && col < G) {
          cpx<T>* zp = Zk + (long)row * ldz + col;
          cpx<T> v = cmma_result<T, BM, BN, G3>(acc, i, j, r);
          if (beta != T(0)) {
            const cpx<T> o = *zp;
            v.x += beta * o.x;
            v.y += beta * o.y;
          }
          *zp = v;
        }
      }
}

// ---------------------------------------------------------------------------
// Grouped launches: the jobs' work items are concatenated; a workgroup finds its
// job by binary search over the jobs' first items and runs the single-job body.
template <typename Job>
__device__ __forceinline__ int group_find(const Job* __restrict__ jobs, int n, long w) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= w) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T>
__global__ void __launch_bounds__(256) k_group_inner(const GroupInner* __restrict__ jobs, int n, long total) {
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const GroupInner& j = jobs[group_find(jobs, n, wid)];
  inner_lds_body<T>(wid - j.first, j.M, j.N, j.K, (const cpx<T>*)j.X, j.ldx, j.sx, (const cpx<T>*)j.Y, j.ldy,
                    j.sy, j.splits, j.kchunk, (T*)j.part, j.ntx, j.nty);
}

template <typename T>
__global__ void __launch_bounds__(256) k_group_reduce(const GroupReduce* __restrict__ jobs, int n, long total) {
  const long b = blockIdx.x;
  if (b >= total) return;
  const GroupReduce& j = jobs[group_find(jobs, n, b)];
  const long local = b - j.first;
  reduce_body<T>(local % j.nblk, (int)(local / j.nblk), j.M, j.N, j.splits, (const T*)j.part, (cpx<T>*)j.C, j.ldc,
                 j.sc, 0);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_group_combine(const GroupCombine* __restrict__ jobs, int n, long total, T beta) {
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const GroupCombine& j = jobs[group_find(jobs, n, wid)];
  combine_lds_body<T>(wid - j.first, j.A, j.G, j.J, (const cpx<T>*)j.U, j.ldu, j.su, 0, (const cpx<T>*)j.Y, j.ldy,
                      j.sy, (cpx<T>*)j.Z, j.ldz, j.sz, beta, j.ntx, j.nty);
}

static int inner_splits(int nk, int M, int N, int K);
void gemm_inner_shape(int nk, int M, int N, int K, int* splits, int* kchunk) {
  const int s = inner_splits(nk, M, N, K);
  int kc = (K + s - 1) / s;
  *splits = s;
  *kchunk = ((kc + kKC - 1) / kKC) * kKC;
}

template <typename T>
int gemm_group_inner(const GroupInner* jobs, int n, long total, hipStream_t st) {
  if (n <= 0 || total <= 0) return GW_OK;
  hipLaunchKernelGGL(k_group_inner<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, jobs, n, total);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
template <typename T>
int gemm_group_reduce(const GroupReduce* jobs, int n, long total, hipStream_t st) {
  if (n <= 0 || total <= 0) return GW_OK;
  hipLaunchKernelGGL(k_group_reduce<T>, dim3((unsigned)total), dim3(256), 0, st, jobs, n, total);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
template <typename T>
int gemm_group_combine(const GroupCombine* jobs, int n, long total, T beta, hipStream_t st) {
  if (n <= 0 || total <= 0) return GW_OK;
  hipLaunchKernelGGL(k_group_combine<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, jobs, n, total,
                     beta);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
template int gemm_group_inner<double>(const GroupInner*, int, long, hipStream_t);
template int gemm_group_inner<float>(const GroupInner*, int, long, hipStream_t);
template int gemm_group_reduce<double>(const GroupReduce*, int, long, hipStream_t);
template int gemm_group_reduce<float>(const GroupReduce*, int, long, hipStream_t);
template int gemm_group_combine<double>(const GroupCombine*, int, long, double, hipStream_t);
template int gemm_group_combine<float>(const GroupCombine*, int, long, float, hipStream_t);

// per precision (GW_FP64, GW_FP32): 0: (2, 2), 1: (4, 2), 2: (2, 4), 3: (4, 4) MFMA blocks per wave
// tile variant per [precision (0 FP64, 1 FP32)][contraction (0 inner, 1 combine)]
static int g_gemm_tile[2][2] = {{0, 0}, {0, 0}};
void set_gemm_tile(int precision, int op, int v) {
  const int t = (v >= 0 && v <= 13) ? v : 0;
  if (op != 1) g_gemm_tile[precision == GW_FP32][0] = t;
  if (op != 0) g_gemm_tile[precision == GW_FP32][1] = t;
}
int gemm_tile(int precision, int op) { return g_gemm_tile[precision == GW_FP32][op == 1]; }
template <typename T> static int tile_of(int op) { return g_gemm_tile[sizeof(T) == 4][op]; }

static bool gemm_simple() {
  static const bool v = [] { const char* e = getenv("GW_GEMM_SIMPLE"); return e && e[0] == '1'; }();
  return v;
}

static int g_target_wgs = 2048;  // split-K aims at this many workgroups (autotuned)
void set_gemm_target_wgs(int v) { g_target_wgs = v > 0 ? v : 2048; }
int gemm_target_wgs() { return g_target_wgs; }

static int inner_splits(int nk, int M, int N, int K) {
  const long tiles = (long)((2 * N + kTile - 1) / kTile) * ((M + kTile - 1) / kTile) * nk;
  long s = (g_target_wgs + tiles - 1) / tiles;
  const long smax = std::max<long>(1, (K + 4 * kStepC - 1) / (4 * kStepC));  // >= 4 groups per split
  s = std::max<long>(1, std::min<long>(s, smax));
  return (int)std::min<long>(s, 4096);
}

// Workgroup tile (complex rows x complex columns, as 32 BM x 32 BN) of the
// complex-accumulator variants 6-9 of each contraction.
// Variants 6-9 use four real products per complex one, 10-13 three.
static void ctile_shape(int variant, int op, int* bm, int* bn) {
  static const int inner[8][2] = {{2, 2}, {3, 3}, {2, 5}, {4, 3}, {2, 2}, {3, 2}, {2, 3}, {3, 3}};
  static const int comb[8][2] = {{2, 2}, {3, 2}, {5, 2}, {3, 4}, {2, 2}, {3, 2}, {2, 3}, {5, 2}};
  const int v = std::min(std::max(variant, 6), 13) - 6;
  *bm = op == 0 ? inner[v][0] : comb[v][0];
  *bn = op == 0 ? inner[v][1] : comb[v][1];
}

// Split-K count of the complex-accumulator inner product: the workgroup target
// over the tiles, each split at least 4 chunks long, at most 256 splits.
template <typename T>
static int ctile_splits(int nk, int M, int N, int K, int bm, int bn) {
  const long tiles = (long)((M + 32 * bm - 1) / (32 * bm)) * ((N + 32 * bn - 1) / (32 * bn)) * nk;
  long s = (g_target_wgs + tiles - 1) / tiles;
  const long smax = std::max<long>(1, K / (4 * CTileCfg<T>::KC));
  return (int)std::max<long>(1, std::min<long>(std::min<long>(s, smax), 256));
}

template <typename T>
__global__ void k_zero_rows(int M, int N, cpx<T>* __restrict__ C, long ldc, long sc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)M * N) return;
  const int i = (int)(e / N), j = (int)(e - (long)i * N);
  C[(long)blockIdx.y * sc + (long)i * ldc + j] = cpx<T>{T(0), T(0)};
}

// Z = beta Z (beta = 0 writes zeros, whatever Z held): a combine over an empty J
template <typename T>
__global__ void k_scale_rows(int A, int G, cpx<T>* __restrict__ Z, long ldz, long sz, T beta) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)A * G) return;
  const int a = (int)(e / G), g = (int)(e - (long)a * G);
  cpx<T>* z = Z + (long)blockIdx.y * sz + (long)a * ldz + g;
  const cpx<T> v = *z;
  *z = beta != T(0) ? cpx<T>{beta * v.x, beta * v.y} : cpx<T>{T(0), T(0)};
}

static int g_inner_lib[2] = {0, 0};
// Per-thread override: while set, every contraction takes the kernels here and
// none goes to rocBLAS (whose handles and workspaces allocate device memory).
static thread_local int g_kernels_only = 0;
int set_kernels_only(int v) {
  const int old = g_kernels_only;
  g_kernels_only = v ? 1 : 0;
  return old;
}

// K-chunks of the rocBLAS inner product: enough that the strided batch of
// 128x128-tile GEMMs reaches about 2048 tiles, at most 128 (every chunk adds an
// M x N partial to the reduce) and at least 64 K-values long
static int lib_splits(int M, int N, int K) {
  const long tiles = (long)((M + 127) / 128) * ((N + 127) / 128);
  long s = std::min<long>((2048 + tiles - 1) / tiles, 128);
  s = std::min<long>(s, std::max<long>(1, K / 64));
  return (int)std::max<long>(1, s);
}

template <typename T>
size_t gemm_inner_partials_bytes(int nk, int M, int N, int K) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  int s = inner_splits(nk, M, N, K);
  if (g_inner_lib[sizeof(T) == 4] && !g_kernels_only) s = std::max(s, lib_splits(M, N, K));
  if (tile_of<T>(0) >= 6) {
    int bm, bn;
    ctile_shape(tile_of<T>(0), 0, &bm, &bn);
    s = std::max(s, ctile_splits<T>(nk, M, N, K, bm, bn));
  }
  return (size_t)nk * s * M * 2 * N * sizeof(T);
}
template size_t gemm_inner_partials_bytes<double>(int, int, int, int);
template size_t gemm_inner_partials_bytes<float>(int, int, int, int);

void set_inner_lib(int precision, int v) { g_inner_lib[precision == GW_FP32] = v ? 1 : 0; }
int inner_lib(int precision) { return g_inner_lib[precision == GW_FP32]; }
static rocblas_handle blas_handle(hipStream_t st);

// Split-K inner product on rocBLAS: one strided-batched GEMM over the K-chunks
// (plus one for a short last chunk) writes conj(C) partials [chunk][M][N] in the
// compute type, in the layout k_gemm_inner_reduce sums (with conj = 1). Per chunk,
// column-major: conj(C)^T (N x M) = Y_chunk^H (N x Kc) X_chunk (Kc x M).
template <typename T>
static int inner_rocblas(int k, int M, int N, int K, int kc, int nfull, const cpx<T>* X, long ldx,
                         const cpx<T>* Y, long ldy, T* part, hipStream_t st) {
  rocblas_handle h = blas_handle(st);
  GW_REQUIRE(h, GW_ERR_HIP, "rocblas handle unavailable");
  const long slot = (long)M * N;  // complex elements per partial
  const int rem = K - nfull * kc;
  rocblas_status rs = rocblas_status_success;
  for (int pass = 0; pass < 2 && rs == rocblas_status_success; ++pass) {
    const int cnt = pass == 0 ? nfull : (rem > 0 ? 1 : 0);
    const int kk = pass == 0 ? kc : rem;
    if (cnt == 0 || kk == 0) continue;
    const long off = pass == 0 ? 0 : (long)nfull * kc;
    cpx<T>* cp = reinterpret_cast<cpx<T>*>(part) + (pass == 0 ? 0 : (long)nfull * slot);
    if constexpr (sizeof(T) == 8) {
      const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
      rs = rocblas_zgemm_strided_batched(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, N, M, kk,
                                         &one, (const rocblas_double_complex*)(Y + off), (rocblas_int)ldy, kc,
                                         (const rocblas_double_complex*)(X + off), (rocblas_int)ldx, kc, &zero,
                                         (rocblas_double_complex*)cp, N, slot, cnt);
    } else {
      const rocblas_float_complex one{1.0f, 0.0f}, zero{0.0f, 0.0f};
      rs = rocblas_cgemm_strided_batched(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, N, M, kk,
                                         &one, (const rocblas_float_complex*)(Y + off), (rocblas_int)ldy, kc,
                                         (const rocblas_float_complex*)(X + off), (rocblas_int)ldx, kc, &zero,
                                         (rocblas_float_complex*)cp, N, slot, cnt);
    }
  }
  (void)k;
  GW_REQUIRE(rs == rocblas_status_success, GW_ERR_HIP, "rocblas inner gemm failed (%d)", (int)rs);
  return GW_OK;
}

template <typename T>
int gemm_inner(int nk, int M, int N, int K, const cpx<T>* X, long ldx, long sx, const cpx<T>* Y, long ldy,
               long sy, cpx<T>* C, long ldc, long sc, void* partials, size_t part_bytes, hipStream_t st) {
  if (M <= 0 || N <= 0 || nk <= 0) return GW_OK;
  if (K <= 0) {  // an empty contraction: C = 0, by a kernel (safe inside a captured graph)
    hipLaunchKernelGGL(k_zero_rows<T>, dim3((unsigned)(((long)M * N + 255) / 256), nk), dim3(256), 0, st, M, N, C,
                       ldc, sc);
    GW_CHECK_HIP(hipGetLastError());
    return GW_OK;
  }
  const bool use_lib = g_inner_lib[sizeof(T) == 4] && !g_kernels_only && ldx < (1L << 31) && ldy < (1L << 31);
  const int splits = use_lib ? lib_splits(M, N, K) : inner_splits(nk, M, N, K);
  const size_t need = (size_t)nk * splits * M * 2 * N * sizeof(T);
  GW_REQUIRE(partials && part_bytes >= need, GW_ERR_WORKSPACE, "gemm_inner workspace %zu < %zu", part_bytes, need);
  int kchunk = (K + splits - 1) / splits;
  kchunk = ((kchunk + kKC - 1) / kKC) * kKC;  // multiple of both K granules (16, 32)
  if (use_lib) {
    // chunks of kchunk; the partial slots per k are the chunks actually used
    const int used = (K + kchunk - 1) / kchunk, nfull = K / kchunk;
    for (int k = 0; k < nk; ++k) {
      const int rc = inner_rocblas<T>(k, M, N, K, kchunk, nfull, X + (long)k * sx, ldx, Y + (long)k * sy, ldy,
                                      (T*)partials + (long)k * used * M * 2 * N, st);
      if (rc) return rc;
    }
    const long nreal = (long)M * 2 * N;
    hipLaunchKernelGGL(k_gemm_inner_reduce<T>, dim3((unsigned)((nreal + 63) / 64), nk), dim3(256), 0, st, M, N,
                       used, (const T*)partials, C, ldc, sc, 1);
    GW_CHECK_HIP(hipGetLastError());
    return GW_OK;
  }
  if (tile_of<T>(0) >= 6 && !gemm_simple()) {
    int bm, bn;
    ctile_shape(tile_of<T>(0), 0, &bm, &bn);
    const int cs = ctile_splits<T>(nk, M, N, K, bm, bn);
    const size_t cneed = (size_t)nk * cs * M * 2 * N * sizeof(T);
    GW_REQUIRE(part_bytes >= cneed, GW_ERR_WORKSPACE, "gemm_inner workspace %zu < %zu", part_bytes, cneed);
    int cchunk = (K + cs - 1) / cs;
    cchunk = ((cchunk + CTileCfg<T>::KC - 1) / CTileCfg<T>::KC) * CTileCfg<T>::KC;
    const int tx = (N + 32 * bn - 1) / (32 * bn), ty = (M + 32 * bm - 1) / (32 * bm);
    const long tt = (long)tx * ty * nk * cs;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3((unsigned)(8 * ((tt + 7) / 8))), dim3(256), 0, st, M, N, K, X, ldx, sx, Y, ldy,
                         sy, cs, cchunk, (T*)partials, tx, ty, tt);
    };
    switch (tile_of<T>(0)) {
      case 6: go(k_cgemm_inner<T, 2, 2, false>); break;
      case 7: go(k_cgemm_inner<T, 3, 3, false>); break;
      case 8: go(k_cgemm_inner<T, 2, 5, false>); break;
      case 9: go(k_cgemm_inner<T, 4, 3, false>); break;
      case 10: go(k_cgemm_inner<T, 2, 2, true>); break;
      case 11: go(k_cgemm_inner<T, 3, 2, true>); break;
      case 12: go(k_cgemm_inner<T, 2, 3, true>); break;
      default: go(k_cgemm_inner<T, 3, 3, true>); break;
    }
    const long nreal = (long)M * 2 * N;
    hipLaunchKernelGGL(k_gemm_inner_reduce<T>, dim3((unsigned)((nreal + 63) / 64), nk), dim3(256), 0, st, M, N, cs,
                       (const T*)partials, C, ldc, sc, 0);
    GW_CHECK_HIP(hipGetLastError());
    return GW_OK;
  }
  const int ntx = (2 * N + kTile - 1) / kTile, nty = (M + kTile - 1) / kTile;
  const long total = (long)ntx * nty * nk * splits;
  auto launch_tile = [&](auto kern, int wm, int wn) {
    const int tx = (2 * N + 32 * wn - 1) / (32 * wn), ty = (M + 32 * wm - 1) / (32 * wm);
    const long tt = (long)tx * ty * nk * splits;
    hipLaunchKernelGGL(kern, dim3((unsigned)(8 * ((tt + 7) / 8))), dim3(256), 0, st, M, N, K, X, ldx, sx, Y, ldy,
                       sy, splits, kchunk, (T*)partials, tx, ty, tt);
  };
  if (gemm_simple())
    hipLaunchKernelGGL(k_gemm_inner<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, M, N, K, X, ldx,
                       sx, Y, ldy, sy, splits, kchunk, (T*)partials, ntx, nty, total);
  else if (tile_of<T>(0) == 1)
    launch_tile(k_gemm_inner_tile<T, 4, 2>, 4, 2);
  else if (tile_of<T>(0) == 2)
    launch_tile(k_gemm_inner_tile<T, 2, 4>, 2, 4);
  else if (tile_of<T>(0) == 3)
    launch_tile(k_gemm_inner_tile<T, 4, 4>, 4, 4);
  else if (tile_of<T>(0) == 4)
    launch_tile(k_gemm_inner_tile<T, 2, 5>, 2, 5);
  else if (tile_of<T>(0) == 5)
    launch_tile(k_gemm_inner_tile<T, 4, 5>, 4, 5);
  else
    hipLaunchKernelGGL(k_gemm_inner_lds<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, M, N, K, X,
                       ldx, sx, Y, ldy, sy, splits, kchunk, (T*)partials, ntx, nty, total);
  const long nreal = (long)M * 2 * N;
  hipLaunchKernelGGL(k_gemm_inner_reduce<T>, dim3((unsigned)((nreal + 63) / 64), nk), dim3(256), 0, st, M, N,
                     splits, (const T*)partials, C, ldc, sc, 0);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
template int gemm_inner<double>(int, int, int, int, const double2*, long, long, const double2*, long, long,
                                double2*, long, long, void*, size_t, hipStream_t);
template int gemm_inner<float>(int, int, int, int, const float2*, long, long, const float2*, long, long,
                               float2*, long, long, void*, size_t, hipStream_t);

template <typename T>
__global__ void __launch_bounds__(256)
k_gemm_combine(int A, int G, int J, const cpx<T>* __restrict__ U, long ldu, long su, int conj_u,
               const cpx<T>* __restrict__ Y, long ldy, long sy, cpx<T>* __restrict__ Z, long ldz, long sz,
               T beta, int ntx, int nty, long total) {
  using A4 = typename Mfma<T>::acc;
  // work id order: k slowest, then column tile, then row tile, so the row tiles
  // that re-read one block of Y run together on one XCD
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int by = (int)(wid % nty), bx = (int)((wid / nty) % ntx), k = (int)(wid / ((long)ntx * nty));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int a0 = by * kTile + (w >> 1) * 32;
  const int c0 = bx * kTile + (w & 1) * 32;  // real column of Zr
  const cpx<T>* Uk = U + (long)k * su;
  const cpx<T>* Yk = Y + (long)k * sy;
  A4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = A4{0, 0, 0, 0};
  int arow[2];
  bool aok[2];
#pragma unroll
  for (int bm = 0; bm < 2; ++bm) { arow[bm] = a0 + bm * 16 + li; aok[bm] = arow[bm] < A; }
  int gcol[2];
  bool gok[2], gim[2];
#pragma unroll
  for (int bn = 0; bn < 2; ++bn) {
    const int col = c0 + bn * 16 + li;
    gcol[bn] = col >> 1;
    gim[bn] = col & 1;
    gok[bn] = gcol[bn] < G;
  }
  const T usign = conj_u ? T(-1) : T(1);
  for (int j0 = 0; j0 < J; j0 += kStepC) {
    const int jl = j0 + lk * (kV / 2);
    T av[2][kV], bv[2][kV];
#pragma unroll
    for (int v = 0; v < kV / 2; ++v) {
      const int j = jl + v;
      const bool jok = j < J;
#pragma unroll
      for (int bm = 0; bm < 2; ++bm) {
        cpx<T> u = (jok && aok[bm]) ? Uk[(long)j * ldu + arow[bm]] : cpx<T>{0, 0};
        av[bm][2 * v] = u.x;
        av[bm][2 * v + 1] = usign * u.y;
      }
#pragma unroll
      for (int bn = 0; bn < 2; ++bn) {
        cpx<T> y = (jok && gok[bn]) ? Yk[(long)j * ldy + gcol[bn]] : cpx<T>{0, 0};
        if (gim[bn]) { bv[bn][2 * v] = y.y; bv[bn][2 * v + 1] = y.x; }
        else { bv[bn][2 * v] = y.x; bv[bn][2 * v + 1] = -y.y; }
      }
    }
#pragma unroll
    for (int v = 0; v < kV; ++v)
#pragma unroll
      for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
  }
  const int ncol = 2 * G;
  T* Zk = reinterpret_cast<T*>(Z + (long)k * sz);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = a0 + bm * 16 + Mfma<T>::row(lk, r);
        const int col = c0 + bn * 16 + li;
        if (row < A && col < ncol) {
          T* zp = Zk + (long)row * 2 * ldz + col;
          const T v = acc[bm][bn][r];
          *zp = (beta != T(0)) ? v + beta * (*zp) : v;
        }
      }
}

// The combine is a plain complex GEMM, so it may also run on rocBLAS (the
// autotuner picks per precision; default: the kernels above). Row-major
// Z (A x G) = U'^T (A x J) Y (J x G) is column-major Z^T = Y^T U' with
// Y^T read as stored (N) and U' = U^T or U^H (T / C).
static int g_combine_lib[2] = {0, 0};
void set_combine_lib(int precision, int v) { g_combine_lib[precision == GW_FP32] = v ? 1 : 0; }
int combine_lib(int precision) { return g_combine_lib[precision == GW_FP32]; }

// One rocBLAS handle per (device, stream), bound to that stream once: GEMMs
// issued on different streams (the ragged batch's side lanes) never share a
// handle's stream binding or its device workspace.
static rocblas_handle blas_handle(hipStream_t st) {
  struct Entry {
    int dev;
    hipStream_t st;
    rocblas_handle h;
  };
  static std::mutex mu;
  static std::vector<Entry> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  for (const Entry& e : handles)
    if (e.dev == dev && e.st == st) return e.h;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
  if (rocblas_set_stream(h, st) != rocblas_status_success) {
    (void)rocblas_destroy_handle(h);
    return nullptr;
  }
  handles.push_back(Entry{dev, st, h});
  return h;
}

template <typename T>
static int combine_rocblas(int nk, int A, int G, int J, const cpx<T>* U, long ldu, long su, int conj_u,
                           const cpx<T>* Y, long ldy, long sy, cpx<T>* Z, long ldz, long sz, T beta, hipStream_t st) {
  rocblas_handle h = blas_handle(st);
  GW_REQUIRE(h, GW_ERR_HIP, "rocblas handle unavailable");
  const rocblas_operation tb = conj_u ? rocblas_operation_conjugate_transpose : rocblas_operation_transpose;
  rocblas_status rs;
  if constexpr (sizeof(T) == 8) {
    const rocblas_double_complex one{1.0, 0.0}, b{(double)beta, 0.0};
    rs = rocblas_zgemm_strided_batched(h, rocblas_operation_none, tb, G, A, J, &one,
                                       (const rocblas_double_complex*)Y, (rocblas_int)ldy, sy,
                                       (const rocblas_double_complex*)U, (rocblas_int)ldu, su, &b,
                                       (rocblas_double_complex*)Z, (rocblas_int)ldz, sz, nk);
  } else {
    const rocblas_float_complex one{1.0f, 0.0f}, b{(float)beta, 0.0f};
    rs = rocblas_cgemm_strided_batched(h, rocblas_operation_none, tb, G, A, J, &one,
                                       (const rocblas_float_complex*)Y, (rocblas_int)ldy, sy,
                                       (const rocblas_float_complex*)U, (rocblas_int)ldu, su, &b,
                                       (rocblas_float_complex*)Z, (rocblas_int)ldz, sz, nk);
  }
  GW_REQUIRE(rs == rocblas_status_success, GW_ERR_HIP, "rocblas gemm failed (%d)", (int)rs);
  return GW_OK;
}

template <typename T>
int gemm_combine(int nk, int A, int G, int J, const cpx<T>* U, long ldu, long su, int conj_u, const cpx<T>* Y,
                 long ldy, long sy, cpx<T>* Z, long ldz, long sz, T beta, hipStream_t st) {
  if (A <= 0 || G <= 0 || nk <= 0) return GW_OK;
  if (J <= 0) {
    hipLaunchKernelGGL(k_scale_rows<T>, dim3((unsigned)(((long)A * G + 255) / 256), nk), dim3(256), 0, st, A, G, Z,
                       ldz, sz, beta);
    GW_CHECK_HIP(hipGetLastError());
    return GW_OK;
  }
  if (g_combine_lib[sizeof(T) == 4] && !g_kernels_only && J > 0 && ldu < (1L << 31) && ldy < (1L << 31) &&
      ldz < (1L << 31))
    return combine_rocblas<T>(nk, A, G, J, U, ldu, su, conj_u, Y, ldy, sy, Z, ldz, sz, beta, st);
  if (tile_of<T>(1) >= 6 && !gemm_simple()) {
    int bm, bn;
    ctile_shape(tile_of<T>(1), 1, &bm, &bn);
    const int tx = (G + 32 * bn - 1) / (32 * bn), ty = (A + 32 * bm - 1) / (32 * bm);
    const long tt = (long)tx * ty * nk;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3((unsigned)(8 * ((tt + 7) / 8))), dim3(256), 0, st, A, G, J, U, ldu, su, Y, ldy,
                         sy, Z, ldz, sz, beta, tx, ty, tt);
    };
    const int v = tile_of<T>(1);
    if (conj_u) {
      switch (v) {
        case 6: go(k_cgemm_combine<T, 2, 2, false, true>); break;
        case 7: go(k_cgemm_combine<T, 3, 2, false, true>); break;
        case 8: go(k_cgemm_combine<T, 5, 2, false, true>); break;
        case 9: go(k_cgemm_combine<T, 3, 4, false, true>); break;
        case 10: go(k_cgemm_combine<T, 2, 2, true, true>); break;
        case 11: go(k_cgemm_combine<T, 3, 2, true, true>); break;
        case 12: go(k_cgemm_combine<T, 2, 3, true, true>); break;
        default: go(k_cgemm_combine<T, 5, 2, true, true>); break;
      }
    } else {
      switch (v) {
        case 6: go(k_cgemm_combine<T, 2, 2, false, false>); break;
        case 7: go(k_cgemm_combine<T, 3, 2, false, false>); break;
        case 8: go(k_cgemm_combine<T, 5, 2, false, false>); break;
        case 9: go(k_cgemm_combine<T, 3, 4, false, false>); break;
        case 10: go(k_cgemm_combine<T, 2, 2, true, false>); break;
        case 11: go(k_cgemm_combine<T, 3, 2, true, false>); break;
        case 12: go(k_cgemm_combine<T, 2, 3, true, false>); break;
        default: go(k_cgemm_combine<T, 5, 2, true, false>); break;
      }
    }
    GW_CHECK_HIP(hipGetLastError());
    return GW_OK;
  }
  const int ntx = (2 * G + kTile - 1) / kTile, nty = (A + kTile - 1) / kTile;
  const long total = (long)ntx * nty * nk;
  auto launch_tile = [&](auto kern, int wm, int wn) {
    const int tx = (2 * G + 32 * wn - 1) / (32 * wn), ty = (A + 32 * wm - 1) / (32 * wm);
    const long tt = (long)tx * ty * nk;
    hipLaunchKernelGGL(kern, dim3((unsigned)(8 * ((tt + 7) / 8))), dim3(256), 0, st, A, G, J, U, ldu, su, conj_u, Y,
                       ldy, sy, Z, ldz, sz, beta, tx, ty, tt);
  };
  if (gemm_simple())
    hipLaunchKernelGGL(k_gemm_combine<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, A, G, J, U,
                       ldu, su, conj_u, Y, ldy, sy, Z, ldz, sz, beta, ntx, nty, total);
  else if (tile_of<T>(1) == 1)
    launch_tile(k_gemm_combine_tile<T, 4, 2>, 4, 2);
  else if (tile_of<T>(1) == 2)
    launch_tile(k_gemm_combine_tile<T, 2, 4>, 2, 4);
  else if (tile_of<T>(1) == 3)
    launch_tile(k_gemm_combine_tile<T, 4, 4>, 4, 4);
  else if (tile_of<T>(1) == 4)
    launch_tile(k_gemm_combine_tile<T, 5, 2>, 5, 2);
  else if (tile_of<T>(1) == 5)
    launch_tile(k_gemm_combine_tile<T, 5, 4>, 5, 4);
  else
    hipLaunchKernelGGL(k_gemm_combine_lds<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, A, G, J,
                       U, ldu, su, conj_u, Y, ldy, sy, Z, ldz, sz, beta, ntx, nty, total);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
template int gemm_combine<double>(int, int, int, int, const double2*, long, long, int, const double2*, long, long,
                                  double2*, long, long, double, hipStream_t);
template int gemm_combine<float>(int, int, int, int, const float2*, long, long, int, const float2*, long, long,
                                 float2*, long, long, float, hipStream_t);

}  // namespace gw

using namespace gw;

extern "C" size_t gw_zgemm_inner_workspace_bytes(int precision, int nk, int M, int N, int K) {
  return precision == GW_FP64 ? gemm_inner_partials_bytes<double>(nk, M, N, K)
                              : gemm_inner_partials_bytes<float>(nk, M, N, K);
}

extern "C" int gw_zgemm_inner(int precision, int nk, int M, int N, int K, const void* X, long ldx, long sx,
                              const void* Y, long ldy, long sy, void* C, long ldc, long sc, void* ws,
                              size_t ws_bytes, void* stream) {
  GW_REQUIRE(nk >= 0 && M >= 0 && N >= 0 && K >= 0, GW_ERR_ARG, "bad argument");
  GW_REQUIRE((X || (long)M * K == 0) && (Y || (long)N * K == 0) && (C || (long)M * N == 0), GW_ERR_ARG,
             "null operand of non-zero size");
  GW_REQUIRE(ldx >= K && ldy >= K && ldc >= N, GW_ERR_ARG, "leading dimension too small");
  if (precision == GW_FP64)
    return gemm_inner<double>(nk, M, N, K, (const double2*)X, ldx, sx, (const double2*)Y, ldy, sy, (double2*)C,
                              ldc, sc, ws, ws_bytes, (hipStream_t)stream);
  return gemm_inner<float>(nk, M, N, K, (const float2*)X, ldx, sx, (const float2*)Y, ldy, sy, (float2*)C, ldc,
                           sc, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int gw_zgemm_combine(int precision, int nk, int A, int G, int J, const void* U, long ldu, long su,
                                int conj_u, const void* Y, long ldy, long sy, void* Z, long ldz, long sz,
                                double beta, void* stream) {
  GW_REQUIRE(nk >= 0 && A >= 0 && G >= 0 && J >= 0, GW_ERR_ARG, "bad argument");
  GW_REQUIRE((U || (long)J * A == 0) && (Y || (long)J * G == 0) && (Z || (long)A * G == 0), GW_ERR_ARG,
             "null operand of non-zero size");
  GW_REQUIRE(ldu >= A && ldy >= G && ldz >= G, GW_ERR_ARG, "leading dimension too small");
  if (precision == GW_FP64)
    return gemm_combine<double>(nk, A, G, J, (const double2*)U, ldu, su, conj_u, (const double2*)Y, ldy, sy,
                                (double2*)Z, ldz, sz, beta, (hipStream_t)stream);
  return gemm_combine<float>(nk, A, G, J, (const float2*)U, ldu, su, conj_u, (const float2*)Y, ldy, sy,
                             (float2*)Z, ldz, sz, (float)beta, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Real contractions (Gamma layout, core/gamma.py: the real-embedded Davidson):
//   inner   C[i][j] = sum_g X[i][g] Y[j][g]              split-K, fp64 partials
//   combine Z[a][g] = beta Z + alpha sum_j U[j][a] Y[j][g]
// Same structure as the LDS-staged complex kernels: 64x64 workgroup tiles of
// four 32x32 wave tiles, 32-wide K chunks staged through padded LDS with the
// next chunk prefetched into registers.
constexpr int kKR = 32;
template <typename T> constexpr int kLdR = kKR + (sizeof(T) == 8 ? 2 : 4);

namespace {

template <typename T>
__global__ void __launch_bounds__(256)
k_rgemm_inner_lds(int M, int N, int K, const T* __restrict__ X, long ldx, long sx, const T* __restrict__ Y,
                  long ldy, long sy, int splits, int kchunk, T* __restrict__ part, int ntx, int nty,
                  long total) {
  using A4 = typename Mfma<T>::acc;
  constexpr int LD = kLdR<T>;
  __shared__ T As[64 * LD];
  __shared__ T Bs[64 * LD];
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int bx = (int)(wid % ntx), by = (int)((wid / ntx) % nty), bz = (int)(wid / ((long)ntx * nty));
  const int k = bz / splits, sp = bz - k * splits;
  const int gb = sp * kchunk, ge = min(K, gb + kchunk);
  const int row0 = by * 64, col0 = bx * 64;
  const T* Xk = X + (long)k * sx;
  const T* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  T ra[8], rb[8];
  auto gload = [&](int g0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + 256 * i, r = e >> 5, cc = e & 31;
      const int g = g0 + cc;
      const int xr = row0 + r, yr = col0 + r;
      ra[i] = (xr < M && g < ge) ? Xk[(long)xr * ldx + g] : T(0);
      rb[i] = (yr < N && g < ge) ? Yk[(long)yr * ldy + g] : T(0);
    }
  };
  A4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = A4{0, 0, 0, 0};
  if (gb < ge) gload(gb);
  for (int g0 = gb; g0 < ge; g0 += kKR) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + 256 * i;
      As[(e >> 5) * LD + (e & 31)] = ra[i];
      Bs[(e >> 5) * LD + (e & 31)] = rb[i];
    }
    __syncthreads();
    if (g0 + kKR < ge) gload(g0 + kKR);
    T av[2][8], bv[2][8];
#pragma unroll
    for (int bm = 0; bm < 2; ++bm) {
      const T* ap = As + (wm * 32 + bm * 16 + li) * LD + lk * 8;
#pragma unroll
      for (int v = 0; v < 8; ++v) av[bm][v] = ap[v];
    }
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
      const T* bp = Bs + (wn * 32 + bn * 16 + li) * LD + lk * 8;
#pragma unroll
      for (int v = 0; v < 8; ++v) bv[bn][v] = bp[v];
    }
#pragma unroll
    for (int v = 0; v < 8; ++v)
#pragma unroll
      for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
  }
  T* pk = part + ((long)k * splits + sp) * M * N;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * 32 + bm * 16 + Mfma<T>::row(lk, r);
        const int col = col0 + wn * 32 + bn * 16 + li;
        if (row < M && col < N) pk[(long)row * N + col] = acc[bm][bn][r];
      }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_rgemm_inner_reduce(int M, int N, int splits, const T* __restrict__ part, T* __restrict__ C, long ldc, long sc) {
  const int k = blockIdx.y;
  const long tot = (long)M * N;
  const long e = (long)blockIdx.x * 64 + (threadIdx.x & 63);
  const double v = reduce_splits<T>(part + (long)k * splits * tot, tot, e, e < tot, splits);
  if (threadIdx.x < 64 && e < tot) {
    const long i = e / N, j = e - i * N;
    C[(long)k * sc + i * ldc + j] = (T)v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_rgemm_combine_lds(int A, int G, int J, const T* __restrict__ U, long ldu, long su, const T* __restrict__ Y,
                    long ldy, long sy, T* __restrict__ Z, long ldz, long sz, T alpha, T beta, int ntx, int nty,
                    long total) {
  using A4 = typename Mfma<T>::acc;
  constexpr int LD = kLdR<T>;
  __shared__ T Ut[64 * LD];  // [a][j]
  __shared__ T Yt[64 * LD];  // [g][j]
  const long wid = gemm_work_id(total);
  if (wid < 0) return;
  const int by = (int)(wid % nty), bx = (int)((wid / nty) % ntx), k = (int)(wid / ((long)ntx * nty));
  const int a0 = by * 64, g0 = bx * 64;
  const T* Uk = U + (long)k * su;
  const T* Yk = Y + (long)k * sy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w >> 1, wn = w & 1;
  T ru[8], ry[8];
  auto gload = [&](int j0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {  // rows j (32) x columns (64), column fastest
      const int e = tid + 256 * i, jr = e >> 6, c = e & 63;
      const int j = j0 + jr;
      ru[i] = (j < J && a0 + c < A) ? Uk[(long)j * ldu + a0 + c] : T(0);
      ry[i] = (j < J && g0 + c < G) ? Yk[(long)j * ldy + g0 + c] : T(0);
    }
  };
  A4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = A4{0, 0, 0, 0};
  if (J > 0) gload(0);
  for (int j0 = 0; j0 < J; j0 += kKR) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + 256 * i;
      Ut[(e & 63) * LD + (e >> 6)] = ru[i];
      Yt[(e & 63) * LD + (e >> 6)] = ry[i];
    }
    __syncthreads();
    if (j0 + kKR < J) gload(j0 + kKR);
    T av[2][8], bv[2][8];
#pragma unroll
    for (int bm = 0; bm < 2; ++bm) {
      const T* up = Ut + (wm * 32 + bm * 16 + li) * LD + lk * 8;
#pragma unroll
      for (int v = 0; v < 8; ++v) av[bm][v] = up[v];
    }
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
      const T* yp = Yt + (wn * 32 + bn * 16 + li) * LD + lk * 8;
#pragma unroll
      for (int v = 0; v < 8; ++v) bv[bn][v] = yp[v];
    }
#pragma unroll
    for (int v = 0; v < 8; ++v)
#pragma unroll
      for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = Mfma<T>::op(av[bm][v], bv[bn][v], acc[bm][bn]);
  }
  T* Zk = Z + (long)k * sz;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = a0 + wm * 32 + bm * 16 + Mfma<T>::row(lk, r);
        const int col = g0 + wn * 32 + bn * 16 + li;
        if (row < A && col < G) {
          T* zp = Zk + (long)row * ldz + col;
          const T v = alpha * acc[bm][bn][r];
          *zp = (beta != T(0)) ? v + beta * (*zp) : v;
        }
      }
}

int rinner_splits(int nk, int M, int N, int K) {
  const long tiles = (long)((N + 63) / 64) * ((M + 63) / 64) * nk;
  long s = (g_target_wgs + tiles - 1) / tiles;
  const long smax = std::max<long>(1, (K + 4 * kKR - 1) / (4 * kKR));  // >= 4 chunks per split
  s = std::max<long>(1, std::min<long>(s, smax));
  return (int)std::min<long>(s, 4096);
}

template <typename T>
int rgemm_inner(int nk, int M, int N, int K, const T* X, long ldx, long sx, const T* Y, long ldy, long sy, T* C,
                long ldc, long sc, void* partials, size_t part_bytes, hipStream_t st) {
  if (M <= 0 || N <= 0 || nk <= 0) return GW_OK;
  GW_REQUIRE(K > 0, GW_ERR_ARG, "K must be positive");
  const int splits = rinner_splits(nk, M, N, K);
  const size_t need = (size_t)nk * splits * M * N * sizeof(T);
  GW_REQUIRE(partials && part_bytes >= need, GW_ERR_WORKSPACE, "rgemm_inner workspace %zu < %zu", part_bytes,
             need);
  GW_REQUIRE(nk <= 65535, GW_ERR_UNSUPPORTED, "nk too large");
  int kchunk = (K + splits - 1) / splits;
  kchunk = ((kchunk + kKR - 1) / kKR) * kKR;
  const int ntx = (N + 63) / 64, nty = (M + 63) / 64;
  const long total = (long)ntx * nty * nk * splits;
  hipLaunchKernelGGL(k_rgemm_inner_lds<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, M, N, K, X,
                     ldx, sx, Y, ldy, sy, splits, kchunk, (T*)partials, ntx, nty, total);
  const long tot = (long)M * N;
  hipLaunchKernelGGL(k_rgemm_inner_reduce<T>, dim3((unsigned)((tot + 63) / 64), nk), dim3(256), 0, st, M, N,
                     splits, (const T*)partials, C, ldc, sc);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

template <typename T>
int rgemm_combine(int nk, int A, int G, int J, const T* U, long ldu, long su, const T* Y, long ldy, long sy, T* Z,
                  long ldz, long sz, T alpha, T beta, hipStream_t st) {
  if (A <= 0 || G <= 0 || nk <= 0) return GW_OK;
  const int ntx = (G + 63) / 64, nty = (A + 63) / 64;
  const long total = (long)ntx * nty * nk;
  hipLaunchKernelGGL(k_rgemm_combine_lds<T>, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, A, G, J,
                     U, ldu, su, Y, ldy, sy, Z, ldz, sz, alpha, beta, ntx, nty, total);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

}  // namespace

extern "C" size_t gw_rgemm_inner_workspace_bytes(int precision, int nk, int M, int N, int K) {
  if (M <= 0 || N <= 0 || K <= 0 || nk <= 0) return 0;
  return (size_t)nk * rinner_splits(nk, M, N, K) * M * N * (precision == GW_FP64 ? sizeof(double) : sizeof(float));
}

extern "C" int gw_rgemm_inner(int precision, int nk, int M, int N, int K, const void* X, long ldx, long sx,
                              const void* Y, long ldy, long sy, void* C, long ldc, long sc, void* ws,
                              size_t ws_bytes, void* stream) {
  GW_REQUIRE(X && Y && C && nk >= 0 && M >= 0 && N >= 0 && K >= 0, GW_ERR_ARG, "bad argument");
  GW_REQUIRE(ldx >= K && ldy >= K && ldc >= N, GW_ERR_ARG, "leading dimension too small");
  if (precision == GW_FP64)
    return rgemm_inner<double>(nk, M, N, K, (const double*)X, ldx, sx, (const double*)Y, ldy, sy, (double*)C, ldc,
                               sc, ws, ws_bytes, (hipStream_t)stream);
  return rgemm_inner<float>(nk, M, N, K, (const float*)X, ldx, sx, (const float*)Y, ldy, sy, (float*)C, ldc, sc,
                            ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int gw_rgemm_combine(int precision, int nk, int A, int G, int J, const void* U, long ldu, long su,
                                const void* Y, long ldy, long sy, void* Z, long ldz, long sz, double alpha,
                                double beta, void* stream) {
  GW_REQUIRE(U && Y && Z && nk >= 0 && A >= 0 && G >= 0 && J >= 0, GW_ERR_ARG, "bad argument");
  GW_REQUIRE(ldu >= A && ldy >= G && ldz >= G, GW_ERR_ARG, "leading dimension too small");
  if (precision == GW_FP64)
    return rgemm_combine<double>(nk, A, G, J, (const double*)U, ldu, su, (const double*)Y, ldy, sy, (double*)Z,
                                 ldz, sz, alpha, beta, (hipStream_t)stream);
  return rgemm_combine<float>(nk, A, G, J, (const float*)U, ldu, su, (const float*)Y, ldy, sy, (float*)Z, ldz,
                              sz, (float)alpha, (float)beta, (hipStream_t)stream);
}
