// Local + kinetic terms of H|psi> on the FFT box, and the operator plan.
//
// Reference contract: core/batch.py:405-428 (_local_fft_into) and 480-505
// (apply): out = t*c + gather(fftn(V * ifftn(scatter(c)))) per band, mask kept.
//
// MI355X design. The sphere occupies ~45% of the x-planes and ~17% of the
// (x, y) sticks of the box, so the 3-D transforms are pruned:
//   P1  per (x-plane, band): scatter the plane's sphere points into an LDS
//       plane tile, inverse z-FFT on the non-empty sticks only, inverse y-FFT
//       on all columns, store the plane (planes outside the sphere's x-extent
//       are never materialised).
//   P2  per (yz-line chunk, band group): inverse x-FFT, multiply by V(r) held
//       in registers across the band group, forward x-FFT; only the sphere's
//       x-planes are read back and written.
//   P3  per (x-plane, band): forward y-FFT on all columns, forward z-FFT on
//       the sticks, gather the sphere points, add t*c, write H c.
// Each 1-D transform leaves digit-permuted order and the matching forward
// transform consumes it (gw_fft.h), so V(r) is stored pre-permuted and no
// reordering pass exists. When a plane tile does not fit in LDS (fp64 boxes
// above ~95x95 per plane) P1/P3 split into a stick kernel (z) and a
// column-chunk plane kernel (y) that exchange z-transformed sticks through HBM.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <numeric>
#include <vector>

#include "gw_internal.h"
#include "gw_regfft.h"

namespace gw {

static thread_local char g_err[1024] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// Preferred first factor per box length, set by the autotuner (0: none).
static int g_radix_pref[1024] = {0};
// x-pass launch shape of the wave-line path, per precision (0 = FP64, 1 = FP32):
// mode (0 = lines straight from HBM, 1 = planes staged through LDS), waves per
// workgroup and bands per workgroup.
struct XpassShape { int mode, waves, bgroup; };
// Defaults from timing Si64 (90^3) and Si216 (120^3) Gamma applies: FP32 gains from
// staging (-16% / -12%), FP64 (twice the bytes per lane, half the tiles per CU) does not.
static XpassShape g_xpass[2] = {{0, 4, 8}, {1, 4, 16}};
// Bands per chunk of the local term, per precision (0 = the whole block in one
// pass sequence). A chunk's x-plane box is written by P1, read and rewritten by
// the x pass and read by P3 back to back, so a chunk small enough stays in the
// memory-side cache between the passes; the box workspace shrinks with it.
static int g_band_chunk[2] = {0, 0};
static int xp_index(int precision) { return precision == GW_FP64 ? 0 : 1; }
static int xpass_encode(const XpassShape& x) { return x.mode * 10000 + x.waves * 100 + x.bgroup; }

bool split_length(int n, int* a, int* b) {
  if (n > 0 && n < 1024 && g_radix_pref[n] > 0) {
    const int x = g_radix_pref[n];
    if (n % x == 0 && codelet_ok(x) && codelet_ok(n / x)) {
      *a = x;
      *b = n / x;
      return true;
    }
  }
  // most balanced split with both factors in the codelet set
  int best_a = -1, best_b = -1;
  for (int x = 1; x <= 16; ++x) {
    if (n % x) continue;
    int y = n / x;
    if (!codelet_ok(x) || !codelet_ok(y)) continue;
    if (best_a < 0 || std::abs(x - y) < std::abs(best_a - best_b) ||
        (std::abs(x - y) == std::abs(best_a - best_b) && x > best_a)) {
      best_a = x;
      best_b = y;
    }
  }
  if (best_a < 0) return false;
  *a = best_a;
  *b = best_b;
  return true;
}

bool split_length3(int n, int f[3]) {
  if (split_length(n, &f[0], &f[1])) {
    f[2] = 1;
    return true;
  }
  // three factors: the smallest largest factor, then the most balanced
  int best[3] = {0, 0, 0};
  for (int a = 2; a <= 16; ++a) {
    if (n % a || !codelet_ok(a)) continue;
    for (int b = 2; b <= 16; ++b) {
      if ((n / a) % b || !codelet_ok(b)) continue;
      const int c = n / a / b;
      if (!codelet_ok(c) || c < 2) continue;
      const int mx = std::max(a, std::max(b, c)), mn = std::min(a, std::min(b, c));
      const int bmx = std::max(best[0], std::max(best[1], best[2]));
      const int bmn = std::min(best[0], std::min(best[1], best[2]));
      if (best[0] == 0 || mx < bmx || (mx == bmx && mn > bmn)) {
        best[0] = a;
        best[1] = b;
        best[2] = c;
      }
    }
  }
  if (best[0] == 0) return false;
  f[0] = best[0];
  f[1] = best[1];
  f[2] = best[2];
  return true;
}

// Three-factor box lengths are switched off: plans take two-codelet splits only.
// The three-factor passes (gw_fft.h, kernels built with RMAX = kRmaxF3) are
// kept, and their index maps are checked on the host, but their GPU runs on a
// 125 x 24 x 30 box faulted and no cause has been shown.
constexpr bool kThreeFactorEnabled = false;

// The split a plan takes for box length n (f[2] = 1 for two factors).
static bool plan_split(int n, int f[3]) {
  if (kThreeFactorEnabled) return split_length3(n, f);
  f[2] = 1;
  return split_length(n, &f[0], &f[1]);
}

extern "C" int gw_box_split(int n, int* f) {
  GW_REQUIRE(f && n > 0, GW_ERR_ARG, "bad argument");
  if (!plan_split(n, f)) {
    f[0] = f[1] = f[2] = 0;
    set_error("box length %d has no split into two codelets <= 16 (three-factor splits are not enabled)", n);
    return GW_ERR_UNSUPPORTED;
  }
  return GW_OK;
}

// Entry points that run the box FFTs refuse a plan whose box has a length with
// no codelet split.
static int require_box_fft(const Plan& P, const char* what) {
  if (!P.nosplit_len) return GW_OK;
  set_error("%s needs the box FFT kernels, and box length %d (box %d x %d x %d) has no split into two "
            "codelets <= 16; this plan serves the kinetic and nonlocal terms and the sphere/box maps",
            what, P.nosplit_len, P.n[0], P.n[1], P.n[2]);
  return GW_ERR_UNSUPPORTED;
}

constexpr int kMaxLds = 160 * 1024;
constexpr int kXTilePad = 0;  // extra elements per row of the staged x-pass tiles (1, 2, 4 measured slower)
constexpr int kPlaneThreads = 512;
constexpr int kXThreads = 256;
constexpr int kXVpt = 16;      // V values cached per thread in the x pass
constexpr int kXBandGroup = 4; // bands per x-pass workgroup (V reuse)
constexpr int kStickChunk = 8; // sticks per workgroup in the unfused z pass
constexpr int kColChunk = 32;  // columns per workgroup in the unfused y pass

// Largest codelet the plan's axis splits use: kernels built for radices <= 10
// need far fewer registers than the full set (up to 16).
static int plan_rmax(const Plan* p) {
  int m = 1;
  for (int ax = 0; ax < 3; ++ax) {
    if (p->f[ax][2] > 1) return kRmaxF3;  // a three-factor axis: the kernels that carry that split
    m = std::max(m, std::max(p->f[ax][0], p->f[ax][1]));
  }
  return m <= 10 ? 10 : 16;
}

// Columns per workgroup of the split y passes: kColChunk, fewer when a
// column chunk of n2 points would not fit in LDS.
static int col_chunk(const Plan* p, size_t e) {
  const size_t cap = (size_t)std::min(p->lds_max, kMaxLds);
  return (int)std::max<size_t>(1, std::min<size_t>(kColChunk, cap / ((size_t)p->n[1] * e)));
}


static size_t csize(int precision) { return precision == GW_FP64 ? 16 : 8; }

bool plane_fits_lds(const Plan* p, int precision) {
  if (p->flags & GW_PLAN_FORCE_UNFUSED) return false;
  return (size_t)p->n[1] * p->n[2] * csize(precision) <= (size_t)std::min(p->lds_max, kMaxLds);
}

// y-lines per x chunk of the pencil slab kernel, 0 when the pencil path is off
// or a slab of the active planes does not fit LDS: two workgroups per CU when
// that leaves at least 8 lines, else one.
static int pencil_yc(const Plan* p, int precision) {
  if (!(p->flags & GW_PLAN_PENCIL) || p->nosplit_len) return 0;
  const long e = (long)csize(precision);
  const long cap = std::min(p->lds_max, kMaxLds);
  const long slab = (long)p->na * p->n[1] * e;
  const long line = (long)p->n[0] * e;
  const long two = (cap / 2 - slab) / line, one = (cap - slab) / line;
  const long yc = two >= 8 ? two : one;
  return yc >= 1 ? (int)std::min<long>(yc, p->n[1]) : 0;
}

static int x_lines(const Plan* p) {
  int n23 = p->n[1] * p->n[2];
  int L = (kXVpt * kXThreads) / p->n[0];
  return std::max(1, std::min(L, n23));
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

ApplyLayout apply_layout(const Plan* p, int nb, int precision, bool chunked) {
  ApplyLayout L{};
  const size_t e = csize(precision);
  const size_t nbt = (size_t)p->nk * nb;
  L.fused = plane_fits_lds(p, precision);
  L.pencil_yc = chunked ? pencil_yc(p, precision) : 0;  // the apply's layout only
  size_t off = 0;
  L.box_off = off;
  const int chunk = chunked ? g_band_chunk[precision == GW_FP64 ? 0 : 1] : 0;
  // a plan without box FFTs (a length with no codelet split) needs no box
  const size_t nbt_box = p->nosplit_len ? 0 : (size_t)p->nk * (chunk > 0 && chunk < nb ? chunk : nb);
  L.box_bytes = L.pencil_yc ? 0 : nbt_box * p->na * p->n[1] * p->n[2] * e;
  off = align256(off + L.box_bytes);
  L.stick_off = off;
  L.stick_bytes = L.fused && !L.pencil_yc ? 0 : nbt_box * p->nst_max * p->n[2] * e;
  off = align256(off + L.stick_bytes);
  L.becp_off = off;
  L.becp_bytes = (size_t)p->nk * p->nproj * nb * e;
  off = align256(off + L.becp_bytes);
  L.u_off = off;
  L.u_bytes = L.becp_bytes;
  off = align256(off + L.u_bytes);
  L.part_off = off;
  L.part_bytes = 0;
  if (p->nproj) {
    L.part_bytes = precision == GW_FP64
                       ? gemm_inner_partials_bytes<double>(p->nk, p->nproj, nb, p->npw_max)
                       : gemm_inner_partials_bytes<float>(p->nk, p->nproj, nb, p->npw_max);
  }
  off = align256(off + L.part_bytes);
  L.total = off;
  return L;
}

// ------------------------------------------------------------------ kernels

struct Geom {
  int n1, n2, n3, na, nst_max, npw_max, nb;
  int xlo, xhi;  // active x-planes = [0, xlo) U [xhi, n1), or xlo = -1
  int st_max_plane;
  int fx0, fx1, fy0, fy1, fz0, fz1;
  int fx2, fy2, fz2;  // third factor of each axis split (1: two factors)
  const int* planes;
  const int* st_plane_ptr;
  const int* st_i1;
  const int* st_pt_ptr;
  const int* pt_src;
  const int* pt_yz;
  const int* npw;
};

static Geom geom_of(const Plan* p, int nb) {
  Geom g;
  g.n1 = p->n[0]; g.n2 = p->n[1]; g.n3 = p->n[2];
  g.na = p->na; g.nst_max = p->nst_max; g.npw_max = p->npw_max; g.nb = nb;
  g.xlo = p->xlo; g.xhi = p->xhi; g.st_max_plane = std::max(1, p->st_max_plane);
  g.fx0 = p->f[0][0]; g.fx1 = p->f[0][1];
  g.fy0 = p->f[1][0]; g.fy1 = p->f[1][1];
  g.fz0 = p->f[2][0]; g.fz1 = p->f[2][1];
  g.fx2 = p->f[0][2]; g.fy2 = p->f[1][2]; g.fz2 = p->f[2][2];
  g.planes = p->planes; g.st_plane_ptr = p->st_plane_ptr; g.st_i1 = p->st_i1;
  g.st_pt_ptr = p->st_pt_ptr; g.pt_src = p->pt_src; g.pt_yz = p->pt_yz; g.npw = p->npw;
  return g;
}

template <typename T> __device__ __forceinline__ cpx<T> czero() { cpx<T> z; z.x = 0; z.y = 0; return z; }

// Range-checked buffer access (raw buffer resource, stride 0): a lane whose byte
// offset lies past `bytes` reads zero and its store is discarded by the memory
// unit, so masked loads and stores need no branch and every lane issues the
// same instruction stream (the wait counts before dependent uses stay exact).
// The base must be wave-uniform; bytes < 2^31.
constexpr int kBufOff = 0x7fffff00;  // offset of a masked lane
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
using u32x2 = unsigned int __attribute__((ext_vector_type(2)));
using u32x4 = unsigned int __attribute__((ext_vector_type(4)));
template <typename T> __device__ __forceinline__ cpx<T> buf_load(__amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(cpx<T>) == 8)
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
// two adjacent FP32 complex values, moved as one 16-byte access
struct alignas(16) cpx2f { float2 a, b; };
__device__ __forceinline__ cpx2f buf_load2(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(cpx2f, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store2(cpx2f v, __amdgpu_buffer_rsrc_t r, int off) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off, 0, 0);
}
template <typename T> __device__ __forceinline__ void buf_store(cpx<T> v, __amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(cpx<T>) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off, 0, 0);
}

// Linear workgroup id -> work id so that consecutive work ids share an XCD
// (dispatch places workgroup L on XCD L % 8). Returns -1 for the padding ids.
__device__ __forceinline__ long xcd_work_id(long total) {
  const long L = blockIdx.x;
  const long per = (total + 7) / 8;
  const long w = (L % 8) * per + L / 8;
  return w < total ? w : -1;
}


// P1 fused: scatter -> inverse z (sticks) -> inverse y (all columns) -> store plane,
// for x-plane a of band row bi (= k * nb + b); `tile` holds n2 * n3 elements.
template <typename T, int RMAX = 16>
__device__ void plane_inv_body(const Geom& g, int a, int bi, const cpx<T>* __restrict__ c, long ldc, long sc,
                               cpx<T>* __restrict__ box, const cpx<T>* __restrict__ twy,
                               const cpx<T>* __restrict__ twz, cpx<T>* tile) {
  const int k = bi / g.nb, b = bi - k * g.nb;
  const int n23 = g.n2 * g.n3;
  for (int e = threadIdx.x; e < n23; e += blockDim.x) tile[e] = czero<T>();
  const int* sp = g.st_plane_ptr + (long)k * (g.na + 1);
  const int s0 = sp[a], s1 = sp[a + 1];
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const int p0 = pp[s0], p1 = pp[s1];
  const cpx<T>* cb = c + (long)k * sc + (long)b * ldc;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  __syncthreads();
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) tile[yz[q]] = cb[src[q]];
  __syncthreads();
  if (s1 > s0) {
    line_fft_inverse<T, false, RMAX>(tile, twz, g.fz0, g.fz1, s1 - s0, g.n3, 1,
                               g.st_i1 + (long)k * g.nst_max + s0, g.fz2);
    __syncthreads();
    line_fft_inverse<T, true, RMAX>(tile, twy, g.fy0, g.fy1, g.n3, 1, g.n3, nullptr, g.fy2);
    __syncthreads();
  }
  cpx<T>* dst = box + ((long)bi * g.na + a) * n23;
  for (int e = threadIdx.x; e < n23; e += blockDim.x) dst[e] = tile[e];
}

template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(kPlaneThreads)
k_plane_inv(Geom g, const cpx<T>* __restrict__ c, long ldc, long sc, cpx<T>* __restrict__ box,
            const cpx<T>* __restrict__ twy, const cpx<T>* __restrict__ twz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  plane_inv_body<T, RMAX>(g, blockIdx.x, blockIdx.y, c, ldc, sc, box, twy, twz, reinterpret_cast<cpx<T>*>(smem_raw));
}

// P2: inverse x-FFT, x V(r), forward x-FFT over chunk `chunk` of L yz-lines for
// band group `bgrp` (kXBandGroup bands); `tile` holds n1 * L elements.
template <typename T, int RMAX = 16>
__device__ void xpass_body(const Geom& g, int L, int nbt, int chunk, int bgrp, cpx<T>* __restrict__ box,
                           const T* __restrict__ vp, const cpx<T>* __restrict__ twx, cpx<T>* tile) {
  const int n23 = g.n2 * g.n3;
  const int j0 = chunk * L;
  const int nl = min(L, n23 - j0);
  const int ne = g.n1 * L;
  T vr[kXVpt];
#pragma unroll
  for (int i = 0; i < kXVpt; ++i) {
    const int e = threadIdx.x + i * blockDim.x;
    T v = 0;
    if (e < ne) {
      const int px = e / L, l = e - px * L;
      if (l < nl) v = vp[(long)px * n23 + j0 + l];
    }
    vr[i] = v;
  }
  // The first PF passes of a band's active-plane elements are loaded one band
  // ahead, so their latency overlaps the previous band's transforms (none for
  // the 16-radix codelets, whose registers are spent already).
  constexpr int PF = RMAX <= 10 ? 8 : 0;
  const int nal = g.na * L;
  cpx<T> pf[PF > 0 ? PF : 1];
  auto prefetch = [&](int bi) {
    const cpx<T>* bb = box + (long)bi * g.na * n23 + j0;
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      const int a = e / L, l = e - a * L;
      pf[i] = (e < nal && l < nl) ? bb[(long)a * n23 + l] : czero<T>();
    }
  };
  if (bgrp * kXBandGroup < nbt) prefetch(bgrp * kXBandGroup);
  for (int bg = 0; bg < kXBandGroup; ++bg) {
    const int bi = bgrp * kXBandGroup + bg;
    if (bi >= nbt) break;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) tile[e] = czero<T>();
    __syncthreads();
    cpx<T>* bb = box + (long)bi * g.na * n23 + j0;
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      const int a = e / L, l = e - a * L;
      if (e < nal && l < nl) tile[g.planes[a] * L + l] = pf[i];
    }
    for (int e = threadIdx.x + PF * blockDim.x; e < nal; e += blockDim.x) {
      const int a = e / L, l = e - a * L;
      if (l < nl) tile[g.planes[a] * L + l] = bb[(long)a * n23 + l];
    }
    __syncthreads();
    if (bg + 1 < kXBandGroup && bi + 1 < nbt) prefetch(bi + 1);
    line_fft_inverse<T, true, RMAX>(tile, twx, g.fx0, g.fx1, L, 1, L, nullptr, g.fx2);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kXVpt; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < ne) { tile[e].x *= vr[i]; tile[e].y *= vr[i]; }
    }
    __syncthreads();
    line_fft_forward<T, true, RMAX>(tile, twx, g.fx0, g.fx1, L, 1, L, nullptr, g.fx2);
    __syncthreads();
    for (int e = threadIdx.x; e < g.na * L; e += blockDim.x) {
      const int a = e / L, l = e - a * L;
      if (l < nl) bb[(long)a * n23 + l] = tile[g.planes[a] * L + l];
    }
    __syncthreads();
  }
}

template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(kXThreads)
k_xpass(Geom g, int L, int nbt, cpx<T>* __restrict__ box, const T* __restrict__ vp,
        const cpx<T>* __restrict__ twx) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  xpass_body<T, RMAX>(g, L, nbt, blockIdx.x, blockIdx.y, box, vp, twx, reinterpret_cast<cpx<T>*>(smem_raw));
}

// P2 with wave-cooperative lines (gw_regfft.h): each x-line is spread over S
// lanes, the transpose between the two DFT passes stays inside the wave's own
// LDS area, V(r) (natural x order) is held in registers across the band group,
// and only the sphere's x-planes are loaded and stored.
template <typename T, int N, int M, int S, int NWX = 4>
__global__ void __launch_bounds__(64 * NWX)
k_xpass_wave(Geom g, int nbt, int bgroup, cpx<T>* __restrict__ box, const T* __restrict__ vx,
             const cpx<T>* __restrict__ twx, int nchunk, long total) {
  using W = WaveLine<T, N, M, S>;
  constexpr int LW = W::LW;
  __shared__ cpx<T> E[NWX * W::E_PER_WAVE];
  __shared__ cpx<T> tw[N];
  // neighbouring line chunks share cache lines of every plane row: keep them on one XCD
  const long wid = xcd_work_id(total);
  if (wid < 0) return;
  const int chunk = (int)(wid % nchunk), bgi = (int)(wid / nchunk);
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twx[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane / LW, l = lane - t * LW;
  const bool part = t < S;
  const int n23 = g.n2 * g.n3;
  const int j = (chunk * NWX + wave) * LW + l;
  const bool jok = j < n23;
  cpx<T>* Ew = E + wave * W::E_PER_WAVE;
  // the sphere's x-planes are [0, xlo) U [xhi, N)
  auto slot = [&](int i0) { return i0 < g.xlo ? i0 : (i0 >= g.xhi ? g.xlo + i0 - g.xhi : -1); };
  T vr[S];
#pragma unroll
  for (int s = 0; s < S; ++s) vr[s] = (t < M && jok) ? vx[(long)W::rpos(t, s) * n23 + j] : T(0);
  // byte offsets of this lane's elements inside one band's planes (kBufOff: none)
  int off[M];
#pragma unroll
  for (int q = 0; q < M; ++q) {
    const int sl = part ? slot(W::gpos(t, q)) : -1;
    off[q] = (jok && sl >= 0) ? (sl * n23 + j) * (int)sizeof(cpx<T>) : kBufOff;
  }
  const unsigned band_bytes = (unsigned)((long)g.na * n23 * sizeof(cpx<T>));
  auto load = [&](int bi, cpx<T> (&v)[M]) {
    const auto r = buf_rsrc(box + (long)bi * g.na * n23, band_bytes);
#pragma unroll
    for (int q = 0; q < M; ++q) v[q] = buf_load<T>(r, off[q]);
  };
  const int b0 = bgi * bgroup;
  const int nbg = min(bgroup, nbt - b0);
  if (nbg <= 0) return;
  cpx<T> x[M], xn[M], y[S];
  // The first band's loads complete before the loop, so no load is pending at
  // the loop head; inside, the next band's loads (an empty range after the last
  // band: zeros, no traffic) overlap this band's transforms, and the copy at the
  // end waits for them only, not for this band's stores.
  load(b0, x);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  for (int bg = 0; bg < nbg; ++bg) {
    const int bi = b0 + bg;
    {
      const bool more = bg + 1 < nbg;
      const auto rn = buf_rsrc(box + (long)(more ? bi + 1 : bi) * g.na * n23, more ? band_bytes : 0u);
#pragma unroll
      for (int q = 0; q < M; ++q) xn[q] = buf_load<T>(rn, off[q]);
    }
    W::inverse(x, y, t, l, part, Ew, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) { y[s].x *= vr[s]; y[s].y *= vr[s]; }
    W::forward(y, x, t, l, part, Ew, tw);
    const auto r = buf_rsrc(box + (long)bi * g.na * n23, band_bytes);
#pragma unroll
    for (int q = 0; q < M; ++q) buf_store<T>(x[q], r, off[q]);
#pragma unroll
    for (int q = 0; q < M; ++q) x[q] = xn[q];
  }
}

// P2 with the band's sphere planes staged through LDS. A workgroup of NWX waves
// owns a chunk of LC = NWX*LW yz-lines; one band's na planes of that chunk form
// a tile of na rows x LC contiguous elements, which arrives with row-contiguous
// range-checked loads (each wave instruction covers LC-element row segments
// instead of S planes x LW elements) and leaves the same way. Each wave reads its
// LW lines out of the tile (a zero row stands in for the planes outside the
// sphere), transforms them in registers as in k_xpass_wave and writes them back
// in place (a discard row takes the outside planes). Two tiles alternate, so the
// next band's loads are in flight during this band's transforms.
template <typename T, int N, int M, int S, int NWX>
__global__ void __launch_bounds__(64 * NWX)
k_xpass_stage(Geom g, int nbt, int bgroup, cpx<T>* __restrict__ box, const T* __restrict__ vx,
              const cpx<T>* __restrict__ twx, int nchunk, long total) {
  using W = WaveLine<T, N, M, S>;
  constexpr int LW = W::LW;
  constexpr int LC = NWX * LW;                  // lines per workgroup
  constexpr int NT = 64 * NWX;
  // FP32 tiles move in pairs of adjacent lines (16-byte accesses): every tile row
  // starts at an even line (LC, j0 and n2*n3 are even), so a pair never straddles
  // a row or the last line
  constexpr int PK = (sizeof(T) == 4 && LC % 2 == 0 && N % 2 == 0) ? 2 : 1;
  using SV = typename std::conditional<PK == 2, cpx2f, cpx<T>>::type;
  constexpr int LCU = LC / PK;                  // staging units per tile row
  constexpr int KS = (N * LCU + NT - 1) / NT;   // staged units per thread (na <= N rows)
  constexpr int LDT = LC + kXTilePad;           // tile row stride
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* E = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* tw = E + NWX * W::E_PER_WAVE;
  cpx<T>* tiles = tw + N;                        // 2 x (na + 2) rows of LC
  const long wid = xcd_work_id(total);
  if (wid < 0) return;
  const int chunk = (int)(wid % nchunk), bgi = (int)(wid / nchunk);
  const int b0 = bgi * bgroup;
  const int nbg = min(bgroup, nbt - b0);
  if (nbg <= 0) return;
  const int na = g.na, n23 = g.n2 * g.n3;
  const int tstride = (na + 2) * LDT;            // one tile: rows 0..na-1, zero row na, discard row na+1
  if constexpr (!W::PFA) {
    for (int i = threadIdx.x; i < N; i += NT) tw[i] = twx[i];
  }
  for (int i = threadIdx.x; i < LC; i += NT) {
    tiles[na * LDT + i] = czero<T>();
    tiles[tstride + na * LDT + i] = czero<T>();
  }
  const int j0 = chunk * LC;
  // staging map: thread unit k is tile unit e = threadIdx.x + k*NT = a*LCU + jj/PK
  // <-> box[band][a][j0 + jj .. + PK-1]; units past the tile or the last line are masked
  const int ne = na * LCU;
  const int ks = (ne + NT - 1) / NT;  // staged units per thread actually needed (uniform)
  int soff[KS], tix[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int e = threadIdx.x + k * NT;
    const int a = e / LCU, jj = (e - a * LCU) * PK;
    soff[k] = (e < ne && j0 + jj < n23) ? (a * n23 + j0 + jj) * (int)sizeof(cpx<T>) : kBufOff;
    tix[k] = e < ne ? a * LDT + jj : 0;
  }
  auto sload = [](__amdgpu_buffer_rsrc_t r, int off) -> SV {
    if constexpr (PK == 2) return buf_load2(r, off);
    else return buf_load<T>(r, off);
  };
  auto sstore = [](SV v, __amdgpu_buffer_rsrc_t r, int off) {
    if constexpr (PK == 2) buf_store2(v, r, off);
    else buf_store<T>(v, r, off);
  };
  const unsigned band_bytes = (unsigned)((long)na * n23 * sizeof(cpx<T>));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane / LW, l = lane - t * LW;
  const bool part = t < S;
  const int col = wave * LW + l;
  const bool jok = j0 + col < n23;
  cpx<T>* Ew = E + wave * W::E_PER_WAVE;
  auto slot = [&](int i0) { return i0 < g.xlo ? i0 : (i0 >= g.xhi ? g.xlo + i0 - g.xhi : -1); };
  // tile entries of this lane's elements: read from the zero row, write to the
  // discard row for planes outside the sphere (and idle lanes)
  int rd[M], wr[M];
#pragma unroll
  for (int q = 0; q < M; ++q) {
    const int sl = part ? slot(W::gpos(t, q)) : -1;
    rd[q] = (sl >= 0 ? sl : na) * LDT + min(col, LC - 1);
    wr[q] = (sl >= 0 ? sl : na + 1) * LDT + min(col, LC - 1);
  }
  T vr[S];
#pragma unroll
  for (int s = 0; s < S; ++s) vr[s] = (t < M && jok) ? vx[(long)W::rpos(t, s) * n23 + j0 + col] : T(0);
  SV st[KS];
  {
    const auto r = buf_rsrc(box + (long)b0 * na * n23, band_bytes);
#pragma unroll
    for (int k = 0; k < KS; ++k)
      if (k < ks) st[k] = sload(r, soff[k]);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int e = threadIdx.x + k * NT;
      if (k < ks && e < ne) *reinterpret_cast<SV*>(tiles + tix[k]) = st[k];
    }
  }
  // V(r) and the first tile have landed on every path: no load is pending at the
  // loop head, so the waits inside the loop concern the next band's tile only
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __syncthreads();
  cpx<T> x[M], y[S];
  for (int bg = 0; bg < nbg; ++bg) {
    const int bi = b0 + bg;
    cpx<T>* cur = tiles + (bg & 1) * tstride;
    cpx<T>* nxt = tiles + ((bg + 1) & 1) * tstride;
    {  // next band's tile entries (an empty range after the last band: zeros, no traffic)
      const bool more = bg + 1 < nbg;
      const auto r = buf_rsrc(box + (long)(more ? bi + 1 : bi) * na * n23, more ? band_bytes : 0u);
#pragma unroll
      for (int k = 0; k < KS; ++k)
        if (k < ks) st[k] = sload(r, soff[k]);
    }
#pragma unroll
    for (int q = 0; q < M; ++q) x[q] = cur[rd[q]];
    W::inverse(x, y, t, l, part, Ew, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) { y[s].x *= vr[s]; y[s].y *= vr[s]; }
    W::forward(y, x, t, l, part, Ew, tw);
#pragma unroll
    for (int q = 0; q < M; ++q) cur[wr[q]] = x[q];
    __syncthreads();
    {
      const auto r = buf_rsrc(box + (long)bi * na * n23, band_bytes);
#pragma unroll
      for (int k = 0; k < KS; ++k)
        if (k < ks) sstore(*reinterpret_cast<const SV*>(cur + tix[k]), r, soff[k]);
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int e = threadIdx.x + k * NT;
      if (k < ks && e < ne) *reinterpret_cast<SV*>(nxt + tix[k]) = st[k];
    }
    __syncthreads();
  }
}

// P3 fused: load plane -> forward y (all columns) -> forward z (sticks) -> gather (+ t c).
template <typename T, bool KIN, int RMAX = 16>
__device__ void plane_fwd_body(const Geom& g, int a, int bi, const cpx<T>* __restrict__ box,
                               const cpx<T>* __restrict__ c, long ldc, long sc, const T* __restrict__ t,
                               cpx<T>* __restrict__ out, long ldo, long so, T scale, const cpx<T>* __restrict__ twy,
                               const cpx<T>* __restrict__ twz, cpx<T>* tile) {
  const int k = bi / g.nb, b = bi - k * g.nb;
  const int n23 = g.n2 * g.n3;
  const int* sp = g.st_plane_ptr + (long)k * (g.na + 1);
  const int s0 = sp[a], s1 = sp[a + 1];
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const int p0 = pp[s0], p1 = pp[s1];
  cpx<T>* ob = out + (long)k * so + (long)b * ldo;
  if (a == 0) {  // padded slots of this band are zero (reference: out * mask)
    for (int q = g.npw[k] + threadIdx.x; q < g.npw_max; q += blockDim.x) ob[q] = czero<T>();
  }
  if (s1 == s0) return;
  const cpx<T>* srcb = box + ((long)bi * g.na + a) * n23;
  for (int e = threadIdx.x; e < n23; e += blockDim.x) tile[e] = srcb[e];
  __syncthreads();
  line_fft_forward<T, true, RMAX>(tile, twy, g.fy0, g.fy1, g.n3, 1, g.n3, nullptr, g.fy2);
  __syncthreads();
  line_fft_forward<T, false, RMAX>(tile, twz, g.fz0, g.fz1, s1 - s0, g.n3, 1,
                             g.st_i1 + (long)k * g.nst_max + s0, g.fz2);
  __syncthreads();
  const cpx<T>* cb = c + (long)k * sc + (long)b * ldc;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const T* tk = t + (long)k * g.npw_max;
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    const int s = src[q];
    cpx<T> v = tile[yz[q]];
    v.x *= scale; v.y *= scale;
    if constexpr (KIN) {
      const cpx<T> cv = cb[s];
      const T ts = tk[s];
      v.x += ts * cv.x; v.y += ts * cv.y;
    }
    ob[s] = v;
  }
}

template <typename T, bool KIN, int RMAX = 16>
__global__ void __launch_bounds__(kPlaneThreads)
k_plane_fwd(Geom g, const cpx<T>* __restrict__ box, const cpx<T>* __restrict__ c, long ldc,
            long sc, const T* __restrict__ t, cpx<T>* __restrict__ out, long ldo, long so,
            T scale, const cpx<T>* __restrict__ twy, const cpx<T>* __restrict__ twz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  plane_fwd_body<T, KIN, RMAX>(g, blockIdx.x, blockIdx.y, box, c, ldc, sc, t, out, ldo, so, scale, twy, twz,
                               reinterpret_cast<cpx<T>*>(smem_raw));
}

// ---- unfused path (plane tile larger than LDS) ----

// z inverse on a chunk of sticks: scatter points into stick rows, transform, store.
template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(256)
k_sticks_inv(Geom g, const cpx<T>* __restrict__ c, long ldc, long sc, cpx<T>* __restrict__ sticks,
             const cpx<T>* __restrict__ twz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);  // [kStickChunk][n3]
  const int bi = blockIdx.y;
  const int k = bi / g.nb, b = bi - k * g.nb;
  const int nst = g.st_plane_ptr[(long)k * (g.na + 1) + g.na];
  const int s0 = blockIdx.x * kStickChunk;
  if (s0 >= nst) return;
  const int s1 = min(s0 + kStickChunk, nst);
  for (int e = threadIdx.x; e < kStickChunk * g.n3; e += blockDim.x) tile[e] = czero<T>();
  __syncthreads();
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const cpx<T>* cb = c + (long)k * sc + (long)b * ldc;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = blockDim.x / 64;
  for (int s = s0 + wave; s < s1; s += nw)
    for (int q = pp[s] + lane; q < pp[s + 1]; q += 64)
      tile[(s - s0) * g.n3 + yz[q] % g.n3] = cb[src[q]];
  __syncthreads();
  line_fft_inverse<T, false, RMAX>(tile, twz, g.fz0, g.fz1, s1 - s0, g.n3, 1, nullptr, g.fz2);
  __syncthreads();
  cpx<T>* dst = sticks + ((long)bi * g.nst_max + s0) * g.n3;
  for (int e = threadIdx.x; e < (s1 - s0) * g.n3; e += blockDim.x) dst[e] = tile[e];
}

// y inverse on a chunk of CC columns of one plane, reading z-transformed sticks.
template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(256)
k_plane_inv_y(Geom g, int CC, const cpx<T>* __restrict__ sticks, cpx<T>* __restrict__ box,
              const cpx<T>* __restrict__ twy) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);  // [n2][CC]
  const int a = blockIdx.x, bi = blockIdx.y, c0 = blockIdx.z * CC;
  const int k = bi / g.nb;
  const int cw = min(CC, g.n3 - c0);
  for (int e = threadIdx.x; e < g.n2 * CC; e += blockDim.x) tile[e] = czero<T>();
  __syncthreads();
  const int* sp = g.st_plane_ptr + (long)k * (g.na + 1);
  const int s0 = sp[a], s1 = sp[a + 1];
  const int* i1s = g.st_i1 + (long)k * g.nst_max;
  for (int e = threadIdx.x; e < (s1 - s0) * CC; e += blockDim.x) {
    const int s = s0 + e / CC, cc = e % CC;
    if (cc < cw) tile[i1s[s] * CC + cc] = sticks[((long)bi * g.nst_max + s) * g.n3 + c0 + cc];
  }
  __syncthreads();
  if (s1 > s0) {
    line_fft_inverse<T, true, RMAX>(tile, twy, g.fy0, g.fy1, CC, 1, CC, nullptr, g.fy2);
    __syncthreads();
  }
  cpx<T>* dst = box + ((long)bi * g.na + a) * g.n2 * g.n3 + c0;
  for (int e = threadIdx.x; e < g.n2 * CC; e += blockDim.x) {
    const int i1 = e / CC, cc = e % CC;
    if (cc < cw) dst[(long)i1 * g.n3 + cc] = tile[e];
  }
}

// y forward on a chunk of CC columns of one plane, writing the sticks' columns.
template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(256)
k_plane_fwd_y(Geom g, int CC, const cpx<T>* __restrict__ box, cpx<T>* __restrict__ sticks,
              const cpx<T>* __restrict__ twy) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);
  const int a = blockIdx.x, bi = blockIdx.y, c0 = blockIdx.z * CC;
  const int k = bi / g.nb;
  const int cw = min(CC, g.n3 - c0);
  const int* sp = g.st_plane_ptr + (long)k * (g.na + 1);
  const int s0 = sp[a], s1 = sp[a + 1];
  if (s1 == s0) return;
  const cpx<T>* srcb = box + ((long)bi * g.na + a) * g.n2 * g.n3 + c0;
  for (int e = threadIdx.x; e < g.n2 * CC; e += blockDim.x) {
    const int i1 = e / CC, cc = e % CC;
    tile[e] = cc < cw ? srcb[(long)i1 * g.n3 + cc] : czero<T>();
  }
  __syncthreads();
  line_fft_forward<T, true, RMAX>(tile, twy, g.fy0, g.fy1, CC, 1, CC, nullptr, g.fy2);
  __syncthreads();
  const int* i1s = g.st_i1 + (long)k * g.nst_max;
  for (int e = threadIdx.x; e < (s1 - s0) * CC; e += blockDim.x) {
    const int s = s0 + e / CC, cc = e % CC;
    if (cc < cw) sticks[((long)bi * g.nst_max + s) * g.n3 + c0 + cc] = tile[i1s[s] * CC + cc];
  }
}

// z forward on a chunk of sticks and gather of the sphere points (+ t c).
template <typename T, bool KIN, int RMAX = 16>
__global__ void __launch_bounds__(256)
k_sticks_fwd(Geom g, const cpx<T>* __restrict__ sticks, const cpx<T>* __restrict__ c, long ldc,
             long sc, const T* __restrict__ t, cpx<T>* __restrict__ out, long ldo, long so,
             T scale, const cpx<T>* __restrict__ twz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);
  const int bi = blockIdx.y;
  const int k = bi / g.nb, b = bi - k * g.nb;
  cpx<T>* ob = out + (long)k * so + (long)b * ldo;
  if (blockIdx.x == 0) {
    for (int q = g.npw[k] + threadIdx.x; q < g.npw_max; q += blockDim.x) ob[q] = czero<T>();
  }
  const int nst = g.st_plane_ptr[(long)k * (g.na + 1) + g.na];
  const int s0 = blockIdx.x * kStickChunk;
  if (s0 >= nst) return;
  const int s1 = min(s0 + kStickChunk, nst);
  const cpx<T>* srcb = sticks + ((long)bi * g.nst_max + s0) * g.n3;
  for (int e = threadIdx.x; e < (s1 - s0) * g.n3; e += blockDim.x) tile[e] = srcb[e];
  __syncthreads();
  line_fft_forward<T, false, RMAX>(tile, twz, g.fz0, g.fz1, s1 - s0, g.n3, 1, nullptr, g.fz2);
  __syncthreads();
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const cpx<T>* cb = c + (long)k * sc + (long)b * ldc;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const T* tk = t + (long)k * g.npw_max;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = blockDim.x / 64;
  for (int s = s0 + wave; s < s1; s += nw)
    for (int q = pp[s] + lane; q < pp[s + 1]; q += 64) {
      const int si = src[q];
      cpx<T> v = tile[(s - s0) * g.n3 + yz[q] % g.n3];
      v.x *= scale; v.y *= scale;
      if constexpr (KIN) {
        const cpx<T> cv = cb[si];
        v.x += tk[si] * cv.x; v.y += tk[si] * cv.y;
      }
      ob[si] = v;
    }
}

// ---- pencil path: z sticks -> (y, x) slabs -> z sticks ----
// The z transforms run on chunks of sticks and store them z-major,
// zs[bi][z][s] (s < nst_max), so that one slab kernel per (z, band) reads and
// writes its nst contiguous values; the slab kernel does the y inverse on the
// active planes, the x inverse / V / x forward in chunks of y-lines, and the y
// forward, all in LDS. Between kernels only sticks move (nst * n3 values per
// band, against na * n2 * n3 for the plane pipeline).
constexpr int kPencilChunk = 16;    // sticks per z-pass workgroup
constexpr int kPencilThreads = 512;

template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(256)
k_pencil_z_inv(Geom g, const cpx<T>* __restrict__ c, long ldc, long sc, cpx<T>* __restrict__ zs,
               const cpx<T>* __restrict__ twz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);  // [kPencilChunk][n3]
  const int bi = blockIdx.y;
  const int k = bi / g.nb, b = bi - k * g.nb;
  const int nst = g.st_plane_ptr[(long)k * (g.na + 1) + g.na];
  const int s0 = blockIdx.x * kPencilChunk;
  if (s0 >= nst) return;
  const int ns = min(kPencilChunk, nst - s0);
  for (int e = threadIdx.x; e < kPencilChunk * g.n3; e += blockDim.x) tile[e] = czero<T>();
  __syncthreads();
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const cpx<T>* cb = c + (long)k * sc + (long)b * ldc;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = blockDim.x / 64;
  for (int j = wave; j < ns; j += nw)
    for (int q = pp[s0 + j] + lane; q < pp[s0 + j + 1]; q += 64) tile[j * g.n3 + yz[q] % g.n3] = cb[src[q]];
  __syncthreads();
  line_fft_inverse<T, false, RMAX>(tile, twz, g.fz0, g.fz1, ns, g.n3, 1, nullptr, g.fz2);
  __syncthreads();
  cpx<T>* dst = zs + (long)bi * g.n3 * g.nst_max + s0;
  for (int e = threadIdx.x; e < ns * g.n3; e += blockDim.x) {
    const int z = e / ns, j = e - z * ns;
    dst[(long)z * g.nst_max + j] = tile[j * g.n3 + z];
  }
}

template <typename T, bool KIN, int RMAX = 16>
__global__ void __launch_bounds__(256)
k_pencil_z_fwd(Geom g, const cpx<T>* __restrict__ zs, const cpx<T>* __restrict__ c, long ldc, long sc,
               const T* __restrict__ t, cpx<T>* __restrict__ out, long ldo, long so, T scale,
               const cpx<T>* __restrict__ twz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);
  const int bi = blockIdx.y;
  const int k = bi / g.nb, b = bi - k * g.nb;
  cpx<T>* ob = out + (long)k * so + (long)b * ldo;
  if (blockIdx.x == 0) {
    for (int q = g.npw[k] + threadIdx.x; q < g.npw_max; q += blockDim.x) ob[q] = czero<T>();
  }
  const int nst = g.st_plane_ptr[(long)k * (g.na + 1) + g.na];
  const int s0 = blockIdx.x * kPencilChunk;
  if (s0 >= nst) return;
  const int ns = min(kPencilChunk, nst - s0);
  const cpx<T>* srcb = zs + (long)bi * g.n3 * g.nst_max + s0;
  for (int e = threadIdx.x; e < ns * g.n3; e += blockDim.x) {
    const int z = e / ns, j = e - z * ns;
    tile[j * g.n3 + z] = srcb[(long)z * g.nst_max + j];
  }
  __syncthreads();
  line_fft_forward<T, false, RMAX>(tile, twz, g.fz0, g.fz1, ns, g.n3, 1, nullptr, g.fz2);
  __syncthreads();
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const cpx<T>* cb = c + (long)k * sc + (long)b * ldc;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const T* tk = t + (long)k * g.npw_max;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = blockDim.x / 64;
  for (int j = wave; j < ns; j += nw)
    for (int q = pp[s0 + j] + lane; q < pp[s0 + j + 1]; q += 64) {
      const int si = src[q];
      cpx<T> v = tile[j * g.n3 + yz[q] % g.n3];
      v.x *= scale; v.y *= scale;
      if constexpr (KIN) {
        const cpx<T> cv = cb[si];
        v.x += tk[si] * cv.x; v.y += tk[si] * cv.y;
      }
      ob[si] = v;
    }
}

// plane slot a of stick s of one k: sp[a] <= s < sp[a + 1] (sp has na + 1 entries)
__device__ __forceinline__ int stick_plane(const int* sp, int na, int s) {
  int lo = 0, hi = na;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (sp[mid] <= s) lo = mid; else hi = mid;
  }
  return lo;
}

// One (z, band) slab: y inverse on the na active planes' lines, then per chunk
// of YC y-lines x inverse, x V(r) (vz: z-major, permuted positions), x forward,
// then y forward; reads and writes the band's nst stick values at z.
template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(kPencilThreads)
k_pencil_yx(Geom g, int nbt, int YC, cpx<T>* __restrict__ zs, const T* __restrict__ vz,
            const cpx<T>* __restrict__ twx, const cpx<T>* __restrict__ twy) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n1 = g.n1, n2 = g.n2, na = g.na;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);  // [na][n2]
  cpx<T>* X = A + (long)na * n2;                     // [YC][n1]
  const long w = xcd_work_id((long)g.n3 * nbt);     // consecutive work ids share z (V reuse in L2)
  if (w < 0) return;
  const int z = (int)(w / nbt), bi = (int)(w - (long)z * nbt);
  const int k = bi / g.nb;
  const int* sp = g.st_plane_ptr + (long)k * (na + 1);
  const int* i1s = g.st_i1 + (long)k * g.nst_max;
  const int nst = sp[na];
  cpx<T>* row = zs + ((long)bi * g.n3 + z) * g.nst_max;
  for (int e = threadIdx.x; e < na * n2; e += blockDim.x) A[e] = czero<T>();
  __syncthreads();
  for (int s = threadIdx.x; s < nst; s += blockDim.x) A[stick_plane(sp, na, s) * n2 + i1s[s]] = row[s];
  __syncthreads();
  line_fft_inverse<T, false, RMAX>(A, twy, g.fy0, g.fy1, na, n2, 1, nullptr, g.fy2);
  __syncthreads();
  const T* vzz = vz + (long)z * n2 * n1;
  for (int y0 = 0; y0 < n2; y0 += YC) {
    const int yc = min(YC, n2 - y0);
    for (int e = threadIdx.x; e < yc * n1; e += blockDim.x) X[e] = czero<T>();
    __syncthreads();
    for (int e = threadIdx.x; e < yc * na; e += blockDim.x) {
      const int yy = e / na, a = e - yy * na;
      X[yy * n1 + g.planes[a]] = A[a * n2 + y0 + yy];
    }
    __syncthreads();
    line_fft_inverse<T, false, RMAX>(X, twx, g.fx0, g.fx1, yc, n1, 1, nullptr, g.fx2);
    __syncthreads();
    for (int e = threadIdx.x; e < yc * n1; e += blockDim.x) {
      const T v = vzz[(long)y0 * n1 + e];
      X[e].x *= v; X[e].y *= v;
    }
    __syncthreads();
    line_fft_forward<T, false, RMAX>(X, twx, g.fx0, g.fx1, yc, n1, 1, nullptr, g.fx2);
    __syncthreads();
    for (int e = threadIdx.x; e < yc * na; e += blockDim.x) {
      const int yy = e / na, a = e - yy * na;
      A[a * n2 + y0 + yy] = X[yy * n1 + g.planes[a]];
    }
    __syncthreads();
  }
  line_fft_forward<T, false, RMAX>(A, twy, g.fy0, g.fy1, na, n2, 1, nullptr, g.fy2);
  __syncthreads();
  for (int s = threadIdx.x; s < nst; s += blockDim.x) row[s] = A[stick_plane(sp, na, s) * n2 + i1s[s]];
}

// V(r) for the pencil slabs: vz[(zp * n2 + yp) * n1 + xp] = V at the permuted
// positions (xp, yp, zp), i.e. the transpose of vp.
template <typename T>
__global__ void k_permute_v_zyx(int n1, int n2, int n3, const int* __restrict__ px, const int* __restrict__ py,
                                const int* __restrict__ pz, const double* __restrict__ v, T* __restrict__ vz) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)n1 * n2 * n3;
  if (e >= n) return;
  const int xp = (int)(e % n1);
  const int yp = (int)((e / n1) % n2);
  const int zp = (int)(e / ((long)n1 * n2));
  vz[e] = (T)v[((long)px[xp] * n2 + py[yp]) * n3 + pz[zp]];
}

// kinetic only: out = t*c (0 in padding)
template <typename T>
__global__ void k_kinetic(int nb, int npw_max, const int* __restrict__ npw, const cpx<T>* __restrict__ c,
                          long ldc, long sc, const T* __restrict__ t, cpx<T>* __restrict__ out,
                          long ldo, long so) {
  const int bi = blockIdx.y, k = bi / nb, b = bi - k * nb;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npw_max) return;
  cpx<T> v = czero<T>();
  if (q < npw[k]) {
    const cpx<T> cv = c[(long)k * sc + (long)b * ldc + q];
    const T tv = t[(long)k * npw_max + q];
    v.x = tv * cv.x; v.y = tv * cv.y;
  }
  out[(long)k * so + (long)b * ldo + q] = v;
}

// U[k][q][b] = sum_p D[p][q] * becp[k][p][b]   (D sparse, CSR by column q)
template <typename T>
__global__ void k_dmul(int nproj, int nb, const int* __restrict__ colptr, const int* __restrict__ row,
                       const double* __restrict__ val, const cpx<T>* __restrict__ becp,
                       cpx<T>* __restrict__ u) {
  const int k = blockIdx.y;
  const long tot = (long)nproj * nb;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= tot) return;
  const int q = (int)(e / nb), b = (int)(e - (long)q * nb);
  const cpx<T>* bk = becp + (long)k * tot;
  cpx<T> acc = czero<T>();
  for (int i = colptr[q]; i < colptr[q + 1]; ++i) {
    const T d = (T)val[i];
    const cpx<T> v = bk[(long)row[i] * nb + b];
    acc.x += d * v.x; acc.y += d * v.y;
  }
  u[(long)k * tot + e] = acc;
}

// the D multiply of many cells in one launch: job j covers blocks first ..
// first + nk*nblk - 1, block b of it is k = b / nblk, elements 256 (b % nblk) + tid
struct GroupDmul {
  const int* colptr; const int* row; const double* val; const void* becp; void* u;
  long first;
  int nproj, nb, nblk;
};
template <typename T>
__global__ void __launch_bounds__(256) k_group_dmul(const GroupDmul* __restrict__ jobs, int n, long total) {
  const long blk = blockIdx.x;
  if (blk >= total) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= blk) lo = mid;
    else hi = mid - 1;
  }
  const GroupDmul& d = jobs[lo];
  const long local = blk - d.first;
  const int k = (int)(local / d.nblk);
  const long tot = (long)d.nproj * d.nb;
  const long e = (local % d.nblk) * 256 + threadIdx.x;
  if (e >= tot) return;
  const int q = (int)(e / d.nb), b = (int)(e - (long)q * d.nb);
  const cpx<T>* bk = (const cpx<T>*)d.becp + (long)k * tot;
  cpx<T> acc = czero<T>();
  for (int i = d.colptr[q]; i < d.colptr[q + 1]; ++i) {
    const T dv = (T)d.val[i];
    const cpx<T> v = bk[(long)d.row[i] * d.nb + b];
    acc.x += dv * v.x; acc.y += dv * v.y;
  }
  ((cpx<T>*)d.u)[(long)k * tot + e] = acc;
}

// ---- wave-line path: natural order on every axis, cubic y/z ----


// P1: scatter -> z inverse on compact stick rows (LDS) -> y inverse on columns,
// natural-order rows stored straight to the box.
// waves per workgroup of the wave-line plane kernels: kPW, or 4 when the
// compact stick rows plus kPW exchange areas exceed LDS (fp64, 120^3)
constexpr int kPW = 8;
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr int kAdjBandGroup = 16;  // bands reduced per x-pass workgroup in the V adjoint

// Coefficients c[si[i]] of one batch of a plane's points, as separate re / im
// registers (an array of FP64 complex was placed in scratch by the compiler).
template <typename T, int KB>
__device__ __forceinline__ void gather_points(const cpx<T>* __restrict__ cb, const int (&si)[KB], T (&re)[KB],
                                              T (&im)[KB]) {
#pragma unroll
  for (int i = 0; i < KB; ++i) {
    const cpx<T> v = cb[si[i]];
    re[i] = v.x;
    im[i] = v.y;
  }
}

// A plane work item is x-plane a of the bands [b0, b1) of one k-point: the
// plane's geometry (stick map, point indices, twiddles) is set up once, and
// while one band's transforms run the next band's points are already loaded
// and the previous band's stores drain.
// LDS row stride of the wave-line plane kernels' stick rows: N + 1 points (odd,
// N is even), so rows of different sticks start in different bank slots for
// 16-byte FP64 and 8-byte FP32 elements alike (a stride of N = 120 repeats
// every 2 rows in FP64 slots). The padding point is never read.
template <int N> constexpr int kStickRowStride = N + 1;

struct PlaneItem {
  int a, k, b0, b1;
};
__device__ __forceinline__ PlaneItem plane_item(long wid, int na, int nb, int bpw) {
  const int ngrp = (nb + bpw - 1) / bpw;
  const long rest = wid / na;
  PlaneItem it;
  it.a = (int)(wid % na);
  it.k = (int)(rest / ngrp);
  it.b0 = (int)(rest % ngrp) * bpw;
  it.b1 = min(nb, it.b0 + bpw);
  return it;
}

template <typename T, int N, int M, int S, int NW, bool LOOP>
__global__ void __launch_bounds__(64 * NW)
k_plane_inv_w(Geom g, long total, int bpw, const cpx<T>* __restrict__ c, long ldc, long sc,
              cpx<T>* __restrict__ box, const cpx<T>* __restrict__ twg) {
  using W = WaveLine<T, N, M, S>;
  constexpr int LW = W::LW;
  constexpr int RS = kStickRowStride<N>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* E = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* tw = E + NW * W::E_PER_WAVE;
  cpx<T>* rows = tw + N;
  int* map = reinterpret_cast<int*>(rows + (long)g.st_max_plane * RS);
  const long wid = xcd_work_id(total);
  if (wid < 0) return;
  const PlaneItem it = plane_item(wid, g.na, g.nb, bpw);
  const int a = it.a, k = it.k;
  const int* sp = g.st_plane_ptr + (long)k * (g.na + 1);
  const int s0 = sp[a], s1 = sp[a + 1], ns = s1 - s0;
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const int* i1s = g.st_i1 + (long)k * g.nst_max;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const cpx<T>* ck = c + (long)k * sc;
  // The plane's points, one per thread and pass over all threads, in batches of
  // KB passes whose loads are all in flight together (from clamped in-range
  // slots; an empty plane reads slot 0 of the point tables and places nothing).
  // The first batch's indices are the same for every band of the item.
  constexpr int NT = 64 * NW;
  constexpr int KB = NW >= 16 ? 2 : (40 + NW - 1) / NW;
  const int p0 = pp[s0], p1 = pp[s1];
  const int q0 = p0 + threadIdx.x;
  int pv[KB], si[KB];
#pragma unroll
  for (int i = 0; i < KB; ++i) {
    const int q = p1 > p0 ? min(q0 + i * NT, p1 - 1) : 0;
    pv[i] = yz[q];
    si[i] = src[q];
  }
  T pr[KB], pim[KB];
  gather_points<T, KB>(ck + (long)it.b0 * ldc, si, pr, pim);
  for (int i = threadIdx.x; i < N; i += blockDim.x) { tw[i] = twg[i]; map[i] = -1; }
  __syncthreads();
  for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) map[i1s[s]] = s - s0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane / LW, l = lane - t * LW;
  const bool part = t < S;
  cpx<T>* Ew = E + wave * W::E_PER_WAVE;
  // stick row of each of the lane's M y-positions (the same for every column)
  int st[M];
#pragma unroll
  for (int q = 0; q < M; ++q) st[q] = part ? map[W::gpos(t, q)] : -1;
  cpx<T> x[M], y[S];
  const int b_end = LOOP ? it.b1 : it.b0 + 1;  // one band per item unless LOOP
  for (int b = it.b0; b < b_end; ++b) {
    const cpx<T>* cb = ck + (long)b * ldc;
    for (int e = threadIdx.x; e < ns * RS; e += blockDim.x) rows[e] = czero<T>();
    __syncthreads();
    // point (y, z) of the plane -> stick row map[y], column z
#pragma unroll
    for (int i = 0; i < KB; ++i)
      if (q0 + i * NT < p1) rows[map[pv[i] / N] * RS + pv[i] % N] = cpx<T>{pr[i], pim[i]};
    for (int qq = q0 + KB * NT; qq < p1; qq += KB * NT) {  // planes of more than KB * NT points
      int v2[KB], s2[KB];
#pragma unroll
      for (int i = 0; i < KB; ++i) {
        const int q = min(qq + i * NT, p1 - 1);
        v2[i] = yz[q];
        s2[i] = src[q];
      }
      T r2[KB], i2[KB];
      gather_points<T, KB>(cb, s2, r2, i2);
#pragma unroll
      for (int i = 0; i < KB; ++i)
        if (qq + i * NT < p1) rows[map[v2[i] / N] * RS + v2[i] % N] = cpx<T>{r2[i], i2[i]};
    }
    __syncthreads();
    if (LOOP && b + 1 < b_end) gather_points<T, KB>(cb + ldc, si, pr, pim);  // the next band's first batch
    // z inverse on the plane's sticks (rows of `rows`)
    for (int r0 = wave * LW; r0 < ns; r0 += NW * LW) {
      const int row = r0 + l;
      const bool ok = part && row < ns;
#pragma unroll
      for (int q = 0; q < M; ++q) x[q] = ok ? rows[row * RS + W::gpos(t, q)] : czero<T>();
      W::inverse(x, y, t, l, part, Ew, tw);
      if (t < M && row < ns) {
#pragma unroll
        for (int u = 0; u < S; ++u) rows[row * RS + W::rpos(t, u)] = y[u];
      }
    }
    __syncthreads();
    // y inverse on all columns; column input is non-zero only at stick rows
    cpx<T>* dst = box + (((long)k * g.nb + b) * g.na + a) * N * N;
    for (int c0 = wave * LW; c0 < N; c0 += NW * LW) {
      const int col = c0 + l;
      const bool ok = col < N;
#pragma unroll
      for (int q = 0; q < M; ++q) x[q] = (ok && st[q] >= 0) ? rows[st[q] * RS + col] : czero<T>();
      W::inverse(x, y, t, l, part, Ew, tw);
      if (t < M && col < N) {
#pragma unroll
        for (int u = 0; u < S; ++u) dst[(long)W::rpos(t, u) * N + col] = y[u];
      }
    }
    __syncthreads();  // the rows are rewritten for the next band
  }
}

// P3: y forward on columns from the box -> stick rows (LDS) -> z forward -> gather (+ t c).
template <typename T, int N, int M, int S, bool KIN, int NW, bool LOOP>
__global__ void __launch_bounds__(64 * NW)
k_plane_fwd_w(Geom g, long total, int bpw, const cpx<T>* __restrict__ box, const cpx<T>* __restrict__ c, long ldc,
              long sc, const T* __restrict__ tk_all, cpx<T>* __restrict__ out, long ldo, long so, T scale,
              const cpx<T>* __restrict__ twg) {
  using W = WaveLine<T, N, M, S>;
  constexpr int LW = W::LW;
  constexpr int RS = kStickRowStride<N>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* E = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* tw = E + NW * W::E_PER_WAVE;
  cpx<T>* rows = tw + N;
  int* map = reinterpret_cast<int*>(rows + (long)g.st_max_plane * RS);
  const long wid = xcd_work_id(total);
  if (wid < 0) return;
  const PlaneItem it = plane_item(wid, g.na, g.nb, bpw);
  const int a = it.a, k = it.k;
  if (a == 0) {  // padded slots of these bands are zero (reference: out * mask)
    for (int b = it.b0; b < it.b1; ++b) {
      cpx<T>* ob = out + (long)k * so + (long)b * ldo;
      for (int q = g.npw[k] + threadIdx.x; q < g.npw_max; q += blockDim.x) ob[q] = czero<T>();
    }
  }
  const int* sp = g.st_plane_ptr + (long)k * (g.na + 1);
  const int s0 = sp[a], s1 = sp[a + 1], ns = s1 - s0;
  if (ns == 0) return;
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const int* i1s = g.st_i1 + (long)k * g.nst_max;
  const cpx<T>* ck = c + (long)k * sc;
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const T* tk = tk_all + (long)k * g.npw_max;
  constexpr int NT = 64 * NW;
  const int p0 = pp[s0], p1 = pp[s1];  // p1 > p0: the plane has sticks, each with points
  // passes per batch, all loads in flight together; with 8 or more waves the
  // workgroup's registers, not its LDS, set the occupancy, so the batch stays short
  constexpr int KB = NW >= 8 ? 2 : (40 + NW - 1) / NW;
  // the first batch's point indices (the same for every band of the item) and
  // the first band's t*c, loaded now so their latency overlaps the transforms
  const int q0 = p0 + threadIdx.x;
  int gs[KB], gv[KB];
#pragma unroll
  for (int i = 0; i < KB; ++i) {
    const int q = min(q0 + i * NT, p1 - 1);
    gs[i] = src[q];
    gv[i] = yz[q];
  }
  T gr[KB], gi[KB], tt[KB];
  if constexpr (KIN) {
#pragma unroll
    for (int i = 0; i < KB; ++i) tt[i] = tk[gs[i]];
    gather_points<T, KB>(ck + (long)it.b0 * ldc, gs, gr, gi);
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) { tw[i] = twg[i]; map[i] = -1; }
  __syncthreads();
  for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) map[i1s[s]] = s - s0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane / LW, l = lane - t * LW;
  const bool part = t < S;
  cpx<T>* Ew = E + wave * W::E_PER_WAVE;
  // the stick row of each of the lane's M y-positions (the same for every column;
  // kept in registers by the 4-wave workgroups only, see KB)
  constexpr bool HOIST = NW < 8;
  int st[HOIST ? M : 1];
  if constexpr (HOIST) {
#pragma unroll
    for (int q = 0; q < M; ++q) st[q] = part ? map[W::gpos(t, q)] : -1;
  }
  cpx<T> x[M], y[S], yn[S];
  const int b_end = LOOP ? it.b1 : it.b0 + 1;  // one band per item unless LOOP
  for (int b = it.b0; b < b_end; ++b) {
    const cpx<T>* cb = ck + (long)b * ldc;
    cpx<T>* ob = out + (long)k * so + (long)b * ldo;
    const cpx<T>* srcp = box + (((long)k * g.nb + b) * g.na + a) * N * N;
    auto load_cols = [&](int c0, cpx<T> (&v)[S]) {
      const int col = c0 + l;
#pragma unroll
      for (int u = 0; u < S; ++u) v[u] = (col < N && t < M) ? srcp[(long)W::rpos(t, u) * N + col] : czero<T>();
    };
    if (wave * LW < N) load_cols(wave * LW, yn);
    for (int c0 = wave * LW; c0 < N; c0 += NW * LW) {
      const int col = c0 + l;
      const bool ok = col < N;
#pragma unroll
      for (int u = 0; u < S; ++u) y[u] = yn[u];
      if (c0 + NW * LW < N) load_cols(c0 + NW * LW, yn);  // next columns' loads overlap this transform
      W::forward(y, x, t, l, part, Ew, tw);
      if (ok && part) {
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const int sq = HOIST ? st[HOIST ? q : 0] : map[W::gpos(t, q)];
          if (sq >= 0) rows[sq * RS + col] = x[q];
        }
      }
    }
    __syncthreads();
    for (int r0 = wave * LW; r0 < ns; r0 += NW * LW) {
      const int row = r0 + l;
      const bool ok = row < ns;
#pragma unroll
      for (int u = 0; u < S; ++u) y[u] = (ok && t < M) ? rows[row * RS + W::rpos(t, u)] : czero<T>();
      W::forward(y, x, t, l, part, Ew, tw);
      if (ok && part) {
#pragma unroll
        for (int q = 0; q < M; ++q) rows[row * RS + W::gpos(t, q)] = x[q];
      }
    }
    __syncthreads();
    // gather: point (y, z) of the plane sits in stick row map[y], column z
    auto put = [&](int si, int v, cpx<T> kin) {
      cpx<T> r = rows[map[v / N] * RS + v % N];
      r.x = r.x * scale + kin.x;
      r.y = r.y * scale + kin.y;
      ob[si] = r;
    };
#pragma unroll
    for (int i = 0; i < KB; ++i) {
      if (q0 + i * NT < p1) {
        cpx<T> kin = czero<T>();
        if constexpr (KIN) {
          kin.x = tt[i] * gr[i];
          kin.y = tt[i] * gi[i];
        }
        put(gs[i], gv[i], kin);
      }
    }
    if constexpr (KIN) {
      if (LOOP && b + 1 < b_end) gather_points<T, KB>(cb + ldc, gs, gr, gi);  // the next band's first batch
    }
    for (int q = q0 + KB * NT; q < p1; q += NT) {  // planes of more than KB * NT points
      const int si = src[q];
      cpx<T> kin = czero<T>();
      if constexpr (KIN) { const cpx<T> cv = cb[si]; const T tt = tk[si]; kin.x = tt * cv.x; kin.y = tt * cv.y; }
      put(si, yz[q], kin);
    }
    __syncthreads();  // the rows are rewritten for the next band
  }
}

// wave-line configurations: N = M*S with M <= S <= 16
template <int N> struct WaveCfg { static constexpr int M = 0, S = 0; };
template <> struct WaveCfg<24> { static constexpr int M = 4, S = 6; };
template <> struct WaveCfg<36> { static constexpr int M = 6, S = 6; };
template <> struct WaveCfg<90> { static constexpr int M = 9, S = 10; };
template <> struct WaveCfg<120> { static constexpr int M = 10, S = 12; };

static bool wave_length(int n) { return n == 24 || n == 36 || n == 90 || n == 120; }

template <typename T, int N, int NW>
static size_t plane_w_lds(const Plan& P) {
  using C = WaveCfg<N>;
  using W = WaveLine<T, N, C::M, C::S>;
  return (size_t)(NW * W::E_PER_WAVE + N + (size_t)std::max(1, P.st_max_plane) * kStickRowStride<N>) *
             sizeof(cpx<T>) +
         N * sizeof(int);
}

// waves per plane workgroup for this plan: 16 when one 16-wave workgroup keeps
// more waves resident than kPW-wave ones (large boxes, where the stick rows
// leave room for one workgroup per CU only), else kPW, else 4; 0 if none fits
template <typename T, int N>
static int plane_waves(const Plan& P) {
  const size_t cap = (size_t)std::min(P.lds_max, kMaxLds);
  auto resident = [&](int nw, size_t lds) { return lds <= cap ? nw * (int)(kLdsPerCu / lds) : 0; };
  const int w16 = resident(16, plane_w_lds<T, N, 16>(P)), w8 = resident(kPW, plane_w_lds<T, N, kPW>(P));
  if (sizeof(T) == 4 && w16 > w8) return 16;
  if (w8 > 0) return kPW;
  if (plane_w_lds<T, N, 4>(P) <= cap) return 4;
  return 0;
}

// Launches P1/P2/P3 of the wave-line path; false if the box is not handled.
// Bands per plane work item, per precision and pass (0 P1, 1 P3), from
// GW_TUNE_PLANE_BANDS; 0 picks 4 where one workgroup per CU is all the plane's
// LDS allows (P1 of both precisions, P3 in FP64: one band's stores drain and the
// next band's points load while a band transforms), else 1. Items of one band
// run the kernels built without the band loop (fewer live registers).
static int g_plane_bpw[2][2] = {{0, 0}, {0, 0}};
template <typename T, int N, int NW>
static int plane_bpw(const Plan& P, bool inv) {
  const int v = g_plane_bpw[sizeof(T) == 4][inv ? 0 : 1];
  if (v > 0) return v;
  const bool one_wg = kLdsPerCu / plane_w_lds<T, N, NW>(P) <= 1;
  return one_wg && (inv || sizeof(T) == 8) ? 4 : 1;
}

// `total` = active planes x bands (P.na * nbt, nbt = nk * g.nb)
template <typename T, int N, int NW>
static void launch_planes(bool inv, Plan& P, const Geom& g, long total, const cpx<T>* c, long ldc, long sc,
                          cpx<T>* box, const T* t, cpx<T>* out, long ldo, long so, T scale, bool kin, hipStream_t st) {
  using C = WaveCfg<N>;
  const int bpw = std::min(plane_bpw<T, N, NW>(P, inv), std::max(1, g.nb));
  const long nk = total / ((long)P.na * std::max(1, g.nb));
  const long items = (long)P.na * nk * ((g.nb + bpw - 1) / bpw);
  const int nblk = (int)(8 * ((items + 7) / 8));
  const size_t lds = plane_w_lds<T, N, NW>(P);
  const cpx<T>* tw = tw_of<T>(&P, 1);
  auto go = [&](auto loop) {
    constexpr bool L = decltype(loop)::value;
    if (inv)
      hipLaunchKernelGGL((k_plane_inv_w<T, N, C::M, C::S, NW, L>), dim3(nblk), dim3(64 * NW), lds, st, g, items, bpw,
                         c, ldc, sc, box, tw);
    else if (kin)
      hipLaunchKernelGGL((k_plane_fwd_w<T, N, C::M, C::S, true, NW, L>), dim3(nblk), dim3(64 * NW), lds, st, g,
                         items, bpw, box, c, ldc, sc, t, out, ldo, so, scale, tw);
    else
      hipLaunchKernelGGL((k_plane_fwd_w<T, N, C::M, C::S, false, NW, L>), dim3(nblk), dim3(64 * NW), lds, st, g,
                         items, bpw, box, c, ldc, sc, t, out, ldo, so, scale, tw);
  };
  if (bpw > 1) go(std::true_type{});
  else go(std::false_type{});
}

// Launches P1/P2/P3 of the wave-line path.
template <typename T, int N>
static int launch_wave_path_n(Plan& P, const Geom& g, int nbt, const cpx<T>* c, long ldc, long sc, cpx<T>* box,
                              const T* vn, const T* t, cpx<T>* out, long ldo, long so, T scale, bool kin,
                              StageTimer& tm, hipStream_t st) {
  using C = WaveCfg<N>;
  using W = WaveLine<T, N, C::M, C::S>;
  const long total = (long)P.na * nbt;
  const int nw = plane_waves<T, N>(P);
  GW_REQUIRE(nw > 0, GW_ERR_UNSUPPORTED, "x-plane sticks exceed LDS");
  // the x-pass addresses one band's planes through a buffer resource
  GW_REQUIRE((long)P.na * N * N * (long)sizeof(cpx<T>) <= (long)kBufOff, GW_ERR_UNSUPPORTED,
             "one band's x-planes exceed 2 GiB");
  hipEvent_t ev = tm.begin();
  if (nw == 16) launch_planes<T, N, 16>(true, P, g, total, c, ldc, sc, box, t, out, ldo, so, scale, kin, st);
  else if (nw == kPW) launch_planes<T, N, kPW>(true, P, g, total, c, ldc, sc, box, t, out, ldo, so, scale, kin, st);
  else launch_planes<T, N, 4>(true, P, g, total, c, ldc, sc, box, t, out, ldo, so, scale, kin, st);
  tm.end(ST_PLANE_INV, ev);
  ev = tm.begin();
  const int n23 = N * N;
  const XpassShape xs = g_xpass[xp_index(sizeof(T) == 8 ? GW_FP64 : GW_FP32)];
  const int bgroup = xs.bgroup;
  const int nwx = xs.waves;
  const int lines_per_wg = nwx * W::LW;
  const int nchunk = (n23 + lines_per_wg - 1) / lines_per_wg;
  const long xtotal = (long)nchunk * ((nbt + bgroup - 1) / bgroup);
  const dim3 xgrid((unsigned)(8 * ((xtotal + 7) / 8)));
  const size_t stage_lds =
      ((size_t)nwx * W::E_PER_WAVE + N + 2 * (size_t)(P.na + 2) * (lines_per_wg + kXTilePad)) * sizeof(cpx<T>);
  if (xs.mode == 1 && stage_lds <= (size_t)std::min(P.lds_max, kMaxLds)) {
    if (nwx == 8)
      hipLaunchKernelGGL((k_xpass_stage<T, N, C::M, C::S, 8>), xgrid, dim3(512), stage_lds, st, g, nbt, bgroup,
                         box, vn, tw_of<T>(&P, 0), nchunk, xtotal);
    else if (nwx == 2)
      hipLaunchKernelGGL((k_xpass_stage<T, N, C::M, C::S, 2>), xgrid, dim3(128), stage_lds, st, g, nbt, bgroup,
                         box, vn, tw_of<T>(&P, 0), nchunk, xtotal);
    else
      hipLaunchKernelGGL((k_xpass_stage<T, N, C::M, C::S, 4>), xgrid, dim3(256), stage_lds, st, g, nbt, bgroup,
                         box, vn, tw_of<T>(&P, 0), nchunk, xtotal);
  } else if (nwx == 8)
    hipLaunchKernelGGL((k_xpass_wave<T, N, C::M, C::S, 8>), xgrid, dim3(512), 0, st, g, nbt, bgroup, box, vn,
                       tw_of<T>(&P, 0), nchunk, xtotal);
  else if (nwx == 2)
    hipLaunchKernelGGL((k_xpass_wave<T, N, C::M, C::S, 2>), xgrid, dim3(128), 0, st, g, nbt, bgroup, box, vn,
                       tw_of<T>(&P, 0), nchunk, xtotal);
  else
    hipLaunchKernelGGL((k_xpass_wave<T, N, C::M, C::S, 4>), xgrid, dim3(256), 0, st, g, nbt, bgroup, box, vn,
                       tw_of<T>(&P, 0), nchunk, xtotal);
  tm.end(ST_XPASS, ev);
  ev = tm.begin();
  if (nw == 16) launch_planes<T, N, 16>(false, P, g, total, c, ldc, sc, box, t, out, ldo, so, scale, kin, st);
  else if (nw == kPW) launch_planes<T, N, kPW>(false, P, g, total, c, ldc, sc, box, t, out, ldo, so, scale, kin, st);
  else launch_planes<T, N, 4>(false, P, g, total, c, ldc, sc, box, t, out, ldo, so, scale, kin, st);
  tm.end(ST_PLANE_FWD, ev);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}


// ---- adjoint of the local term with respect to V(r) ----
// grad_V(r) += (1/N) sum_{k,b} Re[conj(phi(r)) psi(r)], psi = N ifftn(scatter c),
// phi = N ifftn(scatter g): both boxes come from P1, this x-pass finishes the
// two inverse transforms and reduces over a group of bands into partial sums
// [group][x][y*n3 + z] (natural order), summed by k_adjv_reduce.
template <typename T, int N, int M, int S>
__global__ void __launch_bounds__(256)
k_xpass_adjv_wave(Geom g, int nbt, int bgroup, const cpx<T>* __restrict__ boxc, const cpx<T>* __restrict__ boxg,
                  const cpx<T>* __restrict__ twx, int nchunk, long total, double* __restrict__ part) {
  using W = WaveLine<T, N, M, S>;
  constexpr int LW = W::LW;
  __shared__ cpx<T> E[4 * W::E_PER_WAVE];
  __shared__ cpx<T> tw[N];
  const long wid = xcd_work_id(total);
  if (wid < 0) return;
  const int chunk = (int)(wid % nchunk), bgi = (int)(wid / nchunk);
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twx[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane / LW, l = lane - t * LW;
  const bool part_ = t < S;
  const int n23 = g.n2 * g.n3;
  const int j = (chunk * 4 + wave) * LW + l;
  const bool jok = j < n23;
  cpx<T>* Ew = E + wave * W::E_PER_WAVE;
  auto slot = [&](int i0) { return i0 < g.xlo ? i0 : (i0 >= g.xhi ? g.xlo + i0 - g.xhi : -1); };
  double acc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s] = 0.0;
  const int b0 = bgi * bgroup;
  const int nbg = min(bgroup, nbt - b0);
  for (int bg = 0; bg < nbg; ++bg) {
    const long off = (long)(b0 + bg) * g.na * n23 + j;
    cpx<T> x[M], yc[S], yg[S];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int sl = part_ ? slot(W::gpos(t, q)) : -1;
      x[q] = (jok && sl >= 0) ? boxc[off + (long)sl * n23] : czero<T>();
    }
    W::inverse(x, yc, t, l, part_, Ew, tw);
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int sl = part_ ? slot(W::gpos(t, q)) : -1;
      x[q] = (jok && sl >= 0) ? boxg[off + (long)sl * n23] : czero<T>();
    }
    W::inverse(x, yg, t, l, part_, Ew, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] += (double)yg[s].x * yc[s].x + (double)yg[s].y * yc[s].y;
  }
  if (t < M && jok) {
    double* pp = part + (long)bgi * N * n23;
#pragma unroll
    for (int s = 0; s < S; ++s) pp[(long)W::rpos(t, s) * n23 + j] = acc[s];
  }
}

__global__ void k_adjv_reduce(long n, int ngroups, double scale, const double* __restrict__ part,
                              double* __restrict__ gv) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0;
  for (int i = 0; i < ngroups; ++i) s += part[(long)i * n + e];
  gv[e] += scale * s;
}

template <typename T, int N>
static int adjv_wave_n(Plan& P, const Geom& g, int nbt, const cpx<T>* c, long ldc, long sc, const cpx<T>* gr,
                       long ldg, long sg, cpx<T>* boxc, cpx<T>* boxg, double* part, double* gv, hipStream_t st) {
  using C = WaveCfg<N>;
  using W = WaveLine<T, N, C::M, C::S>;
  const long total = (long)P.na * nbt;
  const int nw = plane_waves<T, N>(P);
  GW_REQUIRE(nw > 0, GW_ERR_UNSUPPORTED, "x-plane sticks exceed LDS");
  cpx<T>* nullout = nullptr;
  if (nw == 16) {
    launch_planes<T, N, 16>(true, P, g, total, c, ldc, sc, boxc, nullptr, nullout, 0, 0, T(1), false, st);
    launch_planes<T, N, 16>(true, P, g, total, gr, ldg, sg, boxg, nullptr, nullout, 0, 0, T(1), false, st);
  } else if (nw == kPW) {
    launch_planes<T, N, kPW>(true, P, g, total, c, ldc, sc, boxc, nullptr, nullout, 0, 0, T(1), false, st);
    launch_planes<T, N, kPW>(true, P, g, total, gr, ldg, sg, boxg, nullptr, nullout, 0, 0, T(1), false, st);
  } else {
    launch_planes<T, N, 4>(true, P, g, total, c, ldc, sc, boxc, nullptr, nullout, 0, 0, T(1), false, st);
    launch_planes<T, N, 4>(true, P, g, total, gr, ldg, sg, boxg, nullptr, nullout, 0, 0, T(1), false, st);
  }
  const int n23 = N * N;
  const int lines_per_wg = 4 * W::LW;
  const int nchunk = (n23 + lines_per_wg - 1) / lines_per_wg;
  const int bgroup = kAdjBandGroup;
  const int ngroups = (nbt + bgroup - 1) / bgroup;
  const long xtotal = (long)nchunk * ngroups;
  hipLaunchKernelGGL((k_xpass_adjv_wave<T, N, C::M, C::S>), dim3((unsigned)(8 * ((xtotal + 7) / 8))), dim3(256), 0,
                     st, g, nbt, bgroup, boxc, boxg, tw_of<T>(&P, 0), nchunk, xtotal, part);
  const long n = (long)N * n23;
  hipLaunchKernelGGL(k_adjv_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ngroups,
                     1.0 / (double)n, part, gv);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

// ---- band density: rho(r) += sum_{k,b} w_kb |psi_kb(r)|^2, psi = N ifftn(scatter c) ----
// (core/batch.py:562-587 without the final 1/volume.) P1 fills the box, this
// x-pass finishes the inverse transform of a group of bands and reduces
// w |psi|^2 in FP64 into partials [group][x][y*n3 + z], summed by k_adjv_reduce.
template <typename T, int N, int M, int S>
__global__ void __launch_bounds__(256)
k_xpass_density_wave(Geom g, int nbt, int bgroup, const cpx<T>* __restrict__ box, const double* __restrict__ wts,
                     const cpx<T>* __restrict__ twx, int nchunk, long total, double* __restrict__ part) {
  using W = WaveLine<T, N, M, S>;
  constexpr int LW = W::LW;
  __shared__ cpx<T> E[4 * W::E_PER_WAVE];
  __shared__ cpx<T> tw[N];
  const long wid = xcd_work_id(total);
  if (wid < 0) return;
  const int chunk = (int)(wid % nchunk), bgi = (int)(wid / nchunk);
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twx[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane / LW, l = lane - t * LW;
  const bool part_ = t < S;
  const int n23 = g.n2 * g.n3;
  const int j = (chunk * 4 + wave) * LW + l;
  const bool jok = j < n23;
  cpx<T>* Ew = E + wave * W::E_PER_WAVE;
  auto slot = [&](int i0) { return i0 < g.xlo ? i0 : (i0 >= g.xhi ? g.xlo + i0 - g.xhi : -1); };
  double acc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s] = 0.0;
  const int b0 = bgi * bgroup;
  const int nbg = min(bgroup, nbt - b0);
  for (int bg = 0; bg < nbg; ++bg) {
    const long off = (long)(b0 + bg) * g.na * n23 + j;
    const double wb = wts[b0 + bg];
    cpx<T> x[M], y[S];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int sl = part_ ? slot(W::gpos(t, q)) : -1;
      x[q] = (jok && sl >= 0) ? box[off + (long)sl * n23] : czero<T>();
    }
    W::inverse(x, y, t, l, part_, Ew, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] += wb * ((double)y[s].x * y[s].x + (double)y[s].y * y[s].y);
  }
  if (t < M && jok) {
    double* pp = part + (long)bgi * N * n23;
#pragma unroll
    for (int s = 0; s < S; ++s) pp[(long)W::rpos(t, s) * n23 + j] = acc[s];
  }
}

template <typename T, int N>
static int density_wave_n(Plan& P, const Geom& g, int nbt, const cpx<T>* c, long ldc, long sc, const double* wts,
                          cpx<T>* box, double* part, double* rho, hipStream_t st) {
  using C = WaveCfg<N>;
  using W = WaveLine<T, N, C::M, C::S>;
  const long total = (long)P.na * nbt;
  const int nw = plane_waves<T, N>(P);
  GW_REQUIRE(nw > 0, GW_ERR_UNSUPPORTED, "x-plane sticks exceed LDS");
  cpx<T>* nullout = nullptr;
  if (nw == 16)
    launch_planes<T, N, 16>(true, P, g, total, c, ldc, sc, box, nullptr, nullout, 0, 0, T(1), false, st);
  else if (nw == kPW)
    launch_planes<T, N, kPW>(true, P, g, total, c, ldc, sc, box, nullptr, nullout, 0, 0, T(1), false, st);
  else
    launch_planes<T, N, 4>(true, P, g, total, c, ldc, sc, box, nullptr, nullout, 0, 0, T(1), false, st);
  const int n23 = N * N;
  const int lines_per_wg = 4 * W::LW;
  const int nchunk = (n23 + lines_per_wg - 1) / lines_per_wg;
  const int bgroup = kAdjBandGroup;
  const int ngroups = (nbt + bgroup - 1) / bgroup;
  const long xtotal = (long)nchunk * ngroups;
  hipLaunchKernelGGL((k_xpass_density_wave<T, N, C::M, C::S>), dim3((unsigned)(8 * ((xtotal + 7) / 8))), dim3(256),
                     0, st, g, nbt, bgroup, box, wts, tw_of<T>(&P, 0), nchunk, xtotal, part);
  const long n = (long)N * n23;
  hipLaunchKernelGGL(k_adjv_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ngroups, 1.0, part,
                     rho);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

template <int N>
static bool plane_w_fits(const Plan& P, int precision) {
  return (precision == GW_FP64 ? plane_waves<double, N>(P) : plane_waves<float, N>(P)) > 0;
}

static bool wave_path_ok(const Plan& P, int precision) {
  if (P.flags & GW_PLAN_LEGACY) return false;
  if (P.n[0] != P.n[1] || P.n[1] != P.n[2] || !wave_length(P.n[0]) || P.xlo < 0) return false;
  switch (P.n[0]) {
    case 24: return plane_w_fits<24>(P, precision);
    case 36: return plane_w_fits<36>(P, precision);
    case 90: return plane_w_fits<90>(P, precision);
    case 120: return plane_w_fits<120>(P, precision);
    default: return false;
  }
}

template <typename T>
static int launch_wave_path(Plan& P, const Geom& g, int nbt, const cpx<T>* c, long ldc, long sc, cpx<T>* box,
                            const T* vn, const T* t, cpx<T>* out, long ldo, long so, T scale, bool kin,
                            StageTimer& tm, hipStream_t st) {
  switch (P.n[0]) {
    case 24: return launch_wave_path_n<T, 24>(P, g, nbt, c, ldc, sc, box, vn, t, out, ldo, so, scale, kin, tm, st);
    case 36: return launch_wave_path_n<T, 36>(P, g, nbt, c, ldc, sc, box, vn, t, out, ldo, so, scale, kin, tm, st);
    case 90: return launch_wave_path_n<T, 90>(P, g, nbt, c, ldc, sc, box, vn, t, out, ldo, so, scale, kin, tm, st);
    case 120: return launch_wave_path_n<T, 120>(P, g, nbt, c, ldc, sc, box, vn, t, out, ldo, so, scale, kin, tm, st);
    default: set_error("no wave-line configuration for n=%d", P.n[0]); return GW_ERR_UNSUPPORTED;
  }
}

template <typename T, int N, int NW>
static void set_wave_attr_nw(int want) {
  using C = WaveCfg<N>;
  constexpr auto A = hipFuncAttributeMaxDynamicSharedMemorySize;
  (void)hipFuncSetAttribute((const void*)k_plane_inv_w<T, N, C::M, C::S, NW, false>, A, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv_w<T, N, C::M, C::S, NW, true>, A, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_w<T, N, C::M, C::S, true, NW, false>, A, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_w<T, N, C::M, C::S, true, NW, true>, A, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_w<T, N, C::M, C::S, false, NW, false>, A, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_w<T, N, C::M, C::S, false, NW, true>, A, want);
}
template <typename T, int N>
static void set_wave_attr_n(int want) {
  set_wave_attr_nw<T, N, 16>(want);
  set_wave_attr_nw<T, N, kPW>(want);
  set_wave_attr_nw<T, N, 4>(want);
}
static void set_wave_lds_attrs(int want) {
  set_wave_attr_n<double, 24>(want); set_wave_attr_n<float, 24>(want);
  set_wave_attr_n<double, 36>(want); set_wave_attr_n<float, 36>(want);
  set_wave_attr_n<double, 90>(want); set_wave_attr_n<float, 90>(want);
  set_wave_attr_n<double, 120>(want); set_wave_attr_n<float, 120>(want);
}

// ---- set-up kernels ----
template <typename T>
__global__ void k_permute_v(int n1, int n2, int n3, const int* __restrict__ px, const int* __restrict__ py,
                            const int* __restrict__ pz, const double* __restrict__ v, T* __restrict__ vp) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)n1 * n2 * n3;
  if (e >= n) return;
  const int i2 = (int)(e % n3);
  const int i1 = (int)((e / n3) % n2);
  const int i0 = (int)(e / ((long)n2 * n3));
  vp[e] = (T)v[((long)px[i0] * n2 + py[i1]) * n3 + pz[i2]];
}

__global__ void k_cast_r(long n, const double* __restrict__ a, float* __restrict__ b) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) b[e] = (float)a[e];
}

// copy projectors (nk, nproj, npw_max), zero padding, write both precisions
__global__ void k_proj_copy(int nproj, int npw_max, const int* __restrict__ npw,
                            const double2* __restrict__ p, double2* __restrict__ p64,
                            float2* __restrict__ p32) {
  const int row = blockIdx.y;  // k*nproj + j
  const int k = row / nproj;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npw_max) return;
  const long o = (long)row * npw_max + q;
  double2 v = p[o];
  if (q >= npw[k]) { v.x = 0; v.y = 0; }
  p64[o] = v;
  p32[o] = make_float2((float)v.x, (float)v.y);
}

template <typename T>
__global__ void k_fill_zero(long n, cpx<T>* __restrict__ a) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = czero<T>();
}

template <typename T>
__global__ void k_pack_box(Geom g, const cpx<T>* __restrict__ c, cpx<T>* __restrict__ box, int unpack) {
  const int a = blockIdx.x, bi = blockIdx.y;
  const int k = bi / g.nb;
  const long n = (long)g.n1 * g.n2 * g.n3;
  const int* sp = g.st_plane_ptr + (long)k * (g.na + 1);
  const int* pp = g.st_pt_ptr + (long)k * (g.nst_max + 1);
  const int p0 = pp[sp[a]], p1 = pp[sp[a + 1]];
  const int* src = g.pt_src + (long)k * g.npw_max;
  const int* yz = g.pt_yz + (long)k * g.npw_max;
  const long base = (long)g.planes[a] * g.n2 * g.n3;
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    if (unpack) {
      const_cast<cpx<T>*>(c)[(long)bi * g.npw_max + src[q]] = box[(long)bi * n + base + yz[q]];
    } else {
      box[(long)bi * n + base + yz[q]] = c[(long)bi * g.npw_max + src[q]];
    }
  }
}

// psi(r) in natural order from the P1 planes: inverse x and unpermute
template <typename T>
__global__ void __launch_bounds__(kXThreads)
k_x_to_real(Geom g, int L, const cpx<T>* __restrict__ box, const cpx<T>* __restrict__ twx,
            const int* __restrict__ px, const int* __restrict__ py, const int* __restrict__ pz,
            cpx<T>* __restrict__ psi) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);
  const int n23 = g.n2 * g.n3;
  const int j0 = blockIdx.x * L;
  const int nl = min(L, n23 - j0);
  const int ne = g.n1 * L;
  const int bi = blockIdx.y;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) tile[e] = czero<T>();
  __syncthreads();
  const cpx<T>* bb = box + (long)bi * g.na * n23 + j0;
  for (int e = threadIdx.x; e < g.na * L; e += blockDim.x) {
    const int a = e / L, l = e - a * L;
    if (l < nl) tile[g.planes[a] * L + l] = bb[(long)a * n23 + l];
  }
  __syncthreads();
  line_fft_inverse<T, true>(tile, twx, g.fx0, g.fx1, L, 1, L, nullptr, g.fx2);
  __syncthreads();
  const long nbox = (long)g.n1 * n23;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int x = e / L, l = e - x * L;
    if (l >= nl) continue;
    const int j = j0 + l;
    const int y = j / g.n3, z = j - y * g.n3;
    psi[(long)bi * nbox + ((long)px[x] * g.n2 + py[y]) * g.n3 + pz[z]] = tile[e];
  }
}

}  // namespace gw

// ------------------------------------------------------------------ C ABI

using namespace gw;

struct gw_plan {
  Plan p;
  int* perm[3];  // device permutation per axis: position -> index
  int* ident;    // identity maps of length n1, n2, n3 (natural order)
  int* ident_y;
  int* ident_z;
  StageTimer timer;
};

static int grid_blocks(long n, int bs) { return (int)((n + bs - 1) / bs); }

extern "C" const char* gw_last_error(void) { return gw::g_err; }

// Tuning knobs written by the autotuner (gradwave_amd/autotune.py).
extern "C" int gw_tune_set(int key, int arg, int value) {
  if (key == GW_TUNE_RADIX) {
    GW_REQUIRE(arg > 0 && arg < 1024, GW_ERR_ARG, "box length %d out of range", arg);
    GW_REQUIRE(value == 0 || (arg % value == 0 && codelet_ok(value) && codelet_ok(arg / value)), GW_ERR_ARG,
               "%d does not split %d into two codelets", value, arg);
    g_radix_pref[arg] = value;
    return GW_OK;
  }
  if (key == GW_TUNE_GEMM_WGS) {
    GW_REQUIRE(value >= 0 && value <= (1 << 20), GW_ERR_ARG, "bad workgroup target %d", value);
    set_gemm_target_wgs(value);
    return GW_OK;
  }
  if (key == GW_TUNE_XPASS) {  // value = mode * 10000 + waves * 100 + bands
    const int mode = value / 10000, w = (value / 100) % 100, bgp = value % 100;
    GW_REQUIRE((mode == 0 || mode == 1) && (w == 2 || w == 4 || w == 8) && bgp >= 1 && bgp <= 64, GW_ERR_ARG,
               "bad x-pass shape %d", value);
    const XpassShape xs{mode, w, bgp};
    if (arg == GW_FP64 || arg == GW_FP32) g_xpass[xp_index(arg)] = xs;
    else g_xpass[0] = g_xpass[1] = xs;
    return GW_OK;
  }
  if (key == GW_TUNE_COMBINE_LIB) {
    GW_REQUIRE(arg == GW_FP64 || arg == GW_FP32, GW_ERR_ARG, "bad precision %d", arg);
    GW_REQUIRE(value == 0 || value == 1, GW_ERR_ARG, "bad combine choice %d", value);
    set_combine_lib(arg, value);
    return GW_OK;
  }
  if (key == GW_TUNE_PLANE_BANDS) {  // arg = precision + 256 * (0 both passes, 1 P1, 2 P3)
    const int prec = arg & 255, pass = arg >> 8;
    GW_REQUIRE(prec == GW_FP64 || prec == GW_FP32, GW_ERR_ARG, "plane bands: bad precision %d", prec);
    GW_REQUIRE(pass >= 0 && pass <= 2, GW_ERR_ARG, "plane bands: bad pass selector %d", pass);
    GW_REQUIRE(value >= 0 && value <= 64, GW_ERR_ARG, "plane bands %d outside 0..64", value);
    if (pass != 2) g_plane_bpw[prec == GW_FP32][0] = value;
    if (pass != 1) g_plane_bpw[prec == GW_FP32][1] = value;
    return GW_OK;
  }
  if (key == GW_TUNE_BAND_CHUNK) {
    GW_REQUIRE(arg == GW_FP64 || arg == GW_FP32, GW_ERR_ARG, "band chunk: bad precision %d", arg);
    GW_REQUIRE(value >= 0, GW_ERR_ARG, "band chunk %d < 0", value);
    g_band_chunk[arg == GW_FP64 ? 0 : 1] = value;
    return GW_OK;
  }
  if (key == GW_TUNE_INNER_LIB) {
    GW_REQUIRE(arg == GW_FP64 || arg == GW_FP32, GW_ERR_ARG, "bad precision %d", arg);
    GW_REQUIRE(value == 0 || value == 1, GW_ERR_ARG, "bad inner-product choice %d", value);
    set_inner_lib(arg, value);
    return GW_OK;
  }
  if (key == GW_TUNE_GEMM_TILE) {  // arg = precision + 256 * (0 both, 1 inner, 2 combine)
    const int prec = arg & 255, op = arg >> 8;
    GW_REQUIRE(value >= 0 && value <= 13, GW_ERR_ARG, "bad tile variant %d", value);
    GW_REQUIRE(prec == GW_FP64 || prec == GW_FP32, GW_ERR_ARG, "bad precision %d", prec);
    GW_REQUIRE(op >= 0 && op <= 2, GW_ERR_ARG, "bad contraction selector %d", op);
    set_gemm_tile(prec, op == 0 ? -1 : op - 1, value);
    return GW_OK;
  }
  GW_REQUIRE(false, GW_ERR_ARG, "unknown tuning key %d", key);
}

extern "C" int gw_tune_get(int key, int arg) {
  if (key == GW_TUNE_RADIX) return (arg > 0 && arg < 1024) ? g_radix_pref[arg] : -1;
  if (key == GW_TUNE_GEMM_WGS) return gemm_target_wgs();
  if (key == GW_TUNE_GEMM_TILE) {
    const int prec = arg & 255, op = arg >> 8;
    return (prec == GW_FP64 || prec == GW_FP32) && op >= 0 && op <= 2 ? gemm_tile(prec, op == 2 ? 1 : 0) : -1;
  }
  if (key == GW_TUNE_XPASS) return xpass_encode(g_xpass[xp_index(arg == GW_FP64 ? GW_FP64 : GW_FP32)]);
  if (key == GW_TUNE_COMBINE_LIB) return (arg == GW_FP64 || arg == GW_FP32) ? combine_lib(arg) : -1;
  if (key == GW_TUNE_INNER_LIB) return (arg == GW_FP64 || arg == GW_FP32) ? inner_lib(arg) : -1;
  if (key == GW_TUNE_BAND_CHUNK) return (arg == GW_FP64 || arg == GW_FP32) ? g_band_chunk[arg == GW_FP64 ? 0 : 1] : -1;
  if (key == GW_TUNE_PLANE_BANDS) {
    const int prec = arg & 255, pass = arg >> 8;
    return (prec == GW_FP64 || prec == GW_FP32) && pass >= 0 && pass <= 2
               ? g_plane_bpw[prec == GW_FP32][pass == 2 ? 1 : 0] : -1;
  }
  return -1;
}
extern "C" int gw_version(void) { return 1; }

static int noalloc_scratch_setup();
template <typename T>
static void set_generic_lds_attrs(int want);

extern "C" int gw_plan_create(gw_plan** out, int n1, int n2, int n3, int nk, int npw_max,
                              const int64_t* flat_idx, const int64_t* npw) {
  GW_REQUIRE(out && flat_idx && npw, GW_ERR_ARG, "null argument");
  GW_REQUIRE(n1 > 0 && n2 > 0 && n3 > 0 && nk > 0 && npw_max > 0, GW_ERR_ARG, "bad sizes");
  gw_plan* h = new gw_plan();  // value-initialised: every pointer null
  Plan& P = h->p;
  P.n[0] = n1; P.n[1] = n2; P.n[2] = n3;
  P.nk = nk; P.npw_max = npw_max;
  for (int ax = 0; ax < 3; ++ax) {
    if (!plan_split(P.n[ax], P.f[ax])) {
      // No codelet split: the plan is still built for the kinetic and nonlocal
      // terms and the sphere<->box maps; every box-FFT entry point refuses it.
      P.f[ax][0] = P.n[ax];
      P.f[ax][1] = P.f[ax][2] = 1;
      if (!P.nosplit_len) P.nosplit_len = P.n[ax];
    }
  }
  const long N = (long)n1 * n2 * n3;
  const long n23 = (long)n2 * n3;
  // decode points, active planes
  std::vector<char> plane_on(n1, 0);
  for (int k = 0; k < nk; ++k) {
    if (npw[k] < 0 || npw[k] > npw_max) { delete h; set_error("npw[%d] out of range", k); return GW_ERR_ARG; }
    for (long q = 0; q < npw[k]; ++q) {
      const int64_t f = flat_idx[(long)k * npw_max + q];
      if (f < 0 || f >= N) { delete h; set_error("flat_idx out of box at k=%d q=%ld", k, q); return GW_ERR_ARG; }
      plane_on[f / n23] = 1;
    }
  }
  std::vector<int> planes, slot(n1, -1);
  for (int i = 0; i < n1; ++i)
    if (plane_on[i]) { slot[i] = (int)planes.size(); planes.push_back(i); }
  P.na = (int)planes.size();
  if (P.na == 0) { planes.push_back(0); slot[0] = 0; P.na = 1; }
  // per-k sticks
  std::vector<std::vector<int>> kst_i1(nk), kst_ptr(nk), kpl_ptr(nk), kpt_src(nk), kpt_yz(nk);
  int nst_max = 1, st_max_plane = 0, max_pts_plane = 0;
  for (int k = 0; k < nk; ++k) {
    const long nq = npw[k];
    std::vector<int> order(nq);
    std::iota(order.begin(), order.end(), 0);
    auto key = [&](int q) { const int64_t f = flat_idx[(long)k * npw_max + q];
                            return std::make_pair(slot[f / n23], f % n23); };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
    std::vector<int>& sti = kst_i1[k];
    std::vector<int>& stp = kst_ptr[k];
    std::vector<int> st_slot;
    for (long j = 0; j < nq; ++j) {
      const int q = order[j];
      const int64_t f = flat_idx[(long)k * npw_max + q];
      const int sl = slot[f / n23];
      const int i1 = (int)((f % n23) / n3);
      if (sti.empty() || st_slot.back() != sl || sti.back() != i1) {
        sti.push_back(i1); st_slot.push_back(sl); stp.push_back((int)j);
      }
      kpt_src[k].push_back(q);
      kpt_yz[k].push_back((int)(f % n23));
    }
    stp.push_back((int)nq);
    std::vector<int>& plp = kpl_ptr[k];
    plp.assign(P.na + 1, 0);
    int s = 0;
    for (int a = 0; a <= P.na; ++a) {
      while (s < (int)st_slot.size() && st_slot[s] < a) ++s;
      plp[a] = s;
    }
    for (int a = 0; a < P.na; ++a) {
      st_max_plane = std::max(st_max_plane, plp[a + 1] - plp[a]);
      max_pts_plane = std::max(max_pts_plane, stp[plp[a + 1]] - stp[plp[a]]);
    }
    nst_max = std::max(nst_max, (int)sti.size());
  }
  P.nst_max = nst_max;
  P.st_max_plane = st_max_plane;
  P.max_pts_plane = max_pts_plane;
  std::vector<int> h_plp((long)nk * (P.na + 1)), h_sti((long)nk * nst_max, 0),
      h_stp((long)nk * (nst_max + 1), 0), h_src((long)nk * npw_max, 0), h_yz((long)nk * npw_max, 0),
      h_npw(nk);
  for (int k = 0; k < nk; ++k) {
    std::copy(kpl_ptr[k].begin(), kpl_ptr[k].end(), h_plp.begin() + (long)k * (P.na + 1));
    std::copy(kst_i1[k].begin(), kst_i1[k].end(), h_sti.begin() + (long)k * nst_max);
    std::copy(kst_ptr[k].begin(), kst_ptr[k].end(), h_stp.begin() + (long)k * (nst_max + 1));
    for (long s = (long)kst_ptr[k].size(); s <= nst_max; ++s)
      h_stp[(long)k * (nst_max + 1) + s] = kst_ptr[k].back();
    std::copy(kpt_src[k].begin(), kpt_src[k].end(), h_src.begin() + (long)k * npw_max);
    std::copy(kpt_yz[k].begin(), kpt_yz[k].end(), h_yz.begin() + (long)k * npw_max);
    h_npw[k] = (int)npw[k];
  }
  std::vector<int> h_slot(n1, -1);
  for (int a = 0; a < P.na; ++a) h_slot[planes[a]] = a;
  {  // wrapped contiguous range [0, lo) U [hi, n1) with slots in that order?
    int lo = 0;
    while (lo < n1 && plane_on[lo]) ++lo;
    int hi = n1;
    while (hi > lo && plane_on[hi - 1]) --hi;
    bool ok = true;
    for (int i = lo; i < hi; ++i) ok = ok && !plane_on[i];
    P.xlo = ok ? lo : -1;
    P.xhi = ok ? hi : n1;
  }
  auto up = [&](int** d, const std::vector<int>& v) -> int {
    GW_CHECK_HIP(hipMalloc(d, std::max<size_t>(1, v.size()) * sizeof(int)));
    GW_CHECK_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
    return GW_OK;
  };
  int rc;
  if ((rc = up(&P.planes, planes)) || (rc = up(&P.st_plane_ptr, h_plp)) || (rc = up(&P.st_i1, h_sti)) ||
      (rc = up(&P.st_pt_ptr, h_stp)) || (rc = up(&P.pt_src, h_src)) || (rc = up(&P.pt_yz, h_yz)) ||
      (rc = up(&P.npw, h_npw)) || (rc = up(&P.slot_of, h_slot))) {
    gw_plan_destroy(h);
    return rc;
  }
  // twiddles e^{+2 pi i t/n} and axis permutations
  for (int ax = 0; ax < 3; ++ax) {
    const int n = P.n[ax], f0 = P.f[ax][0], f1 = P.f[ax][1], f2 = P.f[ax][2];
    std::vector<double2> t64(n);
    std::vector<float2> t32(n);
    for (int t = 0; t < n; ++t) {
      const long double ang = 2.0L * 3.14159265358979323846264338327950288L * t / n;
      t64[t] = make_double2((double)cosl(ang), (double)sinl(ang));
      t32[t] = make_float2((float)cosl(ang), (float)sinl(ang));
    }
    // position -> index after the inverse split: M*k0 + f2*k1 + k2 holds k0 + f0*k1 + f0*f1*k2
    std::vector<int> perm(n);
    for (int k0 = 0; k0 < f0; ++k0)
      for (int k1 = 0; k1 < f1; ++k1)
        for (int k2 = 0; k2 < f2; ++k2) perm[f1 * f2 * k0 + f2 * k1 + k2] = k0 + f0 * k1 + f0 * f1 * k2;
    if (hipMalloc(&P.tw64[ax], n * sizeof(double2)) != hipSuccess ||
        hipMalloc(&P.tw32[ax], n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&h->perm[ax], n * sizeof(int)) != hipSuccess) {
      set_error("hipMalloc twiddles failed"); gw_plan_destroy(h); return GW_ERR_HIP;
    }
    hipMemcpy(P.tw64[ax], t64.data(), n * sizeof(double2), hipMemcpyHostToDevice);
    hipMemcpy(P.tw32[ax], t32.data(), n * sizeof(float2), hipMemcpyHostToDevice);
    hipMemcpy(h->perm[ax], perm.data(), n * sizeof(int), hipMemcpyHostToDevice);
  }
  if (hipMalloc(&P.t64, (size_t)nk * npw_max * sizeof(double)) != hipSuccess ||
      hipMalloc(&P.t32, (size_t)nk * npw_max * sizeof(float)) != hipSuccess ||
      hipMalloc(&P.vp64, (size_t)N * sizeof(double)) != hipSuccess ||
      hipMalloc(&P.vp32, (size_t)N * sizeof(float)) != hipSuccess ||
      hipMalloc(&P.vx64, (size_t)N * sizeof(double)) != hipSuccess ||
      hipMalloc(&P.vx32, (size_t)N * sizeof(float)) != hipSuccess ||
      hipMalloc(&P.vz64, (size_t)N * sizeof(double)) != hipSuccess ||
      hipMalloc(&P.vz32, (size_t)N * sizeof(float)) != hipSuccess ||
      hipMalloc(&h->ident, (size_t)n1 * sizeof(int)) != hipSuccess ||
      hipMalloc(&h->ident_y, (size_t)n2 * sizeof(int)) != hipSuccess ||
      hipMalloc(&h->ident_z, (size_t)n3 * sizeof(int)) != hipSuccess) {
    set_error("hipMalloc tables failed"); gw_plan_destroy(h); return GW_ERR_HIP;
  }
  hipMemset(P.t64, 0, (size_t)nk * npw_max * sizeof(double));
  hipMemset(P.t32, 0, (size_t)nk * npw_max * sizeof(float));
  hipMemset(P.vp64, 0, (size_t)N * sizeof(double));
  hipMemset(P.vp32, 0, (size_t)N * sizeof(float));
  hipMemset(P.vx64, 0, (size_t)N * sizeof(double));
  hipMemset(P.vx32, 0, (size_t)N * sizeof(float));
  hipMemset(P.vz64, 0, (size_t)N * sizeof(double));
  hipMemset(P.vz32, 0, (size_t)N * sizeof(float));
  {
    std::vector<int> id(std::max(n1, std::max(n2, n3)));
    std::iota(id.begin(), id.end(), 0);
    hipMemcpy(h->ident, id.data(), n1 * sizeof(int), hipMemcpyHostToDevice);
    hipMemcpy(h->ident_y, id.data(), n2 * sizeof(int), hipMemcpyHostToDevice);
    hipMemcpy(h->ident_z, id.data(), n3 * sizeof(int), hipMemcpyHostToDevice);
  }
  GW_CHECK_HIP(hipDeviceSynchronize());
  int dev = 0, lds = 65536;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
    P.lds_max = lds;
  else
    P.lds_max = 65536;
  // plane kernels may take more than 64 KiB of dynamic LDS on gfx950
  const int want = std::min(P.lds_max, kMaxLds);
  (void)hipFuncSetAttribute((const void*)k_plane_inv<double>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv<float>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass<double>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass<float>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv<double, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv<float, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<double, true, 10>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<double, false, 10>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<float, true, 10>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<float, false, 10>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_xpass<double, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass<float, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_x_to_real<double>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_x_to_real<float>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  set_generic_lds_attrs<double>(want);
  set_generic_lds_attrs<float>(want);
  set_wave_lds_attrs(want);
  (void)noalloc_scratch_setup();  // a failure is reported by gw_hpsi_apply_noalloc itself
  *out = h;
  return GW_OK;
}

extern "C" int gw_plan_set_flags(gw_plan* h, int flags) {
  GW_REQUIRE(h, GW_ERR_ARG, "null plan");
  h->p.flags = flags;
  return GW_OK;
}

extern "C" int gw_plan_stage_times(gw_plan* h, double* ms, long* calls, int n, int reset) {
  GW_REQUIRE(h && ms && calls && n >= 0, GW_ERR_ARG, "bad argument");
  h->timer.collect();
  for (int i = 0; i < n && i < ST_COUNT; ++i) { ms[i] = h->timer.ms[i]; calls[i] = h->timer.calls[i]; }
  if (reset)
    for (int i = 0; i < ST_COUNT; ++i) { h->timer.ms[i] = 0; h->timer.calls[i] = 0; }
  return GW_OK;
}

extern "C" int gw_plan_info(const gw_plan* h, int precision, int* info, int n) {
  GW_REQUIRE(h && info, GW_ERR_ARG, "bad argument");
  const Plan& P = h->p;
  const int v[] = {P.na, P.nst_max, wave_path_ok(P, precision) ? 2 : (plane_fits_lds(&P, precision) ? 1 : 0), P.lds_max, P.nproj,
                   P.f[0][0], P.f[0][1], P.f[1][0], P.f[1][1], P.f[2][0], P.f[2][1],
                   P.f[0][2], P.f[1][2], P.f[2][2], P.nosplit_len};
  for (int i = 0; i < n && i < (int)(sizeof(v) / sizeof(v[0])); ++i) info[i] = v[i];
  return GW_OK;
}

extern "C" int gw_plan_destroy(gw_plan* h) {
  if (!h) return GW_OK;
  Plan& P = h->p;
  void* ptrs[] = {P.planes, P.st_plane_ptr, P.st_i1, P.st_pt_ptr, P.pt_src, P.pt_yz, P.npw,
                  P.tw64[0], P.tw64[1], P.tw64[2], P.tw32[0], P.tw32[1], P.tw32[2],
                  P.t64, P.t32, P.vp64, P.vp32, P.vx64, P.vx32, P.vz64, P.vz32, P.slot_of, P.p64, P.p32, P.d_colptr,
                  P.d_row, P.d_val, h->perm[0], h->perm[1], h->perm[2], h->ident, h->ident_y,
                  h->ident_z};
  hipDeviceSynchronize();
  h->timer.collect();
  for (hipEvent_t e : h->timer.pool) hipEventDestroy(e);
  for (void* q : ptrs)
    if (q) hipFree(q);
  delete h;
  return GW_OK;
}

extern "C" int gw_plan_set_kinetic(gw_plan* h, const double* t, void* stream) {
  GW_REQUIRE(h && t, GW_ERR_ARG, "null argument");
  Plan& P = h->p;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)P.nk * P.npw_max;
  GW_CHECK_HIP(hipMemcpyAsync(P.t64, t, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_cast_r, dim3(grid_blocks(n, 256)), dim3(256), 0, st, n, P.t64, P.t32);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_plan_set_potential(gw_plan* h, const double* v, void* stream) {
  GW_REQUIRE(h && v, GW_ERR_ARG, "null argument");
  Plan& P = h->p;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)P.n[0] * P.n[1] * P.n[2];
  hipLaunchKernelGGL(k_permute_v<double>, dim3(grid_blocks(n, 256)), dim3(256), 0, st, P.n[0], P.n[1],
                     P.n[2], h->perm[0], h->perm[1], h->perm[2], v, P.vp64);
  hipLaunchKernelGGL(k_permute_v<float>, dim3(grid_blocks(n, 256)), dim3(256), 0, st, P.n[0], P.n[1],
                     P.n[2], h->perm[0], h->perm[1], h->perm[2], v, P.vp32);
  hipLaunchKernelGGL(k_permute_v<double>, dim3(grid_blocks(n, 256)), dim3(256), 0, st, P.n[0], P.n[1],
                     P.n[2], h->ident, h->ident_y, h->ident_z, v, P.vx64);
  hipLaunchKernelGGL(k_permute_v<float>, dim3(grid_blocks(n, 256)), dim3(256), 0, st, P.n[0], P.n[1],
                     P.n[2], h->ident, h->ident_y, h->ident_z, v, P.vx32);
  hipLaunchKernelGGL(k_permute_v_zyx<double>, dim3(grid_blocks(n, 256)), dim3(256), 0, st, P.n[0], P.n[1],
                     P.n[2], h->perm[0], h->perm[1], h->perm[2], v, P.vz64);
  hipLaunchKernelGGL(k_permute_v_zyx<float>, dim3(grid_blocks(n, 256)), dim3(256), 0, st, P.n[0], P.n[1],
                     P.n[2], h->perm[0], h->perm[1], h->perm[2], v, P.vz32);
  GW_CHECK_HIP(hipGetLastError());
  P.have_v = true;
  return GW_OK;
}

extern "C" int gw_plan_set_projectors(gw_plan* h, int nproj, const void* p, const double* dij,
                                      void* stream) {
  GW_REQUIRE(h && nproj >= 0, GW_ERR_ARG, "bad argument");
  GW_REQUIRE(nproj == 0 || (p && dij), GW_ERR_ARG, "null projector table");
  Plan& P = h->p;
  hipStream_t st = (hipStream_t)stream;
  GW_CHECK_HIP(hipStreamSynchronize(st));
  for (void** q : {(void**)&P.p64, (void**)&P.p32, (void**)&P.d_colptr, (void**)&P.d_row, (void**)&P.d_val})
    if (*q) { hipFree(*q); *q = nullptr; }
  P.nproj = nproj;
  if (!nproj) return GW_OK;
  const size_t n = (size_t)P.nk * nproj * P.npw_max;
  GW_CHECK_HIP(hipMalloc(&P.p64, n * sizeof(double2)));
  GW_CHECK_HIP(hipMalloc(&P.p32, n * sizeof(float2)));
  hipLaunchKernelGGL(k_proj_copy, dim3(grid_blocks(P.npw_max, 256), P.nk * nproj), dim3(256), 0, st, nproj,
                     P.npw_max, P.npw, (const double2*)p, P.p64, P.p32);
  GW_CHECK_HIP(hipGetLastError());
  std::vector<int> colptr(nproj + 1, 0), row;
  std::vector<double> val;
  for (int q = 0; q < nproj; ++q) {
    for (int r = 0; r < nproj; ++r) {
      const double d = dij[(long)r * nproj + q];
      if (d != 0.0) { row.push_back(r); val.push_back(d); }
    }
    colptr[q + 1] = (int)row.size();
  }
  P.d_nnz = (int)row.size();
  GW_CHECK_HIP(hipMalloc(&P.d_colptr, colptr.size() * sizeof(int)));
  GW_CHECK_HIP(hipMalloc(&P.d_row, std::max<size_t>(1, row.size()) * sizeof(int)));
  GW_CHECK_HIP(hipMalloc(&P.d_val, std::max<size_t>(1, val.size()) * sizeof(double)));
  GW_CHECK_HIP(hipMemcpy(P.d_colptr, colptr.data(), colptr.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!row.empty()) {
    GW_CHECK_HIP(hipMemcpy(P.d_row, row.data(), row.size() * sizeof(int), hipMemcpyHostToDevice));
    GW_CHECK_HIP(hipMemcpy(P.d_val, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return GW_OK;
}

extern "C" size_t gw_hpsi_workspace_bytes(const gw_plan* h, int nb, int precision) {
  if (!h || nb <= 0) return 0;
  return apply_layout(&h->p, nb, precision, true).total;
}

// The generic-size local term (any box length the plan splits): x-plane inverse
// (fused LDS plane kernel, or stick and column-chunk kernels when a plane
// exceeds LDS), x pass, x-plane forward; kernels built for codelets <= RMAX.
template <typename T, int RMAX>
static int generic_local(const Plan& P, const ApplyLayout& L, const Geom& g, int nbt, const cpx<T>* c, long ldc,
                         long sc, cpx<T>* box, const T* vp, const T* t, cpx<T>* out, long ldo, long so, T scale,
                         bool kin, unsigned char* ws, StageTimer& tm, hipStream_t st) {
  const size_t plane_lds = (size_t)P.n[1] * P.n[2] * sizeof(cpx<T>);
  const cpx<T>* twx = tw_of<T>(&P, 0);
  const cpx<T>* twy = tw_of<T>(&P, 1);
  const cpx<T>* twz = tw_of<T>(&P, 2);
  const int CC = col_chunk(&P, sizeof(cpx<T>));
  const dim3 gcol(P.na, nbt, (P.n[2] + CC - 1) / CC);
  const size_t col_lds = (size_t)P.n[1] * CC * sizeof(cpx<T>);
  const size_t stick_lds = (size_t)kStickChunk * P.n[2] * sizeof(cpx<T>);
  if (L.pencil_yc > 0) {
    cpx<T>* zs = (cpx<T>*)(ws + L.stick_off);
    const T* vz = sizeof(T) == 8 ? (const T*)P.vz64 : (const T*)P.vz32;
    const size_t zlds = (size_t)kPencilChunk * P.n[2] * sizeof(cpx<T>);
    const dim3 zgrid((P.nst_max + kPencilChunk - 1) / kPencilChunk, nbt);
    hipEvent_t ev = tm.begin();
    hipLaunchKernelGGL((k_pencil_z_inv<T, RMAX>), zgrid, dim3(256), zlds, st, g, c, ldc, sc, zs, twz);
    tm.end(ST_PLANE_INV, ev);
    const long total = (long)P.n[2] * nbt;
    const size_t slab_lds = ((size_t)P.na * P.n[1] + (size_t)L.pencil_yc * P.n[0]) * sizeof(cpx<T>);
    ev = tm.begin();
    hipLaunchKernelGGL((k_pencil_yx<T, RMAX>), dim3((unsigned)(8 * ((total + 7) / 8))), dim3(kPencilThreads),
                       slab_lds, st, g, nbt, L.pencil_yc, zs, vz, twx, twy);
    tm.end(ST_XPASS, ev);
    ev = tm.begin();
    if (kin)
      hipLaunchKernelGGL((k_pencil_z_fwd<T, true, RMAX>), zgrid, dim3(256), zlds, st, g, zs, c, ldc, sc, t, out,
                         ldo, so, scale, twz);
    else
      hipLaunchKernelGGL((k_pencil_z_fwd<T, false, RMAX>), zgrid, dim3(256), zlds, st, g, zs, c, ldc, sc, t, out,
                         ldo, so, scale, twz);
    tm.end(ST_PLANE_FWD, ev);
    GW_CHECK_HIP(hipGetLastError());
    return GW_OK;
  }
  hipEvent_t ev = tm.begin();
  if (L.fused) {
    hipLaunchKernelGGL((k_plane_inv<T, RMAX>), dim3(P.na, nbt), dim3(kPlaneThreads), plane_lds, st, g, c, ldc, sc,
                       box, twy, twz);
  } else {
    cpx<T>* sticks = (cpx<T>*)(ws + L.stick_off);
    hipLaunchKernelGGL((k_sticks_inv<T, RMAX>), dim3((P.nst_max + kStickChunk - 1) / kStickChunk, nbt), dim3(256),
                       stick_lds, st, g, c, ldc, sc, sticks, twz);
    hipLaunchKernelGGL((k_plane_inv_y<T, RMAX>), gcol, dim3(256), col_lds, st, g, CC, sticks, box, twy);
  }
  tm.end(ST_PLANE_INV, ev);
  const int XL = x_lines(&P);
  ev = tm.begin();
  hipLaunchKernelGGL((k_xpass<T, RMAX>), dim3((P.n[1] * P.n[2] + XL - 1) / XL, (nbt + kXBandGroup - 1) / kXBandGroup),
                     dim3(kXThreads), (size_t)P.n[0] * XL * sizeof(cpx<T>), st, g, XL, nbt, box, vp, twx);
  tm.end(ST_XPASS, ev);
  ev = tm.begin();
  if (L.fused) {
    if (kin)
      hipLaunchKernelGGL((k_plane_fwd<T, true, RMAX>), dim3(P.na, nbt), dim3(kPlaneThreads), plane_lds, st, g, box,
                         c, ldc, sc, t, out, ldo, so, scale, twy, twz);
    else
      hipLaunchKernelGGL((k_plane_fwd<T, false, RMAX>), dim3(P.na, nbt), dim3(kPlaneThreads), plane_lds, st, g, box,
                         c, ldc, sc, t, out, ldo, so, scale, twy, twz);
  } else {
    cpx<T>* sticks = (cpx<T>*)(ws + L.stick_off);
    hipLaunchKernelGGL((k_plane_fwd_y<T, RMAX>), gcol, dim3(256), col_lds, st, g, CC, box, sticks, twy);
    const int nsc = std::max(1, (P.nst_max + kStickChunk - 1) / kStickChunk);
    if (kin)
      hipLaunchKernelGGL((k_sticks_fwd<T, true, RMAX>), dim3(nsc, nbt), dim3(256), stick_lds, st, g, sticks, c, ldc,
                         sc, t, out, ldo, so, scale, twz);
    else
      hipLaunchKernelGGL((k_sticks_fwd<T, false, RMAX>), dim3(nsc, nbt), dim3(256), stick_lds, st, g, sticks, c,
                         ldc, sc, t, out, ldo, so, scale, twz);
  }
  tm.end(ST_PLANE_FWD, ev);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

// Dynamic-LDS limits of the generic local-term kernels (every codelet class).
template <typename T, int RMAX>
static void set_generic_lds_attrs_r(int want) {
  (void)hipFuncSetAttribute((const void*)k_plane_inv<T, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<T, true, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd<T, false, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_xpass<T, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_sticks_inv<T, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_sticks_fwd<T, true, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_sticks_fwd<T, false, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv_y<T, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_y<T, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_pencil_yx<T, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_pencil_z_inv<T, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_pencil_z_fwd<T, true, RMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_pencil_z_fwd<T, false, RMAX>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, want);
}
template <typename T>
static void set_generic_lds_attrs(int want) {
  set_generic_lds_attrs_r<T, 10>(want);
  set_generic_lds_attrs_r<T, 16>(want);
  set_generic_lds_attrs_r<T, kRmaxF3>(want);
}

template <typename T>
static int apply_t(gw_plan* h, int terms, const cpx<T>* c, long ldc, long sc, cpx<T>* out, long ldo,
                   long so, int nb, unsigned char* ws, hipStream_t st) {
  Plan& P = h->p;
  const int prec = sizeof(T) == 8 ? GW_FP64 : GW_FP32;
  const ApplyLayout L = apply_layout(&P, nb, prec, true);
  const Geom g = geom_of(&P, nb);
  const int nbt = P.nk * nb;
  const T* t = prec == GW_FP64 ? (const T*)P.t64 : (const T*)P.t32;
  const T* vp = prec == GW_FP64 ? (const T*)P.vp64 : (const T*)P.vp32;
  const cpx<T>* pr = prec == GW_FP64 ? (const cpx<T>*)P.p64 : (const cpx<T>*)P.p32;
  cpx<T>* box = (cpx<T>*)(ws + L.box_off);
  const bool kin = terms & GW_TERM_KINETIC;
  const T scale = T(1.0 / ((double)P.n[0] * P.n[1] * P.n[2]));
  StageTimer& tm = h->timer;
  tm.on = (P.flags & GW_PLAN_TIMING) != 0;
  tm.st = st;
  hipEvent_t ev;
  if (terms & GW_TERM_LOCAL) {
    GW_REQUIRE(P.have_v, GW_ERR_ARG, "local term requested but no potential set");
    if (int rc = require_box_fft(P, "the local term")) return rc;
    const int chunk0 = g_band_chunk[prec == GW_FP64 ? 0 : 1];
    const int chunk = chunk0 > 0 && chunk0 < nb ? chunk0 : nb;
    for (int b0 = 0; b0 < nb; b0 += chunk) {  // band chunks (one pass sequence each)
      const int nbc = std::min(chunk, nb - b0);
      const Geom gc = geom_of(&P, nbc);
      const cpx<T>* cc = c + (long)b0 * ldc;
      cpx<T>* oc = out + (long)b0 * ldo;
      tm.count_calls = b0 == 0;
      int rc;
      if (wave_path_ok(P, prec) && !L.pencil_yc) {
        const T* vn = prec == GW_FP64 ? (const T*)P.vx64 : (const T*)P.vx32;
        rc = launch_wave_path<T>(P, gc, P.nk * nbc, cc, ldc, sc, box, vn, t, oc, ldo, so, scale, kin, tm, st);
      } else {
        rc = plan_rmax(&P) <= 10 ? generic_local<T, 10>(P, L, gc, P.nk * nbc, cc, ldc, sc, box, vp, t, oc, ldo, so,
                                                        scale, kin, ws, tm, st)
             : plan_rmax(&P) <= 16 ? generic_local<T, 16>(P, L, gc, P.nk * nbc, cc, ldc, sc, box, vp, t, oc, ldo,
                                                          so, scale, kin, ws, tm, st)
                                   : generic_local<T, kRmaxF3>(P, L, gc, P.nk * nbc, cc, ldc, sc, box, vp, t, oc,
                                                               ldo, so, scale, kin, ws, tm, st);
      }
      tm.count_calls = true;
      if (rc) return rc;
    }
  } else if (kin) {
    hipLaunchKernelGGL(k_kinetic<T>, dim3(grid_blocks(P.npw_max, 256), nbt), dim3(256), 0, st, nb, P.npw_max,
                       P.npw, c, ldc, sc, t, out, ldo, so);
  }
  GW_CHECK_HIP(hipGetLastError());
  const bool base = terms & (GW_TERM_KINETIC | GW_TERM_LOCAL);
  if ((terms & GW_TERM_NONLOCAL) && P.nproj) {
    cpx<T>* becp = (cpx<T>*)(ws + L.becp_off);
    cpx<T>* u = (cpx<T>*)(ws + L.u_off);
    const long pstride = (long)P.nproj * P.npw_max;
    ev = tm.begin();
    int rc = gemm_inner<T>(P.nk, P.nproj, nb, P.npw_max, pr, P.npw_max, pstride, c, ldc, sc, becp, nb,
                           (long)P.nproj * nb, ws + L.part_off, L.part_bytes, st);
    if (rc) return rc;
    tm.end(ST_BECP, ev);
    ev = tm.begin();
    hipLaunchKernelGGL(k_dmul<T>, dim3(grid_blocks((long)P.nproj * nb, 256), P.nk), dim3(256), 0, st, P.nproj,
                       nb, P.d_colptr, P.d_row, P.d_val, becp, u);
    GW_CHECK_HIP(hipGetLastError());
    tm.end(ST_DMUL, ev);
    ev = tm.begin();
    rc = gemm_combine<T>(P.nk, nb, P.npw_max, P.nproj, u, nb, (long)P.nproj * nb, 0, pr, P.npw_max, pstride,
                         out, ldo, so, base ? T(1) : T(0), st);
    if (rc) return rc;
    tm.end(ST_BACKPROJ, ev);
  } else if (!base) {
    for (int k = 0; k < P.nk; ++k)
      hipLaunchKernelGGL(k_fill_zero<T>, dim3(grid_blocks((long)nb * ldo, 256)), dim3(256), 0, st,
                         (long)(nb - 1) * ldo + P.npw_max, out + (long)k * so);
  }
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_hpsi_apply_strided(gw_plan* h, int precision, int terms, const void* c, long ldc,
                                     long sc, void* out, long ldo, long so, int nb, void* ws,
                                     size_t ws_bytes, void* stream) {
  GW_REQUIRE(h && c && out && nb > 0, GW_ERR_ARG, "bad argument");
  GW_REQUIRE(precision == GW_FP64 || precision == GW_FP32, GW_ERR_ARG, "bad precision");
  const ApplyLayout L = apply_layout(&h->p, nb, precision, true);
  GW_REQUIRE(ws && ws_bytes >= L.total, GW_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, L.total);
  GW_REQUIRE(ldc >= h->p.npw_max && ldo >= h->p.npw_max, GW_ERR_ARG, "row stride < npw_max");
  if (precision == GW_FP64)
    return apply_t<double>(h, terms, (const double2*)c, ldc, sc, (double2*)out, ldo, so, nb,
                           (unsigned char*)ws, (hipStream_t)stream);
  return apply_t<float>(h, terms, (const float2*)c, ldc, sc, (float2*)out, ldo, so, nb, (unsigned char*)ws,
                        (hipStream_t)stream);
}

// ---- the no-allocation apply ----
// Scratch of gw_hpsi_apply_noalloc: one fixed array per GPU in the code
// object's device memory, held until the process ends. Host code never passes
// the array's name to a kernel (on the host it is only a stand-in symbol):
// each device's address comes from hipGetSymbolAddress, taken once per device
// with that device current (at plan creation, or at the entry point's first
// call on the device) and kept in g_scratch_ptr.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {nullptr};

static int noalloc_scratch_setup() {
  int dev = -1;
  GW_CHECK_HIP(hipGetDevice(&dev));
  GW_REQUIRE(dev >= 0 && dev < kMaxDevices, GW_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return GW_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  GW_REQUIRE(e == hipSuccess && p, GW_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
             hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
  return GW_OK;
}

extern "C" size_t gw_hpsi_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int gw_hpsi_apply_noalloc(gw_plan* h, int precision, int terms, const void* c, long ldc, long sc,
                                     void* out, long ldo, long so, int nb, void* stream) {
  GW_REQUIRE(h && c && out && nb > 0, GW_ERR_ARG, "bad argument");
  GW_REQUIRE(precision == GW_FP64 || precision == GW_FP32, GW_ERR_ARG, "bad precision");
  // contractions on the kernels here only (rocBLAS handles and workspaces allocate)
  const int prev = set_kernels_only(1);
  const size_t need = apply_layout(&h->p, nb, precision, true).total;
  int rc = GW_OK;
  if (need > kNoAllocScratchBytes) {
    set_error("no-allocation apply: %d bands need %zu bytes of scratch, above the fixed maximum of %zu "
              "(use gw_hpsi_apply_strided with a workspace)", nb, need, kNoAllocScratchBytes);
    rc = GW_ERR_WORKSPACE;
  }
  if (!rc) rc = noalloc_scratch_setup();
  if (!rc) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    void* ws = g_scratch_ptr[dev];
    const int flags = h->p.flags;
    h->p.flags &= ~GW_PLAN_TIMING;  // timing would create events
    rc = gw_hpsi_apply_strided(h, precision, terms, c, ldc, sc, out, ldo, so, nb, ws, kNoAllocScratchBytes,
                               stream);
    h->p.flags = flags;
  }
  set_kernels_only(prev);
  return rc;
}

extern "C" int gw_hpsi_apply(gw_plan* h, int precision, int terms, const void* c, void* out, int nb,
                             void* ws, size_t ws_bytes, void* stream) {
  if (!h) { set_error("null plan"); return GW_ERR_ARG; }
  const long ld = h->p.npw_max, s = (long)nb * h->p.npw_max;
  return gw_hpsi_apply_strided(h, precision, terms, c, ld, s, out, ld, s, nb, ws, ws_bytes, stream);
}

extern "C" size_t gw_wfc_to_real_workspace_bytes(const gw_plan* h, int nb, int precision) {
  if (!h || nb <= 0) return 0;
  return apply_layout(&h->p, nb, precision).total;
}

template <typename T>
static int to_real_t(gw_plan* h, const cpx<T>* c, cpx<T>* psi, int nb, unsigned char* ws, hipStream_t st) {
  Plan& P = h->p;
  const int prec = sizeof(T) == 8 ? GW_FP64 : GW_FP32;
  const ApplyLayout L = apply_layout(&P, nb, prec);
  const Geom g = geom_of(&P, nb);
  const int nbt = P.nk * nb;
  cpx<T>* box = (cpx<T>*)(ws + L.box_off);
  const long ld = P.npw_max, sc = (long)nb * P.npw_max;
  if (L.fused) {
    hipLaunchKernelGGL(k_plane_inv<T>, dim3(P.na, nbt), dim3(kPlaneThreads), (size_t)P.n[1] * P.n[2] * sizeof(cpx<T>),
                       st, g, c, ld, sc, box, tw_of<T>(&P, 1), tw_of<T>(&P, 2));
  } else {
    cpx<T>* sticks = (cpx<T>*)(ws + L.stick_off);
    const int CC = col_chunk(&P, sizeof(cpx<T>));
    hipLaunchKernelGGL(k_sticks_inv<T>, dim3((P.nst_max + kStickChunk - 1) / kStickChunk, nbt), dim3(256),
                       kStickChunk * P.n[2] * sizeof(cpx<T>), st, g, c, ld, sc, sticks, tw_of<T>(&P, 2));
    hipLaunchKernelGGL(k_plane_inv_y<T>, dim3(P.na, nbt, (P.n[2] + CC - 1) / CC), dim3(256),
                       (size_t)P.n[1] * CC * sizeof(cpx<T>), st, g, CC, sticks, box, tw_of<T>(&P, 1));
  }
  const int XL = x_lines(&P);
  hipLaunchKernelGGL(k_x_to_real<T>, dim3((P.n[1] * P.n[2] + XL - 1) / XL, nbt), dim3(kXThreads),
                     (size_t)P.n[0] * XL * sizeof(cpx<T>), st, g, XL, box, tw_of<T>(&P, 0), h->perm[0],
                     h->perm[1], h->perm[2], psi);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_wfc_to_real(gw_plan* h, int precision, const void* c, void* psi, int nb, void* ws,
                              size_t ws_bytes, void* stream) {
  GW_REQUIRE(h && c && psi && nb > 0, GW_ERR_ARG, "bad argument");
  if (int rc = require_box_fft(h->p, "psi(r)")) return rc;
  GW_REQUIRE(plan_rmax(&h->p) < kRmaxF3, GW_ERR_UNSUPPORTED,
             "psi(r) is not available on boxes with a three-factor length (%d x %d x %d)", h->p.n[0], h->p.n[1],
             h->p.n[2]);
  const ApplyLayout L = apply_layout(&h->p, nb, precision);
  GW_REQUIRE(ws && ws_bytes >= L.total, GW_ERR_WORKSPACE, "workspace too small");
  if (precision == GW_FP64)
    return to_real_t<double>(h, (const double2*)c, (double2*)psi, nb, (unsigned char*)ws, (hipStream_t)stream);
  return to_real_t<float>(h, (const float2*)c, (float2*)psi, nb, (unsigned char*)ws, (hipStream_t)stream);
}

template <typename T>
static int pack_t(gw_plan* h, const void* c, void* box, int nb, int unpack, hipStream_t st) {
  Plan& P = h->p;
  const Geom g = geom_of(&P, nb);
  const int nbt = P.nk * nb;
  const long N = (long)P.n[0] * P.n[1] * P.n[2];
  if (!unpack)
    hipLaunchKernelGGL(k_fill_zero<T>, dim3(grid_blocks(N * nbt, 256)), dim3(256), 0, st, N * nbt, (cpx<T>*)box);
  else
    hipLaunchKernelGGL(k_fill_zero<T>, dim3(grid_blocks((long)nbt * P.npw_max, 256)), dim3(256), 0, st,
                       (long)nbt * P.npw_max, (cpx<T>*)c);
  hipLaunchKernelGGL(k_pack_box<T>, dim3(P.na, nbt), dim3(256), 0, st, g, (const cpx<T>*)c, (cpx<T>*)box, unpack);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_sphere_to_box(gw_plan* h, int precision, const void* c, void* box, int nb, void* stream) {
  GW_REQUIRE(h && c && box && nb > 0, GW_ERR_ARG, "bad argument");
  return precision == GW_FP64 ? pack_t<double>(h, c, box, nb, 0, (hipStream_t)stream)
                              : pack_t<float>(h, c, box, nb, 0, (hipStream_t)stream);
}

extern "C" int gw_box_to_sphere(gw_plan* h, int precision, const void* box, void* c_out, int nb,
                                void* stream) {
  GW_REQUIRE(h && c_out && box && nb > 0, GW_ERR_ARG, "bad argument");
  return precision == GW_FP64 ? pack_t<double>(h, c_out, const_cast<void*>(box), nb, 1, (hipStream_t)stream)
                              : pack_t<float>(h, c_out, const_cast<void*>(box), nb, 1, (hipStream_t)stream);
}

extern "C" size_t gw_hpsi_adjoint_v_workspace_bytes(const gw_plan* h, int nb, int precision) {
  if (!h || nb <= 0) return 0;
  const Plan& P = h->p;
  const size_t e = precision == GW_FP64 ? 16 : 8;
  const size_t nbt = (size_t)P.nk * nb;
  const size_t box = nbt * P.na * P.n[1] * P.n[2] * e;
  const size_t ngroups = (nbt + kAdjBandGroup - 1) / kAdjBandGroup;
  const size_t part = ngroups * (size_t)P.n[0] * P.n[1] * P.n[2] * sizeof(double);
  return 2 * align256(box) + align256(part);
}

template <typename T>
static int adjv_t(gw_plan* h, const cpx<T>* c, long ldc, long sc, const cpx<T>* gr, long ldg, long sg, int nb,
                  double* gv, unsigned char* ws, hipStream_t st) {
  Plan& P = h->p;
  const int prec = sizeof(T) == 8 ? GW_FP64 : GW_FP32;
  GW_REQUIRE(wave_path_ok(P, prec), GW_ERR_UNSUPPORTED,
             "V adjoint needs a cubic 24/36/90/120 box with the wave-line kernels");
  const Geom g = geom_of(&P, nb);
  const int nbt = P.nk * nb;
  const size_t e = sizeof(cpx<T>);
  const size_t box = align256((size_t)nbt * P.na * P.n[1] * P.n[2] * e);
  cpx<T>* boxc = (cpx<T>*)ws;
  cpx<T>* boxg = (cpx<T>*)(ws + box);
  double* part = (double*)(ws + 2 * box);
  switch (P.n[0]) {
    case 24: return adjv_wave_n<T, 24>(P, g, nbt, c, ldc, sc, gr, ldg, sg, boxc, boxg, part, gv, st);
    case 36: return adjv_wave_n<T, 36>(P, g, nbt, c, ldc, sc, gr, ldg, sg, boxc, boxg, part, gv, st);
    case 90: return adjv_wave_n<T, 90>(P, g, nbt, c, ldc, sc, gr, ldg, sg, boxc, boxg, part, gv, st);
    case 120: return adjv_wave_n<T, 120>(P, g, nbt, c, ldc, sc, gr, ldg, sg, boxc, boxg, part, gv, st);
    default: return GW_ERR_UNSUPPORTED;
  }
}

extern "C" int gw_hpsi_adjoint_v(gw_plan* h, int precision, const void* c, long ldc, long sc, const void* g,
                                 long ldg, long sg, int nb, double* grad_v, void* ws, size_t ws_bytes,
                                 void* stream) {
  GW_REQUIRE(h && c && g && grad_v && nb > 0, GW_ERR_ARG, "bad argument");
  if (int rc = require_box_fft(h->p, "the V(r) adjoint")) return rc;
  const size_t need = gw_hpsi_adjoint_v_workspace_bytes(h, nb, precision);
  GW_REQUIRE(ws && ws_bytes >= need, GW_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  if (precision == GW_FP64)
    return adjv_t<double>(h, (const double2*)c, ldc, sc, (const double2*)g, ldg, sg, nb, grad_v,
                          (unsigned char*)ws, (hipStream_t)stream);
  return adjv_t<float>(h, (const float2*)c, ldc, sc, (const float2*)g, ldg, sg, nb, grad_v, (unsigned char*)ws,
                       (hipStream_t)stream);
}

// ------------------------------------------------------------------ ragged batches
// Many independent operators (different boxes, spheres, potentials, projector
// counts and band counts) applied together: each pass of the local term is ONE
// launch over the concatenated work items of all instances (a workgroup looks up
// its instance in a device table of descriptors), so a batch of small cells
// fills the chip instead of issuing a few dozen under-occupied launches.

namespace gw {

template <typename T>
struct RagInst {
  Geom g;
  const cpx<T>* c;
  long ldc, sc;
  cpx<T>* out;
  long ldo, so;
  const T* t;
  const T* vp;
  const cpx<T>* twx;
  const cpx<T>* twy;
  const cpx<T>* twz;
  cpx<T>* box;
  T scale;
  int XL, nbt, nchunk, kin;
  long w13, w2;  // first work item of this instance in the plane passes / in the x pass
};

template <typename T>
__device__ __forceinline__ int rag_find(const RagInst<T>* R, int n, long w, bool xpass) {
  int i = 0;
  while (i + 1 < n && (xpass ? R[i + 1].w2 : R[i + 1].w13) <= w) ++i;
  return i;
}

template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(kPlaneThreads) k_plane_inv_rag(const RagInst<T>* __restrict__ R, int n, long total) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const long w = xcd_work_id(total);
  if (w < 0) return;
  const RagInst<T>& d = R[rag_find(R, n, w, false)];
  const long local = w - d.w13;
  const int a = (int)(local % d.g.na), bi = (int)(local / d.g.na);
  plane_inv_body<T, RMAX>(d.g, a, bi, d.c, d.ldc, d.sc, d.box, d.twy, d.twz, reinterpret_cast<cpx<T>*>(smem_raw));
}

template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(kXThreads) k_xpass_rag(const RagInst<T>* __restrict__ R, int n, long total) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const long w = xcd_work_id(total);
  if (w < 0) return;
  const RagInst<T>& d = R[rag_find(R, n, w, true)];
  const long local = w - d.w2;
  const int chunk = (int)(local % d.nchunk), bgrp = (int)(local / d.nchunk);
  xpass_body<T, RMAX>(d.g, d.XL, d.nbt, chunk, bgrp, d.box, d.vp, d.twx, reinterpret_cast<cpx<T>*>(smem_raw));
}

template <typename T, int RMAX = 16>
__global__ void __launch_bounds__(kPlaneThreads) k_plane_fwd_rag(const RagInst<T>* __restrict__ R, int n, long total) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const long w = xcd_work_id(total);
  if (w < 0) return;
  const RagInst<T>& d = R[rag_find(R, n, w, false)];
  const long local = w - d.w13;
  const int a = (int)(local % d.g.na), bi = (int)(local / d.g.na);
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem_raw);
  if (d.kin)
    plane_fwd_body<T, true, RMAX>(d.g, a, bi, d.box, d.c, d.ldc, d.sc, d.t, d.out, d.ldo, d.so, d.scale, d.twy, d.twz,
                                  tile);
  else
    plane_fwd_body<T, false, RMAX>(d.g, a, bi, d.box, d.c, d.ldc, d.sc, d.t, d.out, d.ldo, d.so, d.scale, d.twy,
                                   d.twz, tile);
}

}  // namespace gw

// Launch parameters of one bound batched call and the device copy of its
// descriptor tables: [RagInst x n | GroupInner | GroupReduce | GroupDmul |
// GroupCombine] (gn each). The tables reach the device by hipMemcpyAsync on the
// launch stream from pinned staging, so an upload is ordered after every
// earlier launch on that stream and no host-synchronous copy is needed.
struct BatchBound {
  const void* c; long ldc, sc;
  void* out; long ldo, so;
  int nb;
  unsigned char* ws;
};
struct BatchBind {
  int prec = -1, terms = 0, rmax = 16;
  long tot13 = 0, tot2 = 0;
  size_t lds13 = 0, lds2 = 0;
  std::vector<BatchBound> bound;
  bool grouped = false;
  int gn = 0;                                         // jobs per group table
  long gt_inner = 0, gt_reduce = 0, gt_dmul = 0, gt_combine = 0;  // work items / blocks
  size_t gtab_off = 0;                                // byte offset of the group tables
  unsigned char* dev = nullptr;                       // device tables
  unsigned char* host = nullptr;                      // pinned staging of the same bytes
  size_t cap = 0;
  hipEvent_t done = nullptr;  // recorded after the last launch that read `dev` (ring slots)
  bool pending = false;
};

static size_t batch_table_cap(int n) {
  return align256((size_t)n * sizeof(RagInst<double>)) +
         (size_t)n * (sizeof(GroupInner) + sizeof(GroupReduce) + sizeof(GroupDmul) + sizeof(GroupCombine));
}

static bool bind_alloc(BatchBind& s, int n) {
  s.cap = batch_table_cap(n);
  if (hipMalloc(&s.dev, s.cap) != hipSuccess) return false;
  if (hipHostMalloc(&s.host, s.cap, hipHostMallocDefault) != hipSuccess) return false;
  return hipEventCreateWithFlags(&s.done, hipEventDisableTiming) == hipSuccess;
}

static void bind_free(BatchBind& s) {
  if (s.done && s.pending) (void)hipEventSynchronize(s.done);
  if (s.dev) (void)hipFree(s.dev);
  if (s.host) (void)hipHostFree(s.host);
  if (s.done) (void)hipEventDestroy(s.done);
  s.dev = s.host = nullptr;
  s.done = nullptr;
}

constexpr int kBatchLanes = 4;  // the launch stream + 3 side streams (4 hardware queues per process)
constexpr int kBatchRing = 4;   // table slots of eager (unbound) calls

struct gw_batch {
  std::vector<gw_plan*> plans;
  // the instances' nonlocal chains (small, latency-bound GEMMs) are spread over
  // the launch stream and kBatchLanes - 1 side streams of default priority,
  // forked from and joined back into the launch stream by events
  hipStream_t side[kBatchLanes - 1] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[kBatchLanes - 1] = {nullptr, nullptr, nullptr};
  BatchBind ring[kBatchRing];
  int next = 0;
};

// A bound call with tables of its own (never rewritten by later calls), so a
// HIP graph that recorded gw_batch_apply_bound stays valid until destroyed.
struct gw_batch_binding {
  gw_batch* b;
  BatchBind s;
};

static size_t batch_ws_bytes(const gw_batch* b, const int* nb, int precision) {
  size_t tot = 0;
  for (size_t i = 0; i < b->plans.size(); ++i)
    if (nb[i] > 0) tot += align256(apply_layout(&b->plans[i]->p, nb[i], precision).total);
  return tot;
}

extern "C" int gw_batch_create(gw_batch** out, gw_plan* const* plans, int n) {
  GW_REQUIRE(out && plans && n > 0, GW_ERR_ARG, "bad argument");
  auto* b = new gw_batch();
  for (int i = 0; i < n; ++i) {
    if (!plans[i]) {
      delete b;
      GW_REQUIRE(false, GW_ERR_ARG, "plan %d is null", i);
    }
    if (int rc = require_box_fft(plans[i]->p, "a ragged batch")) {
      delete b;
      return rc;
    }
    b->plans.push_back(plans[i]);
  }
  bool ok = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
  for (int j = 0; j < kBatchLanes - 1 && ok; ++j)
    ok = hipStreamCreateWithFlags(&b->side[j], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->ev_join[j], hipEventDisableTiming) == hipSuccess;
  for (int j = 0; j < kBatchRing && ok; ++j) ok = bind_alloc(b->ring[j], n);
  if (!ok) {
    gw_batch_destroy(b);
    GW_REQUIRE(false, GW_ERR_HIP, "allocating the batch's tables, side streams or events failed");
  }
  const int want = kMaxLds;
  (void)hipFuncSetAttribute((const void*)k_plane_inv_rag<double>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv_rag<float>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_rag<double>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_rag<float>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass_rag<double>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass_rag<float>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv_rag<double, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv_rag<float, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_rag<double, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_rag<float, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass_rag<double, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass_rag<float, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv_rag<double, kRmaxF3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_inv_rag<float, kRmaxF3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_rag<double, kRmaxF3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_plane_fwd_rag<float, kRmaxF3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            want);
  (void)hipFuncSetAttribute((const void*)k_xpass_rag<double, kRmaxF3>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  (void)hipFuncSetAttribute((const void*)k_xpass_rag<float, kRmaxF3>, hipFuncAttributeMaxDynamicSharedMemorySize, want);
  *out = b;
  return GW_OK;
}

extern "C" int gw_batch_destroy(gw_batch* b) {
  if (!b) return GW_OK;
  for (int j = 0; j < kBatchLanes - 1; ++j) {
    if (b->side[j]) (void)hipStreamSynchronize(b->side[j]);
    if (b->side[j]) (void)hipStreamDestroy(b->side[j]);
    if (b->ev_join[j]) (void)hipEventDestroy(b->ev_join[j]);
  }
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  for (int j = 0; j < kBatchRing; ++j) bind_free(b->ring[j]);
  delete b;
  return GW_OK;
}

extern "C" size_t gw_batch_workspace_bytes(const gw_batch* b, const int* nb, int precision) {
  if (!b || !nb) return 0;
  return batch_ws_bytes(b, nb, precision);
}

// Fill s's tables for this call on the host and upload them on `st`.
template <typename T>
static int batch_bind_t(gw_batch* b, BatchBind& s, int terms, const void* const* c, const long* ldc, const long* sc,
                        void* const* out, const long* ldo, const long* so, const int* nb, unsigned char* ws,
                        hipStream_t st) {
  const int n = (int)b->plans.size();
  const int prec = sizeof(T) == 8 ? GW_FP64 : GW_FP32;
  std::vector<RagInst<T>> R(n);
  std::vector<BatchBound> bound(n, BatchBound{});
  long w13 = 0, w2 = 0;
  size_t lds13 = 0, lds2 = 0, off = 0;
  for (int i = 0; i < n; ++i) {
    Plan& P = b->plans[i]->p;
    GW_REQUIRE(nb[i] >= 0, GW_ERR_ARG, "instance %d: negative band count", i);
    if (nb[i] == 0) {  // instance sits this call out: an empty work range
      R[i] = RagInst<T>{};
      R[i].g.na = 1;
      R[i].nchunk = 1;
      R[i].w13 = w13;
      R[i].w2 = w2;
      continue;
    }
    GW_REQUIRE(c[i] && out[i], GW_ERR_ARG, "instance %d: bad block", i);
    GW_REQUIRE(ldc[i] >= P.npw_max && ldo[i] >= P.npw_max, GW_ERR_ARG, "instance %d: row stride < npw_max", i);
    GW_REQUIRE(P.have_v, GW_ERR_ARG, "instance %d: no potential set", i);
    GW_REQUIRE(plane_fits_lds(&P, prec), GW_ERR_UNSUPPORTED, "instance %d: x-plane exceeds LDS", i);
    const ApplyLayout L = apply_layout(&P, nb[i], prec);
    RagInst<T>& d = R[i];
    d.g = geom_of(&P, nb[i]);
    d.c = (const cpx<T>*)c[i];
    d.ldc = ldc[i];
    d.sc = sc[i];
    d.out = (cpx<T>*)out[i];
    d.ldo = ldo[i];
    d.so = so[i];
    d.t = prec == GW_FP64 ? (const T*)P.t64 : (const T*)P.t32;
    d.vp = prec == GW_FP64 ? (const T*)P.vp64 : (const T*)P.vp32;
    d.twx = tw_of<T>(&P, 0);
    d.twy = tw_of<T>(&P, 1);
    d.twz = tw_of<T>(&P, 2);
    d.box = (cpx<T>*)(ws + off + L.box_off);
    d.scale = T(1.0 / ((double)P.n[0] * P.n[1] * P.n[2]));
    d.XL = x_lines(&P);
    d.nbt = P.nk * nb[i];
    d.nchunk = (P.n[1] * P.n[2] + d.XL - 1) / d.XL;
    d.kin = (terms & GW_TERM_KINETIC) ? 1 : 0;
    d.w13 = w13;
    d.w2 = w2;
    w13 += (long)P.na * d.nbt;
    w2 += (long)d.nchunk * ((d.nbt + kXBandGroup - 1) / kXBandGroup);
    lds13 = std::max(lds13, (size_t)P.n[1] * P.n[2] * sizeof(cpx<T>));
    lds2 = std::max(lds2, (size_t)P.n[0] * d.XL * sizeof(cpx<T>));
    bound[i] = BatchBound{c[i], ldc[i], sc[i], out[i], ldo[i], so[i], nb[i], ws + off};
    off += align256(L.total);
  }
  std::vector<GroupInner> gi;
  std::vector<GroupReduce> gr;
  std::vector<GroupDmul> gd;
  std::vector<GroupCombine> gc;
  long ti = 0, tr = 0, td = 0, tc = 0;
  const bool grouped = (terms & GW_TERM_NONLOCAL) && !inner_lib(prec) && !combine_lib(prec);
  if (grouped) {
    for (int i = 0; i < n; ++i) {
      Plan& P = b->plans[i]->p;
      const BatchBound& B = bound[i];
      if (!P.nproj || B.nb == 0) continue;
      const ApplyLayout L = apply_layout(&P, B.nb, prec);
      const void* pr = prec == GW_FP64 ? (const void*)P.p64 : (const void*)P.p32;
      unsigned char* becp = B.ws + L.becp_off;
      unsigned char* u = B.ws + L.u_off;
      const long pstride = (long)P.nproj * P.npw_max;
      int splits = 1, kchunk = 1;
      gemm_inner_shape(P.nk, P.nproj, B.nb, P.npw_max, &splits, &kchunk);
      GW_REQUIRE((size_t)P.nk * splits * P.nproj * 2 * B.nb * sizeof(T) <= L.part_bytes, GW_ERR_WORKSPACE,
                 "instance %d: split-K partials exceed the workspace", i);
      const int intx = (2 * B.nb + 63) / 64, inty = (P.nproj + 63) / 64;
      gi.push_back(GroupInner{pr, B.c, B.ws + L.part_off, P.npw_max, pstride, B.ldc, B.sc, ti, P.nproj, B.nb,
                              P.npw_max, splits, kchunk, intx, inty});
      ti += (long)intx * inty * P.nk * splits;
      const int nblk = (int)(((long)P.nproj * 2 * B.nb + 63) / 64);
      gr.push_back(GroupReduce{B.ws + L.part_off, becp, B.nb, (long)P.nproj * B.nb, tr, P.nproj, B.nb, splits,
                               nblk});
      tr += (long)nblk * P.nk;
      const int dblk = (int)(((long)P.nproj * B.nb + 255) / 256);
      gd.push_back(GroupDmul{P.d_colptr, P.d_row, P.d_val, becp, u, td, P.nproj, B.nb, dblk});
      td += (long)dblk * P.nk;
      const int cntx = (2 * P.npw_max + 63) / 64, cnty = (B.nb + 63) / 64;
      gc.push_back(GroupCombine{u, pr, B.out, B.nb, (long)P.nproj * B.nb, P.npw_max, pstride, B.ldo, B.so, tc,
                                B.nb, P.npw_max, P.nproj, cntx, cnty});
      tc += (long)cntx * cnty * P.nk;
    }
  }
  const int gn = (int)gi.size();
  const size_t gtab_off = align256((size_t)n * sizeof(RagInst<T>));
  const size_t bytes =
      gtab_off + (size_t)gn * (sizeof(GroupInner) + sizeof(GroupReduce) + sizeof(GroupDmul) + sizeof(GroupCombine));
  GW_REQUIRE(s.dev && s.host && bytes <= s.cap, GW_ERR_HIP, "batch tables exceed their allocation (%zu > %zu)",
             bytes, s.cap);
  // the staging buffer may still feed an earlier upload of this slot
  if (s.pending) GW_CHECK_HIP(hipEventSynchronize(s.done));
  unsigned char* h = s.host;
  std::memcpy(h, R.data(), (size_t)n * sizeof(RagInst<T>));
  size_t o = gtab_off;
  std::memcpy(h + o, gi.data(), gn * sizeof(GroupInner)); o += gn * sizeof(GroupInner);
  std::memcpy(h + o, gr.data(), gn * sizeof(GroupReduce)); o += gn * sizeof(GroupReduce);
  std::memcpy(h + o, gd.data(), gn * sizeof(GroupDmul)); o += gn * sizeof(GroupDmul);
  std::memcpy(h + o, gc.data(), gn * sizeof(GroupCombine));
  GW_CHECK_HIP(hipMemcpyAsync(s.dev, h, bytes, hipMemcpyHostToDevice, st));
  s.bound = std::move(bound);
  s.grouped = grouped;
  s.gn = gn;
  s.gtab_off = gtab_off;
  s.gt_inner = ti; s.gt_reduce = tr; s.gt_dmul = td; s.gt_combine = tc;
  s.prec = prec;
  s.terms = terms;
  s.rmax = 1;
  for (gw_plan* pl : b->plans) s.rmax = std::max(s.rmax, plan_rmax(&pl->p));
  s.tot13 = w13;
  s.tot2 = w2;
  s.lds13 = lds13;
  s.lds2 = lds2;
  return GW_OK;
}

template <typename T>
static int batch_apply_t(gw_batch* b, const BatchBind& s, hipStream_t st) {
  const int n = (int)b->plans.size();
  const RagInst<T>* R = (const RagInst<T>*)s.dev;
  if (s.tot13 == 0) return GW_OK;
  const dim3 g13((unsigned)(8 * ((s.tot13 + 7) / 8))), g2((unsigned)(8 * ((s.tot2 + 7) / 8)));
  if (s.rmax <= 10) {
    hipLaunchKernelGGL((k_plane_inv_rag<T, 10>), g13, dim3(kPlaneThreads), s.lds13, st, R, n, s.tot13);
    hipLaunchKernelGGL((k_xpass_rag<T, 10>), g2, dim3(kXThreads), s.lds2, st, R, n, s.tot2);
    hipLaunchKernelGGL((k_plane_fwd_rag<T, 10>), g13, dim3(kPlaneThreads), s.lds13, st, R, n, s.tot13);
  } else if (s.rmax >= kRmaxF3) {
    hipLaunchKernelGGL((k_plane_inv_rag<T, kRmaxF3>), g13, dim3(kPlaneThreads), s.lds13, st, R, n, s.tot13);
    hipLaunchKernelGGL((k_xpass_rag<T, kRmaxF3>), g2, dim3(kXThreads), s.lds2, st, R, n, s.tot2);
    hipLaunchKernelGGL((k_plane_fwd_rag<T, kRmaxF3>), g13, dim3(kPlaneThreads), s.lds13, st, R, n, s.tot13);
  } else {
    hipLaunchKernelGGL(k_plane_inv_rag<T>, g13, dim3(kPlaneThreads), s.lds13, st, R, n, s.tot13);
    hipLaunchKernelGGL(k_xpass_rag<T>, g2, dim3(kXThreads), s.lds2, st, R, n, s.tot2);
    hipLaunchKernelGGL(k_plane_fwd_rag<T>, g13, dim3(kPlaneThreads), s.lds13, st, R, n, s.tot13);
  }
  GW_CHECK_HIP(hipGetLastError());
  if (!(s.terms & GW_TERM_NONLOCAL)) return GW_OK;
  const int prec = sizeof(T) == 8 ? GW_FP64 : GW_FP32;
  if (s.grouped) {  // one launch per contraction over every cell
    if (s.gn == 0) return GW_OK;
    const int gn = s.gn;
    const GroupInner* gi = (const GroupInner*)(s.dev + s.gtab_off);
    const GroupReduce* gr = (const GroupReduce*)(gi + gn);
    const GroupDmul* gd = (const GroupDmul*)(gr + gn);
    const GroupCombine* gc = (const GroupCombine*)(gd + gn);
    int rc = gemm_group_inner<T>(gi, gn, s.gt_inner, st);
    if (!rc) rc = gemm_group_reduce<T>(gr, gn, s.gt_reduce, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_group_dmul<T>, dim3((unsigned)s.gt_dmul), dim3(256), 0, st, gd, gn, s.gt_dmul);
    GW_CHECK_HIP(hipGetLastError());
    return gemm_group_combine<T>(gc, gn, s.gt_combine, T(1), st);
  }
  GW_CHECK_HIP(hipEventRecord(b->ev_fork, st));
  for (int j = 0; j < kBatchLanes - 1; ++j) GW_CHECK_HIP(hipStreamWaitEvent(b->side[j], b->ev_fork, 0));
  auto join = [&]() {  // the launch stream waits for every side stream's chains
    for (int j = 0; j < kBatchLanes - 1; ++j) {
      (void)hipEventRecord(b->ev_join[j], b->side[j]);
      (void)hipStreamWaitEvent(st, b->ev_join[j], 0);
    }
  };
  int lane = 0;
  for (int i = 0; i < n; ++i) {
    Plan& P = b->plans[i]->p;
    const BatchBound& B = s.bound[i];
    if (!P.nproj || B.nb == 0) continue;
    // instance i's chain on one lane; the chains write disjoint workspaces and outputs
    const hipStream_t ls = lane == 0 ? st : b->side[lane - 1];
    lane = (lane + 1) % kBatchLanes;
    const ApplyLayout L = apply_layout(&P, B.nb, prec);
    const cpx<T>* pr = prec == GW_FP64 ? (const cpx<T>*)P.p64 : (const cpx<T>*)P.p32;
    cpx<T>* becp = (cpx<T>*)(B.ws + L.becp_off);
    cpx<T>* u = (cpx<T>*)(B.ws + L.u_off);
    const long pstride = (long)P.nproj * P.npw_max;
    int rc = gemm_inner<T>(P.nk, P.nproj, B.nb, P.npw_max, pr, P.npw_max, pstride, (const cpx<T>*)B.c, B.ldc, B.sc,
                           becp, B.nb, (long)P.nproj * B.nb, B.ws + L.part_off, L.part_bytes, ls);
    if (rc == GW_OK) {
      hipLaunchKernelGGL(k_dmul<T>, dim3(grid_blocks((long)P.nproj * B.nb, 256), P.nk), dim3(256), 0, ls, P.nproj,
                         B.nb, P.d_colptr, P.d_row, P.d_val, becp, u);
      rc = gemm_combine<T>(P.nk, B.nb, P.npw_max, P.nproj, u, B.nb, (long)P.nproj * B.nb, 0, pr, P.npw_max,
                           pstride, (cpx<T>*)B.out, B.ldo, B.so, T(1), ls);
    }
    if (rc) {
      join();
      return rc;
    }
  }
  join();
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

static int batch_bind(gw_batch* b, BatchBind& s, int precision, int terms, const void* const* c, const long* ldc,
                      const long* sc, void* const* out, const long* ldo, const long* so, const int* nb, void* ws,
                      size_t ws_bytes, hipStream_t st) {
  GW_REQUIRE(b && c && ldc && sc && out && ldo && so && nb && ws, GW_ERR_ARG, "bad argument");
  GW_REQUIRE(precision == GW_FP64 || precision == GW_FP32, GW_ERR_ARG, "bad precision");
  GW_REQUIRE(terms & GW_TERM_LOCAL, GW_ERR_UNSUPPORTED, "batched apply needs the local term");
  const size_t need = batch_ws_bytes(b, nb, precision);
  GW_REQUIRE(ws_bytes >= need, GW_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  if (precision == GW_FP64)
    return batch_bind_t<double>(b, s, terms, c, ldc, sc, out, ldo, so, nb, (unsigned char*)ws, st);
  return batch_bind_t<float>(b, s, terms, c, ldc, sc, out, ldo, so, nb, (unsigned char*)ws, st);
}

static int batch_launch(gw_batch* b, const BatchBind& s, hipStream_t st) {
  GW_REQUIRE(s.prec >= 0, GW_ERR_ARG, "batch not bound");
  return s.prec == GW_FP64 ? batch_apply_t<double>(b, s, st) : batch_apply_t<float>(b, s, st);
}

extern "C" int gw_batch_bind(gw_batch* b, int precision, int terms, const void* const* c, const long* ldc,
                             const long* sc, void* const* out, const long* ldo, const long* so, const int* nb,
                             void* ws, size_t ws_bytes, void* stream, gw_batch_binding** binding) {
  GW_REQUIRE(b && binding, GW_ERR_ARG, "bad argument");
  auto* bd = new gw_batch_binding();
  bd->b = b;
  if (!bind_alloc(bd->s, (int)b->plans.size())) {
    bind_free(bd->s);
    delete bd;
    GW_REQUIRE(false, GW_ERR_HIP, "allocating the binding's tables failed");
  }
  const int rc = batch_bind(b, bd->s, precision, terms, c, ldc, sc, out, ldo, so, nb, ws, ws_bytes,
                            (hipStream_t)stream);
  if (rc) {
    bind_free(bd->s);
    delete bd;
    return rc;
  }
  *binding = bd;
  return GW_OK;
}

extern "C" int gw_batch_apply_bound(gw_batch_binding* bd, void* stream) {
  GW_REQUIRE(bd && bd->b, GW_ERR_ARG, "null binding");
  return batch_launch(bd->b, bd->s, (hipStream_t)stream);
}

extern "C" int gw_batch_binding_destroy(gw_batch_binding* bd) {
  if (!bd) return GW_OK;
  bind_free(bd->s);
  delete bd;
  return GW_OK;
}

extern "C" int gw_batch_apply(gw_batch* b, int precision, int terms, const void* const* c, const long* ldc,
                              const long* sc, void* const* out, const long* ldo, const long* so, const int* nb,
                              void* ws, size_t ws_bytes, void* stream) {
  GW_REQUIRE(b, GW_ERR_ARG, "null batch");
  hipStream_t st = (hipStream_t)stream;
  BatchBind& s = b->ring[b->next];
  b->next = (b->next + 1) % kBatchRing;
  // this slot's device tables may still be read by a launch on another stream
  if (s.pending) GW_CHECK_HIP(hipStreamWaitEvent(st, s.done, 0));
  int rc = batch_bind(b, s, precision, terms, c, ldc, sc, out, ldo, so, nb, ws, ws_bytes, st);
  if (!rc) rc = batch_launch(b, s, st);
  if (rc) return rc;
  GW_CHECK_HIP(hipEventRecord(s.done, st));
  s.pending = true;
  return GW_OK;
}

extern "C" size_t gw_density_workspace_bytes(const gw_plan* h, int nb, int precision) {
  if (!h || nb <= 0) return 0;
  const Plan& P = h->p;
  const size_t e = precision == GW_FP64 ? 16 : 8;
  const size_t nbt = (size_t)P.nk * nb;
  const size_t box = nbt * P.na * P.n[1] * P.n[2] * e;
  const size_t ngroups = (nbt + kAdjBandGroup - 1) / kAdjBandGroup;
  return align256(box) + align256(ngroups * (size_t)P.n[0] * P.n[1] * P.n[2] * sizeof(double));
}

template <typename T>
static int density_t(gw_plan* h, const cpx<T>* c, long ldc, long sc, int nb, const double* wts, double* rho,
                     unsigned char* ws, hipStream_t st) {
  Plan& P = h->p;
  const int prec = sizeof(T) == 8 ? GW_FP64 : GW_FP32;
  GW_REQUIRE(wave_path_ok(P, prec), GW_ERR_UNSUPPORTED,
             "band density needs a cubic 24/36/90/120 box with the wave-line kernels");
  const Geom g = geom_of(&P, nb);
  const int nbt = P.nk * nb;
  const size_t box = align256((size_t)nbt * P.na * P.n[1] * P.n[2] * sizeof(cpx<T>));
  cpx<T>* bx = (cpx<T>*)ws;
  double* part = (double*)(ws + box);
  switch (P.n[0]) {
    case 24: return density_wave_n<T, 24>(P, g, nbt, c, ldc, sc, wts, bx, part, rho, st);
    case 36: return density_wave_n<T, 36>(P, g, nbt, c, ldc, sc, wts, bx, part, rho, st);
    case 90: return density_wave_n<T, 90>(P, g, nbt, c, ldc, sc, wts, bx, part, rho, st);
    case 120: return density_wave_n<T, 120>(P, g, nbt, c, ldc, sc, wts, bx, part, rho, st);
    default: return GW_ERR_UNSUPPORTED;
  }
}

extern "C" int gw_density(gw_plan* h, int precision, const void* c, long ldc, long sc, int nb, const double* weights,
                          double* rho, void* ws, size_t ws_bytes, void* stream) {
  GW_REQUIRE(h && c && weights && rho && nb > 0, GW_ERR_ARG, "bad argument");
  if (int rc = require_box_fft(h->p, "the band density")) return rc;
  const size_t need = gw_density_workspace_bytes(h, nb, precision);
  GW_REQUIRE(ws && ws_bytes >= need, GW_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  if (precision == GW_FP64)
    return density_t<double>(h, (const double2*)c, ldc, sc, nb, weights, rho, (unsigned char*)ws,
                             (hipStream_t)stream);
  return density_t<float>(h, (const float2*)c, ldc, sc, nb, weights, rho, (unsigned char*)ws, (hipStream_t)stream);
}
