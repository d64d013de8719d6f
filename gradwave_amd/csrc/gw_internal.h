// Internal declarations shared by the HIP translation units of libgradwave_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gradwave_hip.h"
#include "gw_fft.h"

namespace gw {

void set_error(const char* fmt, ...);

#define GW_CHECK_HIP(expr)                                                          \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::gw::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
      return GW_ERR_HIP;                                                            \
    }                                                                               \
  } while (0)

#define GW_REQUIRE(cond, code, ...)      \
  do {                                   \
    if (!(cond)) {                       \
      ::gw::set_error(__VA_ARGS__);      \
      return (code);                     \
    }                                    \
  } while (0)

// One factorisation n = a*b with a, b in the codelet set (<= 16).
bool split_length(int n, int* a, int* b);
// n = f[0]*f[1]*f[2] with every factor in the codelet set: two factors (f[2] = 1)
// when split_length finds them, else three (7-smooth lengths up to 16^3).
bool split_length3(int n, int f[3]);
inline bool codelet_ok(int n) {
  return n == 1 || n == 2 || n == 3 || n == 4 || n == 5 || n == 6 || n == 7 || n == 8 ||
         n == 9 || n == 10 || n == 12 || n == 14 || n == 15 || n == 16;
}

// Geometry + tables of one batched H|psi> operator (all device pointers owned).
struct Plan {
  int n[3];       // box dims (n1, n2, n3); flat = (i0*n2 + i1)*n3 + i2
  int f[3][3];    // per axis n = f[a][0] * f[a][1] * f[a][2] (f[a][2] = 1: two factors)
  int nk, npw_max, na, nst_max;
  int* planes;        // (na) i0 of each active plane slot (union over k)
  int* st_plane_ptr;  // (nk, na+1) stick range of each plane slot
  int* st_i1;         // (nk, nst_max) i1 of each stick
  int* st_pt_ptr;     // (nk, nst_max+1) point range of each stick
  int* pt_src;        // (nk, npw_max) sphere index of each point (stick order)
  int* pt_yz;         // (nk, npw_max) i1*n3 + i2 of each point
  int* npw;           // (nk)
  int* st_count;      // host copy: sticks per k
  int st_max_plane;   // max sticks of one (k, plane)
  int max_pts_plane;  // max points of one (k, plane)
  double2* tw64[3];
  float2* tw32[3];
  // kinetic diagonal and permuted local potential, both precisions
  double* t64;
  float* t32;
  double* vp64;
  float* vp32;
  double* vx64;     // V with natural x order, permuted y/z (register x pass)
  double* vz64;     // V z-major at the permuted positions (pencil slabs)
  float* vz32;
  float* vx32;
  int* slot_of;     // (n1) active-plane slot of each x index, -1 if inactive
  int xlo, xhi;     // active planes = [0, xlo) U [xhi, n1) when xlo >= 0
  bool have_v;
  // 0, or the first box length with no codelet split: the plan then serves the
  // kinetic and nonlocal terms and the sphere<->box maps, not the box FFTs
  int nosplit_len;
  // nonlocal: projectors (nk, nproj, npw_max) both precisions, D as CSR by column
  int nproj;
  double2* p64;
  float2* p32;
  int* d_colptr;  // (nproj+1)
  int* d_row;     // nnz
  double* d_val;  // nnz
  int d_nnz;
  int flags;     // GW_PLAN_* bits
  int lds_max;   // dynamic LDS bytes available per workgroup
};

// Per-stage HIP-event timing of gw_hpsi_apply (enabled by GW_PLAN_TIMING).
enum Stage { ST_PLANE_INV = 0, ST_XPASS, ST_PLANE_FWD, ST_BECP, ST_DMUL, ST_BACKPROJ, ST_COUNT };
struct StageTimer {
  bool on = false;
  hipStream_t st = nullptr;
  struct Pending { int stage; hipEvent_t b, e; bool count; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double ms[ST_COUNT] = {0};
  long calls[ST_COUNT] = {0};
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e; (void)hipEventCreate(&e); return e;
  }
  hipEvent_t begin() { if (!on) return nullptr; hipEvent_t e = get(); (void)hipEventRecord(e, st); return e; }
  // With count_calls false, end() adds the time to the stage without counting a
  // call (the later band chunks of one apply).
  bool count_calls = true;
  void end(int stage, hipEvent_t b) {
    if (!on) return;
    hipEvent_t e = get(); (void)hipEventRecord(e, st);
    pending.push_back({stage, b, e, count_calls});
  }
  void collect() {
    for (auto& p : pending) {
      float t = 0; (void)hipEventSynchronize(p.e);
      (void)hipEventElapsedTime(&t, p.b, p.e);
      ms[p.stage] += t; calls[p.stage] += p.count ? 1 : 0;
      pool.push_back(p.b); pool.push_back(p.e);
    }
    pending.clear();
  }
};

template <typename T> inline const cpx<T>* tw_of(const Plan* p, int axis);
template <> inline const double2* tw_of<double>(const Plan* p, int axis) { return p->tw64[axis]; }
template <> inline const float2* tw_of<float>(const Plan* p, int axis) { return p->tw32[axis]; }

// Workspace layout of one apply (bytes), shared by the size query and the apply.
struct ApplyLayout {
  size_t box_off, box_bytes;      // (nbt, na, n2, n3) complex
  size_t stick_off, stick_bytes;  // (nbt, nst_max, n3) complex (unfused path only)
  size_t becp_off, becp_bytes;    // (nk, nproj, nb) complex
  size_t u_off, u_bytes;          // (nk, nproj, nb) complex
  size_t part_off, part_bytes;    // split-K partials of the projection GEMM
  size_t total;
  bool fused;
  int pencil_yc;  // > 0: the pencil path, y-lines per x chunk of its slab kernel
};
// chunked: the layout of gw_hpsi_apply, whose box holds one band chunk (GW_TUNE_BAND_CHUNK)
ApplyLayout apply_layout(const Plan* p, int nb, int precision, bool chunked = false);
bool plane_fits_lds(const Plan* p, int precision);
void set_gemm_target_wgs(int v);
int gemm_target_wgs();
void set_gemm_tile(int precision, int op, int v);  // op 0 inner, 1 combine, other: both
int gemm_tile(int precision, int op);
void set_combine_lib(int precision, int v);
int combine_lib(int precision);
void set_inner_lib(int precision, int v);
int inner_lib(int precision);
// While set on the calling thread, contractions never go to rocBLAS; returns the previous value.
int set_kernels_only(int v);

// Grouped launches (the cells of a ragged batch): one launch per contraction over
// the concatenated work items of many independent jobs; `first` is each job's
// first work item (strictly increasing), workgroups find their job by search.
struct GroupInner {  // C_partials = X^H Y per K-split (the kernel path of gemm_inner)
  const void* X; const void* Y; void* part;
  long ldx, sx, ldy, sy, first;
  int M, N, K, splits, kchunk, ntx, nty;
};
struct GroupReduce {  // C = sum of the splits' partials; nblk blocks of 64 reals per k
  const void* part; void* C;
  long ldc, sc, first;
  int M, N, splits, nblk;
};
struct GroupCombine {  // Z = beta Z + U^T Y
  const void* U; const void* Y; void* Z;
  long ldu, su, ldy, sy, ldz, sz, first;
  int A, G, J, ntx, nty;
};
// split count and K-chunk of gemm_inner's kernel path for this shape
void gemm_inner_shape(int nk, int M, int N, int K, int* splits, int* kchunk);
template <typename T>
int gemm_group_inner(const GroupInner* jobs, int n, long total, hipStream_t st);
template <typename T>
int gemm_group_reduce(const GroupReduce* jobs, int n, long total, hipStream_t st);
template <typename T>
int gemm_group_combine(const GroupCombine* jobs, int n, long total, T beta, hipStream_t st);

// Complex GEMM launchers (gemm.hip). All matrices complex interleaved, row-major.
//   inner:   C[k](M x N) = alpha * X[k]^H-rows . Y[k]-rows  i.e. C[i][j] = sum_g conj(X[i][g]) Y[j][g]
//   combine: Z[k](A x G) = beta * Z + sum_j U[k][j][a] * Y[k][j][g]   (U optionally conjugated)
template <typename T>
int gemm_inner(int nk, int M, int N, int K, const cpx<T>* X, long ldx, long sx, const cpx<T>* Y,
               long ldy, long sy, cpx<T>* C, long ldc, long sc, void* partials, size_t part_bytes,
               hipStream_t st);
template <typename T>
size_t gemm_inner_partials_bytes(int nk, int M, int N, int K);
template <typename T>
int gemm_combine(int nk, int A, int G, int J, const cpx<T>* U, long ldu, long su, int conj_u,
                 const cpx<T>* Y, long ldy, long sy, cpx<T>* Z, long ldz, long sz, T beta,
                 hipStream_t st);

}  // namespace gw
