// Column kernels of the conduction-projected Sternheimer conjugate gradient
// (the solve of postscf/_response.py:502-690). The CG state of every (k, band)
// column lives in the padded (nk * nocc, ld) layout on the device; each launch
// works on a list of active columns only, one workgroup per listed column, and
// every inner product and norm is accumulated in fp64 whatever the vector
// precision. The Teter coefficient (solvers/precond.py:20-39) is formed in fp64.
#include "gw_internal.h"

namespace gw {
namespace {

__device__ __forceinline__ double wg_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

__device__ __forceinline__ double teter_coef(double t, double inv_tb) {
  const double x = t * inv_tb, x2 = x * x;
  const double num = 27.0 + 18.0 * x + 12.0 * x2 + 8.0 * x2 * x;
  return num / (num + 16.0 * x2 * x2);
}

// x_c += a p_c ; r_c -= a ap_c ; z_c = K(t_k / tb_c) r_c ; |r_c|^2 and Re<r_c, z_c>.
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_update(int npw, const int64_t* __restrict__ rows, const int64_t* __restrict__ kcol,
            const double* __restrict__ alpha, cpx<T>* __restrict__ x, const cpx<T>* __restrict__ p,
            cpx<T>* __restrict__ r, const cpx<T>* __restrict__ ap, long ld, const double* __restrict__ t,
            long ldt, const double* __restrict__ t_band, cpx<T>* __restrict__ z, double* __restrict__ rn2,
            double* __restrict__ rz) {
  __shared__ double sh[4];
  const int j = blockIdx.x;
  const long c = rows[j];
  const long off = c * ld;
  const double a = alpha[j];
  const T at = (T)a;
  const double* tk = t + kcol[j] * ldt;
  const double inv_tb = 1.0 / fmax(t_band[c], 1e-12);
  double nn = 0, rzz = 0;
  for (int g = threadIdx.x; g < npw; g += blockDim.x) {
    cpx<T> rv = r[off + g];
    if (a != 0.0) {
      const cpx<T> pv = p[off + g], av = ap[off + g];
      cpx<T> xv = x[off + g];
      xv.x += at * pv.x;
      xv.y += at * pv.y;
      x[off + g] = xv;
      rv.x -= at * av.x;
      rv.y -= at * av.y;
      r[off + g] = rv;
    }
    const T k = (T)teter_coef(tk[g], inv_tb);
    cpx<T> zv;
    zv.x = k * rv.x;
    zv.y = k * rv.y;
    z[off + g] = zv;
    nn += (double)rv.x * rv.x + (double)rv.y * rv.y;
    rzz += (double)rv.x * zv.x + (double)rv.y * zv.y;
  }
  nn = wg_sum(nn, sh);
  rzz = wg_sum(rzz, sh);
  if (threadIdx.x == 0) {
    rn2[j] = nn;
    rz[j] = rzz;
  }
}

// p_c = z_c + b p_c, and a copy of p_c into the compact row j of pc (when given).
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_direction(int npw, const int64_t* __restrict__ rows, const double* __restrict__ beta,
               const cpx<T>* __restrict__ z, cpx<T>* __restrict__ p, long ld, cpx<T>* __restrict__ pc, long ldc) {
  const int j = blockIdx.x;
  const long off = rows[j] * ld;
  const T b = (T)beta[j];
  cpx<T>* out = pc ? pc + (long)j * ldc : nullptr;
  for (int g = threadIdx.x; g < npw; g += blockDim.x) {
    const cpx<T> zv = z[off + g];
    cpx<T> pv = p[off + g];
    pv.x = zv.x + b * pv.x;
    pv.y = zv.y + b * pv.y;
    p[off + g] = pv;
    if (out) out[g] = pv;
  }
}

// out_j = Re sum_g conj(a_c) w_k[g] b_c  (w = 1 when null).
template <typename T>
__global__ void __launch_bounds__(256)
k_row_dots(int npw, const int64_t* __restrict__ rows, const cpx<T>* __restrict__ a, const cpx<T>* __restrict__ b,
           long ld, const double* __restrict__ w, long ldw, const int64_t* __restrict__ wrow,
           double* __restrict__ out) {
  __shared__ double sh[4];
  const int j = blockIdx.x;
  const long off = (rows ? rows[j] : (long)j) * ld;
  const double* wk = w ? w + (wrow ? wrow[j] : 0) * ldw : nullptr;
  double acc = 0;
  for (int g = threadIdx.x; g < npw; g += blockDim.x) {
    const cpx<T> av = a[off + g], bv = b[off + g];
    const double d = (double)av.x * bv.x + (double)av.y * bv.y;
    acc += wk ? wk[g] * d : d;
  }
  acc = wg_sum(acc, sh);
  if (threadIdx.x == 0) out[j] = acc;
}

}  // namespace
}  // namespace gw

using namespace gw;

extern "C" int gw_cg_update(int precision, int nrows, int npw, const int64_t* rows, const int64_t* kcol,
                            const double* alpha, void* x, const void* p, void* r, const void* ap, long ld,
                            const double* t, long ldt, const double* t_band, void* z, double* rn2, double* rz,
                            void* stream) {
  GW_REQUIRE(nrows >= 0 && npw >= 0 && ld >= npw && ldt >= npw, GW_ERR_ARG,
             "cg_update: bad sizes (nrows %d npw %d ld %ld ldt %ld)", nrows, npw, ld, ldt);
  if (!nrows) return GW_OK;
  GW_REQUIRE(rows && kcol && alpha && x && p && r && ap && t && t_band && z && rn2 && rz, GW_ERR_ARG,
             "cg_update: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_cg_update<double>, dim3(nrows), dim3(256), 0, st, npw, rows, kcol, alpha, (double2*)x,
                       (const double2*)p, (double2*)r, (const double2*)ap, ld, t, ldt, t_band, (double2*)z, rn2, rz);
  else
    hipLaunchKernelGGL(k_cg_update<float>, dim3(nrows), dim3(256), 0, st, npw, rows, kcol, alpha, (float2*)x,
                       (const float2*)p, (float2*)r, (const float2*)ap, ld, t, ldt, t_band, (float2*)z, rn2, rz);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_cg_direction(int precision, int nrows, int npw, const int64_t* rows, const double* beta,
                               const void* z, void* p, long ld, void* pc, long ldc, void* stream) {
  GW_REQUIRE(nrows >= 0 && npw >= 0 && ld >= npw && (!pc || ldc >= npw), GW_ERR_ARG,
             "cg_direction: bad sizes (nrows %d npw %d ld %ld ldc %ld)", nrows, npw, ld, ldc);
  if (!nrows) return GW_OK;
  GW_REQUIRE(rows && beta && z && p, GW_ERR_ARG, "cg_direction: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_cg_direction<double>, dim3(nrows), dim3(256), 0, st, npw, rows, beta, (const double2*)z,
                       (double2*)p, ld, (double2*)pc, ldc);
  else
    hipLaunchKernelGGL(k_cg_direction<float>, dim3(nrows), dim3(256), 0, st, npw, rows, beta, (const float2*)z,
                       (float2*)p, ld, (float2*)pc, ldc);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_row_dots(int precision, int nrows, int npw, const int64_t* rows, const void* a, const void* b,
                           long ld, const double* w, long ldw, const int64_t* wrow, double* out, void* stream) {
  GW_REQUIRE(nrows >= 0 && npw >= 0 && ld >= npw && (!w || ldw >= npw), GW_ERR_ARG,
             "row_dots: bad sizes (nrows %d npw %d ld %ld ldw %ld)", nrows, npw, ld, ldw);
  if (!nrows) return GW_OK;
  GW_REQUIRE(a && b && out, GW_ERR_ARG, "row_dots: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_row_dots<double>, dim3(nrows), dim3(256), 0, st, npw, rows, (const double2*)a,
                       (const double2*)b, ld, w, ldw, wrow, out);
  else
    hipLaunchKernelGGL(k_row_dots<float>, dim3(nrows), dim3(256), 0, st, npw, rows, (const float2*)a,
                       (const float2*)b, ld, w, ldw, wrow, out);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
