// Band-vector kernels of the batched block Davidson (solvers/davidson.py:656-743,
// solvers/precond.py:20-39). One workgroup per (k, band) row; every norm and
// the Teter coefficient are accumulated in fp64 whatever the vector precision
// (the reference forms the coefficient in fp64 to avoid inf/inf in fp32,
// precond.py:26-33, and keeps convergence decisions in fp64).
#include "gw_internal.h"

namespace gw {

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    sh[16] = s;
  }
  __syncthreads();
  s = sh[16];
  __syncthreads();
  return s;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_residual(int nb, int npw, const cpx<T>* __restrict__ x, const cpx<T>* __restrict__ hx, long ld, long s,
           const double* __restrict__ eig, cpx<T>* __restrict__ r, double* __restrict__ rn) {
  __shared__ double sh[17];
  const int k = blockIdx.y, b = blockIdx.x;
  const long off = (long)k * s + (long)b * ld;
  const T e = (T)eig[(long)k * nb + b];
  double acc = 0;
  for (int g = threadIdx.x; g < npw; g += blockDim.x) {
    const cpx<T> xv = x[off + g], hv = hx[off + g];
    cpx<T> rv;
    rv.x = hv.x - e * xv.x;
    rv.y = hv.y - e * xv.y;
    r[off + g] = rv;
    acc += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) rn[(long)k * nb + b] = sqrt(acc);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_teter_expand(int nb, int nsel, int npw, const cpx<T>* __restrict__ x, const cpx<T>* __restrict__ r, long ld,
               long s, const double* __restrict__ t, long ldt, const int64_t* __restrict__ sel,
               cpx<T>* __restrict__ d, long ldd, long sd) {
  __shared__ double sh[17];
  const int k = blockIdx.y, j = blockIdx.x;
  const int b = (int)sel[(long)k * nsel + j];
  const long off = (long)k * s + (long)b * ld;
  const double* tk = t + (long)k * ldt;
  double tb = 0;
  for (int g = threadIdx.x; g < npw; g += blockDim.x) {
    const cpx<T> xv = x[off + g];
    tb += tk[g] * ((double)xv.x * xv.x + (double)xv.y * xv.y);
  }
  tb = block_sum(tb, sh);
  const double inv = 1.0 / fmax(tb, 1e-12);
  cpx<T>* dk = d + (long)k * sd + (long)j * ldd;
  double nrm = 0;
  for (int g = threadIdx.x; g < npw; g += blockDim.x) {
    const double xx = tk[g] * inv;
    const double x2 = xx * xx;
    const double num = 27.0 + 18.0 * xx + 12.0 * x2 + 8.0 * x2 * xx;
    const T coef = (T)(num / (num + 16.0 * x2 * x2));
    const cpx<T> rv = r[off + g];
    cpx<T> dv;
    dv.x = rv.x * coef;
    dv.y = rv.y * coef;
    dk[g] = dv;
    nrm += (double)dv.x * dv.x + (double)dv.y * dv.y;
  }
  nrm = sqrt(block_sum(nrm, sh));
  if (nrm > 1e-300) {
    const T sc = (T)(1.0 / nrm);
    for (int g = threadIdx.x; g < npw; g += blockDim.x) {
      cpx<T> dv = dk[g];
      dv.x *= sc;
      dv.y *= sc;
      dk[g] = dv;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_row_norms(int m, int npw, const cpx<T>* __restrict__ v, long ld, long s, double* __restrict__ out) {
  __shared__ double sh[17];
  const int k = blockIdx.y, i = blockIdx.x;
  const cpx<T>* row = v + (long)k * s + (long)i * ld;
  double acc = 0;
  for (int g = threadIdx.x; g < npw; g += blockDim.x) {
    const cpx<T> a = row[g];
    acc += (double)a.x * a.x + (double)a.y * a.y;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) out[(long)k * m + i] = sqrt(acc);
}

}  // namespace gw

using namespace gw;

extern "C" int gw_residual(int precision, int nk, int nb, int npw, const void* x, const void* hx, long ld, long s,
                           const double* eig, void* r, double* rn, void* stream) {
  GW_REQUIRE(x && hx && eig && r && rn && nk >= 0 && nb >= 0 && npw >= 0 && ld >= npw, GW_ERR_ARG, "bad argument");
  if (!nk || !nb) return GW_OK;
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_residual<double>, dim3(nb, nk), dim3(256), 0, st, nb, npw, (const double2*)x,
                       (const double2*)hx, ld, s, eig, (double2*)r, rn);
  else
    hipLaunchKernelGGL(k_residual<float>, dim3(nb, nk), dim3(256), 0, st, nb, npw, (const float2*)x,
                       (const float2*)hx, ld, s, eig, (float2*)r, rn);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_teter_expand(int precision, int nk, int nb, int nsel, int npw, const void* x, const void* r,
                               long ld, long s, const double* t, long ldt, const int64_t* sel, void* d, long ldd,
                               long sd, void* stream) {
  GW_REQUIRE(x && r && t && sel && d, GW_ERR_ARG, "teter_expand: null pointer (x %p r %p t %p sel %p d %p)", x, r,
             t, sel, d);
  GW_REQUIRE(ld >= npw && ldd >= npw && ldt >= npw, GW_ERR_ARG,
             "teter_expand: row stride below npw (ld %ld ldd %ld ldt %ld npw %d)", ld, ldd, ldt, npw);
  if (!nk || !nsel) return GW_OK;
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_teter_expand<double>, dim3(nsel, nk), dim3(256), 0, st, nb, nsel, npw,
                       (const double2*)x, (const double2*)r, ld, s, t, ldt, sel, (double2*)d, ldd, sd);
  else
    hipLaunchKernelGGL(k_teter_expand<float>, dim3(nsel, nk), dim3(256), 0, st, nb, nsel, npw, (const float2*)x,
                       (const float2*)r, ld, s, t, ldt, sel, (float2*)d, ldd, sd);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_row_norms(int precision, int nk, int m, int npw, const void* v, long ld, long s, double* out,
                            void* stream) {
  GW_REQUIRE(v && out && ld >= npw, GW_ERR_ARG, "bad argument");
  if (!nk || !m) return GW_OK;
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_row_norms<double>, dim3(m, nk), dim3(256), 0, st, m, npw, (const double2*)v, ld, s, out);
  else
    hipLaunchKernelGGL(k_row_norms<float>, dim3(m, nk), dim3(256), 0, st, m, npw, (const float2*)v, ld, s, out);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
