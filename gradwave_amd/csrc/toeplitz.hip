// Small-cell local term as a dense matrix (core/batch.py:333-464):
// on the wavefunction sphere V(r) psi(r) is exactly the convolution
//   out(G_i) = sum_j Vhat(G_i - G_j) c(G_j),   Vhat = fftn(V) / N,
// so the local term is one complex GEMM per k with M[i][j] = Vhat(G_i - G_j).
// M is generated straight from Vhat and the sphere's box indices (no npw^2
// index table in HBM); padded rows and columns are zero.
#include "gw_internal.h"

namespace gw {

// grid (ceil(npw_max / 256), npw_max, nk)
template <typename T>
__global__ void k_toeplitz_build(int npw_max, int n1, int n2, int n3, const int* __restrict__ npw,
                                 const int* __restrict__ flat, const cpx<T>* __restrict__ vhat,
                                 cpx<T>* __restrict__ M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y, k = blockIdx.z;
  if (j >= npw_max) return;
  const int n = npw[k];
  cpx<T> v;
  v.x = T(0);
  v.y = T(0);
  if (i < n && j < n) {
    const int n23 = n2 * n3;
    const int fi = flat[(long)k * npw_max + i], fj = flat[(long)k * npw_max + j];
    const int i0 = fi / n23, i1 = (fi / n3) % n2, i2 = fi % n3;
    const int j0 = fj / n23, j1 = (fj / n3) % n2, j2 = fj % n3;
    const int d0 = (i0 - j0 + n1) % n1, d1 = (i1 - j1 + n2) % n2, d2 = (i2 - j2 + n3) % n3;
    v = vhat[(long)d0 * n23 + d1 * n3 + d2];
  }
  M[((long)k * npw_max + i) * npw_max + j] = v;
}

}  // namespace gw

using namespace gw;

extern "C" int gw_toeplitz_build(int precision, int nk, int npw_max, int n1, int n2, int n3, const int* npw,
                                 const int* flat_idx, const void* vhat, void* M, void* stream) {
  GW_REQUIRE(npw && flat_idx && vhat && M && nk > 0 && npw_max > 0 && n1 > 0 && n2 > 0 && n3 > 0, GW_ERR_ARG,
             "bad argument");
  GW_REQUIRE(npw_max <= 65535 && nk <= 65535, GW_ERR_UNSUPPORTED, "npw_max or nk too large for one launch");
  const dim3 grid((npw_max + 255) / 256, npw_max, nk);
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_toeplitz_build<double>, grid, dim3(256), 0, st, npw_max, n1, n2, n3, npw, flat_idx,
                       (const double2*)vhat, (double2*)M);
  else
    hipLaunchKernelGGL(k_toeplitz_build<float>, grid, dim3(256), 0, st, npw_max, n1, n2, n3, npw, flat_idx,
                       (const float2*)vhat, (float2*)M);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
