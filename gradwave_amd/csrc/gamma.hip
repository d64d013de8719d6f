// Gamma-point real layout (core/gamma.py:1-290) on the complex apply.
//
// At k = 0 a wavefunction with c(-G) = conj(c(G)) is real in r, and H maps such
// vectors to such vectors (t(G) = t(-G), real V(r), projectors with
// p(-G) = conj(p(G)), real D). Two of them, a and b, therefore share one complex
// apply: z = full(a) + i full(b) gives H z = H full(a) + i H full(b), and on the
// half sphere
//   (H a)(G) = [Hz(G) + conj(Hz(-G))] / 2,   (H b)(G) = [Hz(G) - conj(Hz(-G))] / (2i).
// Half-sphere vectors travel in the reference's real embedding (core/gamma.py:242-280):
//   phi[0] = Re c(0), phi[2h-1] = sqrt2 Re c(h), phi[2h] = sqrt2 Im c(h)   (h >= 1),
// where the metric Re c0* d0 + 2 sum_{h>=1} Re c* d is the plain dot product.
#include "gw_internal.h"

namespace gw {

template <typename T>
__device__ __forceinline__ void half_value(const T* __restrict__ phi, int h, T& re, T& im) {
  constexpr T r = T(0.70710678118654752440);
  if (h == 0) {
    re = phi[0];
    im = T(0);
  } else {
    re = r * phi[2 * h - 1];
    im = r * phi[2 * h];
  }
}

// grid (ceil(npw_max / 256), ceil(nb / 2)); z row j holds bands 2j and 2j+1
template <typename T>
__global__ void k_gamma_pack(int nb, int npw, int npw_max, const T* __restrict__ phi, long ldp,
                             const int* __restrict__ src, const unsigned char* __restrict__ cj,
                             cpx<T>* __restrict__ z, long ldz) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (g >= npw_max) return;
  cpx<T> v;
  v.x = T(0);
  v.y = T(0);
  if (g < npw) {
    const int h = src[g];
    const T s = cj[g] ? T(-1) : T(1);
    T ra, ia, rb = T(0), ib = T(0);
    half_value(phi + (long)(2 * j) * ldp, h, ra, ia);
    if (2 * j + 1 < nb) half_value(phi + (long)(2 * j + 1) * ldp, h, rb, ib);
    v.x = ra - s * ib;
    v.y = s * ia + rb;
  }
  z[(long)j * ldz + g] = v;
}

// grid (ceil(nhalf / 256), ceil(nb / 2))
template <typename T>
__global__ void k_gamma_unpack(int nb, int nhalf, const cpx<T>* __restrict__ hz, long ldz,
                               const int* __restrict__ rep, const int* __restrict__ partner,
                               T* __restrict__ out, long ldo) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (h >= nhalf) return;
  const int g = rep[h];
  const cpx<T> a = hz[(long)j * ldz + g];
  const cpx<T> b = hz[(long)j * ldz + partner[g]];
  // A = (a + conj b) / 2 ; B = (a - conj b) / (2i)
  const T ax = T(0.5) * (a.x + b.x), ay = T(0.5) * (a.y - b.y);
  const T bx = T(0.5) * (a.y + b.y), by = -T(0.5) * (a.x - b.x);
  constexpr T r2 = T(1.41421356237309504880);
  T* oa = out + (long)(2 * j) * ldo;
  T* ob = out + (long)(2 * j + 1) * ldo;
  const bool has_b = 2 * j + 1 < nb;
  if (h == 0) {
    oa[0] = ax;
    if (has_b) ob[0] = bx;
  } else {
    oa[2 * h - 1] = r2 * ax;
    oa[2 * h] = r2 * ay;
    if (has_b) {
      ob[2 * h - 1] = r2 * bx;
      ob[2 * h] = r2 * by;
    }
  }
}

}  // namespace gw

using namespace gw;

extern "C" int gw_gamma_pack(int precision, int nb, int npw, int npw_max, const void* phi, long ldp,
                             const int* src_half, const unsigned char* conj_full, void* z, long ldz,
                             void* stream) {
  GW_REQUIRE(phi && src_half && conj_full && z && nb > 0 && npw > 0 && npw <= npw_max && ldz >= npw_max &&
                 ldp >= npw,
             GW_ERR_ARG, "bad argument");
  const int nbh = (nb + 1) / 2;
  GW_REQUIRE(nbh <= 65535, GW_ERR_UNSUPPORTED, "too many bands for one launch");
  const dim3 grid((npw_max + 255) / 256, nbh);
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_gamma_pack<double>, grid, dim3(256), 0, st, nb, npw, npw_max, (const double*)phi, ldp,
                       src_half, conj_full, (double2*)z, ldz);
  else
    hipLaunchKernelGGL(k_gamma_pack<float>, grid, dim3(256), 0, st, nb, npw, npw_max, (const float*)phi, ldp,
                       src_half, conj_full, (float2*)z, ldz);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_gamma_unpack(int precision, int nb, int nhalf, const void* hz, long ldz, const int* rep_full,
                               const int* partner, void* out, long ldo, void* stream) {
  GW_REQUIRE(hz && rep_full && partner && out && nb > 0 && nhalf > 0 && ldo >= 2 * nhalf - 1, GW_ERR_ARG,
             "bad argument");
  const int nbh = (nb + 1) / 2;
  GW_REQUIRE(nbh <= 65535, GW_ERR_UNSUPPORTED, "too many bands for one launch");
  const dim3 grid((nhalf + 255) / 256, nbh);
  hipStream_t st = (hipStream_t)stream;
  if (precision == GW_FP64)
    hipLaunchKernelGGL(k_gamma_unpack<double>, grid, dim3(256), 0, st, nb, nhalf, (const double2*)hz, ldz,
                       rep_full, partner, (double*)out, ldo);
  else
    hipLaunchKernelGGL(k_gamma_unpack<float>, grid, dim3(256), 0, st, nb, nhalf, (const float2*)hz, ldz,
                       rep_full, partner, (float*)out, ldo);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
