// Position dependence of H and its adjoints.
//
// Projector phases, forward (core/batch.py:229-235, core/hamiltonian.py:108-112):
//   p[k][q][g] = pf[k][q][g] * exp(-i (k+G)_g . tau_{a(q)})
// Adjoint (torch's complex-gradient convention, dL = Re sum conj(grad) dz):
//   dp = -i ((k+G) . dtau) p  =>  grad_tau[a][x] = sum_{k, q in a, g} Im[conj(grad_p) p] (k+G)_x
// The row reduction writes per-(k, q) partial sums; a second kernel sums the
// rows of each atom in a fixed order (deterministic).
//
// Local potential through the structure factor (core/energies/local_pp.py:22-38):
//   V(G) = (1/Omega) sum_a v_{s(a)}(G) exp(-i G . tau_a)       on the full box
//   grad_tau[a][x] = (1/Omega) sum_G v_{s(a)}(G) G_x Im[exp(-i G . tau_a) conj(grad_V(G))]
#include "gw_internal.h"

namespace gw {

__global__ void k_proj_phase(int nproj, int npw_max, const int* __restrict__ npw, const double2* __restrict__ pf,
                             const double* __restrict__ kpg, const double* __restrict__ pos,
                             const int64_t* __restrict__ atom, double2* __restrict__ p) {
  const int row = blockIdx.y;  // k*nproj + q
  const int k = row / nproj, q = row - k * nproj;
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= npw_max) return;
  const long o = (long)row * npw_max + gi;
  double2 v = make_double2(0.0, 0.0);
  if (gi < npw[k]) {
    const double* kg = kpg + ((long)k * npw_max + gi) * 3;
    const double* t = pos + atom[q] * 3;
    const double arg = kg[0] * t[0] + kg[1] * t[1] + kg[2] * t[2];
    double s, c;
    sincos(arg, &s, &c);
    const double2 f = pf[o];
    // f * (c - i s)
    v.x = f.x * c + f.y * s;
    v.y = f.y * c - f.x * s;
  }
  p[o] = v;
}

__global__ void __launch_bounds__(256)
k_proj_phase_grad_rows(int nproj, int npw_max, const double2* __restrict__ gp, const double2* __restrict__ p,
                       const double* __restrict__ kpg, double* __restrict__ rows) {
  __shared__ double sh[3][4];
  const int row = blockIdx.x;
  const int k = row / nproj;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int gi = threadIdx.x; gi < npw_max; gi += blockDim.x) {
    const long o = (long)row * npw_max + gi;
    const double2 a = gp[o], b = p[o];
    const double im = a.x * b.y - a.y * b.x;  // Im(conj(a) b)
    const double* kg = kpg + ((long)k * npw_max + gi) * 3;
    a0 += im * kg[0];
    a1 += im * kg[1];
    a2 += im * kg[2];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_down(a0, off, 64);
    a1 += __shfl_down(a1, off, 64);
    a2 += __shfl_down(a2, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sh[0][w] = a0; sh[1][w] = a1; sh[2][w] = a2; }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[threadIdx.x][i];
    rows[(long)row * 3 + threadIdx.x] = s;
  }
}

__global__ void k_proj_phase_grad_atoms(int nk, int nproj, int na, const int64_t* __restrict__ atom,
                                        const double* __restrict__ rows, double* __restrict__ gtau) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // a*3 + x
  if (e >= na * 3) return;
  const int a = e / 3, x = e - a * 3;
  double s = 0;
  for (int k = 0; k < nk; ++k)
    for (int q = 0; q < nproj; ++q)
      if (atom[q] == a) s += rows[((long)k * nproj + q) * 3 + x];
  gtau[e] += s;
}

// One thread per box point G; every lane reads the same atom at a time.
__global__ void k_vloc_g(int na, long npts, const double* __restrict__ pos, const int64_t* __restrict__ sidx,
                         const double* __restrict__ tab, const double* __restrict__ gc, double inv_vol,
                         double2* __restrict__ vg) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= npts) return;
  const double g0 = gc[e * 3], g1 = gc[e * 3 + 1], g2 = gc[e * 3 + 2];
  double re = 0, im = 0;
  for (int a = 0; a < na; ++a) {
    const double arg = g0 * pos[a * 3] + g1 * pos[a * 3 + 1] + g2 * pos[a * 3 + 2];
    double s, c;
    sincos(arg, &s, &c);
    const double v = tab[sidx[a] * npts + e];
    re += v * c;
    im -= v * s;
  }
  vg[e] = make_double2(re * inv_vol, im * inv_vol);
}

constexpr int kVlocChunks = 64;  // G-chunks per atom in the structure-factor adjoint

// grid (kVlocChunks, na): partial[a][chunk][x]
__global__ void __launch_bounds__(256)
k_vloc_g_grad_part(long npts, const double* __restrict__ pos, const int64_t* __restrict__ sidx,
                   const double* __restrict__ tab, const double* __restrict__ gc, const double2* __restrict__ gv,
                   double* __restrict__ part) {
  __shared__ double sh[3][4];
  const int a = blockIdx.y;
  const long per = (npts + gridDim.x - 1) / gridDim.x;
  const long e0 = (long)blockIdx.x * per, e1 = min(npts, e0 + per);
  const double t0 = pos[a * 3], t1 = pos[a * 3 + 1], t2 = pos[a * 3 + 2];
  const double* tb = tab + sidx[a] * npts;
  double a0 = 0, a1 = 0, a2 = 0;
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const double g0 = gc[e * 3], g1 = gc[e * 3 + 1], g2 = gc[e * 3 + 2];
    double s, c;
    sincos(g0 * t0 + g1 * t1 + g2 * t2, &s, &c);
    const double2 f = gv[e];
    // S = c - i s;  Im[S conj(f)] = -s f.x - c f.y
    const double w = tb[e] * (-s * f.x - c * f.y);
    a0 += w * g0;
    a1 += w * g1;
    a2 += w * g2;
  }
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_down(a0, off, 64);
    a1 += __shfl_down(a1, off, 64);
    a2 += __shfl_down(a2, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sh[0][w] = a0; sh[1][w] = a1; sh[2][w] = a2; }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[threadIdx.x][i];
    part[((long)a * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = s;
  }
}

__global__ void k_vloc_g_grad_sum(int na, int nchunk, double inv_vol, const double* __restrict__ part,
                                  double* __restrict__ gtau) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // a*3 + x
  if (e >= na * 3) return;
  const int a = e / 3, x = e - a * 3;
  double s = 0;
  for (int i = 0; i < nchunk; ++i) s += part[((long)a * nchunk + i) * 3 + x];
  gtau[e] += s * inv_vol;
}

}  // namespace gw

using namespace gw;

extern "C" int gw_proj_phase(int nk, int nproj, int npw_max, const int* npw, const void* pf, const double* kpg,
                             const double* positions, const int64_t* atom_index, void* p, void* stream) {
  GW_REQUIRE(pf && kpg && positions && atom_index && p && npw && nk > 0 && nproj > 0 && npw_max > 0, GW_ERR_ARG,
             "bad argument");
  GW_REQUIRE((long)nk * nproj <= 65535, GW_ERR_UNSUPPORTED, "nk*nproj = %ld exceeds one launch", (long)nk * nproj);
  hipLaunchKernelGGL(k_proj_phase, dim3((npw_max + 255) / 256, nk * nproj), dim3(256), 0, (hipStream_t)stream,
                     nproj, npw_max, npw, (const double2*)pf, kpg, positions, atom_index, (double2*)p);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" size_t gw_proj_phase_grad_workspace_bytes(int nk, int nproj) {
  return (size_t)nk * nproj * 3 * sizeof(double);
}

extern "C" int gw_proj_phase_grad(int nk, int nproj, int npw_max, int na, const void* grad_p, const void* p,
                                  const double* kpg, const int64_t* atom_index, double* grad_tau, void* ws,
                                  size_t ws_bytes, void* stream) {
  GW_REQUIRE(grad_p && p && kpg && atom_index && grad_tau && nk > 0 && nproj > 0 && na > 0, GW_ERR_ARG,
             "bad argument");
  GW_REQUIRE(ws && ws_bytes >= gw_proj_phase_grad_workspace_bytes(nk, nproj), GW_ERR_WORKSPACE,
             "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_proj_phase_grad_rows, dim3(nk * nproj), dim3(256), 0, st, nproj, npw_max,
                     (const double2*)grad_p, (const double2*)p, kpg, (double*)ws);
  hipLaunchKernelGGL(k_proj_phase_grad_atoms, dim3((na * 3 + 63) / 64), dim3(64), 0, st, nk, nproj, na, atom_index,
                     (const double*)ws, grad_tau);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" int gw_vloc_g(int na, long npts, const double* positions, const int64_t* species_index,
                         const double* tables, const double* g_cart, double volume, void* vloc_g, void* stream) {
  GW_REQUIRE(positions && species_index && tables && g_cart && vloc_g && na > 0 && npts > 0 && volume > 0,
             GW_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_vloc_g, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream, na, npts,
                     positions, species_index, tables, g_cart, 1.0 / volume, (double2*)vloc_g);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}

extern "C" size_t gw_vloc_g_grad_workspace_bytes(int na) { return (size_t)na * kVlocChunks * 3 * sizeof(double); }

extern "C" int gw_vloc_g_grad(int na, long npts, const double* positions, const int64_t* species_index,
                              const double* tables, const double* g_cart, double volume, const void* grad_vloc_g,
                              double* grad_tau, void* ws, size_t ws_bytes, void* stream) {
  GW_REQUIRE(positions && species_index && tables && g_cart && grad_vloc_g && grad_tau && na > 0 && npts > 0 &&
                 volume > 0,
             GW_ERR_ARG, "bad argument");
  GW_REQUIRE(ws && ws_bytes >= gw_vloc_g_grad_workspace_bytes(na), GW_ERR_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_vloc_g_grad_part, dim3(kVlocChunks, na), dim3(256), 0, st, npts, positions, species_index,
                     tables, g_cart, (const double2*)grad_vloc_g, (double*)ws);
  hipLaunchKernelGGL(k_vloc_g_grad_sum, dim3((na * 3 + 63) / 64), dim3(64), 0, st, na, kVlocChunks, 1.0 / volume,
                     (const double*)ws, grad_tau);
  GW_CHECK_HIP(hipGetLastError());
  return GW_OK;
}
