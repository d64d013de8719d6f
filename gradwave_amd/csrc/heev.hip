// Batched Hermitian eigensolver for the small subspace matrices of Davidson:
// the lowest `nev` eigenpairs of many complex128 Hermitian matrices of
// different sizes (n <= kHeevMaxN) in ONE launch, one workgroup per matrix.
//
// Reference contract: the subspace eigensolve of davidson_batched
// (solvers/davidson.py:640-652, torch.linalg.eigh of the (nk, m, m) Hermitised
// subspace matrix; only the first nb Ritz pairs are used, davidson.py:651-652).
//
// Each workgroup:
//   1. copies A (row-major, Hermitised as (A + A^H)/2) into its workspace in
//      column-major order;
//   2. reduces it to a real symmetric tridiagonal T = Q^H A Q by Householder
//      reflectors H_j = I - tau_j v_j v_j^H (v_j[0] = 1) chosen so that
//      H_j^H x = beta e_1 with beta real (beta = -sign(Re a) |x|,
//      tau = (beta - a) / beta, v = x / (a - beta)); the trailing block is
//      updated as A -= v w^H + w v^H with w = y - (tau/2)(y^H v) v, y = tau A v;
//   3. finds the nev lowest eigenvalues of T by multisection on Sturm counts
//      (three interior points per step, one per thread and eigenvalue);
//   4. computes T's eigenvectors by inverse iteration (LU of T - lambda I with
//      partial pivoting, tiny pivots raised to eps |T|, three steps from a
//      hashed start vector), orthogonalising each vector against the earlier
//      members of its cluster (eigenvalues closer than 1e-3 |T|) in every step;
//      clusters are processed member by member, all clusters at once;
//   5. back-transforms u = H_0 H_1 ... H_{n-2} z into the caller's output.
// Every loop has a fixed bound (n, nev, 64 multisection steps, 3 inverse
// iteration steps), so the kernel ends for any input, NaN included.
#include <cfloat>

#include "gw_internal.h"

namespace gw {

constexpr int kHeevThreads = 1024;
constexpr int kHeevMaxN = 512;
constexpr int kHeevPerLaunch = 32;

struct HeevArgs {
  gw_heev_job j[kHeevPerLaunch];
  int count;
};

__host__ __device__ static size_t heev_align(size_t x) { return (x + 255) & ~size_t(255); }

// workspace: W (n*n complex) | lu (5 n nev doubles) | zr (n nev doubles)
struct HeevWs {
  size_t w_off, lu_off, zr_off, total;
};
__host__ __device__ static HeevWs heev_ws(int n, int nev) {
  HeevWs L;
  L.w_off = 0;
  L.lu_off = heev_align((size_t)n * n * sizeof(double2));
  L.zr_off = L.lu_off + heev_align((size_t)5 * n * nev * sizeof(double));
  L.total = L.zr_off + heev_align((size_t)n * nev * sizeof(double));
  return L;
}

__device__ __forceinline__ double2 hc_mul(double2 a, double2 b) {
  return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ double2 hc_cmul(double2 a, double2 b) {  // conj(a) b
  return double2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}

// Block reductions over the waves; every thread gets the result. Two barriers.
constexpr int kHeevWaves = kHeevThreads / 64;
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < kHeevWaves; ++i) r += red[i];
  __syncthreads();
  return r;
}
__device__ double2 block_sum2(double2 v, double* red) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[w] = v.x; red[kHeevWaves + w] = v.y; }
  __syncthreads();
  double2 r{0.0, 0.0};
#pragma unroll
  for (int i = 0; i < kHeevWaves; ++i) { r.x += red[i]; r.y += red[kHeevWaves + i]; }
  __syncthreads();
  return r;
}
__device__ double block_max(double v, double* red) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < kHeevWaves; ++i) r = fmax(r, red[i]);
  __syncthreads();
  return r;
}
__device__ double block_min(double v, double* red) {
  v = wave_min(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < kHeevWaves; ++i) r = fmin(r, red[i]);
  __syncthreads();
  return r;
}

// a -= vi conj(wk) + wi conj(vk)
__device__ __forceinline__ void rank2_sub(double2& a, double2 vi, double2 wi, double2 vk, double2 wk) {
  a.x -= vi.x * wk.x + vi.y * wk.y + wi.x * vk.x + wi.y * vk.y;
  a.y -= vi.y * wk.x - vi.x * wk.y + wi.y * vk.x - wi.x * vk.y;
}

// Reflector for x = src[0..L-1] (LDS): H^H x = beta e_1 with H = I - tau v v^H,
// v[0] = 1, beta real. v goes to vdst[0..L-1] (LDS) and to gdst[1..L-1]
// (the workspace column kept for the back-transformation); the results are
// the same in every thread. Ends with a barrier.
__device__ void house(const double2* src, double2* gdst, int L, double2* vdst, double* red, double2& tau,
                      double& beta, bool& skip) {
  double p = 0;
  for (int i = 1 + (int)threadIdx.x; i < L; i += kHeevThreads) { const double2 x = src[i]; p += x.x * x.x + x.y * x.y; }
  const double xn2 = block_sum(p, red);
  const double2 al = src[0];
  skip = xn2 == 0.0 && al.y == 0.0;
  beta = al.x;
  tau = double2{0, 0};
  double2 sc{1, 0};
  if (!skip) {
    beta = -copysign(sqrt(al.x * al.x + al.y * al.y + xn2), al.x);
    tau = double2{(beta - al.x) / beta, -al.y / beta};
    const double2 den{al.x - beta, al.y};
    const double dd = den.x * den.x + den.y * den.y;
    sc = double2{den.x / dd, -den.y / dd};
  }
  for (int i = threadIdx.x; i < L; i += kHeevThreads) {
    double2 v{1, 0};
    if (i > 0) { v = hc_mul(src[i], sc); gdst[i] = v; }
    vdst[i] = v;
  }
  __syncthreads();
}

constexpr int kHeevU = 4;  // loads in flight per thread in the passes over the trailing block

// One pass over rows and columns [first, L) of the trailing block a (element
// (i, k) at a[i + k n]): with UPD, a -= v w^H + w v^H (v = vc, w = sw, indexed
// like the rows); then y[i - first] = sum_k a(i, k) vn[k - first]. Rows are
// spread over the threads and, when there are fewer rows than threads, the
// columns over NT / rows groups whose partial sums meet in sred. Ends with a
// barrier. Column `first` after the pass also goes to scol[i - first]: it is
// the first column of the next trailing block.
template <bool UPD>
__device__ void trail_pass(double2* a22, int n, int first, int L, const double2* vc, const double2* sw,
                           const double2* vn, double2* sy, double2* sred, double2* scol) {
  constexpr int NT = kHeevThreads, U = kHeevU;
  const int R = L - first, tid = threadIdx.x;
  if (R <= 0) return;
  const int ng = R >= NT ? 1 : NT / R;
  if (tid < (R >= NT ? NT : ng * R)) {
    const int s = R >= NT ? 0 : tid / R;
    for (int ii = R >= NT ? tid : tid % R; ii < R; ii += NT) {
      const int i = first + ii;
      double2 vi{0, 0}, wi{0, 0};
      if (UPD) { vi = vc[i]; wi = sw[i]; }
      double2 acc{0, 0};
      double2* a = a22 + i;
      for (int k0 = first + s; k0 < L; k0 += ng * U) {
        double2 x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = a[(long)min(k0 + u * ng, L - 1) * n];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + u * ng;
          if (k < L) {
            if (UPD) {
              rank2_sub(x[u], vi, wi, vc[k], sw[k]);
              a[(long)k * n] = x[u];
            }
            if (k == first) scol[ii] = x[u];
            const double2 v = vn[k - first];
            acc.x += x[u].x * v.x - x[u].y * v.y;
            acc.y += x[u].x * v.y + x[u].y * v.x;
          }
        }
      }
      if (ng == 1) sy[ii] = acc;
      else sred[s * R + ii] = acc;
    }
  }
  __syncthreads();
  if (ng > 1) {
    for (int ii = tid; ii < R; ii += NT) {
      double2 sum{0, 0};
      for (int g = 0; g < ng; ++g) { sum.x += sred[g * R + ii].x; sum.y += sred[g * R + ii].y; }
      sy[ii] = sum;
    }
    __syncthreads();
  }
}

// Number of eigenvalues of T (diagonal d, off-diagonal e) below x: signs of
// the pivots of T - x I, tiny pivots replaced by -pivmin. The division by
// the previous pivot uses the hardware reciprocal refined by one Newton step.
__device__ __forceinline__ int sturm_count(const double* d, const double* e, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0;
  for (int i = 1; i < n; ++i) {
    double r = __builtin_amdgcn_rcp(q);
    r = fma(r, fma(-q, r, 1.0), r);
    const double ei = e[i - 1];
    q = fma(-ei * ei, r, d[i] - x);
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

// Deterministic start vector entry in [-1, 1) for eigenvector t, row k.
__device__ __forceinline__ double heev_start(int t, int k) {
  unsigned s = (unsigned)t * 2654435761u ^ ((unsigned)k * 40503u + 0x9e3779b9u);
  s ^= s >> 15;
  s *= 0x2c1b3c6du;
  s ^= s >> 12;
  s *= 0x297a2d39u;
  s ^= s >> 15;
  return (double)(s & 0xffffffu) / 8388608.0 - 1.0;
}

__global__ void __launch_bounds__(kHeevThreads) k_heev(HeevArgs args) {
  const gw_heev_job J = args.j[blockIdx.x];
  const int n = J.n, nev = J.nev, tid = threadIdx.x;
  constexpr int NT = kHeevThreads;
  const HeevWs WL = heev_ws(n, nev);
  unsigned char* ws = reinterpret_cast<unsigned char*>(J.ws);
  double2* W = reinterpret_cast<double2*>(ws + WL.w_off);
  double* __restrict__ lu = reinterpret_cast<double*>(ws + WL.lu_off);
  double* __restrict__ zr = reinterpret_cast<double*>(ws + WL.zr_off);
  const double2* A = reinterpret_cast<const double2*>(J.a);
  double2* Z = reinterpret_cast<double2*>(J.z);
  double* wout = reinterpret_cast<double*>(J.w);

  __shared__ double2 sv[kHeevMaxN], svn[kHeevMaxN], sw[kHeevMaxN], sy[kHeevMaxN], scol[kHeevMaxN], stau[kHeevMaxN];
  __shared__ double sd[kHeevMaxN], se[kHeevMaxN], slam[kHeevMaxN];
  __shared__ int scs[kHeevMaxN];
  __shared__ double2 sred[NT];
  __shared__ double red[2 * kHeevWaves];

  // 1. W (column-major) = (A + A^H) / 2
  for (long idx = tid; idx < (long)n * n; idx += NT) {
    const int r = (int)(idx % n), c = (int)(idx / n);
    const double2 arc = A[(long)r * J.lda + c], acr = A[(long)c * J.lda + r];
    const double2 h{0.5 * (arc.x + acr.x), 0.5 * (arc.y - acr.y)};
    W[idx] = h;
    if (c == 0 && r > 0) scol[r - 1] = h;
    if (idx == 0) sd[0] = h.x;
  }
  __syncthreads();

  // 2. Householder tridiagonalisation. The rank-2 update of step j and the
  // product A v of step j+1 share one pass over the trailing block: its first
  // column is updated first, which gives the reflector of step j+1.
  constexpr int U = kHeevU;
  double2 *vc = sv, *vn = svn;
  bool skip = true;
  double2 tau{0, 0};
  if (n > 1) {
    double beta;
    house(scol, W + 1, n - 1, vc, red, tau, beta, skip);
    if (tid == 0) { se[0] = beta; stau[0] = tau; }
    trail_pass<false>(W + n + 1, n, 0, n - 1, vc, sw, vc, sy, sred, scol);
  } else if (tid == 0) {
    se[0] = 0.0;
  }
  for (int j = 0; j < n - 1; ++j) {
    const int L = n - j - 1;
    double2* a22 = W + (long)(j + 1) * n + (j + 1);
    if (!skip) {  // w = tau y - (tau/2) (tau y)^H v v
      double2 pd{0, 0};
      for (int i = tid; i < L; i += NT) {
        const double2 ty = hc_mul(tau, sy[i]);
        sw[i] = ty;
        const double2 t = hc_cmul(ty, vc[i]);
        pd.x += t.x;
        pd.y += t.y;
      }
      const double2 dot = block_sum2(pd, red);
      const double2 td = hc_mul(tau, dot);
      const double2 a2{-0.5 * td.x, -0.5 * td.y};
      for (int i = tid; i < L; i += NT) {
        const double2 t = hc_mul(a2, vc[i]);
        sw[i] = double2{sw[i].x + t.x, sw[i].y + t.y};
      }
      __syncthreads();
    }
    if (L == 1) {  // last diagonal element
      if (tid == 0) {
        double2 a = scol[0];
        if (!skip) a.x -= 2.0 * (vc[0].x * sw[0].x + vc[0].y * sw[0].y);
        sd[n - 1] = a.x;
        se[n - 1] = 0.0;
      }
      __syncthreads();
      break;
    }
    // (a) first column of the trailing block (kept in LDS), diagonal included
    if (!skip)
      for (int i = tid; i < L; i += NT) {
        double2 a = scol[i];
        rank2_sub(a, vc[i], sw[i], vc[0], sw[0]);
        scol[i] = a;
      }
    __syncthreads();
    if (tid == 0) sd[j + 1] = scol[0].x;
    // (b) reflector of step j + 1 from rows 1..L-1 of that column
    double2 tau1;
    double beta1;
    bool skip1;
    house(scol + 1, a22 + 1, L - 1, vn, red, tau1, beta1, skip1);
    if (tid == 0) { se[j + 1] = beta1; stau[j + 1] = tau1; }
    // (c) columns 1..L-1, rows 1..L-1: update of step j, then y = A v of step j + 1
    if (skip) trail_pass<false>(a22, n, 1, L, vc, sw, vn, sy, sred, scol);
    else trail_pass<true>(a22, n, 1, L, vc, sw, vn, sy, sred, scol);
    double2* t = vc;
    vc = vn;
    vn = t;
    skip = skip1;
    tau = tau1;
  }
  __syncthreads();

  // 3. lowest nev eigenvalues of T by multisection
  double lo_i = DBL_MAX, hi_i = -DBL_MAX, nrm_i = 0, e2max_i = 0;
  for (int i = tid; i < n; i += NT) {
    const double el = i > 0 ? fabs(se[i - 1]) : 0.0, er = i < n - 1 ? fabs(se[i]) : 0.0;
    lo_i = fmin(lo_i, sd[i] - el - er);
    hi_i = fmax(hi_i, sd[i] + el + er);
    nrm_i = fmax(nrm_i, fabs(sd[i]) + el + er);
    e2max_i = fmax(e2max_i, er * er);
  }
  double gl = block_min(lo_i, red), gu = block_max(hi_i, red);
  double tn = block_max(nrm_i, red);
  const double e2max = block_max(e2max_i, red);
  const double eps = DBL_EPSILON * 0.5;
  const double pivmin = DBL_MIN * fmax(1.0, e2max);
  if (!(tn > 0.0)) tn = 1.0;  // zero (or non-finite) matrix: tolerances relative to 1
  const double pad = 2.1 * eps * tn * n + 4.0 * pivmin;
  gl -= pad;
  gu += pad;
  // P lanes of one wave per eigenvalue, each counting at one of P interior
  // points; the group keeps the sub-interval where the count first passes t.
  int P = 1;
  while (P < 16 && 2 * P * nev <= NT) P *= 2;
  const double inv_p1 = 1.0 / (P + 1);
  for (int t0 = 0; t0 < nev; t0 += NT / P) {
    const int t = t0 + tid / P, p = tid % P;
    const bool act = t < nev;
    const int base = (tid & 63) / P * P;
    const unsigned long long gmask = P >= 64 ? ~0ull : ((1ull << P) - 1ull);
    double lo = gl, hi = gu;
    for (int it = 0; it < 64; ++it) {
      const bool go = act && hi - lo > 2.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin;
      if (!__any(go)) break;  // wave-uniform exit
      const double h = (hi - lo) * inv_p1;
      const int c = sturm_count(sd, se, n, lo + (p + 1) * h, pivmin);
      const unsigned long long m = (__ballot(act && c > t) >> base) & gmask;
      if (go) {
        if (m) {
          const int f = __builtin_ctzll(m);
          const double nlo = f == 0 ? lo : lo + f * h;
          hi = lo + (f + 1) * h;
          lo = nlo;
        } else {
          lo = lo + P * h;
        }
      }
    }
    if (act && p == 0) {
      const double lam = 0.5 * (lo + hi);
      slam[t] = lam;
      wout[t] = lam;
    }
  }
  __syncthreads();

  // 4. eigenvectors of T: inverse iteration, clusters member by member
  const double ortol = 1e-3 * tn, ptol = eps * tn;
  double pmax = 0;
  for (int t = tid; t < nev; t += NT) {
    int cs = t;
    while (cs > 0 && slam[cs] - slam[cs - 1] < ortol) --cs;
    scs[t] = cs;
    pmax = fmax(pmax, (double)(t - cs));
  }
  const int maxp = (int)block_max(pmax, red);  // barrier inside: scs visible
  for (int p = 0; p <= maxp; ++p) {
    for (int t = tid; t < nev; t += NT) {
      const int cs = scs[t];
      if (t - cs != p) continue;
      const double lam = slam[t];
      // LU of T - lam I with partial pivoting; per row k: m, piv, 1/u0, u1, u2
      double alpha = sd[0] - lam, beta = n > 1 ? se[0] : 0.0;
      for (int k = 0; k < n - 1; ++k) {
        const double sub = se[k], a1 = sd[k + 1] - lam, b1 = k + 1 < n - 1 ? se[k + 1] : 0.0;
        double u0, u1, u2, m, piv;
        if (fabs(alpha) >= fabs(sub)) {
          piv = 0.0; u0 = alpha; u1 = beta; u2 = 0.0;
          m = alpha != 0.0 ? sub / alpha : 0.0;
          alpha = a1 - m * beta;
          beta = b1;
        } else {
          piv = 1.0; u0 = sub; u1 = a1; u2 = b1;
          m = alpha / sub;
          alpha = beta - m * a1;
          beta = -m * b1;
        }
        if (fabs(u0) < ptol) u0 = u0 < 0.0 ? -ptol : ptol;
        double* l = lu + (long)5 * k * nev + t;
        l[0] = m;
        l[(long)nev] = piv;
        l[(long)2 * nev] = 1.0 / u0;
        l[(long)3 * nev] = u1;
        l[(long)4 * nev] = u2;
      }
      {
        double u0 = alpha;
        if (fabs(u0) < ptol) u0 = u0 < 0.0 ? -ptol : ptol;
        lu[(long)5 * (n - 1) * nev + 2 * (long)nev + t] = 1.0 / u0;
      }
      double* z = zr + t;  // z[k * nev]
      for (int k = 0; k < n; ++k) z[(long)k * nev] = heev_start(t, k);
      for (int step = 0; step < 3; ++step) {
        // forward: apply the row operations to the right-hand side
        double r = z[0];
        for (int k = 0; k < n - 1; ++k) {
          const double* l = lu + (long)5 * k * nev + t;
          const double bk1 = z[(long)(k + 1) * nev];
          if (l[(long)nev] == 0.0) { z[(long)k * nev] = r; r = bk1 - l[0] * r; }
          else { z[(long)k * nev] = bk1; r = r - l[0] * bk1; }
        }
        z[(long)(n - 1) * nev] = r;
        // back substitution with U (diagonal, two superdiagonals)
        double z1 = 0.0, z2 = 0.0;  // z[k+1], z[k+2]
        for (int k = n - 1; k >= 0; --k) {
          const double* l = lu + (long)5 * k * nev + t;
          const double zk = (z[(long)k * nev] - l[(long)3 * nev] * z1 - l[(long)4 * nev] * z2) * l[(long)2 * nev];
          z[(long)k * nev] = zk;
          z2 = z1;
          z1 = zk;
        }
        // orthogonalise against the cluster's earlier (final) vectors
        for (int i = cs; i < t; ++i) {
          const double* q = zr + i;
          double dt = 0.0;
          for (int k = 0; k < n; ++k) dt += q[(long)k * nev] * z[(long)k * nev];
          for (int k = 0; k < n; ++k) z[(long)k * nev] -= dt * q[(long)k * nev];
        }
        double nn = 0.0;
        for (int k = 0; k < n; ++k) nn += z[(long)k * nev] * z[(long)k * nev];
        const double s = nn > 0.0 && nn <= DBL_MAX ? 1.0 / sqrt(nn) : 0.0;
        for (int k = 0; k < n; ++k) z[(long)k * nev] *= s;
      }
    }
    __threadfence_block();
    __syncthreads();
  }

  // 5. u = H_0 ... H_{n-2} z into the output (row-major n x ldz)
  for (long idx = tid; idx < (long)n * nev; idx += NT) {
    const int r = (int)(idx / nev), c = (int)(idx % nev);
    Z[(long)r * J.ldz + c] = double2{zr[idx], 0.0};
  }
  __threadfence_block();
  __syncthreads();
  // The update with H_i and the projection v_{i-1}^H z of the next reflector
  // share one pass over rows i..n-1 (H_i leaves row i alone).
  const int nc = nev < NT ? nev : NT;
  const int G = NT / nc;
  for (int cb = 0; cb < nev; cb += nc) {
    const int cl = tid % nc, g = tid / nc, c = cb + cl;
    const bool act = g < G && c < nev;
    if (n >= 2) {  // f = tau v^H z of the last reflector (v = [1] on row n-1)
      if (act && g == 0) sw[cl] = hc_mul(stau[n - 2], Z[(long)(n - 1) * J.ldz + c]);
      __syncthreads();
    }
    for (int i = n - 2; i >= 0; --i) {
      const int nr = n - i;  // rows i..n-1; v_i covers rows i+1.., v_{i-1} rows i..
      const double2* wi = W + (long)i * n + i;            // v_i[q - 1] = wi[q] (q >= 2), 1 at q = 1
      const double2* wp = W + (long)(i > 0 ? i - 1 : 0) * n + i;  // v_{i-1}[q] = wp[q] (q >= 1), 1 at q = 0
      double2 part{0, 0};
      if (act) {
        const double2 f = sw[cl];
        for (int q0 = g; q0 < nr; q0 += G * U) {
          double2 x[U], va[U], vb[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int q = min(q0 + u * G, nr - 1);
            x[u] = Z[(long)(i + q) * J.ldz + c];
            va[u] = wi[q];
            vb[u] = wp[q];
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int q = q0 + u * G;
            if (q < nr) {
              if (q > 0) {
                const double2 v = q == 1 ? double2{1, 0} : va[u];
                const double2 t = hc_mul(v, f);
                x[u].x -= t.x;
                x[u].y -= t.y;
                Z[(long)(i + q) * J.ldz + c] = x[u];
              }
              if (i >= 1) {
                const double2 t = hc_cmul(q == 0 ? double2{1, 0} : vb[u], x[u]);
                part.x += t.x;
                part.y += t.y;
              }
            }
          }
        }
      }
      if (i >= 1) {
        sred[tid] = part;
        __syncthreads();
        if (act && g == 0) {
          double2 sum{0, 0};
          for (int gg = 0; gg < G; ++gg) { sum.x += sred[gg * nc + cl].x; sum.y += sred[gg * nc + cl].y; }
          sw[cl] = hc_mul(stau[i - 1], sum);
        }
        __syncthreads();
      }
    }
  }
}

}  // namespace gw

using namespace gw;

extern "C" size_t gw_heev_workspace_bytes(int n, int nev) {
  if (n < 1 || n > kHeevMaxN || nev < 1 || nev > n) return 0;
  return heev_ws(n, nev).total;
}

extern "C" int gw_heev_max_n(void) { return kHeevMaxN; }

extern "C" int gw_heev_batched(int count, const gw_heev_job* jobs, void* stream) {
  GW_REQUIRE(count >= 0, GW_ERR_ARG, "negative job count");
  if (count == 0) return GW_OK;
  GW_REQUIRE(jobs, GW_ERR_ARG, "null job array");
  for (int i = 0; i < count; ++i) {
    const gw_heev_job& j = jobs[i];
    GW_REQUIRE(j.n >= 1 && j.n <= kHeevMaxN, GW_ERR_UNSUPPORTED, "job %d: n = %d outside [1, %d]", i, j.n,
               kHeevMaxN);
    GW_REQUIRE(j.nev >= 1 && j.nev <= j.n, GW_ERR_ARG, "job %d: nev = %d outside [1, n = %d]", i, j.nev, j.n);
    GW_REQUIRE(j.lda >= j.n && j.ldz >= j.nev, GW_ERR_ARG, "job %d: lda %d < n or ldz %d < nev", i, j.lda, j.ldz);
    GW_REQUIRE(j.a && j.w && j.z && j.ws, GW_ERR_ARG, "job %d: null pointer", i);
    GW_REQUIRE(((uintptr_t)j.a | (uintptr_t)j.z | (uintptr_t)j.ws) % 16 == 0 && (uintptr_t)j.w % 8 == 0,
               GW_ERR_ARG, "job %d: misaligned pointer", i);
  }
  hipStream_t st = (hipStream_t)stream;
  for (int b = 0; b < count; b += kHeevPerLaunch) {
    HeevArgs args{};
    args.count = std::min(kHeevPerLaunch, count - b);
    for (int i = 0; i < args.count; ++i) args.j[i] = jobs[b + i];
    hipLaunchKernelGGL(k_heev, dim3(args.count), dim3(kHeevThreads), 0, st, args);
    GW_CHECK_HIP(hipGetLastError());
  }
  return GW_OK;
}
