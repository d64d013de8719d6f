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
//      (P = 2..16 lanes of one wave per eigenvalue, one interior point each,
//      the interval shrinking by P + 1 per step);
//   4. computes T's eigenvectors by inverse iteration (LU of T - lambda I with
//      partial pivoting, tiny pivots raised to eps |T|, three steps from a
//      hashed start vector), orthogonalising each vector against the earlier
//      members of its cluster (eigenvalues closer than 1e-5 |T|) in every step;
//      clusters are processed member by member, all clusters at once;
//   5. back-transforms u = H_0 H_1 ... H_{n-2} z into the caller's output.
// Every loop has a fixed bound (n, nev, 64 multisection steps, 3 inverse
// iteration steps), so the kernel ends for any input, NaN included.
// The file also holds the batched Cholesky factor inverse of the CholQR step
// (k_chol_inv, below).
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

// A double moved between lanes by a DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double read_lane_d(double v, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// Sum over the 64 lanes (all active), the same in every lane: butterflies inside
// each row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror), then the four row sums read from lanes 0, 16, 32, 48.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  return (read_lane_d(v, 0) + read_lane_d(v, 16)) + (read_lane_d(v, 32) + read_lane_d(v, 48));
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

// Reflector scalars for x with x[0] = al and |x[1:]|^2 = xn2: H^H x = beta e_1
// with H = I - tau v v^H, v = (1, x[1:] sc), beta real, sc = 1 / (al - beta).
// x = 0 (or x = al e_1 with al real) gives tau = 0, H = I.
__device__ __forceinline__ void house_scalars(double2 al, double xn2, double2& tau, double& beta, double2& sc) {
  beta = al.x;
  tau = double2{0, 0};
  sc = double2{1, 0};
  if (xn2 == 0.0 && al.y == 0.0) return;
  beta = -copysign(sqrt(al.x * al.x + al.y * al.y + xn2), al.x);
  const double rb = 1.0 / beta;
  tau = double2{(beta - al.x) * rb, -al.y * rb};
  const double dx = al.x - beta, dy = al.y;
  const double rd = 1.0 / (dx * dx + dy * dy);
  sc = double2{dx * rd, -dy * rd};
}

constexpr int kHeevU = 4;  // loads in flight per thread in the passes over the trailing block

// One pass over rows and columns [first, L) of the trailing block a (element
// (i, k) at a[i + k n]): with UPD, a -= v w^H + w v^H (v = vc, w = sw, indexed
// like the rows); then sy[i - first] = tau1 sum_k a(i, k) vn[k - first] and
// the partial sums of sy^H vn per wave go to red2 (re, then im). Column
// `first` after the update goes to scol[i - first]: it is the first column of
// the next trailing block. Rows are spread over the threads and, when there
// are fewer rows than threads, the columns over NT / rows groups whose partial
// sums meet in sred. Ends with a barrier.
template <bool UPD>
__device__ void trail_pass(double2* a22, int n, int first, int L, const double2* vc, const double2* sw,
                           const double2* vn, double2 tau1, double2* sy, double2* sred, double2* scol,
                           double* red2) {
  constexpr int NT = kHeevThreads, U = kHeevU;
  const int R = L - first, tid = threadIdx.x;
  const int ng = R >= NT ? 1 : min(8, NT / R);  // column groups (a short final sum)
  double2 dp{0, 0};
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
      if (ng == 1) {
        const double2 yv = hc_mul(tau1, acc);
        sy[ii] = yv;
        const double2 t = hc_cmul(yv, vn[ii]);
        dp.x += t.x;
        dp.y += t.y;
      } else {
        sred[s * R + ii] = acc;
      }
    }
  }
  if (ng > 1) {
    __syncthreads();
    for (int ii = tid; ii < R; ii += NT) {
      double2 sum{0, 0};
      for (int g = 0; g < ng; ++g) { sum.x += sred[g * R + ii].x; sum.y += sred[g * R + ii].y; }
      const double2 yv = hc_mul(tau1, sum);
      sy[ii] = yv;
      const double2 t = hc_cmul(yv, vn[ii]);
      dp.x += t.x;
      dp.y += t.y;
    }
  }
  dp.x = wave_sum(dp.x);
  dp.y = wave_sum(dp.y);
  if ((tid & 63) == 0) { red2[tid >> 6] = dp.x; red2[kHeevWaves + (tid >> 6)] = dp.y; }
  __syncthreads();
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
  __shared__ double red[2 * kHeevWaves], red2[2 * kHeevWaves];

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

  // 2. Householder tridiagonalisation. Per step j: (1) w = tau y - (tau/2)
  // (tau y)^H v v from the pass's tau y and partial dot products, the update
  // of the trailing block's first column (kept in LDS) and the norm of the
  // next x; (2) the reflector of step j + 1; (3) one pass over the rest of
  // the block: rank-2 update of step j and tau y of step j + 1. Three or four
  // barriers per step.
  constexpr int U = kHeevU;
  double2 *vc = sv, *vn = svn;
  double2 tau{0, 0};
  if (n > 1) {
    double p = 0;
    for (int i = 1 + tid; i < n - 1; i += NT) { const double2 x = scol[i]; p += x.x * x.x + x.y * x.y; }
    const double xn2 = block_sum(p, red);
    double beta;
    double2 sc;
    house_scalars(scol[0], xn2, tau, beta, sc);
    for (int i = tid; i < n - 1; i += NT) {
      double2 v{1, 0};
      if (i > 0) { v = hc_mul(scol[i], sc); W[1 + i] = v; }
      vc[i] = v;
    }
    if (tid == 0) { se[0] = beta; stau[0] = tau; }
    __syncthreads();
    trail_pass<false>(W + n + 1, n, 0, n - 1, vc, sw, vc, tau, sy, sred, scol, red2);
  } else if (tid == 0) {
    se[0] = 0.0;
  }
  for (int j = 0; j < n - 1; ++j) {
    const int L = n - j - 1;
    double2* a22 = W + (long)(j + 1) * n + (j + 1);
    // (1)
    double2 dot{0, 0};
#pragma unroll
    for (int w = 0; w < kHeevWaves; ++w) { dot.x += red2[w]; dot.y += red2[kHeevWaves + w]; }
    const double2 td = hc_mul(tau, dot);
    const double2 a2{-0.5 * td.x, -0.5 * td.y};
    const double2 v0 = vc[0], y0 = sy[0];
    const double2 w0{y0.x + a2.x * v0.x - a2.y * v0.y, y0.y + a2.x * v0.y + a2.y * v0.x};
    double p = 0;
    for (int i = tid; i < L; i += NT) {
      const double2 vi = vc[i], yi = sy[i];
      const double2 wi{yi.x + a2.x * vi.x - a2.y * vi.y, yi.y + a2.x * vi.y + a2.y * vi.x};
      sw[i] = wi;
      double2 a = scol[i];
      rank2_sub(a, vi, wi, v0, w0);
      scol[i] = a;
      if (i >= 2) p += a.x * a.x + a.y * a.y;  // |x[1:]|^2 of the next reflector's x = scol[1..]
    }
    if (L == 1) {  // last diagonal element
      if (tid == 0) { sd[n - 1] = scol[0].x; se[n - 1] = 0.0; }
      __syncthreads();
      break;
    }
    const double xn2 = block_sum(p, red);
    // (2)
    if (tid == 0) sd[j + 1] = scol[0].x;
    double2 tau1, sc;
    double beta1;
    house_scalars(scol[1], xn2, tau1, beta1, sc);
    for (int i = tid; i < L - 1; i += NT) {
      double2 v{1, 0};
      if (i > 0) { v = hc_mul(scol[1 + i], sc); a22[1 + i] = v; }
      vn[i] = v;
    }
    if (tid == 0) { se[j + 1] = beta1; stau[j + 1] = tau1; }
    __syncthreads();
    // (3)
    trail_pass<true>(a22, n, 1, L, vc, sw, vn, tau1, sy, sred, scol, red2);
    double2* t = vc;
    vc = vn;
    vn = t;
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
  // Clusters: eigenvalues closer than 1e-5 |T| (inverse-iteration vectors of
  // eigenvalues further apart are orthogonal to ~eps |T| / gap <= 1e-11). The
  // 1e-3 |T| of LAPACK's dstein would put all wanted eigenvalues of a
  // Davidson subspace matrix (|T| set by high-kinetic correction directions,
  // 1e2-1e4 eV, against a few eV between bands) into one cluster, processed
  // member by member.
  const double ortol = 1e-5 * tn, ptol = eps * tn;
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
  const int G = min(16, NT / nc);  // row groups per column (a short final sum)
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

// ---- batched Cholesky factor inverse (the CholQR2 step of the orthonormalisation)
//
// For each Hermitian positive definite G (n <= kCholMaxN, lower triangle read):
// G = L L^H, and X = L^-1 (lower triangular) is written; info = 0, or the
// 1-based index of the first non-positive pivot (the factor is then not
// usable, as for torch.linalg.cholesky_ex). One workgroup per matrix, both
// triangles in LDS: right-looking factorisation (three barriers per column),
// then each thread forms one column of L^-1 by forward substitution.
constexpr int kCholMaxN = 64;
constexpr int kCholThreads = 256;
constexpr int kCholPerLaunch = 32;

struct CholArgs {
  gw_chol_job j[kCholPerLaunch];
  int count;
};

__global__ void __launch_bounds__(kCholThreads) k_chol_inv(CholArgs args) {
  const gw_chol_job J = args.j[blockIdx.x];
  const int n = J.n, tid = threadIdx.x;
  constexpr int NT = kCholThreads;
  __shared__ double2 La[kCholMaxN * kCholMaxN];
  __shared__ double2 Xs[kCholMaxN * kCholMaxN];
  __shared__ double rdiag;
  __shared__ int sinfo;
  const double2* A = reinterpret_cast<const double2*>(J.a);
  double2* Out = reinterpret_cast<double2*>(J.linv);
  for (int idx = tid; idx < n * n; idx += NT) {
    const int i = idx / n, k = idx - i * n;
    La[idx] = k <= i ? A[(long)i * J.lda + k] : double2{0, 0};
  }
  if (tid == 0) sinfo = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      double d = La[j * n + j].x;
      if (!(d > 0.0)) {
        if (sinfo == 0) sinfo = j + 1;
        d = 1.0;
      }
      const double l = sqrt(d);
      La[j * n + j] = double2{l, 0.0};
      rdiag = 1.0 / l;
    }
    __syncthreads();
    const double r = rdiag;
    for (int i = j + 1 + tid; i < n; i += NT) { La[i * n + j].x *= r; La[i * n + j].y *= r; }
    __syncthreads();
    const int m = n - j - 1;
    for (int q = tid; q < m * m; q += NT) {
      const int ii = q / m, kk = q - ii * m;
      if (kk > ii) continue;
      const int i = j + 1 + ii, k = j + 1 + kk;
      const double2 a = La[i * n + j], b = La[k * n + j];
      La[i * n + k].x -= a.x * b.x + a.y * b.y;
      La[i * n + k].y -= a.y * b.x - a.x * b.y;
    }
    __syncthreads();
  }
  for (int c = tid; c < n; c += NT) {
    Xs[c * n + c] = double2{1.0 / La[c * n + c].x, 0.0};
    for (int i = c + 1; i < n; ++i) {
      double2 acc{0, 0};
      for (int k = c; k < i; ++k) {
        const double2 l = La[i * n + k], x = Xs[k * n + c];
        acc.x += l.x * x.x - l.y * x.y;
        acc.y += l.x * x.y + l.y * x.x;
      }
      const double inv = 1.0 / La[i * n + i].x;
      Xs[i * n + c] = double2{-acc.x * inv, -acc.y * inv};
    }
  }
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += NT) {
    const int i = idx / n, k = idx - i * n;
    Out[(long)i * J.ldl + k] = k <= i ? Xs[idx] : double2{0, 0};
  }
  if (tid == 0) *reinterpret_cast<int*>(J.info) = sinfo;
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

extern "C" int gw_chol_max_n(void) { return kCholMaxN; }

extern "C" int gw_chol_inv_batched(int count, const gw_chol_job* jobs, void* stream) {
  GW_REQUIRE(count >= 0, GW_ERR_ARG, "negative job count");
  if (count == 0) return GW_OK;
  GW_REQUIRE(jobs, GW_ERR_ARG, "null job array");
  for (int i = 0; i < count; ++i) {
    const gw_chol_job& j = jobs[i];
    GW_REQUIRE(j.n >= 1 && j.n <= kCholMaxN, GW_ERR_UNSUPPORTED, "job %d: n = %d outside [1, %d]", i, j.n,
               kCholMaxN);
    GW_REQUIRE(j.lda >= j.n && j.ldl >= j.n, GW_ERR_ARG, "job %d: lda %d or ldl %d < n", i, j.lda, j.ldl);
    GW_REQUIRE(j.a && j.linv && j.info, GW_ERR_ARG, "job %d: null pointer", i);
    GW_REQUIRE(((uintptr_t)j.a | (uintptr_t)j.linv) % 16 == 0 && (uintptr_t)j.info % 4 == 0, GW_ERR_ARG,
               "job %d: misaligned pointer", i);
  }
  hipStream_t st = (hipStream_t)stream;
  for (int b = 0; b < count; b += kCholPerLaunch) {
    CholArgs args{};
    args.count = std::min(kCholPerLaunch, count - b);
    for (int i = 0; i < args.count; ++i) args.j[i] = jobs[b + i];
    hipLaunchKernelGGL(k_chol_inv, dim3(args.count), dim3(kCholThreads), 0, st, args);
    GW_CHECK_HIP(hipGetLastError());
  }
  return GW_OK;
}
