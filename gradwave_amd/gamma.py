"""Gamma-point real layout on MI355X (core/gamma.py:1-290).

At k = 0 the reference stores the independent half of the plane-wave sphere
(G=0 first, then the lexicographically positive member of each {G, -G} pair)
and solves a real symmetric eigenproblem in the embedding

    phi[0] = Re c(0),  phi[2h-1] = sqrt2 Re c(h),  phi[2h] = sqrt2 Im c(h),

whose dot product is the half-sphere metric. Here the half-sphere apply runs
two real bands per complex apply of the full-sphere HIP pipeline
(``gw_gamma_pack`` -> ``gw_hpsi_apply_strided`` -> ``gw_gamma_unpack``), which
halves the FFT and projector work per band, and the Davidson
(``davidson_gamma``, the reference's single-k solver of
solvers/davidson.py:329-395 in the embedding) runs in real arithmetic: real
subspace matrix, real eigensolve, real GEMMs.
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from types import SimpleNamespace

import numpy as np
import torch

from gradwave_amd import _lib as L
from gradwave_amd import linalg as LA
from gradwave_amd.davidson import buffer_block
from gradwave_amd.hamiltonian import BatchedHamiltonian

HBAR2_2M = 27.211386245988 / 2.0 * 0.529177210903**2


@dataclass
class GammaBasis:
    """Half-sphere maps for one k = 0 sphere (core/gamma.py:40-67), on the device."""

    shape: tuple
    nh3: int
    t_half: torch.Tensor  # (nhalf,) float64
    metric_w: torch.Tensor  # (nhalf,) 1 at G=0, else 2
    full_flat_idx: torch.Tensor  # (npw,) int64
    rep_full_idx: torch.Tensor  # (nhalf,) int64
    src_half: torch.Tensor  # (npw,) int64
    conj_full: torch.Tensor  # (npw,) bool
    back_flat: torch.Tensor  # (nhalf,) int64, rfft half-box index
    back_conj: torch.Tensor  # (nhalf,) bool
    partner: torch.Tensor  # (npw,) int64, sphere index of -G
    t_full: torch.Tensor  # (npw,) float64
    # int32 / uint8 copies for the kernels
    k_src: torch.Tensor
    k_conj: torch.Tensor
    k_rep: torch.Tensor
    k_partner: torch.Tensor

    @property
    def nhalf(self) -> int:
        return int(self.t_half.shape[0])

    @property
    def npw(self) -> int:
        return int(self.src_half.shape[0])

    @property
    def n_real(self) -> int:
        return 2 * self.nhalf - 1


def build_gamma_basis(sphere, shape, device=None) -> GammaBasis:
    """core/gamma.py:70-169. ``sphere`` needs flat_idx and kpg2 (miller is
    recovered from flat_idx); raises ValueError unless the sphere is closed under
    G -> -G with exactly one G = 0."""
    n1, n2, n3 = (int(s) for s in shape)
    flat = sphere.flat_idx.detach().cpu().numpy().astype(np.int64)
    dev = torch.device(device) if device is not None else sphere.flat_idx.device
    i0, rem = flat // (n2 * n3), flat % (n2 * n3)
    i1, i2 = rem // n3, rem % n3
    box = np.stack([i0, i1, i2], axis=1)
    n = np.array([n1, n2, n3])
    m = np.where(box > n // 2, box - n, box)
    neg = (-m) % n
    npw = flat.shape[0]
    where = np.full(n1 * n2 * n3, -1, dtype=np.int64)
    where[flat] = np.arange(npw)
    partner = where[(neg[:, 0] * n2 + neg[:, 1]) * n3 + neg[:, 2]]
    if (partner < 0).any():
        raise ValueError("Gamma basis needs a sphere closed under G -> -G (k must be Gamma)")
    zero = (m == 0).all(axis=1)
    if int(zero.sum()) != 1:
        raise ValueError("Gamma sphere must contain exactly one G=0 vector")
    pos = (m[:, 2] > 0) | ((m[:, 2] == 0) & ((m[:, 1] > 0) | ((m[:, 1] == 0) & (m[:, 0] > 0))))
    rep = np.concatenate([np.flatnonzero(zero), np.flatnonzero(pos)])
    nhalf = rep.shape[0]
    if nhalf != (npw + 1) // 2:
        raise ValueError(f"half-sphere size {nhalf} != {(npw + 1) // 2}")
    is_rep = zero | pos
    slot = np.full(npw, -1, dtype=np.int64)
    slot[rep] = np.arange(nhalf)
    src = np.where(is_rep, slot, slot[partner])
    nh3 = n3 // 2 + 1
    rb = box[rep]
    inside = rb[:, 2] <= n3 // 2
    sb = np.where(inside[:, None], rb, neg[rep])
    back = (sb[:, 0] * n2 + sb[:, 1]) * nh3 + sb[:, 2]
    t_full = HBAR2_2M * sphere.kpg2.detach().cpu().numpy().astype(np.float64)
    w = np.full(nhalf, 2.0)
    w[0] = 1.0

    def dt(a, t):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=t, device=dev)

    return GammaBasis((n1, n2, n3), nh3, dt(t_full[rep], torch.float64), dt(w, torch.float64),
                      dt(flat, torch.int64), dt(rep, torch.int64), dt(src, torch.int64), dt(~is_rep, torch.bool),
                      dt(back, torch.int64), dt(~inside, torch.bool), dt(partner, torch.int64),
                      dt(t_full, torch.float64), dt(src, torch.int32), dt((~is_rep).astype(np.uint8), torch.uint8),
                      dt(rep, torch.int32), dt(partner, torch.int32))


def half_to_full(gb: GammaBasis, ch: torch.Tensor) -> torch.Tensor:
    """(nb, nhalf) -> (nb, npw), Hermitian by construction (core/gamma.py:172-175)."""
    c = ch[:, gb.src_half]
    return torch.where(gb.conj_full[None, :], c.conj(), c)


def full_to_half(gb: GammaBasis, cf: torch.Tensor) -> torch.Tensor:
    """(nb, npw) -> (nb, nhalf) (core/gamma.py:178-180)."""
    return cf[:, gb.rep_full_idx]


def embed_real(gb: GammaBasis, ch: torch.Tensor) -> torch.Tensor:
    """Complex half sphere -> real embedding (core/gamma.py:242-262)."""
    r2 = 2.0 ** 0.5
    rdt = torch.float64 if ch.dtype == torch.complex128 else torch.float32
    phi = torch.empty(ch.shape[0], gb.n_real, dtype=rdt, device=ch.device)
    phi[:, 0] = ch[:, 0].real
    phi[:, 1::2] = r2 * ch[:, 1:].real
    phi[:, 2::2] = r2 * ch[:, 1:].imag
    return phi


def unembed_real(gb: GammaBasis, phi: torch.Tensor) -> torch.Tensor:
    """Inverse of embed_real (core/gamma.py:265-272)."""
    r2i = 2.0 ** -0.5
    cdt = torch.complex128 if phi.dtype == torch.float64 else torch.complex64
    ch = torch.empty(phi.shape[0], gb.nhalf, dtype=cdt, device=phi.device)
    ch[:, 0] = phi[:, 0].to(cdt)
    ch[:, 1:] = r2i * torch.complex(phi[:, 1::2], phi[:, 2::2])
    return ch


def kinetic_phi(gb: GammaBasis) -> torch.Tensor:
    """Kinetic diagonal in the embedding (core/gamma.py:275-280)."""
    t = torch.empty(gb.n_real, dtype=torch.float64, device=gb.t_half.device)
    t[0] = gb.t_half[0]
    t[1::2] = gb.t_half[1:]
    t[2::2] = gb.t_half[1:]
    return t


def metric_inner(gb: GammaBasis, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Real Gram in the half-sphere metric (core/gamma.py:231-239)."""
    return (a.conj() @ (gb.metric_w * b).T).real


class GammaHamiltonian:
    """H at Gamma on the half sphere (core/gamma.py:183-228): ``v_eff_r`` real
    (n1, n2, n3), ``p`` (nproj, npw) full-sphere projectors, ``dij`` (nproj, nproj).

    ``apply`` takes complex half vectors like the reference; ``apply_real``
    takes the real embedding (float64 or float32 rows) and is what the solver
    calls."""

    def __init__(self, gb: GammaBasis, v_eff_r: torch.Tensor, p: torch.Tensor, dij: torch.Tensor) -> None:
        self.gb = gb
        self.shape = gb.shape
        self.v_eff_r = v_eff_r
        self.p = p
        self.dij = dij
        dev = gb.t_half.device
        npw = gb.npw
        bk = SimpleNamespace(npw=torch.tensor([npw], dtype=torch.int64, device=dev),
                             mask=torch.ones(1, npw, dtype=torch.bool, device=dev),
                             flat_idx=gb.full_flat_idx[None], t=gb.t_full[None], dij_full=dij)
        pp = p if p.dim() == 3 else p[None]
        self.h = BatchedHamiltonian(bk, gb.shape, v_eff_r, pp)
        self._buf: dict = {}

    def _buffers(self, nbh: int, cdt: torch.dtype):
        key = cdt
        z = self._buf.get(key)
        if z is None or z[0].shape[1] < nbh:
            dev = self.gb.t_half.device
            z = (torch.empty(1, nbh, self.gb.npw, dtype=cdt, device=dev),
                 torch.empty(1, nbh, self.gb.npw, dtype=cdt, device=dev))
            self._buf[key] = z
        return z[0][:, :nbh], z[1][:, :nbh]

    def apply_real(self, phi: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """H on rows of the real embedding (nb, 2*nhalf-1) -> same shape/dtype."""
        gb = self.gb
        L.require_gpu(phi, "phi")
        if phi.stride(-1) != 1:
            phi = phi.contiguous()
        nb = int(phi.shape[0])
        if out is None:
            out = torch.empty_like(phi)
        if nb == 0:
            return out
        prec = L.precision_of(phi.dtype)
        cdt = torch.complex128 if prec == L.GW_FP64 else torch.complex64
        nbh = (nb + 1) // 2
        z, hz = self._buffers(nbh, cdt)
        st = L.stream_ptr(phi.device)
        L.check(L.lib().gw_gamma_pack(prec, nb, gb.npw, gb.npw, phi.data_ptr(), phi.stride(0),
                                      gb.k_src.data_ptr(), gb.k_conj.data_ptr(), z.data_ptr(), gb.npw, st),
                "gw_gamma_pack")
        self.h.apply_terms(z, L.GW_TERM_ALL, out=hz)
        L.check(L.lib().gw_gamma_unpack(prec, nb, gb.nhalf, hz.data_ptr(), gb.npw, gb.k_rep.data_ptr(),
                                        gb.k_partner.data_ptr(), out.data_ptr(), out.stride(0), st),
                "gw_gamma_unpack")
        return out

    def apply(self, ch: torch.Tensor) -> torch.Tensor:
        """(nb, nhalf) complex half vectors -> H on the half sphere (core/gamma.py:206-228)."""
        return unembed_real(self.gb, self.apply_real(embed_real(self.gb, ch)))

    __call__ = apply


@dataclass
class GammaDavidsonResult:
    """core/gamma.py:283-288."""

    eigenvalues: torch.Tensor  # (nb,) ascending, eV
    eigenvectors: torch.Tensor  # (nb, nhalf) metric-orthonormal complex half sphere
    n_iter: int
    residual_norms: torch.Tensor  # (nb,)
    n_draft_iter: int = 0


def _teter(r: torch.Tensor, t_g: torch.Tensor, t_band: torch.Tensor) -> torch.Tensor:
    """solvers/precond.py:8-17, coefficient in float64."""
    x = t_g[None, :] / torch.clamp(t_band[:, None], min=1e-12)
    x2 = x * x
    num = 27.0 + 18.0 * x + 12.0 * x2 + 8.0 * x2 * x
    return r * (num / (num + 16.0 * x2 * x2)).to(r.dtype)


def _project_out(v: torch.Tensor, against: torch.Tensor, passes: int = 2) -> torch.Tensor:
    """v - (v a^T) a, ``passes`` times (twice in the reference,
    solvers/davidson.py:321-323), on the MFMA kernels."""
    v = v.clone()
    for _ in range(passes):
        c = LA.rinner(against, v)  # (m, k): c[j][a] = against_j . v_a
        LA.rcombine(c, against, out=v, alpha=-1.0, beta=1.0)
    return v


def _orthonormalize(v: torch.Tensor, against: torch.Tensor | None = None, passes: int = 2) -> torch.Tensor:
    """solvers/davidson.py:310-327 in real arithmetic: project out ``against``,
    then orthonormalize the rows, dropping rows whose QR diagonal |R_ii| is at
    most 1e-10.

    Rows are scaled to unit norm and orthonormalized by CholQR2 (MFMA Gram,
    FP64 Cholesky, triangular combination, twice); with row norms n_i the QR
    diagonal of the unscaled rows is n_i L_ii. When that test would drop a row,
    or the Gram is too ill-conditioned for CholQR, Householder QR runs instead
    and applies the reference's test directly."""
    if against is not None and against.shape[0]:
        v = _project_out(v, against, passes)
    k = v.shape[0]
    if k == 0:
        return v
    nrm = torch.linalg.norm(v.double(), dim=1)
    ok = bool((nrm > 1e-10).all())
    if ok:
        vh = v / nrm[:, None].to(v.dtype)
        q = vh
        for it in range(2):
            # FP32 rows are upcast first: exact products, FP64 sums (davidson.py:213-216)
            q64 = q if q.dtype == torch.float64 else q.double()
            g = LA.rinner(q64, q64)
            lo, info = torch.linalg.cholesky_ex(0.5 * (g + g.T))
            d = lo.diagonal()
            if int(info) != 0 or (it == 0 and (bool((d * nrm <= 1e-10).any()) or float(d.min()) < 1e-6)):
                ok = False
                break
            eye = torch.eye(k, dtype=torch.float64, device=v.device)
            linv = torch.linalg.solve_triangular(lo, eye, upper=False)
            q = LA.rcombine(linv.T.contiguous().to(v.dtype), q)
        if ok:
            return q
    qq, r = torch.linalg.qr(v.T.double(), mode="reduced")
    keep = r.diagonal().abs() > 1e-10
    return qq.T[keep].contiguous().to(v.dtype)


class _Timer:
    def __init__(self, out):
        self.out = out
        self.t = None
        self.key = None

    def __call__(self, key):
        if self.out is None:
            return
        torch.cuda.synchronize()
        now = time.perf_counter()
        if self.key is not None:
            self.out[self.key] = self.out.get(self.key, 0.0) + now - self.t
        self.key, self.t = key, now


@torch.no_grad()
def davidson_real(h_apply, x0: torch.Tensor, t_g: torch.Tensor, tol: float = 1e-9, max_iter: int = 40,
                  max_dim_factor: int = 4, profile_out: dict | None = None, nconv: int | None = None,
                  proj_passes: int = 2):
    """The reference's single-k block Davidson (solvers/davidson.py:329-395) for a
    real symmetric operator on the GPU. The basis lives in preallocated buffers
    of x0's precision, every contraction over the long axis is a real MFMA
    kernel, the subspace matrix (FP64) grows by its new rows and columns
    instead of being rebuilt, and its eigensolve is FP64. ``nconv`` (default:
    every band) stops once the first ``nconv`` bands are converged; the others
    are buffer bands, still expanded while unconverged. ``proj_passes``
    projections of the expansion block against the basis (two in the
    reference; the mixed solver's FP32 draft uses one)."""
    nb, n = x0.shape
    ng = nb if nconv is None else max(1, min(int(nconv), nb))
    dev = x0.device
    dt = x0.dtype
    f64 = torch.float64
    max_dim = min(max_dim_factor * nb, n)
    tm = _Timer(profile_out)
    tm("orthonormalize")
    v = _orthonormalize(x0.clone())
    if v.shape[0] < nb:
        gen = torch.Generator().manual_seed(1234)
        pad = torch.randn(nb - v.shape[0], n, generator=gen, dtype=f64)
        v = torch.cat([v, _orthonormalize(pad.to(dev, dt), against=v)])
    V = torch.empty(max_dim, n, dtype=dt, device=dev)
    HV = torch.empty(max_dim, n, dtype=dt, device=dev)
    S = torch.empty(max_dim, max_dim, dtype=f64, device=dev)
    t_g = t_g.to(dev, dt)

    def reset(vv):
        m = vv.shape[0]
        V[:m] = vv
        tm("h_apply")
        HV[:m] = h_apply(vv)
        tm("subspace_gemm")
        S[:m, :m] = LA.rinner(V[:m], HV[:m])
        return m

    m = reset(v)
    eig = torch.zeros(nb, dtype=f64, device=dev)
    rn = torch.full((nb,), float("inf"), dtype=f64, device=dev)
    x = V[:nb]
    for it in range(1, max_iter + 1):
        tm("eigh")
        s = S[:m, :m]
        w, u = torch.linalg.eigh(0.5 * (s + s.T))
        eig = w[:nb]
        ub = u[:, :nb].to(dt)
        tm("ritz_gemm")
        x = LA.rcombine(ub, V[:m])
        hx = LA.rcombine(ub, HV[:m])
        tm("residual")
        r = hx - eig[:, None].to(dt) * x
        rn = torch.linalg.norm(r.to(f64), dim=1)
        bad = rn > tol
        nbad = int(bad[:ng].sum())
        if nbad == 0:
            tm(None)
            return eig, x, it, rn
        tb = (x * x * t_g).sum(1, dtype=f64)
        idx = bad.nonzero().squeeze(1)
        t = _teter(r[idx], t_g.to(f64), tb[idx])
        tm("orthonormalize")
        t = _orthonormalize(t, against=V[:m], passes=proj_passes)
        k = t.shape[0]
        if k == 0:
            tm(None)
            return eig, x, it, rn
        if m + k > max_dim:
            tm("restart")
            m = reset(_orthonormalize(torch.cat([x, t])))
        else:
            V[m:m + k] = t
            tm("h_apply")
            HV[m:m + k] = h_apply(t)
            tm("subspace_gemm")
            S[:m + k, m:m + k] = LA.rinner(V[:m + k], HV[m:m + k])
            S[m:m + k, :m] = LA.rinner(V[m:m + k], HV[:m])
            m += k
    tm(None)
    return eig, x, max_iter, rn


def davidson_gamma(ham: GammaHamiltonian, x0: torch.Tensor, tol: float = 1e-9, max_iter: int = 40,
                   max_dim_factor: int = 4, profile_out: dict | None = None) -> GammaDavidsonResult:
    """core/gamma.py:291-320: real-symmetric block Davidson on the half sphere,
    FP64 throughout; eigenvectors come back as complex half vectors."""
    gb = ham.gb
    eig, x, it, rn = davidson_real(ham.apply_real, embed_real(gb, x0.to(torch.complex128)), kinetic_phi(gb),
                                   tol=tol, max_iter=max_iter, max_dim_factor=max_dim_factor,
                                   profile_out=profile_out)
    return GammaDavidsonResult(eig, unembed_real(gb, x), it, rn)


def refine_real(h_apply, x0: torch.Tensor, t_g: torch.Tensor, sweeps: int = 1):
    """``sweeps`` Davidson sweeps of a real symmetric operator from drafted
    vectors, with no convergence gate: Rayleigh-Ritz of the orthonormalised
    block, then per sweep every band's Teter-preconditioned residual
    (solvers/precond.py:20-39) joins the basis and Rayleigh-Ritz runs over the
    grown basis. Returns (eig, x, sweeps, residual norms) in x0's precision
    with FP64 subspace matrices, eigensolves and norms."""
    nb = x0.shape[0]
    dt = x0.dtype
    f64 = torch.float64
    t_g = t_g.to(x0.device, dt)
    v = _orthonormalize(x0.clone())
    if v.shape[0] < nb:
        raise ValueError("drafted vectors are linearly dependent")
    hv = h_apply(v)
    for sweep in range(sweeps + 1):
        s = LA.rinner(v, hv)
        w, u = torch.linalg.eigh(0.5 * (s + s.T))
        eig = w[:nb]
        ub = u[:, :nb].to(dt)
        x, hx = LA.rcombine(ub, v), LA.rcombine(ub, hv)
        r = hx - eig[:, None].to(dt) * x
        rn = torch.linalg.norm(r.to(f64), dim=1)
        if sweep == sweeps:
            break
        tb = (x * x * t_g).sum(1, dtype=f64)
        d = _orthonormalize(_teter(r, t_g.to(f64), tb), against=v)
        v = torch.cat([v, d])
        hv = torch.cat([hv, h_apply(d)])
    return eig, x, sweeps, rn


def davidson_gamma_mixed(ham: GammaHamiltonian, x0: torch.Tensor, tol: float = 1e-9, draft_tol: float = 1e-4,
                         max_iter: int = 100, max_dim_factor: int = 4, sweeps: int | None = None,
                         profile_out: dict | None = None, buffer: int = 0) -> GammaDavidsonResult:
    """Precision form at Gamma: the draft solve keeps its basis in FP32 and
    applies H in FP32 (FFTs, V multiply, MFMA projector and subspace GEMMs with
    FP32 accumulation and FP64 split-K sums) to ``draft_tol``, with the subspace
    eigensolve, Cholesky factors and residual norms in FP64. With
    ``sweeps=None`` the polish solve restarts from the draft Ritz vectors in
    FP64 and converges on FP64 residuals of the FP64 operator to ``tol``; with
    ``sweeps=n`` n FP64 refinement sweeps (``refine_real``) replace it and
    ``tol`` is unused. ``buffer`` > 0 carries that many extra bands (seeded
    half-sphere start vectors, G = 0 real) and returns the requested ones only,
    as in ``davidson_mixed``."""
    if buffer < 0:
        raise ValueError(f"buffer must be >= 0, got {buffer}")
    gb = ham.gb
    nb = x0.shape[0]
    if buffer:
        xb = buffer_block(1, buffer, x0.shape[1], None, x0.device)[0]
        xb[:, 0] = xb[:, 0].real.to(xb.dtype)
        x0 = torch.cat([x0.to(torch.complex128), xb])
    t_phi = kinetic_phi(gb)
    x32 = embed_real(gb, x0.to(torch.complex128)).to(torch.float32)
    nconv = nb if buffer else None  # buffer bands are expanded but do not gate
    _, xd, itd, _ = davidson_real(ham.apply_real, x32, t_phi, tol=draft_tol, max_iter=max_iter,
                                  max_dim_factor=max_dim_factor, profile_out=profile_out, nconv=nconv,
                                  proj_passes=1)  # the FP64 polish or sweeps re-orthonormalise
    if sweeps is not None:
        if sweeps < 0:
            raise ValueError(f"sweeps must be >= 0, got {sweeps}")
        eig, x, it, rn = refine_real(ham.apply_real, xd.to(torch.float64), t_phi, int(sweeps))
    else:
        eig, x, it, rn = davidson_real(ham.apply_real, xd.to(torch.float64), t_phi, tol=tol, max_iter=max_iter,
                                       max_dim_factor=max_dim_factor, profile_out=profile_out, nconv=nconv)
    if buffer:
        eig, x, rn = eig[:nb], x[:nb], rn[:nb]
    return GammaDavidsonResult(eig, unembed_real(gb, x), it, rn, n_draft_iter=itd)
