"""Drop-in batched block Davidson on MI355X (solvers/davidson.py:527-812).

Same signature, schedule and result type as the reference ``davidson_batched``:
uniform subspace across k, Rayleigh-Ritz every round with the subspace matrix
and its eigensolve in fp64 (davidson.py:640-652), the n_add worst residuals
expanded with the Teter filter and unit-normalised (729-743), two-pass
projection against the basis, restart at max_dim reusing H x through the
triangular factor (745-759), and the fp64-certified fp32-expansion mode
(GRADWAVE_FP32_EXPANSION, 268-308, 658-687, 765-768).

MI355X-native differences that leave the span and the fp64 contract unchanged:
* the basis V and its images HV live in preallocated (nk, max_dim, npw) ring
  buffers instead of being rebuilt by torch.cat every round (763);
* every contraction over npw (V^H HV, Ritz combination, projections, Gram
  matrices) is an MFMA kernel of libgradwave_hip, and V^H HV grows by block
  columns instead of being rebuilt each round;
* residuals, norms and the Teter filter are fused kernels with fp64 reductions;
* orthonormalisation is CholQR2 (the reference's GRADWAVE_CHOLQR path,
  davidson.py:195-222) with Householder QR as the breakdown fallback.
"""

from __future__ import annotations

import os
from collections.abc import Callable
from dataclasses import dataclass
from typing import Any, NamedTuple

import torch

from gradwave_amd import linalg as LA

_FP32_CERT_FLOOR_EPS = 32 * torch.finfo(torch.float32).eps  # davidson.py:258
_FP32_CERT_STALL = 0.7  # davidson.py:259
# Subspace eigensolves at or below this size go to host LAPACK (the reference
# offloads m <= 256 on CUDA, davidson.py:58-75); measured on MI355X, rocSOLVER
# wins from m ~ 160 up, host LAPACK below.
_SUBSPACE_CPU_MAX = int(os.environ.get("GW_EIGH_CPU_MAX", "64"))
# Requests for the lowest nb pairs of GPU matrices up to this size go to the
# batched subspace eigensolver (gw_heev_batched: every k of the batch in one
# launch); GW_SUBSPACE_EIGH=library keeps the policy above for them too.
_HEEV_MAX = int(os.environ.get("GW_HEEV_MAX", "256"))


def _eigh_low(s: torch.Tensor, nev: int):
    """Lowest nev eigenpairs of the (nk, m, m) subspace matrix: the batched
    kernel for GPU matrices with m <= GW_HEEV_MAX, else the first nev pairs of
    ``_eigh``."""
    mode = os.environ.get("GW_SUBSPACE_EIGH", "kernel").strip().lower()
    if mode not in ("kernel", "library"):
        raise ValueError(f"GW_SUBSPACE_EIGH must be kernel|library, got {mode!r}")
    if mode == "kernel" and s.is_cuda and s.shape[-1] <= min(_HEEV_MAX, LA.heev_max_n()):
        return LA.eigh_lowest([s.to(torch.complex128)], [nev])[0]
    w, u = _eigh(s)
    return w[..., :nev], u[..., :nev]


class PhaseTimer:
    """Wall time per Davidson phase (device-synchronising); on when a dict is given."""

    def __init__(self, out: dict | None) -> None:
        self.out = out
        self.t = None
        self.key = None

    def __call__(self, key: str) -> None:
        if self.out is None:
            return
        import time

        torch.cuda.synchronize()
        now = time.perf_counter()
        if self.key is not None:
            self.out[self.key] = self.out.get(self.key, 0.0) + now - self.t
        self.key, self.t = key, now


@dataclass
class BatchedDavidsonResult:
    """solvers/davidson.py:511-524."""

    eigenvalues: torch.Tensor  # (nk, nb) ascending
    eigenvectors: torch.Tensor  # (nk, nb, npw_max), padded slots zero
    n_iter: int
    residual_norms: torch.Tensor  # (nk, nb)
    n_apply_low: int = 0
    n_apply_full: int = 0


def _fp32_expansion_active(x0: torch.Tensor) -> bool:
    """davidson.py:268-284 (the "auto" gate compares fp64/fp32 GEMM speed; on
    MI355X both MFMA peaks are within 2x, so "auto" stays off like the
    reference's measurement for data-center fp64 hardware)."""
    mode = os.environ.get("GRADWAVE_FP32_EXPANSION", "off").strip().lower()
    if mode not in ("auto", "on", "off"):
        raise ValueError(f"GRADWAVE_FP32_EXPANSION must be auto|on|off, got {mode!r}")
    return mode == "on" and x0.dtype == torch.complex128


def _eigh(s: torch.Tensor):
    """Hermitian eigensolve of the (nk, m, m) complex128 subspace matrix on the GPU
    (torch.linalg.eigh -> rocSOLVER); optional CPU LAPACK offload below
    GW_EIGH_CPU_MAX (the reference's davidson.py:61-75 policy)."""
    if s.shape[-1] <= _SUBSPACE_CPU_MAX:
        w, u = torch.linalg.eigh(s.cpu())
        return w.to(s.device), u.to(s.device)
    return torch.linalg.eigh(s)


def _hermitize(s: torch.Tensor) -> torch.Tensor:
    return (0.5 * (s + s.conj().transpose(-1, -2))).to(torch.complex128)


def _gram(v: torch.Tensor) -> torch.Tensor:
    """G[k][i][j] = sum_g v_i conj(v_j) in complex128, Hermitised. A complex64
    block is upcast before the contraction, as davidson.py:213-216 does, so the
    products of its FP32 values are exact and the sums FP64."""
    w = v if v.dtype == torch.complex128 else v.to(torch.complex128)
    g = LA.inner(w, w).conj()
    return 0.5 * (g + g.mH)


def _cholinv_torch(g: torch.Tensor):
    """(L^-1, info) of G = L L^H with torch (rocSOLVER potrf, rocBLAS trsm)."""
    lo, info = torch.linalg.cholesky_ex(g)
    eye = torch.eye(g.shape[-1], dtype=torch.complex128, device=g.device)
    return torch.linalg.solve_triangular(lo, eye.expand_as(lo), upper=False), info


def _cholqr2_steps(v: torch.Tensor):
    """Orthonormal rows spanning v (nk, n, npw) by two rounds of fp64 Gram ->
    Cholesky -> triangular inverse -> combination (davidson.py:195-222 in row form:
    G = v v^H = L L^H, q = L^-1 v). A step generator: each round's L^-1 is a
    ("cholinv", G) request. Returns (q, breakdown flags on the device)."""
    infos = []
    for _ in range(2):
        linv, info = yield ("cholinv", _gram(v))
        infos.append(info.max())
        v = LA.combine(linv.transpose(-1, -2).to(v.dtype).contiguous(), v)
    return v, torch.stack(infos)


def _cholqr2_dev(v: torch.Tensor):
    return drive(_cholqr2_steps(v), None)


def _project(v: torch.Tensor, against: torch.Tensor | None, passes: int = 2) -> torch.Tensor:
    """``passes`` (two in the reference, davidson.py:428-432) of v -= (v a^H) a,
    in place: callers pass a block of their own (the masked copy made by the
    orthonormalisation)."""
    if against is None or against.shape[1] == 0:
        return v
    for _ in range(passes):
        coef = LA.inner(against, v)  # coef[j][i] = sum conj(a_j) v_i = (v a^H)[i][j]
        LA.combine((-coef).contiguous(), against, out=v, beta=1.0)
    return v


def _orthonormalize_steps(v: torch.Tensor, mask: torch.Tensor, against: torch.Tensor | None = None,
                          passes: int = 2):
    """davidson.py:406-457: mask, project, deterministic jitter on degenerate rows,
    orthonormalise (CholQR2 first, Householder QR on breakdown). A step
    generator: its two host reads are ("host", tensor) requests."""
    v = _project((v * mask[:, None, :]).contiguous(), against, passes)
    rn = LA.row_norms(v)
    degenerate = rn < 1e-8
    # CholQR2 runs before the degenerate-row test is read, so both tests cost one host read
    q, infos = yield from _cholqr2_steps(v)
    flags = yield ("host", torch.cat([degenerate.any().reshape(1).to(infos.dtype), infos]))
    if bool(flags[0]):
        gen = torch.Generator(device="cpu").manual_seed(v.shape[1] + 7919)
        noise = torch.randn(*v.shape, 2, generator=gen, dtype=torch.float64)
        jit = torch.view_as_complex(noise).to(v.device).to(v.dtype)
        v = _project(((v + degenerate[..., None] * jit) * mask[:, None, :]).contiguous(), against, passes)
        q, infos = yield from _cholqr2_steps(v)
        flags = yield ("host", torch.cat([flags[:1].to(infos.device, infos.dtype), infos]))
    if bool((flags[1:] != 0).any()):
        qq, _ = torch.linalg.qr(v.transpose(-1, -2), mode="reduced")
        q = qq.transpose(-1, -2).contiguous()
    return q


def _orthonormalize_b(v: torch.Tensor, mask: torch.Tensor,
                      against: torch.Tensor | None = None) -> torch.Tensor:
    return drive(_orthonormalize_steps(v, mask, against), None)


class _Basis:
    """Preallocated (nk, max_dim, npw) basis V, images HV and the fp64 subspace
    matrix s = V^H H V, grown by block columns.

    The reference rebuilds the whole m x m matrix every round
    (solvers/davidson.py:648-649); here only the new block column
    C = V^H (H D) is contracted when directions D are appended, and s is
    completed Hermitian as [[s, C_top], [C_top^H, (C_bot + C_bot^H)/2]]."""

    def __init__(self, nk: int, max_dim: int, m: int, dtype, device) -> None:
        self.v = torch.zeros(nk, max_dim, m, dtype=dtype, device=device)
        self.hv = torch.zeros(nk, max_dim, m, dtype=dtype, device=device)
        self.s = torch.zeros(nk, max_dim, max_dim, dtype=torch.complex128, device=device)
        self.n = 0

    def views(self):
        return self.v[:, : self.n], self.hv[:, : self.n]

    def subspace(self) -> torch.Tensor:
        return self.s[:, : self.n, : self.n]

    def set(self, v: torch.Tensor, hv: torch.Tensor) -> None:
        k = v.shape[1]
        self.v[:, :k].copy_(v)
        self.hv[:, :k].copy_(hv)
        self.n = k
        self.s[:, :k, :k] = _hermitize(LA.inner(self.v[:, :k], self.hv[:, :k]))

    def append(self, v: torch.Tensor, hv: torch.Tensor) -> None:
        k = v.shape[1]
        n0, n1 = self.n, self.n + k
        self.v[:, n0:n1].copy_(v)
        self.hv[:, n0:n1].copy_(hv)
        c = LA.inner(self.v[:, :n1], self.hv[:, n0:n1]).to(torch.complex128)  # (nk, n1, k)
        self.s[:, :n0, n0:n1] = c[:, :n0]
        self.s[:, n0:n1, :n0] = c[:, :n0].mH
        self.s[:, n0:n1, n0:n1] = _hermitize(c[:, n0:])
        self.n = n1


def _certify_steps(x, counts):
    """davidson.py:287-308: re-measure the block with the FP64 operator (a step
    generator: yields ("apply", z) / ("eigh", s) requests)."""
    counts["full"] += x.shape[0] * x.shape[1]
    hx = yield ("apply", x)
    s = _hermitize(LA.inner(x, hx))
    w, u = yield ("eigh", s)
    eig = w.real.to(x.real.dtype)
    u = u.to(x.dtype).contiguous()
    x = LA.combine(u, x)
    hx = LA.combine(u, hx)
    r, rn = LA.residual(x, hx, eig)
    return eig, x, hx, r, rn


def _apply_steps(z, low, counts):
    if not low:
        counts["full"] += z.shape[0] * z.shape[1]
        return (yield ("apply", z))
    counts["low"] += z.shape[0] * z.shape[1]
    hz = yield ("apply", z.to(torch.complex64))
    return hz.to(z.dtype)


def drive(steps, h_apply):
    """Run a step generator to completion, serving its ("apply", block) requests
    with ``h_apply``, its ("eigh", matrix) requests with ``_eigh``, its
    ("eigh_low", (matrix, nev)) requests (only the lowest nev pairs are used)
    with ``_eigh_low``, its ("cholinv", G) requests with
    ``_cholinv_torch`` and its ("host", tensor) requests with a copy to the
    host."""
    try:
        req = next(steps)
        while True:
            kind, arg = req
            if kind == "apply":
                resp = h_apply(arg)
            elif kind == "eigh":
                resp = _eigh(arg)
            elif kind == "eigh_low":
                resp = _eigh_low(arg[0], arg[1])
            elif kind == "cholinv":
                resp = _cholinv_torch(arg)
            else:
                resp = arg.cpu()
            req = steps.send(resp)
    except StopIteration as stop:
        return stop.value


@torch.no_grad()
def davidson_batched(h_apply: Callable[[torch.Tensor], torch.Tensor], x0: torch.Tensor,
                     t: torch.Tensor, mask: torch.Tensor, tol: float = 1e-9, max_iter: int = 40,
                     max_dim_factor: int = 4, sync_free: bool = False,
                     history_out: list | None = None,
                     profile_out: dict | None = None, nconv: int | None = None,
                     proj_passes: int = 2) -> BatchedDavidsonResult:
    """Batched block Davidson over all k at once -- solvers/davidson.py:527-774.

    ``sync_free`` is accepted for signature parity; the synchronous gate is
    always used (the reference measured the delayed gate as a loss,
    davidson.py:545-554). ``profile_out`` (not in the reference) collects wall
    seconds per phase. ``nconv`` (not in the reference; default: every band)
    stops the solve once the first ``nconv`` bands are converged, the rest
    being buffer bands that are still expanded but do not gate.
    ``proj_passes`` (not in the reference, which projects twice) is the number
    of projections of the expansion block against the basis; the mixed
    solvers' FP32 drafts use one, their FP64 polish or sweeps two."""
    return drive(davidson_steps(x0, t, mask, tol=tol, max_iter=max_iter, max_dim_factor=max_dim_factor,
                                history_out=history_out, profile_out=profile_out, nconv=nconv,
                                proj_passes=proj_passes), h_apply)


@torch.no_grad()
def davidson_steps(x0: torch.Tensor, t: torch.Tensor, mask: torch.Tensor, tol: float = 1e-9,
                   max_iter: int = 40, max_dim_factor: int = 4, history_out: list | None = None,
                   profile_out: dict | None = None, nconv: int | None = None, proj_passes: int = 2):
    """The solver of ``davidson_batched`` as a step generator: it yields
    ("apply", block) and ("eigh_low", (subspace matrix, nb)) requests and returns the
    BatchedDavidsonResult, so a driver can serve many independent solves in
    lockstep with one batched apply and one batched eigensolve per round."""
    nk, nb, m = x0.shape
    ng = nb if nconv is None else max(1, min(int(nconv), nb))  # bands that gate convergence
    tm = PhaseTimer(profile_out)
    device = x0.device
    max_dim = min(max_dim_factor * nb, int(mask.sum(dim=1).min()))
    rdtype = x0.real.dtype
    use_fp32_exp = _fp32_expansion_active(x0)
    counts = {"low": 0, "full": 0}

    cert_floor = _FP32_CERT_FLOOR_EPS * max(float(t.max()), 1.0) if use_fp32_exp else 0.0
    prev_rnmax = float("inf")
    # a restart keeps nb rows and appends up to nb more, so the buffers hold
    # max(max_dim, 2 nb) rows: max_dim < 2 nb is legal (max_dim_factor=1, or
    # a k-point with fewer than 2 nb plane waves), and the reference's
    # concatenation grows past max_dim there too (davidson.py:757-759)
    basis = _Basis(nk, max(max_dim, 2 * nb), m, x0.dtype, device)

    v0 = yield from _orthonormalize_steps(x0.contiguous(), mask)
    hv0 = yield from _apply_steps(v0, use_fp32_exp, counts)
    basis.set(v0, hv0)
    eig = torch.zeros(nk, nb, dtype=rdtype, device=device)
    x = v0[:, :nb]
    rn = torch.full((nk, nb), float("inf"), dtype=torch.float64, device=device)

    for it in range(1, max_iter + 1):
        v, hv = basis.views()
        tm("eigh")
        w, u = yield ("eigh_low", (basis.subspace(), nb))
        u = u[:, :, :nb].to(v.dtype).contiguous()
        eig = w[:, :nb].real.to(rdtype)
        tm("ritz_gemm")
        x = LA.combine(u, v)
        hx = LA.combine(u, hv)
        tm("residual")
        r, rn = LA.residual(x, hx, eig)
        if use_fp32_exp:
            rnmax = float(rn[:, :ng].max())
            stalled = it > 1 and rnmax > _FP32_CERT_STALL * prev_rnmax
            prev_rnmax = rnmax
            if rnmax < max(tol, cert_floor) or stalled:
                eig, x, hx, r, rn = yield from _certify_steps(x, counts)
                if float(rn[:, :ng].max()) >= tol:
                    basis.set(x, hx)
                    use_fp32_exp = False
        if history_out is not None:
            history_out.append((it, rn.detach().to("cpu").clone(), eig.detach().to("cpu").clone()))
        gate = yield ("host", torch.stack([rn[:, :ng].max(), (rn > tol).sum(dim=1).max().to(rn.dtype)]))
        if float(gate[0]) < tol:
            tm(None)
            return BatchedDavidsonResult(eig, x, it, rn.to(rdtype), counts["low"], counts["full"])
        n_add = int(gate[1])
        sel = torch.argsort(rn, dim=1, descending=True)[:, :n_add]
        d = LA.teter_expand(x, r, t, sel)
        tm("restart")
        if basis.n + n_add > max_dim:
            # restart on the Ritz block with its drift repaired: x = L q, hx -> L^-1 hx
            linv, info = yield ("cholinv", _gram(x))
            bad = yield ("host", info.max().reshape(1))
            if int(bad[0]) == 0:
                ut = linv.transpose(-1, -2).to(x.dtype).contiguous()
                x_orth, hx_orth = LA.combine(ut, x), LA.combine(ut, hx)
            else:
                q, rmat = torch.linalg.qr(x.transpose(-1, -2), mode="reduced")
                x_orth = q.transpose(-1, -2).contiguous()
                hx_orth = torch.linalg.solve_triangular(rmat.transpose(-1, -2), hx, upper=False)
            basis.set(x_orth, hx_orth)
        v, _ = basis.views()
        tm("orthonormalize")
        d = yield from _orthonormalize_steps(d, mask, against=v, passes=proj_passes)
        tm("h_apply")
        hd = yield from _apply_steps(d, use_fp32_exp, counts)
        tm("subspace_gemm")
        basis.append(d, hd)

    if use_fp32_exp:
        eig, x, hx, r, rn = yield from _certify_steps(x, counts)
    tm(None)
    return BatchedDavidsonResult(eig, x, max_iter, rn.to(rdtype), counts["low"], counts["full"])


LOW = torch.complex64


@torch.no_grad()
def davidson_batched_ms(h_apply, x0, t, mask, tol=1e-9, max_iter=40, max_dim_factor=4,
                        crossover=1e-5, mixed_precision=True,
                        profile_out: dict | None = None) -> BatchedDavidsonResult:
    """Two-stage Davidson (davidson.py:777-812, _ms.py:26-52): complex64 draft to
    ``crossover``, then a complex128 polish to ``tol`` warm-started from it."""
    if (not mixed_precision) or x0.dtype == LOW or crossover >= tol:
        return davidson_batched(h_apply, x0, t, mask, tol=tol, max_iter=max_iter,
                                max_dim_factor=max_dim_factor)
    hi = x0.dtype
    d = davidson_batched(h_apply, x0.to(LOW), t.to(torch.float32), mask, tol=crossover,
                         max_iter=max_iter, max_dim_factor=max_dim_factor, profile_out=profile_out)
    p = davidson_batched(h_apply, d.eigenvectors.to(hi), t, mask, tol=tol, max_iter=max_iter,
                         max_dim_factor=max_dim_factor, profile_out=profile_out)
    p.n_draft_iter = d.n_iter
    p.n_apply_low += d.n_apply_full
    return p


@torch.no_grad()
def _refine_steps(x0: torch.Tensor, t: torch.Tensor, mask: torch.Tensor, sweeps: int):
    """``sweeps`` FP64 Davidson sweeps from drafted vectors, with no convergence
    gate: Rayleigh-Ritz of the orthonormalised block with the FP64 operator, then
    per sweep every band's Teter-preconditioned residual (davidson.py:729-743)
    joins the basis and Rayleigh-Ritz runs over the grown basis. The returned
    residual norms are FP64 norms of the FP64 operator; the last step applies H
    to nothing it does not use."""
    nk, nb, m = x0.shape
    counts = {"low": 0, "full": 0}
    basis = _Basis(nk, nb * (sweeps + 1), m, x0.dtype, x0.device)
    v = yield from _orthonormalize_steps(x0.contiguous(), mask)
    hv = yield from _apply_steps(v, False, counts)
    basis.set(v, hv)
    sel = torch.arange(nb, device=x0.device).expand(nk, nb).contiguous()
    for sweep in range(sweeps + 1):
        v, hv = basis.views()
        w, u = yield ("eigh_low", (basis.subspace(), nb))
        u = u[:, :, :nb].to(v.dtype).contiguous()
        eig = w[:, :nb].real.to(x0.real.dtype)
        x, hx = LA.combine(u, v), LA.combine(u, hv)
        r, rn = LA.residual(x, hx, eig)
        if sweep == sweeps:
            break
        d = LA.teter_expand(x, r, t, sel)
        d = yield from _orthonormalize_steps(d, mask, against=v)
        hd = yield from _apply_steps(d, False, counts)
        basis.append(d, hd)
    return BatchedDavidsonResult(eig, x, sweeps, rn.to(x0.real.dtype), 0, counts["full"])


BUFFER_SEED = 4241


def buffer_block(nk: int, nbuf: int, m: int, mask: torch.Tensor | None, device) -> torch.Tensor:
    """``nbuf`` seeded complex128 start vectors per k (generator seed BUFFER_SEED,
    padded slots zero) for the buffer bands of the precision form."""
    g = torch.Generator().manual_seed(BUFFER_SEED)
    x = torch.randn(nk, nbuf, m, dtype=torch.complex128, generator=g).to(device)
    return x * mask[:, None, :].to(x.dtype) if mask is not None else x


@torch.no_grad()
def davidson_mixed(h_apply, x0, t, mask, tol: float = 1e-9, draft_tol: float = 1e-4,
                   max_iter: int = 100, max_dim_factor: int = 4, sweeps: int | None = None,
                   profile_out: dict | None = None, buffer: int = 0) -> BatchedDavidsonResult:
    """Precision form: the draft solve runs on complex64 blocks -- FP32 FFTs, V(r)
    multiply and MFMA contractions with FP32 accumulation -- to ``draft_tol``,
    keeping the subspace matrix, its eigensolve, the Cholesky factors and every
    residual norm in FP64. With ``sweeps=None`` the polish solve then runs in
    complex128 to ``tol`` from the drafted Ritz vectors, so convergence is
    certified by FP64 residuals of an FP64 operator; with ``sweeps=n`` the draft
    is refined by n FP64 Davidson sweeps instead (no gate; the result carries
    their FP64 residual norms, and ``tol`` is unused). The reference's
    draft/polish scaffold is solvers/_ms.py:26-52 (davidson_batched_ms gates it
    on crossover < tol). ``buffer`` > 0 carries that many extra bands (seeded
    start vectors, ``buffer_block``) through draft and sweeps and returns only
    the requested ones: the last bands of a block then converge like interior
    bands (in the sweep form they otherwise carry the largest eigenvalue error);
    only the requested bands gate the draft's (and the polish's) convergence."""
    if buffer < 0:
        raise ValueError(f"buffer must be >= 0, got {buffer}")
    nb = x0.shape[1]
    if buffer:
        x0 = torch.cat([x0, buffer_block(x0.shape[0], buffer, x0.shape[2], mask, x0.device).to(x0.dtype)], dim=1)
    hi = x0.dtype
    nconv = nb if buffer else None  # buffer bands are expanded but do not gate
    # one projection per round in the FP32 draft: the FP64 polish or sweeps
    # re-orthonormalise against their own basis (Si216: 0.12 s of a 2.8 s
    # sweep-form solve, same draft iterations and eigenvalue error)
    d = davidson_batched(h_apply, x0.to(LOW), t.to(torch.float32), mask, tol=draft_tol,
                         max_iter=max_iter, max_dim_factor=max_dim_factor, profile_out=profile_out, nconv=nconv,
                         proj_passes=1)
    if sweeps is not None:
        if sweeps < 0:
            raise ValueError(f"sweeps must be >= 0, got {sweeps}")
        p = drive(_refine_steps(d.eigenvectors.to(hi), t, mask, int(sweeps)), h_apply)
    else:
        p = davidson_batched(h_apply, d.eigenvectors.to(hi), t, mask, tol=tol, max_iter=max_iter,
                             max_dim_factor=max_dim_factor, profile_out=profile_out, nconv=nconv)
    p.n_apply_low += d.n_apply_full
    p.n_draft_iter = d.n_iter
    if buffer:
        p.eigenvalues, p.eigenvectors = p.eigenvalues[:, :nb], p.eigenvectors[:, :nb]
        p.residual_norms = p.residual_norms[:, :nb]
    return p


class EigResult(NamedTuple):
    """solvers/registry.py:86-93."""

    eigenvalues: torch.Tensor
    eigenvectors: torch.Tensor
    n_iter: int
    residual_norms: torch.Tensor
    diagnostics: dict[str, Any]
