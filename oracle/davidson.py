"""Batched block Davidson and the Teter preconditioner (CPU oracle).

Restates solvers/precond.py:8-39 (teter, teter_b), solvers/davidson.py:195-222
(_cholqr2), 287-308 (_certify_fp64), 406-486 (_orthonormalize_b, _rr, _combine),
511-774 (BatchedDavidsonResult, davidson_batched, synchronous gate), 777-812
(davidson_batched_ms), solvers/_ms.py:26-52 (mixed_precision_solve) and
solvers/registry.py:86-182 (EigResult, register/get, davidson_adapter).

The fp32-expansion mode, which the reference switches with the
GRADWAVE_FP32_EXPANSION environment variable read per solve
(davidson.py:268-284), is a keyword here (``fp32_expansion``); CholQR2
(GRADWAVE_CHOLQR, davidson.py:192) is the keyword ``cholqr``.

Pinned by: dense eigvalsh of explicit Hermitian operators (synthetic and the
plane-wave H), the empty-lattice/Mathieu closed forms, the certification
contract of test_davidson_fp32_expansion.py:45-114 (residuals re-measured in
fp64, orthonormality), and golden Davidson trajectories written from the
reference's davidson_batched (n_iter, eigenvalues, residual norms).
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, field
from typing import Any, NamedTuple

import torch


def teter(residual, t_g, t_band):
    """precond.py:8-17."""
    x = t_g[None, :] / torch.clamp(t_band[:, None], min=1e-12)
    x2 = x * x
    num = 27.0 + 18.0 * x + 12.0 * x2 + 8.0 * x2 * x
    return residual * (num / (num + 16.0 * x2 * x2))


def teter_b(residual, t_g, t_band):
    """precond.py:20-39 -- coefficient formed in float64 whatever the residual dtype."""
    x = t_g.double()[:, None, :] / torch.clamp(t_band.double()[..., None], min=1e-12)
    x2 = x * x
    num = 27.0 + 18.0 * x + 12.0 * x2 + 8.0 * x2 * x
    coeff = num / (num + 16.0 * x2 * x2)
    return residual * coeff.to(residual.real.dtype)


def _cholqr2(x: torch.Tensor) -> torch.Tensor | None:
    """davidson.py:195-222 -- two rounds of fp64 Gram -> Cholesky -> triangular solve."""
    cd = torch.complex128
    for _ in range(2):
        gram = x.mH.to(cd) @ x.to(cd)
        lo, info = torch.linalg.cholesky_ex(gram)
        if int(info.max()) != 0:
            return None
        x = torch.linalg.solve_triangular(lo.mH.to(x.dtype), x, upper=True, left=False)
    return x


def _eigh_subspace(s):
    """davidson.py:66-75 (CPU: no offload)."""
    return torch.linalg.eigh(s)


def _combine(c, block):
    """davidson.py:483-485."""
    return torch.einsum("kja,kjg->kag", c, block)


def _certify_fp64(h_apply, x):
    """davidson.py:287-308."""
    hx = h_apply(x)
    s = torch.matmul(x.conj(), hx.mT)
    s = 0.5 * (s + s.conj().transpose(-1, -2))
    w, u = _eigh_subspace(s)
    eig = w.real.to(x.real.dtype)
    u = u.to(x.dtype)
    x = _combine(u, x)
    hx = _combine(u, hx)
    r = hx - eig[..., None] * x
    rn = torch.linalg.norm(r, dim=-1).real
    return eig, x, hx, r, rn


def _orthonormalize_b(v, mask, against=None, cholqr: bool = False):
    """davidson.py:406-457 (jitter=None path)."""

    def project(x):
        if against is not None and against.shape[1]:
            for _ in range(2):
                x = x - (x @ against.conj().transpose(-1, -2)) @ against
        return x

    def orthonormalize(x):
        if cholqr:
            q = _cholqr2(x.transpose(-1, -2))
            if q is not None:
                return q.transpose(-1, -2)
        q, _ = torch.linalg.qr(x.transpose(-1, -2), mode="reduced")
        return q.transpose(-1, -2)

    v = project(v * mask[:, None, :])
    row_norm = torch.linalg.norm(v, dim=-1, keepdim=True).real
    degenerate = row_norm < 1e-8
    if bool(degenerate.any()):
        gen = torch.Generator(device="cpu").manual_seed(v.shape[1] + 7919)
        noise = torch.randn(*v.shape, 2, generator=gen, dtype=torch.float64)
        jit = torch.view_as_complex(noise).to(v.dtype)
        v = project((v + degenerate * jit) * mask[:, None, :])
    return orthonormalize(v)


_FP32_CERT_FLOOR_EPS = 32 * torch.finfo(torch.float32).eps  # davidson.py:258
_FP32_CERT_STALL = 0.7  # davidson.py:259


@dataclass
class BatchedDavidsonResult:
    """davidson.py:511-524."""

    eigenvalues: torch.Tensor
    eigenvectors: torch.Tensor
    n_iter: int
    residual_norms: torch.Tensor
    n_apply_low: int = 0
    n_apply_full: int = 0
    history: list = field(default_factory=list)


@torch.no_grad()
def davidson_batched(h_apply: Callable, x0: torch.Tensor, t: torch.Tensor, mask: torch.Tensor,
                     tol: float = 1e-9, max_iter: int = 40, max_dim_factor: int = 4,
                     fp32_expansion: bool = False, cholqr: bool = False,
                     history_out: list | None = None) -> BatchedDavidsonResult:
    """davidson.py:527-774, synchronous gate (sync_free=False)."""
    nk, nb, m = x0.shape
    max_dim = min(max_dim_factor * nb, int(mask.sum(dim=1).min()))
    rdtype = x0.real.dtype
    use_fp32_exp = fp32_expansion and x0.dtype == torch.complex128
    counts = {"low": 0, "full": 0}

    def _apply_full(z):
        counts["full"] += z.shape[0] * z.shape[1]
        return h_apply(z)

    def _apply_exp(z):
        if not use_fp32_exp:
            return _apply_full(z)
        counts["low"] += z.shape[0] * z.shape[1]
        return h_apply(z.to(torch.complex64)).to(z.dtype)

    cert_floor = _FP32_CERT_FLOOR_EPS * max(float(t.max()), 1.0) if use_fp32_exp else 0.0
    prev_rnmax = float("inf")

    v = _orthonormalize_b(x0, mask, cholqr=cholqr)
    hv = _apply_exp(v)
    eig = torch.zeros(nk, nb, dtype=rdtype)
    x = v[:, :nb]
    rn = torch.full((nk, nb), float("inf"), dtype=rdtype)

    for it in range(1, max_iter + 1):
        s = torch.matmul(v.conj(), hv.mT)
        s = (0.5 * (s + s.conj().transpose(-1, -2))).to(torch.complex128)
        w, u = _eigh_subspace(s)
        u = u[:, :, :nb].to(v.dtype)
        eig = w[:, :nb].real.to(rdtype)
        x = torch.einsum("kja,kjg->kag", u, v)
        hx = torch.einsum("kja,kjg->kag", u, hv)
        r = hx - eig[..., None] * x
        rn = torch.linalg.norm(r, dim=-1).real
        if use_fp32_exp:
            rnmax = float(rn.max())
            stalled = it > 1 and rnmax > _FP32_CERT_STALL * prev_rnmax
            prev_rnmax = rnmax
            if rnmax < max(tol, cert_floor) or stalled:
                eig, x, hx, r, rn = _certify_fp64(_apply_full, x)
                if float(rn.max()) >= tol:
                    v, hv = x, hx
                    use_fp32_exp = False
        if history_out is not None:
            history_out.append((it, rn.clone(), eig.clone()))
        if float(rn.max()) < tol:
            return BatchedDavidsonResult(eig, x, it, rn, counts["low"], counts["full"])
        n_add = int((rn > tol).sum(dim=1).max())
        sel = torch.argsort(rn, dim=1, descending=True)[:, :n_add]
        r_sel = torch.gather(r, 1, sel[..., None].expand(-1, -1, m))
        t_band = torch.einsum("kbg,kg,kbg->kb", x.conj(), t.to(x.dtype), x).real
        tb_sel = torch.gather(t_band, 1, sel)
        d = teter_b(r_sel, t, tb_sel)
        dn = torch.linalg.norm(d, dim=-1, keepdim=True).real
        d = torch.where(dn > 1e-300, d / dn.clamp_min(1e-300), d)
        if v.shape[1] + n_add > max_dim:
            q, rmat = torch.linalg.qr(x.transpose(-1, -2), mode="reduced")
            x_orth = q.transpose(-1, -2)
            hx_orth = torch.linalg.solve_triangular(rmat.transpose(-1, -2), hx, upper=False)
            d = _orthonormalize_b(d, mask, against=x_orth, cholqr=cholqr)
            v = torch.cat([x_orth, d], dim=1)
            hv = torch.cat([hx_orth, _apply_exp(d)], dim=1)
        else:
            d = _orthonormalize_b(d, mask, against=v, cholqr=cholqr)
            v = torch.cat([v, d], dim=1)
            hv = torch.cat([hv, _apply_exp(d)], dim=1)

    if use_fp32_exp:
        eig, x, hx, r, rn = _certify_fp64(_apply_full, x)
    return BatchedDavidsonResult(eig, x, max_iter, rn, counts["low"], counts["full"])


LOW = torch.complex64  # _ms.py:21


def mixed_precision_solve(x0, tol, crossover, mixed_precision, *, full, make_stages):
    """_ms.py:26-52."""
    if (not mixed_precision) or x0.dtype == LOW or crossover >= tol:
        return full()
    hi_dtype = x0.dtype
    draft, polish = make_stages()
    d = draft(x0.to(LOW))
    return polish(d.eigenvectors.to(hi_dtype))


def davidson_batched_ms(h_apply, x0, t, mask, tol=1e-9, max_iter=40, max_dim_factor=4,
                        crossover=1e-5, mixed_precision=True):
    """davidson.py:777-812 -- complex64 draft to `crossover`, fp64 polish to `tol`."""

    def solve(x, t_, tl):
        return davidson_batched(h_apply, x, t_, mask, tol=tl, max_iter=max_iter,
                                max_dim_factor=max_dim_factor)

    return mixed_precision_solve(
        x0, tol, crossover, mixed_precision,
        full=lambda: solve(x0, t, tol),
        make_stages=lambda: (lambda xlo: solve(xlo, t.to(torch.float32), crossover),
                             lambda x1: solve(x1, t, tol)))


class EigResult(NamedTuple):
    """registry.py:86-93."""

    eigenvalues: torch.Tensor
    eigenvectors: torch.Tensor
    n_iter: int
    residual_norms: torch.Tensor
    diagnostics: dict[str, Any]


def davidson_adapter(apply_H, X0, precond, mask, *, tol, nbands=None, max_iter=40,
                     max_dim_factor=4, **kw) -> EigResult:
    """registry.py:154-182."""
    r = davidson_batched(apply_H, X0, precond, mask, tol=tol, max_iter=max_iter,
                         max_dim_factor=max_dim_factor)
    return EigResult(r.eigenvalues, r.eigenvectors, r.n_iter, r.residual_norms,
                     {"solver": "davidson", "max_dim_factor": max_dim_factor,
                      "max_iter": max_iter, "hit_max_iter": r.n_iter >= max_iter,
                      "n_apply_low": r.n_apply_low, "n_apply_full": r.n_apply_full})
