"""Conduction-projected Sternheimer solves on MI355X (postscf/_response.py:502-681).

``cg_sternheimer(h, bk, c_occ, eps_occ, rhs, x0, shift)`` solves, for every
occupied band n of every k at once,

    (H - eps_n + shift * P_occ) dpsi = rhs,    P_occ = sum_m |c_m><c_m|,

with the Teter-preconditioned conjugate gradient of the reference and returns
(1 - P_occ) dpsi. As in the reference, when the operator has ``apply_cols``
the iteration is per (k, band) column: full-block steps while more than a
quarter of the columns are active, then the surviving columns are gathered
into one compacted batch, and up to three true-residual rounds certify every
column to ``tol``. The H applies are the HIP kernels; the occupied-space
projections of the full block are MFMA contractions (``linalg.inner`` /
``combine``). Norm-conserving operators only (the S-metric branch of the
reference is not implemented here).
"""

from __future__ import annotations

import torch

from gradwave_amd import linalg as LA


def _teter_b(r: torch.Tensor, t: torch.Tensor, t_band: torch.Tensor) -> torch.Tensor:
    """solvers/precond.py:20-39: K(t_G / <T>) r with the coefficient in float64."""
    x = t.double()[:, None, :] / torch.clamp(t_band.double()[..., None], min=1e-12)
    x2 = x * x
    num = 27.0 + 18.0 * x + 12.0 * x2 + 8.0 * x2 * x
    return r * (num / (num + 16.0 * x2 * x2)).to(r.real.dtype)


def _dot(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Re sum_g conj(a) b over the last axis."""
    return (a.conj() * b).real.sum(-1)


@torch.no_grad()
def cg_sternheimer(h, bk, c_occ: torch.Tensor, eps_occ: torch.Tensor, rhs: torch.Tensor, x0: torch.Tensor,
                   shift: float, tol: float = 1e-8, max_iter: int = 400) -> torch.Tensor:
    """Batched conduction-projected Sternheimer solve (postscf/_response.py:502-595);
    c_occ, rhs, x0: (nk, nocc, npw_max) complex, eps_occ (nk, nocc)."""

    def p_occ(x: torch.Tensor) -> torch.Tensor:
        ov = LA.inner(c_occ, x)  # (nk, nocc, nb): ov[n][b] = <c_n|x_b>
        return LA.combine(ov, c_occ)  # sum_n ov[n][b] c_n

    def a_apply(x: torch.Tensor) -> torch.Tensor:
        hx = h.apply(x) - eps_occ[..., None] * x
        return hx - p_occ(hx) + shift * p_occ(x)

    mask = getattr(bk, "mask", None)
    if mask is not None:
        x0 = x0 * mask[:, None, :].to(x0.dtype)
    t_band = torch.clamp((c_occ.conj() * bk.t.to(c_occ.dtype)[:, None, :] * c_occ).real.sum(-1), min=1e-6)
    if hasattr(h, "apply_cols"):
        return _cg_percolumn(h, bk, c_occ, eps_occ, rhs, x0, shift, t_band, tol, max_iter, p_occ, a_apply)

    x = x0.clone()
    r = rhs - a_apply(x)
    z = _teter_b(r, bk.t, t_band)
    p = z
    rz = _dot(r, z)
    for _ in range(max_iter):
        ap = a_apply(p)
        a_cg = rz / torch.clamp(_dot(p, ap), min=1e-300)
        x = x + a_cg[..., None] * p
        r = r - a_cg[..., None] * ap
        if float(torch.linalg.norm(r, dim=-1).max()) < tol:
            break
        z = _teter_b(r, bk.t, t_band)
        rz_new = _dot(r, z)
        p = z + (rz_new / torch.clamp(rz, min=1e-300))[..., None] * p
        rz = rz_new
    return x - p_occ(x)


def _cg_percolumn(h, bk, c_occ, eps_occ, rhs, x0, shift, t_band, tol, max_iter, p_occ, a_apply):
    """postscf/_response.py:598-681: lockstep while > 1/4 of the columns are
    active, then a compacted column batch through ``h.apply_cols``; up to three
    true-residual rounds."""
    nk, nb, _ = rhs.shape
    total = nk * nb
    x = x0.clone()
    used = 0
    for _round in range(3):
        r = rhs - a_apply(x)
        rn = torch.linalg.norm(r, dim=-1)
        if used >= max_iter or float(rn.max()) < tol:
            break
        z = _teter_b(r, bk.t, t_band)
        p = z
        rz = _dot(r, z)
        while used < max_iter and int((rn >= tol).sum()) > total // 4:
            ap = a_apply(p)
            a_cg = rz / torch.clamp(_dot(p, ap), min=1e-300)
            x = x + a_cg[..., None] * p
            r = r - a_cg[..., None] * ap
            used += 1
            rn = torch.linalg.norm(r, dim=-1)
            if float(rn.max()) < tol:
                break
            z = _teter_b(r, bk.t, t_band)
            rz_new = _dot(r, z)
            p = z + (rz_new / torch.clamp(rz, min=1e-300))[..., None] * p
            rz = rz_new
        if used >= max_iter or float(rn.max()) < tol:
            continue
        kcol, bcol = (rn >= tol).nonzero(as_tuple=True)
        xa, ra, pa = x[kcol, bcol], r[kcol, bcol], p[kcol, bcol]
        rza = rz[kcol, bcol]
        c_sel = c_occ[kcol]
        eps_col = eps_occ[kcol, bcol]
        t_sel = bk.t[kcol]
        tb_sel = t_band[kcol, bcol]
        while used < max_iter and kcol.numel():
            hv = h.apply_cols(pa, kcol) - eps_col[:, None] * pa
            ov = torch.einsum("cng,cg->cn", c_sel.conj(), hv)
            hv = hv - torch.einsum("cn,cng->cg", ov, c_sel)
            ov2 = torch.einsum("cng,cg->cn", c_sel.conj(), pa)
            ap_a = hv + shift * torch.einsum("cn,cng->cg", ov2, c_sel)
            al = rza / torch.clamp(_dot(pa, ap_a), min=1e-300)
            xa = xa + al[:, None] * pa
            ra = ra - al[:, None] * ap_a
            used += 1
            done = torch.linalg.norm(ra, dim=-1) < tol
            if bool(done.any()):
                x[kcol[done], bcol[done]] = xa[done]
                keep = ~done
                kcol, bcol = kcol[keep], bcol[keep]
                xa, ra, pa, rza = xa[keep], ra[keep], pa[keep], rza[keep]
                c_sel, eps_col = c_sel[keep], eps_col[keep]
                t_sel, tb_sel = t_sel[keep], tb_sel[keep]
                if not kcol.numel():
                    break
            za = _teter_b(ra[:, None, :], t_sel, tb_sel[:, None])[:, 0]
            rz_new_a = _dot(ra, za)
            pa = za + (rz_new_a / torch.clamp(rza, min=1e-300))[:, None] * pa
            rza = rz_new_a
        if kcol.numel():
            x[kcol, bcol] = xa
    return x - p_occ(x)
