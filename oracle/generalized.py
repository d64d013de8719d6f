"""S-metric generalized block Davidson for H x = e S x (CPU oracle; test infrastructure).

Restates scf/uspp_batch.py:27-72 (BatchedHS: H with the screened D swapped into
the batched operator, S = 1 + sum_pq |beta_p> q_pq <beta_q| as one becp
contraction, masked) and 99-233 (davidson_gen_batched: standard-orthonormal
basis; fp64 subspace matrices; Cholesky reduction L^-1 H L^-H with the oldest
basis rows dropped while the overlap is not positive definite and a 1e-10
shift as the last resort; eigenvectors back-solved with L^H; Ritz rotation of
the cached H and S applies; residual H x - e S x; worst-residual expansion of
the same count on every k; Teter preconditioner on the band kinetic energy;
unit-normalised rows; QR restart in the standard metric rotating the cached
applies with the triangular factor).

Pinned by: the eigenvalues of the dense generalized problem with the explicit
plane-wave H (oracle.dense.dense_h) and S = 1 + p^T q^T conj(p), solved by
scipy.linalg.eigh(H, S) (tests/test_oracle_generalized.py); S applied here
against the dense S; S = 1 for q = 0, where the generalized solve must give
the standard Davidson's eigenvalues.
"""

from __future__ import annotations

import dataclasses

import torch

from oracle.davidson import _orthonormalize_b, teter_b
from oracle.hamiltonian import BatchedHamiltonian, becp_b


class BatchedHS:
    """uspp_batch.py:27-72 (the dual grid and DFT+U through BatchedHamiltonian;
    no meta-GGA term)."""

    def __init__(self, bk, shape, v_eff_r, p, dscr, q_full, hub_sphi=None, hub_d=None, smooth=None) -> None:
        self.bk = bk
        self.shape = shape
        self.p = p
        self.q = q_full.to(p.dtype)
        self.ham = BatchedHamiltonian(dataclasses.replace(bk, dij_full=dscr), shape, v_eff_r, p,
                                      hub_q=hub_sphi, hub_dij=hub_d, smooth=smooth)
        self.t = bk.t

    def h(self, c: torch.Tensor) -> torch.Tensor:
        return self.ham.apply(c)

    def s(self, c: torch.Tensor) -> torch.Tensor:
        p = self.p.to(c.dtype)
        b = becp_b(p, c)
        return (c + torch.einsum("kbp,pq,kqg->kbg", b, self.q.to(c.dtype), p)) * self.bk.mask[:, None, :]


def _contract(x, hx, sx):
    """uspp_batch.py:116-123 -- QR of the Ritz block in the standard metric;
    H x and S x rotate with the triangular factor."""
    q, rmat = torch.linalg.qr(x.transpose(-1, -2), mode="reduced")
    rt = rmat.transpose(-1, -2)
    return (q.transpose(-1, -2).contiguous(),
            torch.linalg.solve_triangular(rt, hx, upper=False),
            torch.linalg.solve_triangular(rt, sx, upper=False))


def _subspace(v, hv, sv):
    """uspp_batch.py:125-138 -- fp64 Hermitian subspace matrices."""
    vc = v.conj()
    h_sub = torch.matmul(vc, hv.mT).to(torch.complex128)
    s_sub = torch.matmul(vc, sv.mT).to(torch.complex128)
    return 0.5 * (h_sub + h_sub.mH), 0.5 * (s_sub + s_sub.mH)


def davidson_gen_batched(hs, x0: torch.Tensor, nbands: int, tol: float, max_iter: int = 60,
                         max_dim_factor: int = 4):
    """uspp_batch.py:99-233. Returns (eig (nk, nbands), x (nk, nbands, npw_max))."""
    nk, nb0, m = x0.shape
    mask = hs.bk.mask
    max_dim = min(int(hs.bk.npw.min()), max(max_dim_factor * nbands, nbands + 24))
    v = _orthonormalize_b(x0, mask)
    hv, sv = hs.h(v), hs.s(v)
    eig = x = None
    for _ in range(max_iter):
        while True:  # drop the oldest rows while the overlap is not positive definite
            h_sub, s_sub = _subspace(v, hv, sv)
            ell, info = torch.linalg.cholesky_ex(s_sub)
            bad = int(info.max()) > 0
            if not bad or v.shape[1] <= nbands + 1:
                break
            v, hv, sv = v[:, 1:], hv[:, 1:], sv[:, 1:]
        if bad:
            eye = torch.eye(s_sub.shape[-1], dtype=s_sub.dtype)
            ell = torch.linalg.cholesky(s_sub + 1e-10 * eye)
        a = torch.linalg.solve_triangular(ell, h_sub, upper=False)
        a = torch.linalg.solve_triangular(ell, a.mH, upper=False).mH
        w, u = torch.linalg.eigh(0.5 * (a + a.mH))
        u = torch.linalg.solve_triangular(ell.mH, u, upper=True)
        eig = w[:, :nbands].real
        u_r = u[:, :, :nbands].transpose(-1, -2).to(x0.dtype)
        x = torch.einsum("kbj,kjg->kbg", u_r, v)
        hx = torch.einsum("kbj,kjg->kbg", u_r, hv)
        sx = torch.einsum("kbj,kjg->kbg", u_r, sv)
        r = hx - eig[..., None].to(x0.dtype) * sx
        rn = torch.linalg.norm(r, dim=-1).real
        if float(rn.max()) < tol:
            return eig, x
        n_add = int((rn > tol).sum(dim=1).max())
        sel = torch.argsort(rn, dim=1, descending=True)[:, :n_add]
        r_sel = torch.gather(r, 1, sel[..., None].expand(-1, -1, m))
        t_band = torch.einsum("kbg,kg,kbg->kb", x.conj(), hs.t.to(x.dtype), x).real
        d = teter_b(r_sel, hs.t.to(r_sel.real.dtype), torch.gather(t_band, 1, sel))
        dn = torch.linalg.norm(d, dim=-1, keepdim=True).real
        d = torch.where(dn > 1e-300, d / dn.clamp_min(1e-300), d)
        if v.shape[1] + n_add > max_dim:
            v, hv, sv = _contract(x, hx, sx)
        d = _orthonormalize_b(d, mask, against=v)
        v = torch.cat([v, d], dim=1)
        hv = torch.cat([hv, hs.h(d)], dim=1)
        sv = torch.cat([sv, hs.s(d)], dim=1)
    return eig, x
