"""S-metric generalized block Davidson on MI355X (scf/uspp_batch.py:27-233).

``BatchedHS`` applies H through the HIP operator built with the screened D, and
S = 1 + sum_pq |beta_p> q_pq <beta_q| through a second HIP operator whose
nonlocal kernels carry q in place of D (the projection <beta|psi> and the
back-projection are the MFMA contractions of H's nonlocal term), so both
applies run in the hand-written kernels. ``davidson_gen_batched`` follows the
reference's reduction step by step: a standard-orthonormal basis (CholQR2 on
the GPU), fp64 subspace matrices V^H H V and V^H S V from the MFMA inner
product, Cholesky of the overlap with the oldest basis rows dropped while it is
not positive definite (a 1e-10 shift as the last resort), L^-1 H L^-H and its
eigensolve, the back-solve with L^H, the Ritz rotation of the cached H and S
applies, residuals H x - e S x, the same worst-residual expansion count on
every k, the Teter-preconditioned unit-normalised directions (one kernel), and
the QR restart in the standard metric. H carries the USPP/PAW dual grid and a
DFT+U term with a real D when given (as BatchedHamiltonian does); the
meta-GGA tau term is refused.
"""

from __future__ import annotations

import copy

import torch

from gradwave_amd import _lib as L
from gradwave_amd import linalg as LA
from gradwave_amd.davidson import _eigh, _orthonormalize_b
from gradwave_amd.hamiltonian import BatchedHamiltonian


class BatchedHS:
    """uspp_batch.py:27-72: H with the screened D ``dscr`` and S = 1 + sum q |beta><beta|
    for all k at once; ``bk``, ``v_eff_r`` and ``p`` as for BatchedHamiltonian,
    ``dscr`` and ``q_full`` (nproj, nproj) real symmetric."""

    def __init__(self, bk, shape, v_eff_r: torch.Tensor, p: torch.Tensor, dscr: torch.Tensor,
                 q_full: torch.Tensor, hub_sphi=None, hub_d=None, smooth=None, metagga_v=None) -> None:
        if metagga_v is not None:
            raise NotImplementedError("the meta-GGA tau term is not part of this operator")
        if torch.is_complex(q_full):
            if float(q_full.imag.abs().max()) != 0.0:
                raise ValueError("q_full must be real symmetric")
            q_full = q_full.real
        self.bk = bk
        self.shape = tuple(int(s) for s in shape)
        self.p = p
        self.t = bk.t
        hb = copy.copy(bk)
        hb.dij_full = dscr
        # H with the dual grid and the DFT+U term as the reference builds it (uspp_batch.py:37-41)
        self.ham = BatchedHamiltonian(hb, self.shape, v_eff_r, p, hub_q=hub_sphi, hub_dij=hub_d, smooth=smooth)
        self.shape = self.ham.shape
        sb = copy.copy(bk)
        sb.dij_full = q_full
        # S needs only the projector contractions (its local-term geometry is H's, unused)
        self._aug = BatchedHamiltonian(sb, shape, v_eff_r, p, smooth=smooth)

    def h(self, c: torch.Tensor) -> torch.Tensor:
        return self.ham.apply(c)

    def s(self, c: torch.Tensor) -> torch.Tensor:
        out = self._aug.apply_terms(c, L.GW_TERM_NONLOCAL)
        return out.add_(c).mul_(self.bk.mask[:, None, :].to(c.real.dtype))


def _contract(x, hx, sx):
    """uspp_batch.py:116-123: the Ritz block re-orthonormalised in the standard
    metric by QR, H x and S x rotated with the triangular factor."""
    q, rmat = torch.linalg.qr(x.transpose(-1, -2), mode="reduced")
    rt = rmat.transpose(-1, -2)
    return (q.transpose(-1, -2).contiguous(),
            torch.linalg.solve_triangular(rt, hx, upper=False).contiguous(),
            torch.linalg.solve_triangular(rt, sx, upper=False).contiguous())


def _hermitian(a: torch.Tensor) -> torch.Tensor:
    return 0.5 * (a + a.mH)


@torch.no_grad()
def davidson_gen_batched(hs, x0: torch.Tensor, nbands: int, tol: float, max_iter: int = 60,
                         max_dim_factor: int = 4):
    """Block Davidson for H x = e S x over (nk, ., npw_max) padded blocks
    (uspp_batch.py:99-233). Returns (eig (nk, nbands), x (nk, nbands, npw_max));
    x is Ritz-rotated from a standard-orthonormal basis (not S-normalised)."""
    L.require_gpu(x0, "x0")
    nk, _, m = x0.shape
    mask = hs.bk.mask
    max_dim = min(int(hs.bk.npw.min()), max(max_dim_factor * nbands, nbands + 24))
    v = _orthonormalize_b(x0.contiguous(), mask)
    hv, sv = hs.h(v), hs.s(v)
    eig = x = None
    for _ in range(max_iter):
        while True:  # drop the oldest rows while the overlap is not positive definite
            h_sub = _hermitian(LA.inner(v, hv).to(torch.complex128))
            s_sub = _hermitian(LA.inner(v, sv).to(torch.complex128))
            ell, info = torch.linalg.cholesky_ex(s_sub)
            bad = int(info.max()) > 0
            if not bad or v.shape[1] <= nbands + 1:
                break
            v, hv, sv = v[:, 1:].contiguous(), hv[:, 1:].contiguous(), sv[:, 1:].contiguous()
        if bad:
            eye = torch.eye(s_sub.shape[-1], dtype=s_sub.dtype, device=s_sub.device)
            ell = torch.linalg.cholesky(s_sub + 1e-10 * eye)
        a = torch.linalg.solve_triangular(ell, h_sub, upper=False)
        a = torch.linalg.solve_triangular(ell, a.mH, upper=False).mH
        w, u = _eigh(_hermitian(a))
        u = torch.linalg.solve_triangular(ell.mH, u, upper=True)
        eig = w[:, :nbands].real
        ub = u[:, :, :nbands].to(x0.dtype).contiguous()  # (nk, m_sub, nbands)
        x, hx, sx = LA.combine(ub, v), LA.combine(ub, hv), LA.combine(ub, sv)
        r = hx - eig[..., None].to(x0.dtype) * sx
        rn = torch.linalg.norm(r, dim=-1).real
        gate = torch.stack([rn.max(), (rn > tol).sum(dim=1).max().to(rn.dtype)]).cpu()
        if float(gate[0]) < tol:
            return eig, x
        n_add = int(gate[1])
        sel = torch.argsort(rn, dim=1, descending=True)[:, :n_add]
        d = LA.teter_expand(x, r.contiguous(), hs.t, sel)  # <T>-scaled Teter, unit rows
        if v.shape[1] + n_add > max_dim:
            v, hv, sv = _contract(x, hx, sx)
        d = _orthonormalize_b(d, mask, against=v)
        v = torch.cat([v, d], dim=1)
        hv = torch.cat([hv, hs.h(d)], dim=1)
        sv = torch.cat([sv, hs.s(d)], dim=1)
    return eig, x
