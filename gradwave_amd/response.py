"""Conduction-projected Sternheimer solves on MI355X.

Contract (postscf/_response.py:502-595, norm-conserving branch): for every
occupied band n of every k, solve

    (H - eps_n + shift * P_occ) dpsi = rhs,    P_occ = sum_m |c_m><c_m|,

each (k, n) column to a residual norm below ``tol``, and return
(1 - P_occ) dpsi; a warm ``x0`` is certified, not trusted; padded plane-wave
slots of ``x0`` are zeroed; ``rhs`` already lies in the conduction space.

Design. The solver is an active-set column CG whose whole state stays on the
device in the padded (nk, nocc, npw_max) layout, one row per column:

* every step works on the list of still-active columns only: their search
  directions are gathered into one compact block by the direction kernel and
  applied with ``h.apply_cols`` (the operator's compacted column apply, which
  groups them by k into one HIP apply); operators without ``apply_cols`` get
  the full block;
* the occupied-space projections of the step, P_occ (H - eps) p and P_occ p,
  are MFMA contractions (``linalg.inner`` / ``combine``) over the full layout:
  P_occ acts column by column, so inactive rows cost a few GEMM flops and
  change nothing;
* x += a p, r -= a Ap, the Teter-filtered z = K r, ||r||^2 and Re<r, z> are one
  fused kernel over the active columns (``gw_cg_update``), p = z + b p another
  (``gw_cg_direction``), Re<p, Ap> a third (``gw_row_dots``); all sums fp64;
* a column leaves the active list as soon as its recurrence residual is below
  ``tol``; when the list is empty the true residual rhs - A x of every column is
  measured with one block apply, and the columns that drifted above ``tol``
  start a fresh CG (at most ``_CERTIFY_ROUNDS`` measurements); ``max_iter``
  bounds the CG steps over all rounds.

S metric (ultrasoft/PAW, postscf/_response.py:691-745): with the overlap
operator ``s_apply`` and ``s_occ = S c_occ`` the solved system is

    [(1 - P_S^H)(H - eps_n S) + shift * Pi] dpsi = rhs,
    P_S = sum_m |c_m><c_m| S,   Pi = sum_m S|c_m><c_m| S,

and the result (1 - P_S) dpsi. The same column CG drives it: only the three
projections change (their bra and ket blocks are c_occ or s_occ), and every
apply is a full-block one, since S is a block operator. With S the identity
and s_occ = c_occ every operation is the norm-conserving one.
"""

from __future__ import annotations

import torch

from gradwave_amd import _lib as L
from gradwave_amd import linalg as LA

_CERTIFY_ROUNDS = 4


def _i64(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.int64).contiguous()


class _ColumnCG:
    """Device state and kernels of one batched Sternheimer solve."""

    def __init__(self, h, t, c_occ, eps_occ, rhs, shift, s_apply=None, s_occ=None):
        self.h, self.c_occ, self.rhs, self.shift = h, c_occ.contiguous(), rhs.contiguous(), float(shift)
        self.s_apply = s_apply
        self.s_occ = self.c_occ if s_occ is None else s_occ.to(rhs.dtype).contiguous()
        self.nk, self.nocc, self.m = rhs.shape
        self.nc = self.nk * self.nocc
        self.dev = rhs.device
        self.prec = L.precision_of(rhs.dtype)
        self.t = t.to(torch.float64).contiguous()
        self.eps = eps_occ.to(rhs.real.dtype).contiguous()
        self.kcol_all = torch.arange(self.nk, device=self.dev).repeat_interleave(self.nocc)
        self.z = torch.zeros_like(self.rhs)
        self.p = torch.zeros_like(self.rhs)
        self.w = torch.zeros_like(self.rhs)  # A p, full layout (valid at active rows)
        rows = torch.arange(self.nc, device=self.dev)
        # <c_n|T|c_n>, the Teter scale of column (k, n) (postscf/_response.py:559-562)
        tb = torch.empty(self.nc, dtype=torch.float64, device=self.dev)
        self._dots(rows, self.c_occ, self.c_occ, tb, weight=self.t, wrow=self.kcol_all)
        self.t_band = tb.clamp_(min=1e-6)

    def _stream(self):
        return L.stream_ptr(self.dev)

    def _dots(self, rows, a, b, out, weight=None, wrow=None):
        r = _i64(rows)
        L.check(L.lib().gw_row_dots(self.prec, int(r.numel()), self.m, r.data_ptr(), a.data_ptr(), b.data_ptr(),
                                    self.m, weight.data_ptr() if weight is not None else None,
                                    weight.stride(0) if weight is not None else 0,
                                    _i64(wrow).data_ptr() if wrow is not None else None, out.data_ptr(),
                                    self._stream()), "gw_row_dots")
        return out

    def project(self, y: torch.Tensor, sign: float, bra: torch.Tensor, ket: torch.Tensor,
                out: torch.Tensor | None = None) -> torch.Tensor:
        """out + sign * sum_n |ket_n><bra_n|y> on the full layout (out = y, in place,
        when None): two MFMA contractions."""
        ov = LA.inner(bra, y)  # ov[n][b] = <bra_n|y_b>
        target = y if out is None else out
        LA.combine((sign * ov).contiguous(), ket, out=target, beta=1.0)
        return target

    def _hs(self, v: torch.Tensor) -> torch.Tensor:
        """(H - eps_n S) v on the full block (S = 1 without ``s_apply``)."""
        hv = self.h.apply(v)
        hv -= self.eps[..., None] * (v if self.s_apply is None else self.s_apply(v))
        return hv

    def _finish_a(self, w: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
        """w <- (1 - P_S^H) w + shift * Pi v, where w = (H - eps S) v."""
        self.project(w, -1.0, self.c_occ, self.s_occ)
        if self.shift != 0.0:
            self.project(v, self.shift, self.s_occ, self.s_occ, out=w)
        return w

    def update(self, x, r, rows, alpha, ap):
        n = int(rows.numel())
        rn2 = torch.empty(n, dtype=torch.float64, device=self.dev)
        rz = torch.empty(n, dtype=torch.float64, device=self.dev)
        L.check(L.lib().gw_cg_update(self.prec, n, self.m, rows.data_ptr(), self.kcol_all[rows].contiguous().data_ptr(),
                                     alpha.data_ptr(), x.data_ptr(), self.p.data_ptr(), r.data_ptr(), ap.data_ptr(),
                                     self.m, self.t.data_ptr(), self.t.stride(0), self.t_band.data_ptr(),
                                     self.z.data_ptr(), rn2.data_ptr(), rz.data_ptr(), self._stream()),
                "gw_cg_update")
        return rn2, rz

    def direction(self, rows, beta, compact: bool):
        n = int(rows.numel())
        pc = torch.empty(n, self.m, dtype=self.rhs.dtype, device=self.dev) if compact else None
        L.check(L.lib().gw_cg_direction(self.prec, n, self.m, rows.data_ptr(), beta.data_ptr(), self.z.data_ptr(),
                                        self.p.data_ptr(), self.m, pc.data_ptr() if pc is not None else None,
                                        self.m, self._stream()), "gw_cg_direction")
        return pc

    def apply_a(self, rows, pc):
        """A p = (1 - P_S^H)(H - eps S) p + shift Pi p into self.w at the listed rows
        (pc: the compacted active directions, norm-conserving case only)."""
        if pc is not None:
            kc = self.kcol_all[rows]
            hp = self.h.apply_cols(pc, kc)
            hp -= self.eps.view(-1)[rows][:, None] * pc
            self.w.view(self.nc, self.m).index_copy_(0, rows, hp)
        else:
            self.w.copy_(self._hs(self.p))
        return self._finish_a(self.w, self.p)

    def true_residual(self, x):
        """rhs - A x for every column (one block apply)."""
        return self.rhs - self._finish_a(self._hs(x), x)


@torch.no_grad()
def cg_sternheimer(h, bk, c_occ: torch.Tensor, eps_occ: torch.Tensor, rhs: torch.Tensor, x0: torch.Tensor,
                   shift: float, tol: float = 1e-8, max_iter: int = 400, s_apply=None,
                   s_occ: torch.Tensor | None = None) -> torch.Tensor:
    """Batched conduction-projected Sternheimer solve (postscf/_response.py:502-595;
    S metric 691-745).

    h: operator with ``apply`` (and optionally ``apply_cols``); bk: object with
    the kinetic table ``t`` (nk, npw_max) and optionally ``mask``;
    c_occ, rhs, x0: (nk, nocc, npw_max) complex on the GPU; eps_occ (nk, nocc).
    ``s_apply`` (a block operator, S c for c of that layout) with ``s_occ`` =
    S c_occ selects the S metric. Returns (1 - P_occ) dpsi, or (1 - P_S) dpsi."""
    if s_apply is not None and s_occ is None:
        raise ValueError("the S-metric cg_sternheimer needs s_occ = S c_occ")
    L.require_gpu(rhs, "rhs")
    nk, nocc, m = rhs.shape
    x = x0.to(rhs.dtype).clone().contiguous()
    mask = getattr(bk, "mask", None)
    if mask is not None:
        x *= mask[:, None, :].to(x.dtype)
    if nk * nocc == 0:
        return x
    cg = _ColumnCG(h, bk.t, c_occ, eps_occ, rhs, shift, s_apply, s_occ)
    compact = hasattr(h, "apply_cols") and s_apply is None
    used = 0
    for _ in range(_CERTIFY_ROUNDS):
        r = cg.true_residual(x).contiguous()
        rn = LA.row_norms(r.view(1, cg.nc, m))[0]
        rows = _i64(torch.nonzero(rn >= tol).flatten())
        if rows.numel() == 0 or used >= max_iter:
            break
        zero = torch.zeros(rows.numel(), dtype=torch.float64, device=rhs.device)
        _, rz = cg.update(x, r, rows, zero, r)  # a = 0: z = K r and Re<r, z> only
        cg.p.view(cg.nc, m).index_fill_(0, rows, 0)
        pc = cg.direction(rows, zero, compact)  # p = z
        while rows.numel() and used < max_iter:
            ap = cg.apply_a(rows, pc)
            pap = cg._dots(rows, cg.p, ap, torch.empty(rows.numel(), dtype=torch.float64, device=rhs.device))
            alpha = rz / pap.clamp(min=1e-300)
            rn2, rz_new = cg.update(x, r, rows, alpha, ap)
            used += 1
            keep = (rn2 >= tol * tol).nonzero().flatten()
            if keep.numel() == 0:
                break
            beta = rz_new[keep] / rz[keep].clamp(min=1e-300)
            rows, rz = rows[keep].contiguous(), rz_new[keep]
            pc = cg.direction(rows, beta.contiguous(), compact)
    return cg.project(x, -1.0, cg.s_occ, cg.c_occ)
