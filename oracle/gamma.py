"""Gamma-point real-wavefunction layout and its Davidson (CPU oracle; test
infrastructure only -- the product never imports this package).

Restates core/gamma.py:
* build_gamma_basis (70-169): half sphere = G=0 first, then the member of each
  {G, -G} pair whose Miller key (m3, m2, m1) is lexicographically positive, in
  sphere order; full <- half source slots with a conjugation flag; rfft
  half-box back-gather; kinetic on the half sphere; metric weights 1 / 2;
* half_to_full / full_to_half (172-180), GammaHamiltonian.apply (183-228:
  kinetic + irfftn/rfftn local term + full-sphere KB projection kept on the
  half sphere), metric_inner (231-239), embed_real / unembed_real /
  kinetic_phi (242-280);
* davidson_gamma (283-290): the single-k solvers/davidson.py davidson
  (329-395, with _orthonormalize 310-327: two projection passes, reduced QR,
  rows kept where |R_ii| > 1e-10; restart collapses to [x, t] and re-applies)
  run in the real embedding.

Pinned by: the reality of N ifftn of a half_to_full vector, equality of the
half-sphere apply with the complex BatchedHamiltonian restricted to a
Hermitian-symmetric block, the embedding round trip and metric, and equality
of the Gamma eigenvalues with the complex batched Davidson and with dense
eigvalsh of the explicit H (tests/test_oracle_gamma.py, mirroring
tests/unit/test_gamma.py:100-174).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from oracle.davidson import teter
from oracle.units import HBAR2_2M


@dataclass
class GammaBasis:
    shape: tuple
    nh3: int
    t_half: torch.Tensor
    metric_w: torch.Tensor
    full_flat_idx: torch.Tensor
    rep_full_idx: torch.Tensor
    src_half: torch.Tensor
    conj_full: torch.Tensor
    back_flat: torch.Tensor
    back_conj: torch.Tensor
    partner: torch.Tensor  # (npw,) sphere index of -G

    @property
    def nhalf(self) -> int:
        return int(self.t_half.shape[0])

    @property
    def npw(self) -> int:
        return int(self.src_half.shape[0])


def build_gamma_basis(sphere, shape) -> GammaBasis:
    n1, n2, n3 = (int(s) for s in shape)
    nh3 = n3 // 2 + 1
    m = sphere.miller.numpy().astype(np.int64)
    npw = m.shape[0]
    pos_box = np.stack([m[:, 0] % n1, m[:, 1] % n2, m[:, 2] % n3], axis=1)
    neg_box = np.stack([(-m[:, 0]) % n1, (-m[:, 1]) % n2, (-m[:, 2]) % n3], axis=1)
    lin = lambda b: (b[:, 0] * n2 + b[:, 1]) * n3 + b[:, 2]  # noqa: E731
    where = np.full(n1 * n2 * n3, -1, dtype=np.int64)
    where[lin(pos_box)] = np.arange(npw)
    partner = where[lin(neg_box)]
    if (partner < 0).any():
        raise ValueError("the sphere is not closed under G -> -G (k is not Gamma)")
    zero = (m == 0).all(axis=1)
    if zero.sum() != 1:
        raise ValueError("Gamma sphere must hold G=0 exactly once")
    positive = (m[:, 2] > 0) | ((m[:, 2] == 0) & ((m[:, 1] > 0) | ((m[:, 1] == 0) & (m[:, 0] > 0))))
    rep = np.concatenate([np.flatnonzero(zero), np.flatnonzero(positive)])
    nhalf = rep.shape[0]
    if nhalf != (npw + 1) // 2:
        raise ValueError("half sphere has the wrong size")
    slot = np.full(npw, -1, dtype=np.int64)
    slot[rep] = np.arange(nhalf)
    is_rep = zero | positive
    src = np.where(is_rep, slot, slot[partner])
    rb = pos_box[rep]
    inside = rb[:, 2] <= n3 // 2
    sb = np.where(inside[:, None], rb, neg_box[rep])
    back = (sb[:, 0] * n2 + sb[:, 1]) * nh3 + sb[:, 2]
    w = np.full(nhalf, 2.0)
    w[0] = 1.0
    t = HBAR2_2M * sphere.kpg2.numpy()[rep]
    as_t = torch.as_tensor
    return GammaBasis((n1, n2, n3), nh3, as_t(t, dtype=torch.float64), as_t(w, dtype=torch.float64),
                      sphere.flat_idx.clone(), as_t(rep), as_t(src), as_t(~is_rep), as_t(back),
                      as_t(~inside), as_t(partner))


def half_to_full(gb: GammaBasis, ch: torch.Tensor) -> torch.Tensor:
    c = ch[:, gb.src_half]
    return torch.where(gb.conj_full[None, :], c.conj(), c)


def full_to_half(gb: GammaBasis, cf: torch.Tensor) -> torch.Tensor:
    return cf[:, gb.rep_full_idx]


class GammaHamiltonian:
    def __init__(self, gb: GammaBasis, v_eff_r, p, dij) -> None:
        self.gb, self.v_eff_r, self.p, self.dij = gb, v_eff_r, p, dij
        self.shape = gb.shape
        self.n = gb.shape[0] * gb.shape[1] * gb.shape[2]

    def apply(self, ch: torch.Tensor) -> torch.Tensor:
        gb = self.gb
        nb = ch.shape[0]
        cf = half_to_full(gb, ch)
        out = gb.t_half * ch
        box = torch.zeros(nb, self.n, dtype=ch.dtype).index_add(1, gb.full_flat_idx, cf)
        half_box = box.reshape(nb, *self.shape)[..., : gb.nh3]
        psi = torch.fft.irfftn(half_box, s=self.shape, dim=(-3, -2, -1))
        vg = torch.fft.rfftn(psi * self.v_eff_r, dim=(-3, -2, -1)).reshape(nb, -1)[:, gb.back_flat]
        out = out + torch.where(gb.back_conj[None, :], vg.conj(), vg)
        if self.p.shape[0]:
            b = cf @ self.p.conj().T
            out = out + ((b @ self.dij.to(cf.dtype)) @ self.p)[:, gb.rep_full_idx]
        return out


def metric_inner(gb: GammaBasis, a, b):
    return (a.conj() @ (gb.metric_w * b).T).real


def embed_real(gb: GammaBasis, ch: torch.Tensor) -> torch.Tensor:
    s = 2.0 ** 0.5
    phi = torch.empty(ch.shape[0], 2 * gb.nhalf - 1, dtype=torch.float64)
    phi[:, 0] = ch[:, 0].real
    phi[:, 1::2] = s * ch[:, 1:].real
    phi[:, 2::2] = s * ch[:, 1:].imag
    return phi


def unembed_real(gb: GammaBasis, phi: torch.Tensor) -> torch.Tensor:
    s = 2.0 ** -0.5
    ch = torch.empty(phi.shape[0], gb.nhalf, dtype=torch.complex128)
    ch[:, 0] = phi[:, 0].to(torch.complex128)
    ch[:, 1:] = s * torch.complex(phi[:, 1::2], phi[:, 2::2])
    return ch


def kinetic_phi(gb: GammaBasis) -> torch.Tensor:
    t = torch.empty(2 * gb.nhalf - 1, dtype=torch.float64)
    t[0] = gb.t_half[0]
    t[1::2] = gb.t_half[1:]
    t[2::2] = gb.t_half[1:]
    return t


@dataclass
class DavidsonResult:
    eigenvalues: torch.Tensor
    eigenvectors: torch.Tensor
    n_iter: int
    residual_norms: torch.Tensor


def _orthonormalize(v: torch.Tensor, against: torch.Tensor | None = None) -> torch.Tensor:
    if against is not None and against.shape[0]:
        for _ in range(2):
            v = v - (v @ against.conj().T) @ against
    q, r = torch.linalg.qr(v.T, mode="reduced")
    return q.T[r.diagonal().abs() > 1e-10]


def davidson(h_apply, x0, t_g, tol=1e-9, max_iter=40, max_dim_factor=4) -> DavidsonResult:
    nb, npw = x0.shape
    max_dim = min(max_dim_factor * nb, npw)
    v = _orthonormalize(x0.clone())
    if v.shape[0] < nb:
        gen = torch.Generator().manual_seed(1234)
        extra = (torch.randn(nb - v.shape[0], npw, generator=gen, dtype=torch.float64)
                 + 1j * torch.randn(nb - v.shape[0], npw, generator=gen, dtype=torch.float64))
        v = torch.cat([v, _orthonormalize(extra.to(x0.dtype), against=v)])
    hv = h_apply(v)
    rn = torch.full((nb,), float("inf"), dtype=torch.float64)
    for it in range(1, max_iter + 1):
        s = v @ hv.conj().T
        s = 0.5 * (s + s.conj().T)
        w, u = torch.linalg.eigh(s)
        eig = w[:nb].real
        ut = u[:, :nb].T.conj()
        x, hx = ut @ v, ut @ hv
        r = hx - eig[:, None] * x
        rn = torch.linalg.norm(r, dim=1).real
        bad = rn > tol
        if not bool(bad.any()):
            return DavidsonResult(eig, x, it, rn)
        tb = torch.einsum("bg,g,bg->b", x.conj(), t_g.to(x.dtype), x).real
        t = _orthonormalize(teter(r[bad], t_g, tb[bad]), against=v)
        if t.shape[0] == 0:
            return DavidsonResult(eig, x, it, rn)
        if v.shape[0] + t.shape[0] > max_dim:
            v = _orthonormalize(torch.cat([x, t]))
            hv = h_apply(v)
        else:
            v = torch.cat([v, t])
            hv = torch.cat([hv, h_apply(t)])
    return DavidsonResult(eig, x, max_iter, rn)


def davidson_gamma(ham: GammaHamiltonian, x0, tol=1e-9, max_iter=40, max_dim_factor=4) -> DavidsonResult:
    gb = ham.gb
    res = davidson(lambda phi: embed_real(gb, ham.apply(unembed_real(gb, phi))), embed_real(gb, x0),
                   kinetic_phi(gb), tol=tol, max_iter=max_iter, max_dim_factor=max_dim_factor)
    return DavidsonResult(res.eigenvalues, unembed_real(gb, res.eigenvectors), res.n_iter,
                          res.residual_norms)
