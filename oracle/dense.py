"""Brute-force dense plane-wave Hamiltonian (closed-form check of the apply).

H[i,j] = t_i delta_ij + Vhat(G_i - G_j) + sum_pq p_q(G_i) D_pq conj(p_p(G_j))
with Vhat = fftn(V)/N on the box and the box-wrapped Miller difference -- the
matrix the batched apply (core/batch.py:480-505) multiplies by, written out
element by element from the definitions in core/fftbox.py:8-10 and
core/hamiltonian.py:8-17. Its eigvalsh is the independent check of the apply
and of Davidson's eigenvalues.
"""

from __future__ import annotations

import torch


def dense_h(system, ik: int) -> torch.Tensor:
    bk = system.bk
    n = int(bk.npw[ik])
    shape = system.shape
    s1s2, s2 = shape[1] * shape[2], shape[2]
    flat = bk.flat_idx[ik, :n]
    g = torch.stack([flat // s1s2, (flat % s1s2) // s2, flat % s2], dim=-1)
    diff = (g[:, None, :] - g[None, :, :]) % torch.tensor(shape)
    idx = diff[..., 0] * s1s2 + diff[..., 1] * s2 + diff[..., 2]
    vhat = (torch.fft.fftn(system.v_r.to(torch.complex128)) / (shape[0] * s1s2)).reshape(-1)
    h = vhat[idx]
    h = h + torch.diag(bk.t[ik, :n].to(torch.complex128))
    p = system.p[ik, :, :n]  # (nproj, n)
    d = bk.dij_full.to(torch.complex128)
    h = h + p.T @ d.T @ p.conj()
    return 0.5 * (h + h.conj().T)
