"""Shared inputs for the generalized (S-metric) eigenproblem tests."""

from __future__ import annotations

import numpy as np
import torch


def q_blocks(proj_atom_index, seed=0, lo=0.1, hi=0.5):
    """A positive semidefinite q: one random SPD block per atom (its s+p projectors),
    eigenvalues in [lo, hi] -- the structure of USPP augmentation integrals."""
    g = np.random.default_rng(seed)
    atoms = np.asarray(proj_atom_index)
    nproj = atoms.shape[0]
    q = np.zeros((nproj, nproj))
    for a in np.unique(atoms):
        idx = np.nonzero(atoms == a)[0]
        n = idx.size
        o, _ = np.linalg.qr(g.standard_normal((n, n)))
        q[np.ix_(idx, idx)] = o @ np.diag(g.uniform(lo, hi, n)) @ o.T
    return torch.as_tensor(q)


def dense_s(p_k, q, n):
    """S = 1 + p^T q^T conj(p) on the first n plane waves of one k."""
    p = p_k[:, :n]
    s = torch.eye(n, dtype=torch.complex128) + p.T @ q.to(torch.complex128).T @ p.conj()
    return 0.5 * (s + s.conj().T)
