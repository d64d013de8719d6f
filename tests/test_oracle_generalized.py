"""The oracle's S-metric generalized Davidson (scf/uspp_batch.py:27-233) pinned by
the dense generalized eigenproblem H x = e S x with the explicit plane-wave H and
S = 1 + p^T q^T conj(p) (scipy.linalg.eigh(H, S)), by S against the dense S, and
by the q = 0 limit, where it must reproduce the standard Davidson's eigenvalues."""

import dataclasses

import numpy as np
import pytest
import scipy.linalg as sl
import torch

from oracle import systems as S
from oracle.davidson import davidson_batched
from oracle.dense import dense_h
from oracle.generalized import BatchedHS, davidson_gen_batched
from oracle.hamiltonian import BatchedHamiltonian
from tests.uspp_cases import dense_s as _dense_s
from tests.uspp_cases import q_blocks as _q_blocks


def q_blocks(bk, seed=0):
    return _q_blocks(bk.proj_atom_index.numpy(), seed)


def dense_s(system, q, ik):
    return _dense_s(system.p[ik], q, int(system.bk.npw[ik]))


def _two_k(system):
    bk = system.bk
    sel = [0, 3]
    bk2 = dataclasses.replace(bk, npw=bk.npw[sel], mask=bk.mask[sel], flat_idx=bk.flat_idx[sel], kpg=bk.kpg[sel],
                              t=bk.t[sel], proj_phase_free=bk.proj_phase_free[sel])
    return dataclasses.replace(system, bk=bk2, p=system.p[sel], kfrac=system.kfrac[sel],
                               kweights=system.kweights[sel], spheres=[system.spheres[i] for i in sel])


@pytest.fixture(scope="module")
def si2():
    return _two_k(S.si2_c1())


def _start(bk, nb, seed=5):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(bk.nk, nb, bk.npw_max, dtype=torch.complex128, generator=g)
    return x * bk.mask[:, None, :]


def test_s_apply_matches_dense_s(si2):
    q = q_blocks(si2.bk)
    hs = BatchedHS(si2.bk, si2.shape, si2.v_r, si2.p, si2.bk.dij_full, q)
    c = _start(si2.bk, 3, seed=2)
    sc = hs.s(c)
    for ik in range(si2.bk.nk):
        n = int(si2.bk.npw[ik])
        want = (dense_s(si2, q, ik) @ c[ik, :, :n].T).T
        assert float((sc[ik, :, :n] - want).abs().max()) < 1e-12 * float(want.abs().max())
        assert float(sc[ik, :, n:].abs().max()) == 0.0


def test_generalized_davidson_matches_dense_eigh(si2):
    q = q_blocks(si2.bk)
    hs = BatchedHS(si2.bk, si2.shape, si2.v_r, si2.p, si2.bk.dij_full, q)
    nb = 8
    eig, x = davidson_gen_batched(hs, _start(si2.bk, nb), nb, tol=1e-9, max_iter=200)
    for ik in range(si2.bk.nk):
        want = sl.eigh(dense_h(si2, ik).numpy(), dense_s(si2, q, ik).numpy(), eigvals_only=True,
                       subset_by_index=[0, nb - 1])
        assert np.abs(eig[ik].numpy() - want).max() < 1e-9
    # residual contract, re-measured: || H x - e S x || with x S-normalised
    hx, sx = hs.h(x), hs.s(x)
    r = hx - eig[..., None].to(x.dtype) * sx
    assert float(torch.linalg.norm(r, dim=-1).max()) < 1e-8
    # S raises the eigenvalues' denominators: the spectrum differs from the q = 0 one
    h0 = BatchedHamiltonian(si2.bk, si2.shape, si2.v_r, si2.p)
    e0 = davidson_batched(h0.apply, _start(si2.bk, nb), si2.bk.t, si2.bk.mask, tol=1e-9, max_iter=200).eigenvalues
    assert float((eig - e0).abs().max()) > 1e-3


def test_q_zero_reduces_to_standard_davidson(si2):
    q = torch.zeros(si2.p.shape[1], si2.p.shape[1], dtype=torch.float64)
    hs = BatchedHS(si2.bk, si2.shape, si2.v_r, si2.p, si2.bk.dij_full, q)
    nb = 8
    eig, _ = davidson_gen_batched(hs, _start(si2.bk, nb), nb, tol=1e-9, max_iter=200)
    h0 = BatchedHamiltonian(si2.bk, si2.shape, si2.v_r, si2.p)
    e0 = davidson_batched(h0.apply, _start(si2.bk, nb), si2.bk.t, si2.bk.mask, tol=1e-9, max_iter=200).eigenvalues
    assert float((eig - e0).abs().max()) < 1e-9
