"""Pins the Gamma-layout oracle (oracle/gamma.py) against the complex path,
mirroring tests/unit/test_gamma.py:100-174 on the analytic-pseudo Si8 cell at
k = 0, plus dense eigvalsh of the explicit H."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import systems as S
from oracle.davidson import davidson_batched
from oracle.dense import dense_h
from oracle.gamma import (GammaHamiltonian, build_gamma_basis, davidson_gamma, embed_real,
                          full_to_half, half_to_full, metric_inner, unembed_real)
from oracle.hamiltonian import BatchedHamiltonian


def si8_gamma(nbands=16):
    cell, pos = S.si_conventional()
    pos = S.jitter(pos, 0.05, 2)
    return S.build_system("si8_gamma", cell, pos, [S.SI] * 8, 14.0, np.zeros((1, 3)), np.ones(1), nbands)


@pytest.fixture(scope="module")
def case():
    s = si8_gamma()
    gb = build_gamma_basis(s.spheres[0], s.shape)
    return s, gb


def _half(gb, nb, seed):
    g = torch.Generator().manual_seed(seed)
    ch = torch.randn(nb, gb.nhalf, dtype=torch.complex128, generator=g)
    ch[:, 0] = ch[:, 0].real.to(torch.complex128)
    return ch


def test_half_sphere_maps(case):
    s, gb = case
    assert gb.nhalf == (gb.npw + 1) // 2 and int(gb.rep_full_idx[0]) == int(
        (s.spheres[0].miller == 0).all(1).nonzero()[0])
    # every full slot is reached exactly from its representative
    assert torch.equal(gb.src_half[gb.rep_full_idx], torch.arange(gb.nhalf))
    assert torch.equal(gb.partner[gb.partner], torch.arange(gb.npw))


def test_half_vector_is_real_in_r(case):
    s, gb = case
    cf = half_to_full(gb, _half(gb, 3, 1))
    n = int(np.prod(s.shape))
    box = torch.zeros(3, n, dtype=torch.complex128).index_add(1, gb.full_flat_idx, cf)
    psi = torch.fft.ifftn(box.reshape(3, *s.shape), dim=(-3, -2, -1))
    assert float(psi.imag.abs().max()) < 1e-14 * float(psi.real.abs().max()) * 10


def test_apply_matches_complex(case):
    s, gb = case
    ch = _half(gb, 6, 4)
    cf = half_to_full(gb, ch)
    gh = GammaHamiltonian(gb, s.v_r, s.p[0], s.bk.dij_full)
    full = BatchedHamiltonian(s.bk, s.shape, s.v_r, s.p).apply(cf[None])[0]
    assert float((full - full[:, gb.partner].conj()).abs().max()) < 1e-11
    assert float((gh.apply(ch) - full_to_half(gb, full)).abs().max()) < 1e-11


def test_embedding_roundtrip_and_metric(case):
    _, gb = case
    a, b = _half(gb, 5, 6), _half(gb, 5, 7)
    assert torch.allclose(unembed_real(gb, embed_real(gb, a)), a, atol=1e-15)
    assert torch.allclose(embed_real(gb, a) @ embed_real(gb, b).T, metric_inner(gb, a, b), atol=1e-12)


def test_gamma_eigenvalues_match_complex_and_dense(case):
    s, gb = case
    nb = s.nbands
    gh = GammaHamiltonian(gb, s.v_r, s.p[0], s.bk.dij_full)
    x0 = torch.zeros(nb, gb.nhalf, dtype=torch.complex128)
    order = torch.argsort(gb.t_half, stable=True)
    for i in range(nb):
        x0[i, order[i]] = 1.0
    gres = davidson_gamma(gh, x0, tol=1e-9, max_iter=200)
    xc = torch.zeros(1, nb, s.bk.npw_max, dtype=torch.complex128)
    tord = torch.argsort(s.bk.t[0], stable=True)
    for i in range(nb):
        xc[0, i, tord[i]] = 1.0
    hb = BatchedHamiltonian(s.bk, s.shape, s.v_r, s.p)
    cres = davidson_batched(hb.apply, xc, s.bk.t, s.bk.mask, tol=1e-9, max_iter=200)
    dense = torch.linalg.eigvalsh(dense_h(s, 0))[:nb]
    assert float(gres.residual_norms.max()) <= 1e-9
    assert float((gres.eigenvalues - cres.eigenvalues[0]).abs().max()) < 1e-8
    assert float((gres.eigenvalues - dense).abs().max()) < 1e-8
    g = metric_inner(gb, gres.eigenvectors, gres.eigenvectors)
    assert float((g - torch.eye(nb, dtype=g.dtype)).abs().max()) < 1e-10
