"""Host-side Gamma layout (gradwave_amd.gamma maps and embedding) against the
oracle's restatement of core/gamma.py: index maps bit-exact."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import gamma as OG
from oracle import systems as S


@pytest.mark.parametrize("name", ["si8_gamma", "si2_gamma"])
def test_gamma_maps_bit_exact(name):
    from gradwave_amd import gamma as G

    if name == "si8_gamma":
        cell, pos = S.si_conventional()
        s = S.build_system(name, cell, S.jitter(pos, 0.05, 2), [S.SI] * 8, 14.0, np.zeros((1, 3)), np.ones(1), 8)
    else:
        cell = S.fcc_primitive(S.A_SI)
        s = S.build_system(name, cell, np.array([[0, 0, 0], [0.25, 0.25, 0.25]]) * S.A_SI, [S.SI] * 2, 30.0,
                           np.zeros((1, 3)), np.ones(1), 8, shape=(24, 24, 24))
    ref = OG.build_gamma_basis(s.spheres[0], s.shape)
    ours = G.build_gamma_basis(s.spheres[0], s.shape, device="cpu")
    for f in ("full_flat_idx", "rep_full_idx", "src_half", "conj_full", "back_flat", "back_conj", "partner"):
        assert torch.equal(getattr(ours, f), getattr(ref, f)), f
    assert torch.equal(ours.t_half, ref.t_half) and torch.equal(ours.metric_w, ref.metric_w)
    assert ours.nh3 == ref.nh3
    g = torch.Generator().manual_seed(0)
    ch = torch.randn(3, ours.nhalf, dtype=torch.complex128, generator=g)
    ch[:, 0] = ch[:, 0].real.to(torch.complex128)
    assert torch.equal(G.embed_real(ours, ch), OG.embed_real(ref, ch))
    assert torch.equal(G.unembed_real(ours, G.embed_real(ours, ch)), OG.unembed_real(ref, OG.embed_real(ref, ch)))
    assert torch.equal(G.half_to_full(ours, ch), OG.half_to_full(ref, ch))


def test_gamma_basis_rejects_non_gamma_sphere():
    from gradwave_amd import gamma as G

    cell = S.fcc_primitive(S.A_SI)
    s = S.build_system("si2_k", cell, np.array([[0, 0, 0], [0.25, 0.25, 0.25]]) * S.A_SI, [S.SI] * 2, 20.0,
                       np.array([[0.125, 0.25, 0.0]]), np.ones(1), 4)
    with pytest.raises(ValueError):
        G.build_gamma_basis(s.spheres[0], s.shape, device="cpu")
