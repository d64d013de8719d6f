"""The product's workload builder reproduces the oracle's setup (CPU only).

Index maps (flat_idx order and content) must be bit-identical; the projector
and potential tables agree to round-off (different but equivalent op order).
"""

import numpy as np
import pytest
import torch

from gradwave_amd import workloads as W
from oracle import systems as S


@pytest.mark.parametrize("wf,of", [(W.si2_fcc, S.si2_c1), (W.si8_mp, S.si8_c4),
                                   (W.si64_gamma, S.si64_c2)])
def test_workload_matches_oracle(wf, of):
    w, o = wf(device="cpu"), of()
    assert torch.equal(w.bk.flat_idx, o.bk.flat_idx)
    assert torch.equal(w.bk.npw, o.bk.npw)
    assert torch.equal(w.bk.mask, o.bk.mask)
    assert np.array_equal(w.kweights, o.kweights)
    assert torch.equal(w.bk.t, o.bk.t)
    assert float((w.p - o.p).abs().max()) <= 1e-13 * float(o.p.abs().max())
    assert float((w.v_r - o.v_r).abs().max()) <= 1e-12 * float(o.v_r.abs().max())
    assert torch.equal(w.bk.dij_full, o.bk.dij_full)


def test_si216_sizes():
    """Gamma Si216 at 30 Ry holds 80,797 plane waves in a 120^3 box."""
    cell, pos = W.si_conventional((3, 3, 3))
    mm, kpg2, flat = W.sphere(cell, (120, 120, 120), 30 * W.RY_EV, [0, 0, 0])
    assert len(flat) == 80797 and len(pos) == 216


def test_si8_mp_kpoints():
    k, w = W.reduce_mesh(W.mp_mesh((4, 4, 4)), W.si_conventional()[0], full_cubic=False)
    assert len(k) == 36 and abs(w.sum() - 1) < 1e-14  # time reversal only: (64 + 8) / 2


def test_ragged32_cells_match_oracle_tables():
    """The 32 ragged cells of the batching form (seed 0): sphere maps bit-exact,
    V(r) and projector tables to 1e-12 against the oracle's build of the same
    cell, cutoff, k-point and box (grids.py:145-184, core/batch.py:174-235)."""
    from gradwave_amd import batching as B
    from oracle import systems as S
    from oracle.pseudo import AL, SI

    wls = B.ragged_instances(device="cpu", n=32, seed=0)
    for w in wls:
        species = [SI if sp.z == 4.0 else AL for sp in w.species]
        s = S.build_system(w.name, w.cell, w.positions.numpy(), species, w.ecut_ry, w.kfrac, w.kweights, w.nbands,
                           shape=w.shape)
        n = int(s.bk.npw[0])
        assert int(w.bk.npw[0]) == n, w.name
        assert torch.equal(w.bk.flat_idx[0, :n], s.bk.flat_idx[0, :n]), w.name
        assert float((w.v_r - s.v_r).abs().max() / s.v_r.abs().max()) < 1e-12, w.name
        assert float((w.p[0, :, :n] - s.p[0, :, :n]).abs().max() / s.p.abs().max()) < 1e-12, w.name
