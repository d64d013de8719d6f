"""The oracle against golden vectors written from the reference's own modules.

tests/golden/make_golden.py wrote these with the reference's grids.build_gsphere,
core.hamiltonian.build_projector_data/projectors, core.batch.BatchedHamiltonian
and solvers.davidson.davidson_batched. Integer maps must match bit for bit; the
floating-point outputs matched bit for bit when the fixture was written (same
torch build), and are held here to 1e-12.
"""

import os

import numpy as np
import pytest
import torch

from oracle import grids as G
from oracle import systems as S
from oracle.davidson import davidson_batched
from oracle.hamiltonian import (BatchedHamiltonian, build_batched, build_projector_data,
                                projectors, projectors_b)
from oracle.pseudo import SI
from oracle.units import RY_EV

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.npz"))


@pytest.mark.parametrize("key", ["si2_k0", "si2_k8", "si8_gen", "tric", "al1_g"])
def test_gsphere_maps_bit_exact(key):
    cell = GOLD[f"gs_{key}_cell"]
    ecut = float(GOLD[f"gs_{key}_ecut_ry"]) * RY_EV
    grid = G.build_fft_grid(cell, ecut, shape_override=tuple(GOLD[f"gs_{key}_shape"]))
    sph = G.build_gsphere(grid, ecut, GOLD[f"gs_{key}_k"])
    assert np.array_equal(sph.flat_idx.numpy(), GOLD[f"gs_{key}_flat"])
    assert np.array_equal(sph.miller.numpy(), GOLD[f"gs_{key}_miller"])
    assert np.array_equal(sph.kpg2.numpy(), GOLD[f"gs_{key}_kpg2"])


def test_projector_tables():
    grid = G.build_fft_grid(S.fcc_primitive(S.A_SI), 30 * RY_EV, shape_override=(24, 24, 24))
    sph = G.build_gsphere(grid, 30 * RY_EV, GOLD["pd_k"])
    pd = build_projector_data(sph, [SI, SI], grid.volume)
    assert np.array_equal(pd.atom_index.numpy(), GOLD["pd_atom_index"])
    assert np.array_equal(pd.dij_full.numpy(), GOLD["pd_dij"])
    assert np.abs(pd.f_ylm_phase_free.numpy() - GOLD["pd_f_ylm"]).max() <= 1e-15
    pos = torch.tensor([[0.0, 0, 0], [0.25, 0.25, 0.25]], dtype=torch.float64) * S.A_SI
    assert np.abs(projectors(pd, pos).numpy() - GOLD["pd_proj"]).max() <= 1e-15


@pytest.mark.parametrize("mode", ["off", "on"])
def test_davidson_trajectory_synthetic(mode):
    h = torch.as_tensor(GOLD["syn_h"])
    x0 = torch.as_tensor(GOLD["syn_x0"])
    t = torch.linspace(1.0, 100.0, 80, dtype=torch.float64).expand(4, 80).contiguous()
    mask = torch.ones(4, 80, dtype=torch.bool)

    def apply(c):
        return torch.einsum("kij,kbj->kbi", h.to(c.dtype), c)

    r = davidson_batched(apply, x0.clone(), t, mask, tol=1e-9, max_iter=100,
                         fp32_expansion=(mode == "on"))
    assert r.n_iter == int(GOLD[f"syn_{mode}_niter"])
    assert [r.n_apply_low, r.n_apply_full] == list(GOLD[f"syn_{mode}_napply"])
    assert np.abs(r.eigenvalues.numpy() - GOLD[f"syn_{mode}_eig"]).max() < 1e-12
    assert np.abs(r.residual_norms.numpy() - GOLD[f"syn_{mode}_rn"]).max() < 1e-12


def test_planewave_apply_and_davidson_si2():
    s = S.si2_c1()
    sel = list(GOLD["pw_ksel"])
    bk = build_batched([s.spheres[i] for i in sel],
                       [build_projector_data(s.spheres[i], s.species, s.grid.volume) for i in sel])
    p = projectors_b(bk, s.positions)
    assert np.array_equal(bk.flat_idx.numpy(), GOLD["pw_flat"])
    assert np.abs(p.numpy() - GOLD["pw_p"]).max() <= 1e-15
    assert np.abs(s.v_r.numpy() - GOLD["pw_v_r"]).max() <= 1e-12
    h = BatchedHamiltonian(bk, s.shape, s.v_r, p)
    x0 = torch.as_tensor(GOLD["pw_x0"])
    assert np.abs(h.apply(x0).numpy() - GOLD["pw_hx0"]).max() < 1e-12
    r = davidson_batched(h.apply, x0.clone(), bk.t, bk.mask, tol=2.7e-9, max_iter=60)
    assert r.n_iter == int(GOLD["pw_niter"])
    assert np.abs(r.eigenvalues.numpy() - GOLD["pw_eig"]).max() < 1e-10
