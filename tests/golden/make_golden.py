"""Write golden vectors from the reference's own leaf modules (run once, here).

Pins what no closed form can: the exact G-sphere ordering and flat box index
(grids.py:145-184), the projector table bits (core/hamiltonian.py:46-105), and
the Davidson trajectory (solvers/davidson.py:527-774: n_iter, apply tallies,
eigenvalues, residual norms), in both the plain fp64 path and the
fp64-certified fp32-expansion path (davidson.py:268-308).

The reference package __init__ imports ase/spglib (absent here), so a bare
``gradwave`` package object whose __path__ points at the reference sources is
registered first; only leaf modules (numpy/torch/scipy imports) are loaded.
Inputs are built with the oracle's setup and stored alongside the outputs, so
the fixture is self-contained data. Usage:

    python tests/golden/make_golden.py /root/reference
"""

from __future__ import annotations

import os
import sys
import types

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def load_reference(ref_root: str):
    pkg = types.ModuleType("gradwave")
    pkg.__path__ = [os.path.join(ref_root, "src", "gradwave")]
    sys.modules["gradwave"] = pkg
    import gradwave.grids as rgrids  # noqa: E402
    import gradwave.core.hamiltonian as rham  # noqa: E402
    import gradwave.core.batch as rbatch  # noqa: E402
    import gradwave.solvers.davidson  # noqa: E402,F401
    # the solvers package re-exports a function named ``davidson`` that shadows
    # the submodule attribute; take the module object itself
    rdav = sys.modules["gradwave.solvers.davidson"]
    return rgrids, rham, rbatch, rdav


def main(ref_root: str) -> None:
    sys.path.insert(0, REPO)
    from oracle import systems as S
    from oracle.pseudo import SI, kb_form_factor
    from oracle.units import RY_EV

    rgrids, rham, rbatch, rdav = load_reference(ref_root)
    torch.set_num_threads(4)

    # 1. G-sphere ordering + flat index (bit-exact integer maps)
    cases = {
        "si2_k0": (S.fcc_primitive(S.A_SI), 30.0, (24, 24, 24), [0.125, 0.125, 0.125]),
        "si2_k8": (S.fcc_primitive(S.A_SI), 30.0, (24, 24, 24), [0.375, 0.375, 0.375]),
        "si8_gen": (S.si_conventional()[0], 30.0, (36, 36, 36), [0.31, -0.17, 0.23]),
        "tric": (np.array([[3.1, 0.0, 0.0], [0.4, 2.7, 0.0], [0.2, -0.3, 3.3]]), 30.0, None,
                 [0.31, 0.17, -0.23]),
        "al1_g": (S.fcc_primitive(S.A_AL), 25.0, None, [0.0, 0.0, 0.0]),
    }
    out = {}
    for key, (cell, ecut_ry, shape, k) in cases.items():
        grid = rgrids.build_fft_grid(cell, ecut_ry * RY_EV, shape_override=shape)
        sph = rgrids.build_gsphere(grid, ecut_ry * RY_EV, k)
        out[f"gs_{key}_cell"] = np.asarray(cell, dtype=np.float64)
        out[f"gs_{key}_ecut_ry"] = np.float64(ecut_ry)
        out[f"gs_{key}_k"] = np.asarray(k, dtype=np.float64)
        out[f"gs_{key}_shape"] = np.asarray(grid.shape, dtype=np.int64)
        out[f"gs_{key}_miller"] = sph.miller.numpy()
        out[f"gs_{key}_flat"] = sph.flat_idx.numpy()
        out[f"gs_{key}_kpg2"] = sph.kpg2.numpy()

    # 2. projector table assembly on analytic Gaussian F_l tables
    grid = rgrids.build_fft_grid(S.fcc_primitive(S.A_SI), 30 * RY_EV, shape_override=(24, 24, 24))
    sph = rgrids.build_gsphere(grid, 30 * RY_EV, [0.125, 0.125, 0.375])
    q = np.sqrt(sph.kpg2.numpy())
    tab = torch.as_tensor(np.stack([kb_form_factor(SI, 0, q), kb_form_factor(SI, 1, q)]))
    pd = rham.build_projector_data(sph, [0, 0], [tab], [[0, 1]], [SI.dij], grid.volume)
    out["pd_ftab"] = tab.numpy()
    out["pd_k"] = np.array([0.125, 0.125, 0.375])
    out["pd_f_ylm"] = pd.f_ylm_phase_free.numpy()
    out["pd_dij"] = pd.dij_full.numpy()
    out["pd_atom_index"] = pd.atom_index.numpy()
    pos = torch.tensor([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]], dtype=torch.float64) * S.A_SI
    out["pd_proj"] = rham.projectors(pd, pos).numpy()

    # 3. Davidson trajectories on the reference's synthetic operator
    #    (test_davidson_fp32_expansion.py:24-42 construction, seed 2, nk=4, npw=80, nb=8)
    def make_operator(nk, npw, seed):
        torch.manual_seed(seed)
        diag = torch.linspace(1.0, 100.0, npw, dtype=torch.float64)
        h = torch.zeros(nk, npw, npw, dtype=torch.complex128)
        for kk in range(nk):
            a = torch.randn(npw, npw, dtype=torch.complex128)
            herm = 0.5 * (a + a.conj().T)
            h[kk] = herm + torch.diag(diag.to(torch.complex128))
            h[kk] = 0.5 * (h[kk] + h[kk].conj().T)
        return h, diag.expand(nk, npw).contiguous()

    h, t = make_operator(4, 80, 2)
    mask = torch.ones(4, 80, dtype=torch.bool)
    torch.manual_seed(99)
    x0 = torch.randn(4, 8, 80, dtype=torch.complex128)

    def apply(c):
        return torch.einsum("kij,kbj->kbi", h.to(c.dtype), c)

    out["syn_h"] = h.numpy()
    out["syn_x0"] = x0.numpy()
    for mode in ("off", "on"):
        os.environ["GRADWAVE_FP32_EXPANSION"] = mode
        r = rdav.davidson_batched(apply, x0.clone(), t, mask, tol=1e-9, max_iter=100)
        out[f"syn_{mode}_eig"] = r.eigenvalues.numpy()
        out[f"syn_{mode}_rn"] = r.residual_norms.numpy()
        out[f"syn_{mode}_niter"] = np.int64(r.n_iter)
        out[f"syn_{mode}_napply"] = np.array([r.n_apply_low, r.n_apply_full], dtype=np.int64)
    os.environ["GRADWAVE_FP32_EXPANSION"] = "off"

    # 4. Davidson on the plane-wave H of Si2 (first 3 irreducible k of the
    #    shifted 4x4x4 mesh), the reference's BatchedHamiltonian as h_apply
    s = S.si2_c1()
    sel = [0, 4, 8]
    spheres = [rgrids.build_gsphere(rgrids.build_fft_grid(s.cell, s.ecut, shape_override=s.shape),
                                    s.ecut, s.kfrac[i]) for i in sel]
    pds = []
    for sp in spheres:
        qq = np.sqrt(sp.kpg2.numpy())
        tt = torch.as_tensor(np.stack([kb_form_factor(SI, 0, qq), kb_form_factor(SI, 1, qq)]))
        pds.append(rham.build_projector_data(sp, [0, 0], [tt], [[0, 1]], [SI.dij],
                                             s.grid.volume))
    bk = rbatch.build_batched(spheres, pds)
    p = rbatch.projectors_b(bk, s.positions)
    os.environ["GRADWAVE_TOEPLITZ"] = "off"
    rbatch._TOEPLITZ_MODE = "off"
    hb = rbatch.BatchedHamiltonian(bk, s.shape, s.v_r, p)
    g = torch.Generator().manual_seed(5)
    x0 = torch.randn(len(sel), 8, bk.npw_max, dtype=torch.complex128, generator=g)
    x0 = x0 * bk.mask[:, None, :]
    hc = hb.apply(x0)
    r = rdav.davidson_batched(hb.apply, x0.clone(), bk.t, bk.mask, tol=2.7e-9, max_iter=60)
    out["pw_ksel"] = np.array(sel, dtype=np.int64)
    out["pw_v_r"] = s.v_r.numpy()
    out["pw_x0"] = x0.numpy()
    out["pw_hx0"] = hc.numpy()
    out["pw_flat"] = bk.flat_idx.numpy()
    out["pw_p"] = p.numpy()
    out["pw_eig"] = r.eigenvalues.numpy()
    out["pw_rn"] = r.residual_norms.numpy()
    out["pw_niter"] = np.int64(r.n_iter)

    path = os.path.join(HERE, "reference_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
