"""GPU parity of the Gamma layout (gradwave_amd.gamma): the two-bands-per-
complex-apply half-sphere H against the oracle's GammaHamiltonian (irfftn /
rfftn restatement of core/gamma.py:183-228), and the real Davidson against
the complex batched Davidson and dense eigvalsh.

Tolerances: apply 1e-12 relative (fp64), 2e-5 (fp32); eigenvalues 1e-10 Ha.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.gpu_cases import dev, say, system

pytestmark = pytest.mark.gpu

HARTREE_EV = 27.211386245988
_CACHE = {}


def _si8_gamma():
    from oracle import systems as S

    if "si8g" not in _CACHE:
        cell, pos = S.si_conventional()
        _CACHE["si8g"] = S.build_system("si8_gamma", cell, S.jitter(pos, 0.05, 2), [S.SI] * 8, 14.0,
                                        np.zeros((1, 3)), np.ones(1), 16)
    return _CACHE["si8g"]


def _pair(s):
    from gradwave_amd import gamma as G
    from oracle import gamma as OG

    d = dev()
    gb = G.build_gamma_basis(s.spheres[0], s.shape, device=d)
    n = s.spheres[0].npw
    p = s.p[0, :, :n]
    ours = G.GammaHamiltonian(gb, s.v_r.to(d), p.to(d), s.bk.dij_full)
    ref = OG.GammaHamiltonian(OG.build_gamma_basis(s.spheres[0], s.shape), s.v_r, p, s.bk.dij_full)
    return gb, ours, ref


def _half(nhalf, nb, seed):
    g = torch.Generator().manual_seed(seed)
    ch = torch.randn(nb, nhalf, dtype=torch.complex128, generator=g)
    ch[:, 0] = ch[:, 0].real.to(torch.complex128)
    return ch


def _rel(a, b):
    return float((a.cpu() - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("name,nb", [("si8_gamma", 6), ("si8_gamma", 5), ("si64", 7)])
def test_gamma_apply_matches_oracle(name, nb):
    say(f"gamma apply {name} nb={nb}")
    s = _si8_gamma() if name == "si8_gamma" else system(name)
    gb, ours, ref = _pair(s)
    ch = _half(gb.nhalf, nb, 3)
    want = ref.apply(ch)
    got = ours.apply(ch.to(dev()))
    assert _rel(got, want) < 1e-12
    got32 = ours.apply(ch.to(dev(), torch.complex64))
    assert _rel(got32.to(torch.complex128), want) < 2e-5


def test_gamma_davidson_matches_dense_and_complex():
    from gradwave_amd import gamma as G
    from gradwave_amd.davidson import davidson_batched
    from gradwave_amd.hamiltonian import BatchedHamiltonian
    from oracle.dense import dense_h
    from types import SimpleNamespace

    say("gamma davidson si8_gamma 16 bands vs dense and complex")
    s = _si8_gamma()
    gb, ours, _ = _pair(s)
    nb = s.nbands
    tol = 1e-10 * HARTREE_EV
    x0 = torch.zeros(nb, gb.nhalf, dtype=torch.complex128)
    order = torch.argsort(gb.t_half.cpu(), stable=True)
    for i in range(nb):
        x0[i, order[i]] = 1.0
    res = G.davidson_gamma(ours, x0.to(dev()), tol=tol, max_iter=300)
    dense = torch.linalg.eigvalsh(dense_h(s, 0))[:nb]
    assert float(res.residual_norms.max()) <= tol
    assert float((res.eigenvalues.cpu() - dense).abs().max()) / HARTREE_EV < 1e-10
    gram = G.metric_inner(gb, res.eigenvectors, res.eigenvectors)
    assert float((gram - torch.eye(nb, dtype=gram.dtype, device=gram.device)).abs().max()) < 1e-10
    mixed = G.davidson_gamma_mixed(ours, x0.to(dev()), tol=tol, draft_tol=1e-3, max_iter=300)
    assert float(mixed.residual_norms.max()) <= tol
    assert float((mixed.eigenvalues.cpu() - dense).abs().max()) / HARTREE_EV < 1e-10
    # draft + FP64 refinement sweeps: upper bounds, non-increasing with sweeps, 1e-6 Ha after one
    t_phi = G.kinetic_phi(gb)
    x32 = G.embed_real(gb, x0.to(dev())).to(torch.float32)
    _, xd, _, _ = G.davidson_real(ours.apply_real, x32, t_phi, tol=1e-3, max_iter=300)
    sw = [G.refine_real(ours.apply_real, xd.to(torch.float64), t_phi, n) for n in (0, 1)]
    for e, x, it, rn in sw:
        assert float((e.cpu() - dense).min()) > -1e-12 * HARTREE_EV
        hx = ours.apply_real(x)
        assert float((torch.linalg.norm(hx - e[:, None] * x, dim=1) - rn).abs().max()) < 1e-10
    assert float((sw[1][0] - sw[0][0]).max()) < 1e-12 * HARTREE_EV
    one = G.davidson_gamma_mixed(ours, x0.to(dev()), draft_tol=1e-3, max_iter=300, sweeps=1)
    assert one.n_iter == 1
    assert float((one.eigenvalues.cpu() - dense).abs().max()) / HARTREE_EV < 1e-6
    # buffer bands: the requested bands only, within the bound, upper bounds
    buf = G.davidson_gamma_mixed(ours, x0.to(dev()), draft_tol=5e-3, max_iter=300, sweeps=1, buffer=4)
    assert buf.eigenvalues.shape == (x0.shape[0],) and buf.eigenvectors.shape == x0.shape
    assert float((buf.eigenvalues.cpu() - dense).abs().max()) / HARTREE_EV < 1e-6
    assert float((buf.eigenvalues.cpu() - dense).min()) > -1e-12 * HARTREE_EV
    # the complex batched solver on the same frozen potential
    d = dev()
    bk = SimpleNamespace(npw=s.bk.npw.to(d), mask=s.bk.mask.to(d), flat_idx=s.bk.flat_idx.to(d), t=s.bk.t.to(d),
                         dij_full=s.bk.dij_full)
    hb = BatchedHamiltonian(bk, s.shape, s.v_r.to(d), s.p.to(d))
    g = torch.Generator().manual_seed(0)
    xc = torch.randn(1, nb, s.bk.npw_max, dtype=torch.complex128, generator=g).to(d)
    cres = davidson_batched(hb.apply, xc, bk.t, bk.mask, tol=tol, max_iter=300)
    assert float((res.eigenvalues - cres.eigenvalues[0]).abs().max()) / HARTREE_EV < 1e-10


def test_gamma_davidson_si64_matches_complex_solver():
    from gradwave_amd import gamma as G
    from gradwave_amd import workloads as W
    from gradwave_amd.davidson import davidson_batched
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    say("gamma davidson si64 160 bands vs complex davidson_batched")
    wl = W.si64_gamma(device=dev())
    sph = type("Sphere", (), {})()
    sph.flat_idx = wl.bk.flat_idx[0]
    sph.kpg2 = (wl.bk.kpg[0] ** 2).sum(-1)
    gb = G.build_gamma_basis(sph, wl.shape)
    gh = G.GammaHamiltonian(gb, wl.v_r, wl.p[0], wl.bk.dij_full)
    tol = 1e-10 * HARTREE_EV
    x0 = W.start_vectors(wl)[0]
    xh = G.full_to_half(gb, x0)
    xh[:, 0] = xh[:, 0].real.to(xh.dtype)
    res = G.davidson_gamma(gh, xh, tol=tol, max_iter=300)
    hb = BatchedHamiltonian(wl.bk, wl.shape, wl.v_r, wl.p)
    cres = davidson_batched(hb.apply, x0[None], wl.bk.t, wl.bk.mask, tol=tol, max_iter=300)
    say(f"gamma iters {res.n_iter} complex iters {cres.n_iter}")
    assert float(res.residual_norms.max()) <= tol
    assert float((res.eigenvalues - cres.eigenvalues[0]).abs().max()) / HARTREE_EV < 1e-10
