"""Pins the gradient oracle: torch autograd through the restated single-k
HamiltonianK.apply, projectors and local_potential_g (core/hamiltonian.py:108-191,
core/energies/local_pp.py:22-38) against central finite differences.

The GPU adjoint kernels are checked against these autograd gradients; here
the autograd gradients themselves are checked against the loss they
differentiate, on Si2 (24^3, two k-points), with the step h = 1e-4 bohr for
positions that the differentiation form prescribes.
"""

from __future__ import annotations

import pytest
import torch

from oracle import systems as S
from oracle.hamiltonian import HamiltonianK, build_projector_data, projectors
from oracle.pseudo import local_potential_g, local_potential_r

BOHR_ANG = 0.529177210903
SEL = (0, 3)


@pytest.fixture(scope="module")
def case():
    s = S.si2_c1()
    g = torch.Generator().manual_seed(11)
    cs, ws, pds = {}, {}, {}
    for ik in SEL:
        n = s.spheres[ik].npw
        cs[ik] = torch.randn(4, n, dtype=torch.complex128, generator=g) / n**0.5
        ws[ik] = torch.randn(4, n, dtype=torch.complex128, generator=g) / n**0.5
        pds[ik] = build_projector_data(s.spheres[ik], s.species, s.grid.volume)
    return s, cs, ws, pds


def _band_energy(s, cs, pds, pos):
    v = local_potential_r(local_potential_g(pos, s.species, s.grid.g_cart, s.grid.volume))
    e = 0.0
    for ik in SEL:
        h = HamiltonianK(s.spheres[ik], s.shape, v, pds[ik], projectors(pds[ik], pos))
        e = e + (cs[ik].conj() * h.apply(cs[ik])).real.sum()
    return e


def _bilinear(s, cs, ws, pds, v, dij, ps):
    """L = Re sum_k <w_k| H_k c_k> with V, D and the projectors as free inputs."""
    e = 0.0
    for ik in SEL:
        pd = build_projector_data(s.spheres[ik], s.species, s.grid.volume)
        pd.dij_full = dij
        h = HamiltonianK(s.spheres[ik], s.shape, v, pd, ps[ik])
        e = e + (ws[ik].conj() * h.apply(cs[ik])).real.sum()
    return e


def test_position_gradient_matches_central_fd(case):
    s, cs, ws, pds = case
    pos = s.positions.clone().requires_grad_(True)
    (grad,) = torch.autograd.grad(_band_energy(s, cs, pds, pos), pos)
    h = 1e-4 * BOHR_ANG
    fd = torch.zeros_like(grad)
    with torch.no_grad():
        for a in range(pos.shape[0]):
            for x in range(3):
                d = torch.zeros_like(pos)
                d[a, x] = h
                fd[a, x] = (_band_energy(s, cs, pds, pos + d) - _band_energy(s, cs, pds, pos - d)) / (2 * h)
    assert float((fd - grad).norm() / grad.norm()) < 1e-6


def test_energies_module_forces_match_central_fd(case):
    """oracle.energies.band_energy_forces (the forces.py:168-213 pattern) on a
    padded block with per-band weights: F = -dE/dtau against central FD."""
    from oracle.energies import band_energy_forces, band_energy_forces_hf, band_energy_psi

    s, cs, _, _ = case
    m = max(int(s.spheres[ik].npw) for ik in SEL)
    c = torch.zeros(len(SEL), 4, m, dtype=torch.complex128)
    for j, ik in enumerate(SEL):
        c[j, :, : cs[ik].shape[1]] = cs[ik]
    w = torch.tensor([[2.0, 2.0, 1.0, 0.5], [0.25, 2.0, 0.0, 1.0]], dtype=torch.float64)
    e, f = band_energy_forces(s, c, w, kpoints=SEL)
    h = 1e-4 * BOHR_ANG
    fd = torch.zeros_like(f)
    with torch.no_grad():
        for a in range(f.shape[0]):
            for x in range(3):
                d = torch.zeros_like(s.positions)
                d[a, x] = h
                ep = band_energy_psi(s, c, w, s.positions + d, kpoints=SEL)
                em = band_energy_psi(s, c, w, s.positions - d, kpoints=SEL)
                fd[a, x] = -(ep - em) / (2 * h)
    assert float((fd - f).norm() / f.norm()) < 1e-6
    assert abs(e - float(band_energy_psi(s, c, w, s.positions, kpoints=SEL))) == 0.0
    # the forces.py route (density local term + <beta|psi> nonlocal term, chunked bands)
    e_hf, f_hf = band_energy_forces_hf(s, c, w, kpoints=SEL, band_chunk=3)
    assert abs(e_hf - e) < 1e-12 * abs(e)
    assert float((f_hf - f).norm() / f.norm()) < 1e-10


def test_operator_input_gradients_match_directional_fd(case):
    s, cs, ws, pds = case
    g = torch.Generator().manual_seed(5)
    v0 = s.v_r.clone()
    d0 = pds[SEL[0]].dij_full.clone()
    p0 = {ik: projectors(pds[ik], s.positions) for ik in SEL}
    v = v0.clone().requires_grad_(True)
    dij = d0.clone().requires_grad_(True)
    ps = {ik: p0[ik].clone().requires_grad_(True) for ik in SEL}
    loss = _bilinear(s, cs, ws, pds, v, dij, ps)
    grads = torch.autograd.grad(loss, [v, dij, *ps.values()])
    dv = torch.randn(v0.shape, dtype=torch.float64, generator=g)
    dd = torch.randn(d0.shape, dtype=torch.float64, generator=g)
    dp = {ik: torch.randn(p0[ik].shape, dtype=torch.complex128, generator=g) for ik in SEL}
    eps = 1e-6
    with torch.no_grad():
        def f(t):
            return _bilinear(s, cs, ws, pds, v0 + t * dv, d0 + t * dd,
                             {ik: p0[ik] + t * dp[ik] for ik in SEL})
        fd = (f(eps) - f(-eps)) / (2 * eps)
    # directional derivative from the gradients (torch convention: dL = Re sum conj(grad) dz)
    ad = (grads[0] * dv).sum() + (grads[1] * dd).sum()
    for gp, ik in zip(grads[2:], SEL):
        ad = ad + (gp.conj() * dp[ik]).real.sum()
    assert abs(float(fd - ad)) < 1e-7 * max(1.0, abs(float(ad)))
