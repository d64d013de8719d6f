"""GPU parity of the hand-written adjoints (gradwave_amd.adjoint) against torch
autograd through the restated reference code (oracle HamiltonianK, projectors,
local_potential_g) and against central finite differences of the HIP forward.

Tolerances (float64): operator-input gradients 1e-10 relative (both sides are
exact up to FFT/GEMM rounding); position gradients 1e-6 relative to the
oracle's autograd and to central differences at h = 1e-4 bohr; complex64
inputs 1e-4 relative to the float64 gradients.
"""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from tests.gpu_cases import dev, say, select_k, system

pytestmark = pytest.mark.gpu

BOHR_ANG = 0.529177210903


def _rel(a, b):
    a = torch.as_tensor(a).detach().cpu()
    b = torch.as_tensor(b).detach().cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def _setup(name, sel, nb, seed=3):
    s = system(name)
    cpu, gpu, p = select_k(s, sel)
    d = dev()
    gpu.kpg = cpu.kpg.to(d)
    gpu.proj_phase_free = cpu.proj_phase_free.to(d)
    gpu.proj_atom_index = cpu.proj_atom_index.to(d)
    g = torch.Generator().manual_seed(seed)
    nk, npw_max = cpu.mask.shape
    c = torch.randn(nk, nb, npw_max, dtype=torch.complex128, generator=g) * cpu.mask[:, None, :]
    w = torch.randn(nk, nb, npw_max, dtype=torch.complex128, generator=g) * cpu.mask[:, None, :]
    c = c / cpu.npw[:, None, None].sqrt()
    w = w / cpu.npw[:, None, None].sqrt()
    return s, cpu, gpu, p, c, w


def _oracle_bilinear_grads(s, sel, cpu, c, w, v0, d0, p0):
    """autograd of L = Re sum_k <w_k|H_k c_k> through oracle HamiltonianK (per k)."""
    from oracle.hamiltonian import HamiltonianK, build_projector_data

    v = v0.clone().requires_grad_(True)
    dij = d0.clone().requires_grad_(True)
    cc = c.clone().requires_grad_(True)
    pp = p0.clone().requires_grad_(True)
    loss = 0.0
    for j, ik in enumerate(sel):
        n = int(cpu.npw[j])
        pd = build_projector_data(s.spheres[ik], s.species, s.grid.volume)
        pd.dij_full = dij
        h = HamiltonianK(s.spheres[ik], s.shape, v, pd, pp[j, :, :n])
        loss = loss + (w[j, :, :n].conj() * h.apply(cc[j, :, :n])).real.sum()
    gc, gv, gd, gp = torch.autograd.grad(loss, [cc, v, dij, pp])
    return float(loss), gc, gv, gd, gp


@pytest.mark.parametrize("name,sel,nb", [("si8", [0, 5], 6), ("si2", [0, 1, 2], 4)])
def test_operator_gradients_match_oracle_autograd(name, sel, nb):
    from gradwave_amd import adjoint as A

    say(f"adjoint operator grads {name} k={sel} nb={nb}")
    s, cpu, gpu, p, c, w = _setup(name, sel, nb)
    d = dev()
    v = s.v_r.to(d).requires_grad_(True)
    dij = cpu.dij_full.to(d).requires_grad_(True)
    pg = p.to(d).requires_grad_(True)
    cg = c.to(d).requires_grad_(True)
    h = A.DifferentiableHamiltonian(gpu, s.shape, v, pg, dij)
    loss = (w.to(d).conj() * h.apply(cg)).real.sum()
    gc, gv, gd, gp = torch.autograd.grad(loss, [cg, v, dij, pg])
    torch.cuda.synchronize()
    ref = _oracle_bilinear_grads(s, sel, cpu, c, w, s.v_r, cpu.dij_full, p)
    assert abs(float(loss) - ref[0]) < 1e-11 * abs(ref[0])
    assert _rel(gc, ref[1]) < 1e-10
    assert _rel(gv, ref[2]) < 1e-10
    assert _rel(gd, ref[3]) < 1e-10
    for j in range(len(sel)):
        n = int(cpu.npw[j])
        assert _rel(gp[j, :, :n], ref[4][j, :, :n]) < 1e-10


def test_nonsymmetric_d_uses_transposed_adjoint():
    from gradwave_amd import adjoint as A

    say("adjoint with non-symmetric D (si2)")
    sel = [0, 4]
    s, cpu, gpu, p, c, w = _setup("si2", sel, 4, seed=9)
    g = torch.Generator().manual_seed(2)
    d_ns = cpu.dij_full + 0.3 * torch.randn(cpu.dij_full.shape, dtype=torch.float64, generator=g)
    d = dev()
    cg = c.to(d).requires_grad_(True)
    h = A.DifferentiableHamiltonian(gpu, s.shape, s.v_r.to(d), p.to(d), d_ns.to(d).requires_grad_(True))
    loss = (w.to(d).conj() * h.apply(cg)).real.sum()
    (gc,) = torch.autograd.grad(loss, [cg])
    ref = _oracle_bilinear_grads(s, sel, cpu, c, w, s.v_r, d_ns, p)
    assert _rel(gc, ref[1]) < 1e-10


def test_complex64_gradients_follow_float64():
    from gradwave_amd import adjoint as A

    say("adjoint complex64 vs complex128 (si8)")
    s, cpu, gpu, p, c, w = _setup("si8", [1, 2], 8, seed=4)
    d = dev()
    out = {}
    for dt in (torch.complex128, torch.complex64):
        v = s.v_r.to(d).requires_grad_(True)
        cg = c.to(d, dt).requires_grad_(True)
        h = A.DifferentiableHamiltonian(gpu, s.shape, v, p.to(d), cpu.dij_full.to(d))
        loss = (w.to(d, dt).conj() * h.apply(cg)).real.sum()
        out[dt] = torch.autograd.grad(loss, [cg, v])
    assert _rel(out[torch.complex64][0].to(torch.complex128), out[torch.complex128][0]) < 1e-4
    assert _rel(out[torch.complex64][1], out[torch.complex128][1]) < 1e-4


def _tables(s):
    from oracle.pseudo import vloc_form_factor

    g_abs = np.sqrt((s.grid.g_cart.numpy() ** 2).sum(-1))
    names = sorted({sp.name for sp in s.species})
    tab = np.stack([vloc_form_factor(next(sp for sp in s.species if sp.name == nm), g_abs) for nm in names])
    sidx = torch.tensor([names.index(sp.name) for sp in s.species], dtype=torch.int64)
    return torch.as_tensor(tab), sidx


def _hip_band_energy(s, gpu, c, tab, sidx, pos):
    from gradwave_amd import adjoint as A

    d = dev()
    vg = A.local_potential_g(pos, sidx.to(d), tab.to(d), s.grid.g_cart.to(d), s.grid.volume)
    v = A.local_potential_r(vg)
    p = A.projectors_b(gpu, pos)
    h = A.DifferentiableHamiltonian(gpu, s.shape, v, p, gpu.dij_full)
    return (c.conj() * h.apply(c)).real.sum()


def _oracle_band_energy_grad(s, sel, c_cpu, npw):
    from oracle.hamiltonian import HamiltonianK, build_projector_data, projectors
    from oracle.pseudo import local_potential_g, local_potential_r

    pos = s.positions.clone().requires_grad_(True)
    v = local_potential_r(local_potential_g(pos, s.species, s.grid.g_cart, s.grid.volume))
    e = 0.0
    for j, ik in enumerate(sel):
        n = int(npw[j])
        pd = build_projector_data(s.spheres[ik], s.species, s.grid.volume)
        h = HamiltonianK(s.spheres[ik], s.shape, v, pd, projectors(pd, pos))
        cj = c_cpu[j, :, :n]
        e = e + (cj.conj() * h.apply(cj)).real.sum()
    (g,) = torch.autograd.grad(e, pos)
    return float(e), g


def test_position_gradient_matches_oracle_and_fd():
    say("position gradient (local + nonlocal) si8 vs oracle autograd and FD h=1e-4 bohr")
    sel = [0, 7]
    s, cpu, gpu, p, c, _ = _setup("si8", sel, 6, seed=21)
    tab, sidx = _tables(s)
    d = dev()
    cg = c.to(d)
    pos = s.positions.to(d).requires_grad_(True)
    e = _hip_band_energy(s, gpu, cg, tab, sidx, pos)
    (grad,) = torch.autograd.grad(e, pos)
    e_ref, g_ref = _oracle_band_energy_grad(s, sel, c, cpu.npw)
    assert abs(float(e) - e_ref) < 1e-11 * abs(e_ref)
    assert float((grad.cpu() - g_ref).norm() / g_ref.norm()) < 1e-6
    h = 1e-4 * BOHR_ANG
    gen = torch.Generator().manual_seed(8)
    with torch.no_grad():
        for trial in range(3):
            dirn = torch.randn(pos.shape, dtype=torch.float64, generator=gen).to(d)
            dirn = dirn / dirn.norm()
            ep = _hip_band_energy(s, gpu, cg, tab, sidx, pos + h * dirn)
            em = _hip_band_energy(s, gpu, cg, tab, sidx, pos - h * dirn)
            fd = float((ep - em) / (2 * h))
            ad = float((grad * dirn).sum())
            assert abs(fd - ad) < 1e-6 * float(grad.norm()), (trial, fd, ad)


def test_projector_phases_match_oracle():
    from gradwave_amd import adjoint as A
    from oracle.hamiltonian import projectors_b

    say("projector phases si8")
    s, cpu, gpu, p, _, _ = _setup("si8", [0, 3, 9], 1)
    pos = s.positions.to(dev())
    ours = A.projectors_b(gpu, pos).cpu()
    ref = projectors_b(cpu, s.positions)
    assert _rel(ours, ref) < 1e-13


def test_local_potential_g_matches_oracle():
    from gradwave_amd import adjoint as A
    from oracle.pseudo import local_potential_g

    say("local_potential_g si8")
    s = system("si8")
    tab, sidx = _tables(s)
    d = dev()
    ours = A.local_potential_g(s.positions.to(d), sidx.to(d), tab.to(d), s.grid.g_cart.to(d), s.grid.volume)
    ref = local_potential_g(s.positions, s.species, s.grid.g_cart, s.grid.volume)
    assert _rel(ours.cpu(), ref) < 1e-12


def test_adjoint_identities_at_si64_size():
    """Full Si64 (90^3) size: <grad V, dV> = Re <w| H_loc[dV] c> and
    <grad D, dD> = Re <w| H_nl[dD] c>, the Jacobian applied by the forward kernels."""
    from gradwave_amd import _lib as L
    from gradwave_amd import adjoint as A
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    say("adjoint identities si64 90^3 nb=32")
    s, cpu, gpu, p, c, w = _setup("si64", None, 32, seed=6)
    d = dev()
    v = s.v_r.to(d).requires_grad_(True)
    dij = cpu.dij_full.to(d).requires_grad_(True)
    cg, wg = c.to(d), w.to(d)
    h = A.DifferentiableHamiltonian(gpu, s.shape, v, p.to(d), dij)
    loss = (wg.conj() * h.apply(cg)).real.sum()
    gv, gd = torch.autograd.grad(loss, [v, dij])
    gen = torch.Generator().manual_seed(1)
    dv = torch.randn(s.shape, dtype=torch.float64, generator=gen).to(d)
    dd = torch.randn(cpu.dij_full.shape, dtype=torch.float64, generator=gen)
    dd = dd * (cpu.dij_full != 0)  # keep the block structure
    hv = BatchedHamiltonian(gpu, s.shape, dv, p.to(d))
    lin_v = float((wg.conj() * hv.apply_terms(cg, L.GW_TERM_LOCAL)).real.sum())
    bk_d = SimpleNamespace(npw=gpu.npw, mask=gpu.mask, flat_idx=gpu.flat_idx, t=gpu.t, dij_full=dd)
    hd = BatchedHamiltonian(bk_d, s.shape, s.v_r.to(d), p.to(d))
    lin_d = float((wg.conj() * hd.apply_terms(cg, L.GW_TERM_NONLOCAL)).real.sum())
    assert abs(float((gv * dv).sum()) - lin_v) < 1e-10 * abs(lin_v)
    assert abs(float((gd.cpu() * dd).sum()) - lin_d) < 1e-10 * abs(lin_d)


def test_band_energy_forces_si8_all_k_against_oracle():
    """C4 shape: Si8, 36^3, all 36 k-points, 24 bands, k-weighted occupations."""
    from gradwave_amd.adjoint import band_energy_forces

    say("band-energy forces si8 36 k x 24 bands vs oracle autograd")
    s, cpu, gpu, p, c, _ = _setup("si8", None, 24, seed=17)
    tab, sidx = _tables(s)
    d = dev()
    nocc = int(s.n_electrons // 2)
    w = torch.zeros(c.shape[0], c.shape[1], dtype=torch.float64)
    w[:, :nocc] = 2.0 * torch.as_tensor(s.kweights)[:, None]
    e, f = band_energy_forces(gpu, s.shape, s.positions.to(d), sidx.to(d), tab.to(d),
                              s.grid.g_cart.to(d), s.grid.volume, c.to(d), w.to(d))
    from oracle.hamiltonian import HamiltonianK, build_projector_data, projectors
    from oracle.pseudo import local_potential_g, local_potential_r

    pos = s.positions.clone().requires_grad_(True)
    v = local_potential_r(local_potential_g(pos, s.species, s.grid.g_cart, s.grid.volume))
    e_ref = 0.0
    for ik in range(c.shape[0]):
        n = int(cpu.npw[ik])
        pd = build_projector_data(s.spheres[ik], s.species, s.grid.volume)
        h = HamiltonianK(s.spheres[ik], s.shape, v, pd, projectors(pd, pos))
        ck = c[ik, :, :n]
        e_ref = e_ref + (w[ik][:, None] * (ck.conj() * h.apply(ck)).real).sum()
    (g_ref,) = torch.autograd.grad(e_ref, pos)
    assert abs(float(e) - float(e_ref)) < 1e-11 * abs(float(e_ref))
    assert float((f.cpu() + g_ref).norm() / g_ref.norm()) < 1e-6


def test_workload_geometry_tables_reproduce_p_and_v():
    """The bench's forces use the workload's phase-free projectors and V_loc tables."""
    from gradwave_amd import adjoint as A
    from gradwave_amd import workloads as W

    say("workload geometry tables si8_mp")
    wl = W.si8_mp(device=dev())
    p = A.projectors_b(wl.bk, wl.positions)
    v = A.local_potential_r(A.local_potential_g(wl.positions, wl.species_index, wl.vloc_tables, wl.g_cart,
                                                wl.volume))
    assert _rel(p, wl.p) < 1e-13
    assert _rel(v, wl.v_r) < 1e-12


def test_single_k_hamiltonian_api_matches_oracle():
    """adjoint.HamiltonianK(sphere, shape, v, pd, p).apply(c (nb, npw)) with gradients,
    against the oracle's HamiltonianK under autograd (core/hamiltonian.py:120-191)."""
    from gradwave_amd import adjoint as A
    from oracle.hamiltonian import HamiltonianK as RefK
    from oracle.hamiltonian import build_projector_data, projectors

    say("single-k HamiltonianK api si8 k=3")
    s = system("si8")
    ik = 3
    sph = s.spheres[ik]
    pd = build_projector_data(sph, s.species, s.grid.volume)
    p = projectors(pd, s.positions)
    g = torch.Generator().manual_seed(12)
    c = torch.randn(5, sph.npw, dtype=torch.complex128, generator=g) / sph.npw**0.5
    w = torch.randn(5, sph.npw, dtype=torch.complex128, generator=g) / sph.npw**0.5
    d = dev()
    sph_d = SimpleNamespace(flat_idx=sph.flat_idx.to(d), kpg2=sph.kpg2.to(d))
    pd_d = SimpleNamespace(dij_full=pd.dij_full.to(d).requires_grad_(True))
    v = s.v_r.to(d).requires_grad_(True)
    pg = p.to(d).requires_grad_(True)
    cg = c.to(d).requires_grad_(True)
    out = A.HamiltonianK(sph_d, s.shape, v, pd_d, pg).apply(cg)
    grads = torch.autograd.grad((w.to(d).conj() * out).real.sum(), [cg, v, pg, pd_d.dij_full])
    vr = s.v_r.clone().requires_grad_(True)
    pr = p.clone().requires_grad_(True)
    cr = c.clone().requires_grad_(True)
    pd.dij_full = pd.dij_full.clone().requires_grad_(True)
    ref_out = RefK(sph, s.shape, vr, pd, pr).apply(cr)
    ref = torch.autograd.grad((w.conj() * ref_out).real.sum(), [cr, vr, pr, pd.dij_full])
    assert _rel(out, ref_out) < 1e-12
    for a, b in zip(grads, ref):
        assert _rel(a, b) < 1e-10
