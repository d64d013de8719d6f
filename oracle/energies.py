"""Band energy and frozen-potential forces (CPU oracle).

TEST INFRASTRUCTURE ONLY (see the package docstring).

* ``band_energy_eigs`` -- E_band = sum_k w_k sum_n f_nk eps_nk, the band sum the
  reference forms at postscf/mae.py:349.
* ``band_energy_psi`` -- the same sum at fixed orbitals as an expectation value,
  sum_k sum_n w_kn <psi_kn|H_k(tau)|psi_kn>, with H_k built from the atomic
  positions tau through the structure factor (local part,
  core/energies/local_pp.py:22-38 restated in ``oracle.pseudo``) and the
  projector phases (core/hamiltonian.py:108-112), so that torch autograd
  differentiates it in tau.
* ``band_energy_forces`` -- F = -dE/dtau at fixed orbitals by autograd of
  ``band_energy_psi`` (positions made leaf tensors, one ``torch.autograd.grad``).
* ``band_energy_forces_hf`` -- the same forces in the form of
  postscf/forces.py:168-213: the position-dependent energy at fixed orbitals is
  the local term from the density, E_loc = Omega sum_G rho*(G) V_loc(G)
  (core/energies/local_pp.py:41-43), plus the nonlocal term
  sum_k sum_b w_kb sum_ij conj(b_bi) D_ij b_bj with b = <beta|psi>
  (core/energies/nl_pp.py:19-31); bands are taken in chunks, so it runs at the
  216-atom size. No net-force removal, symmetrisation or Ewald term (the ions'
  mutual energy is not part of H).

Pinned by: central differences of ``band_energy_psi`` in the positions
(h = 1e-4 bohr, ``tests/test_oracle_adjoint.py``), ``band_energy_forces_hf``
against ``band_energy_forces`` (two different routes to the same derivative,
same file), and the identity band_energy_psi == band_energy_eigs at converged
eigenvectors (``tests/test_oracle_solvers.py``).
"""

from __future__ import annotations

import torch

from oracle.hamiltonian import (HamiltonianK, ProjectorData, build_projector_data, density_b, projectors,
                               r_to_g)
from oracle.pseudo import local_potential_g, local_potential_r


def band_energy_eigs(eig: torch.Tensor, occ: torch.Tensor, kweights: torch.Tensor) -> torch.Tensor:
    """postscf/mae.py:349 -- sum_k w_k sum_n f_nk eps_nk (eig, occ (nk, nb); kweights (nk,))."""
    return (kweights.to(eig.dtype)[:, None] * occ.to(eig.dtype) * eig).sum()


def band_energy_psi(system, coeffs: torch.Tensor, weights: torch.Tensor, positions: torch.Tensor,
                    kpoints=None) -> torch.Tensor:
    """sum_k sum_n weights[k, n] Re <c_kn|H_k(positions)|c_kn> in eV.

    coeffs (nk, nb, npw_max) padded complex128 over the k rows ``kpoints`` of
    ``system`` (all rows when None); weights (nk, nb) real (k weight times
    occupation). Differentiable in ``positions`` (Angstrom).
    """
    ks = list(range(len(system.spheres))) if kpoints is None else list(kpoints)
    v = local_potential_r(local_potential_g(positions, system.species, system.grid.g_cart, system.grid.volume))
    e = positions.new_zeros(())
    for j, ik in enumerate(ks):
        sph = system.spheres[ik]
        pd = build_projector_data(sph, system.species, system.grid.volume)
        h = HamiltonianK(sph, system.shape, v, pd, projectors(pd, positions))
        c = coeffs[j, :, : sph.npw]
        e = e + (weights[j].to(e.dtype)[:, None] * (c.conj() * h.apply(c)).real).sum()
    return e


def band_energy_forces(system, coeffs: torch.Tensor, weights: torch.Tensor, positions: torch.Tensor | None = None,
                       kpoints=None) -> tuple[float, torch.Tensor]:
    """postscf/forces.py:168-213 pattern -- (E_band [eV], F = -dE/dtau (na, 3) [eV/A])."""
    pos = (system.positions if positions is None else positions).detach().clone().requires_grad_(True)
    e = band_energy_psi(system, coeffs.detach(), weights.detach(), pos, kpoints)
    (grad,) = torch.autograd.grad(e, pos)
    return float(e), -grad


def band_energy_forces_hf(system, coeffs: torch.Tensor, weights: torch.Tensor, positions: torch.Tensor | None = None,
                          kpoints=None, band_chunk: int = 32, atoms=None) -> tuple[float, torch.Tensor]:
    """postscf/forces.py:168-213 -- (E_band [eV], F [eV/A]) at fixed orbitals.

    E_band = sum_kb w_kb <c|T|c> + E_loc(tau) + E_nl(tau); only the last two
    depend on the positions. coeffs (nk, nb, npw_max) over the k rows
    ``kpoints`` (all when None), weights (nk, nb). Both position-dependent
    terms are sums over atoms (V_loc = sum_a V_a; D is block-diagonal by atom,
    core/hamiltonian.py:78-93), so the force on atom a is the autograd gradient
    of atom a's terms alone: F is (len(atoms), 3) for the listed atoms (all
    atoms when None), and the autograd graph stays one box in size."""
    ks = list(range(len(system.spheres))) if kpoints is None else list(kpoints)
    pos = (system.positions if positions is None else positions).detach().clone()
    atoms = list(range(pos.shape[0])) if atoms is None else [int(a) for a in atoms]
    coeffs, weights = coeffs.detach(), weights.detach().to(torch.float64)
    grid = system.grid
    nb = coeffs.shape[1]
    # density of the weighted orbitals, band chunk by band chunk (weights as occupations, unit k weights)
    rho = torch.zeros(system.shape, dtype=torch.float64)
    e_kin = 0.0
    for j, ik in enumerate(ks):
        sph = system.spheres[ik]
        n = sph.npw
        one = type("K", (), {})()
        one.flat_idx = sph.flat_idx[None]
        for lo in range(0, nb, band_chunk):
            hi = min(nb, lo + band_chunk)
            c = coeffs[j : j + 1, lo:hi, :n]
            rho += density_b(c, weights[j : j + 1, lo:hi], torch.ones(1, dtype=torch.float64), one, system.shape,
                             grid.volume)
            e_kin += float((weights[j, lo:hi, None] * system.bk.t[ik, None, :n] * c[0].abs() ** 2).sum())
    rho_g_conj = r_to_g(rho.to(torch.complex128)).conj()
    e_loc = float(grid.volume * (rho_g_conj * local_potential_g(pos, system.species, grid.g_cart,
                                                                grid.volume)).sum().real)
    pds = {ik: build_projector_data(system.spheres[ik], system.species, grid.volume) for ik in ks}

    def e_nl_of(p, dij, j):
        e = 0.0
        for lo in range(0, nb, band_chunk):
            hi = min(nb, lo + band_chunk)
            b = coeffs[j, lo:hi, : p.shape[1]] @ p.conj().T  # (nbc, nproj) = <beta|psi>
            quad = torch.einsum("bi,ij,bj->b", b.conj(), dij, b).real
            e = e + (weights[j, lo:hi] * quad).sum()
        return e

    e_nl = 0.0
    with torch.no_grad():
        for j, ik in enumerate(ks):
            e_nl += float(e_nl_of(projectors(pds[ik], pos), pds[ik].dij_full.to(torch.complex128), j))
    forces = torch.zeros(len(atoms), 3, dtype=torch.float64)
    for row, a in enumerate(atoms):
        pa = pos[a : a + 1].clone().requires_grad_(True)
        e_a = grid.volume * (rho_g_conj * local_potential_g(pa, [system.species[a]], grid.g_cart,
                                                            grid.volume)).sum().real
        for j, ik in enumerate(ks):
            pd = pds[ik]
            sel = (pd.atom_index == a).nonzero().flatten()
            one = ProjectorData(atom_index=torch.zeros(len(sel), dtype=torch.int64),
                                f_ylm_phase_free=pd.f_ylm_phase_free[sel], kpg=pd.kpg,
                                dij_full=pd.dij_full[sel][:, sel])
            e_a = e_a + e_nl_of(projectors(one, pa), one.dij_full.to(torch.complex128), j)
        (ga,) = torch.autograd.grad(e_a, pa)
        forces[row] = -ga[0]
    return e_kin + e_loc + e_nl, forces
