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
* ``band_energy_forces`` -- F = -dE/dtau at fixed orbitals, the
  Hellmann-Feynman pattern of postscf/forces.py:168-213 (positions made leaf
  tensors, energy terms built from them, one ``torch.autograd.grad``; no net
  force removal, no symmetrisation, no Ewald term: the ions' mutual energy is
  not part of H).

Pinned by: central differences of ``band_energy_psi`` in the positions
(h = 1e-4 bohr, ``tests/test_oracle_adjoint.py``) and the identity
band_energy_psi == band_energy_eigs at converged eigenvectors
(``tests/test_oracle_solvers.py``).
"""

from __future__ import annotations

import torch

from oracle.hamiltonian import HamiltonianK, build_projector_data, projectors
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
