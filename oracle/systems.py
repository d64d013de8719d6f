"""Synthetic Si/Al workloads on the analytic pseudopotential (CPU oracle side).

Builds cells, atom positions (with seeded jitter), k-lists and the reference's
setup products for them: FFT box (grids.build_fft_grid, optionally pinned with
shape_override as grids.py:73-96 allows), per-k G-spheres, projector data,
BatchedK (core/batch.py:174-205), projectors (core/batch.py:229-235) and the
local potential V(r) (scf/loop.py:532-544).

k-lists: Monkhorst-Pack in the QE convention k_i = (m + s/2)/n folded to
(-1/2, 1/2] (kpoints.py:15-54). The 10-point irreducible set of the shifted
4x4x4 mesh of fcc Si is the O_h + time-reversal reduction, computed here by
star enumeration (the reference can only reduce unshifted meshes, so callers
pass this list explicitly).

Pinned by: the reduced weights (2,6,6,6,6,12,12,6,2,6)/64 and npw counts
computed by brute force in tests/test_oracle_setup.py.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np
import torch

from oracle import grids as G
from oracle.hamiltonian import BatchedK, build_batched, build_projector_data, projectors_b
from oracle.pseudo import AL, SI, AnalyticSpecies, local_potential_g, local_potential_r
from oracle.units import RY_EV

A_SI = 5.43
A_AL = 4.05


def fcc_primitive(a: float) -> np.ndarray:
    return a / 2 * np.array([[0.0, 1, 1], [1, 0, 1], [1, 1, 0]])


def si_conventional(a: float = A_SI, reps=(1, 1, 1)):
    """Diamond Si conventional cubic cell repeated reps times: (cell, positions A)."""
    fcc = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    basis = np.concatenate([fcc, fcc + 0.25])
    frac = []
    for i, j, k in itertools.product(range(reps[0]), range(reps[1]), range(reps[2])):
        frac.append(basis + np.array([i, j, k]))
    frac = np.concatenate(frac) / np.array(reps, dtype=np.float64)
    cell = np.diag(np.array(reps, dtype=np.float64) * a)
    return cell, frac @ cell


def monkhorst_pack(mesh, shift=(0, 0, 0)) -> np.ndarray:
    """kpoints.py:15-37 (full mesh, no reduction), folded to (-1/2, 1/2]."""
    grids = [(np.arange(n) + 0.5 * s) / n for n, s in zip(mesh, shift)]
    k = np.stack(np.meshgrid(*grids, indexing="ij"), axis=-1).reshape(-1, 3)
    return -((-k + 0.5) % 1.0 - 0.5)


def time_reversal_reduce(k: np.ndarray):
    """kpoints.py:38-50 -- k and -k identified; weights sum to 1."""
    fold = lambda x: -((-x + 0.5) % 1.0 - 0.5)  # noqa: E731
    w = np.full(len(k), 1.0 / len(k))
    seen, keep_k, keep_w = {}, [], []
    for ki, wi in zip(k, w):
        key = tuple(np.round(ki, 9))
        neg = tuple(np.round(fold(-ki), 9))
        if neg in seen:
            keep_w[seen[neg]] += wi
        else:
            seen[key] = len(keep_k)
            keep_k.append(ki)
            keep_w.append(wi)
    return np.array(keep_k), np.array(keep_w)


def cubic_rotations() -> list[np.ndarray]:
    """The 48 O_h point operations as signed 3x3 permutation matrices."""
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3))
            for i in range(3):
                m[i, perm[i]] = signs[i]
            ops.append(m)
    return ops


def oh_reduce(k_frac: np.ndarray, cell: np.ndarray):
    """Reduce a k-mesh of a cubic-symmetric crystal by O_h (+ inversion = TR)."""
    b = G.reciprocal_cell(cell)
    binv = np.linalg.inv(b)
    nk = len(k_frac)
    kc = k_frac @ b
    ops = cubic_rotations()
    rep = -np.ones(nk, dtype=np.int64)
    reps, weights = [], []
    for i in range(nk):
        if rep[i] >= 0:
            continue
        star = [j for j in range(nk) if rep[j] < 0 and any(
            np.allclose(np.round(d := (kc[i] @ op.T) @ binv - k_frac[j]), d, atol=1e-8)
            for op in ops)]
        for j in star:
            rep[j] = len(reps)
        reps.append(k_frac[i])
        weights.append(len(star) / nk)
    return np.array(reps), np.array(weights)


def jitter(positions: np.ndarray, delta: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return positions + rng.uniform(-delta, delta, positions.shape)


@dataclass
class PWSystem:
    """Everything BatchedHamiltonian and davidson_batched consume."""

    name: str
    cell: np.ndarray
    positions: torch.Tensor  # (na,3) A
    species: list[AnalyticSpecies]
    ecut: float  # eV
    grid: G.FFTGrid
    kfrac: np.ndarray
    kweights: np.ndarray
    spheres: list
    bk: BatchedK
    p: torch.Tensor  # (nk, nproj, npw_max)
    v_r: torch.Tensor  # (n1,n2,n3)
    nbands: int

    @property
    def shape(self):
        return self.grid.shape

    @property
    def n_electrons(self) -> float:
        return float(sum(sp.z for sp in self.species))


def build_system(name, cell, positions, species, ecut_ry, kfrac, kweights, nbands,
                 shape=None) -> PWSystem:
    ecut = ecut_ry * RY_EV
    grid = G.build_fft_grid(cell, ecut, shape_override=shape)
    spheres = [G.build_gsphere(grid, ecut, k) for k in kfrac]
    pds = [build_projector_data(s, species, grid.volume) for s in spheres]
    bk = build_batched(spheres, pds)
    pos = torch.as_tensor(np.asarray(positions), dtype=torch.float64)
    p = projectors_b(bk, pos)
    v_r = local_potential_r(local_potential_g(pos, species, grid.g_cart, grid.volume))
    return PWSystem(name, np.asarray(cell, dtype=np.float64), pos, list(species), ecut, grid,
                    np.asarray(kfrac, dtype=np.float64), np.asarray(kweights, dtype=np.float64),
                    spheres, bk, p, v_r, nbands)


def si2_c1(nbands: int = 8) -> PWSystem:
    """Bulk Si diamond fcc a=5.43, 30 Ry, 24^3 box, shifted 4x4x4 MP reduced to 10 k."""
    cell = fcc_primitive(A_SI)
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) * A_SI
    k, w = oh_reduce(monkhorst_pack((4, 4, 4), (1, 1, 1)), cell)
    return build_system("si2_c1", cell, pos, [SI, SI], 30.0, k, w, nbands, shape=(24, 24, 24))


def si64_c2(nbands: int = 160, seed: int = 0) -> PWSystem:
    """Si64 Gamma (2x2x2 conventional, L=10.86 A), jitter U(+-0.02 A), 40 Ry, 90^3."""
    cell, pos = si_conventional(reps=(2, 2, 2))
    pos = jitter(pos, 0.02, seed)
    return build_system("si64_c2", cell, pos, [SI] * 64, 40.0, np.zeros((1, 3)), np.ones(1),
                        nbands, shape=(90, 90, 90))


def si8_c4(nbands: int = 24, seed: int = 0) -> PWSystem:
    """Si8 conventional, jitter U(+-0.05 A), 30 Ry, 36^3, 4x4x4 MP (time reversal only)."""
    cell, pos = si_conventional()
    pos = jitter(pos, 0.05, seed)
    k, w = time_reversal_reduce(monkhorst_pack((4, 4, 4)))
    return build_system("si8_c4", cell, pos, [SI] * 8, 30.0, k, w, nbands, shape=(36, 36, 36))


def si216_c5(nbands: int = 480, seed: int = 0) -> PWSystem:
    """Si216 Gamma (3x3x3 conventional, L=16.29 A), jitter U(+-0.03 A), 30 Ry, 120^3."""
    cell, pos = si_conventional(reps=(3, 3, 3))
    pos = jitter(pos, 0.03, seed)
    return build_system("si216_c5", cell, pos, [SI] * 216, 30.0, np.zeros((1, 3)), np.ones(1),
                        nbands, shape=(120, 120, 120))


def al1(ecut_ry: float, kfrac, nbands: int) -> PWSystem:
    cell = fcc_primitive(A_AL)
    return build_system("al1", cell, np.zeros((1, 3)), [AL], ecut_ry, np.atleast_2d(kfrac),
                        np.ones(len(np.atleast_2d(kfrac))), nbands)
