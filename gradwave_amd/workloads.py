"""Synthetic Si/Al plane-wave workloads on the analytic Gaussian pseudopotential.

Builds the inputs gradwave's H-apply and Davidson consume -- the BatchedK maps
(core/batch.py:126-205), projectors (core/batch.py:229-235, core/hamiltonian.py:46-112)
and V(r) (scf/loop.py:532-544) -- for the cells of the workloads:

* si2_fcc   bulk Si, fcc a=5.43 A, 30 Ry, 24^3 box, 10 k (shifted 4x4x4 MP, O_h+TR)
* si64_gamma  Si64 2x2x2 conventional, jitter U(+-0.02 A), 40 Ry, 90^3, Gamma
* si8_mp    Si8 conventional, jitter U(+-0.05 A), 30 Ry, 36^3, 4x4x4 MP (TR only)
* si216_gamma Si216 3x3x3 conventional, jitter U(+-0.03 A), 30 Ry, 120^3, Gamma

The G-sphere order and flat box index must equal the reference's bit for bit,
so the sphere construction keeps the reference's arithmetic (grids.py:145-184:
(m + k) @ b, einsum |k+G|^2, 1e-12 relative cutoff slack, lexsort on
round(|k+G|^2, 10) then Miller indices). Projector and potential tables are
assembled on the target device in torch.
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass

import numpy as np
import torch

HARTREE_EV = 27.211386245988
RY_EV = HARTREE_EV / 2.0
BOHR_ANG = 0.529177210903
HBAR2_2M = RY_EV * BOHR_ANG**2
E2 = HARTREE_EV * BOHR_ANG
A_SI, A_AL = 5.43, 4.05


@dataclass(frozen=True)
class Species:
    z: float
    rc_bohr: float
    w_bohr: float = 1.0
    d_s_ha: float = 1.0
    d_p_ha: float = -0.5


SI = Species(4.0, 0.8)
AL = Species(3.0, 0.6)


@dataclass
class BatchedK:
    """Padded per-k maps (core/batch.py:126-171), tensors on the target device."""

    npw: torch.Tensor
    mask: torch.Tensor
    flat_idx: torch.Tensor
    kpg: torch.Tensor
    t: torch.Tensor
    dij_full: torch.Tensor
    proj_atom_index: torch.Tensor
    proj_phase_free: torch.Tensor | None = None  # (nk, nproj, npw_max), e^{-i(k+G).tau} left out

    @property
    def nk(self) -> int:
        return int(self.npw.shape[0])

    @property
    def npw_max(self) -> int:
        return int(self.mask.shape[1])


@dataclass
class Workload:
    name: str
    cell: np.ndarray
    positions: torch.Tensor  # (na, 3) A, on device
    shape: tuple[int, int, int]
    ecut: float  # eV
    kfrac: np.ndarray
    kweights: np.ndarray
    bk: BatchedK
    p: torch.Tensor  # (nk, nproj, npw_max) complex128
    v_r: torch.Tensor  # (n1, n2, n3) float64
    nbands: int
    n_electrons: float
    # geometry inputs of the position-differentiable path (local_pp.py:22-38)
    volume: float = 0.0
    species_index: torch.Tensor | None = None  # (na,) row of each atom in vloc_tables
    vloc_tables: torch.Tensor | None = None  # (nspecies, n1, n2, n3) eV A^3, G=0 entry alpha
    g_cart: torch.Tensor | None = None  # (n1, n2, n3, 3) A^-1
    ecut_ry: float = 0.0  # the cutoff as given (Ry); ecut = ecut_ry * RY_EV
    species: tuple = ()  # Species of each atom


def reciprocal(cell: np.ndarray) -> np.ndarray:
    return 2.0 * np.pi * np.linalg.inv(cell).T


def sphere(cell: np.ndarray, shape, ecut: float, k_frac) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(miller, |k+G|^2, flat index) in the reference's order (grids.py:145-184)."""
    b = reciprocal(cell)
    k_frac = np.asarray(k_frac, dtype=np.float64)
    gmax = math.sqrt(ecut / HBAR2_2M)
    bound = [int(np.floor(gmax * np.linalg.norm(cell[i]) / (2 * np.pi) + 1)) + 1 for i in range(3)]
    mm = np.stack(np.meshgrid(*[np.arange(-bd, bd + 1) for bd in bound], indexing="ij"),
                  axis=-1).reshape(-1, 3)
    kpg = (mm + k_frac) @ b
    kpg2 = np.einsum("ij,ij->i", kpg, kpg)
    keep = HBAR2_2M * kpg2 <= ecut * (1 + 1e-12)
    mm, kpg2 = mm[keep], kpg2[keep]
    n1, n2, n3 = shape
    if np.any(np.abs(mm) > np.array([n1 // 2, n2 // 2, n3 // 2])):
        raise RuntimeError("wavefunction sphere exceeds FFT box")
    flat = (mm[:, 0] % n1) * (n2 * n3) + (mm[:, 1] % n2) * n3 + (mm[:, 2] % n3)
    order = np.lexsort((mm[:, 2], mm[:, 1], mm[:, 0], np.round(kpg2, 10)))
    return mm[order], kpg2[order], flat[order]


def fcc_primitive(a: float) -> np.ndarray:
    return a / 2 * np.array([[0.0, 1, 1], [1, 0, 1], [1, 1, 0]])


def si_conventional(reps=(1, 1, 1), a: float = A_SI):
    fcc = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    basis = np.concatenate([fcc, fcc + 0.25])
    frac = np.concatenate([basis + np.array(ijk) for ijk in itertools.product(*map(range, reps))])
    frac = frac / np.array(reps, dtype=np.float64)
    cell = np.diag(np.array(reps, dtype=np.float64) * a)
    return cell, frac @ cell


def mp_mesh(mesh, shift=(0, 0, 0)) -> np.ndarray:
    grids = [(np.arange(n) + 0.5 * s) / n for n, s in zip(mesh, shift)]
    k = np.stack(np.meshgrid(*grids, indexing="ij"), axis=-1).reshape(-1, 3)
    return -((-k + 0.5) % 1.0 - 0.5)


def reduce_mesh(k: np.ndarray, cell: np.ndarray, full_cubic: bool):
    """Star reduction of a mesh under O_h (full_cubic) or inversion only (time reversal)."""
    b = reciprocal(cell)
    binv = np.linalg.inv(b)
    kc = k @ b
    if full_cubic:
        ops = []
        for perm in itertools.permutations(range(3)):
            for signs in itertools.product((1, -1), repeat=3):
                m = np.zeros((3, 3))
                m[range(3), perm] = signs
                ops.append(m)
    else:
        ops = [np.eye(3), -np.eye(3)]
    nk = len(k)
    owner = -np.ones(nk, dtype=np.int64)
    reps, weights = [], []
    for i in range(nk):
        if owner[i] >= 0:
            continue
        images = np.stack([(kc[i] @ op.T) @ binv for op in ops])  # (nops, 3) fractional
        cnt = 0
        for j in range(i, nk):
            if owner[j] >= 0:
                continue
            d = images - k[j]
            if np.any(np.all(np.abs(d - np.round(d)) < 1e-8, axis=1)):
                owner[j] = len(reps)
                cnt += 1
        reps.append(k[i])
        weights.append(cnt / nk)
    return np.array(reps), np.array(weights)


def kb_form_factor(sp: Species, l: int, q: torch.Tensor) -> torch.Tensor:
    w = sp.w_bohr * BOHR_ANG
    return math.sqrt(math.pi / math.gamma(l + 1.5)) * w ** (l + 1.5) * q**l * torch.exp(-0.5 * (q * w) ** 2)


def build(name: str, cell, positions, species: list[Species], ecut_ry: float, kfrac, kweights,
          nbands: int, shape, device) -> Workload:
    cell = np.asarray(cell, dtype=np.float64)
    ecut = ecut_ry * RY_EV
    b = reciprocal(cell)
    vol = float(abs(np.linalg.det(cell)))
    kfrac = np.atleast_2d(np.asarray(kfrac, dtype=np.float64))
    sph = [sphere(cell, shape, ecut, k) for k in kfrac]
    nk = len(sph)
    npw = np.array([len(s[0]) for s in sph])
    m = int(npw.max())
    flat = np.zeros((nk, m), dtype=np.int64)
    mill = np.zeros((nk, m, 3), dtype=np.float64)
    mask = np.zeros((nk, m), dtype=bool)
    t = np.zeros((nk, m))
    for i, (mm, kpg2, fl) in enumerate(sph):
        n = len(fl)
        flat[i, :n] = fl
        mill[i, :n] = mm
        mask[i, :n] = True
        t[i, :n] = HBAR2_2M * kpg2
    dev = torch.device(device)
    f64 = torch.float64
    kpg = torch.as_tensor((mill + kfrac[:, None, :]) @ b, dtype=f64, device=dev)
    kpg = kpg * torch.as_tensor(mask, device=dev)[..., None]
    pos = torch.as_tensor(np.asarray(positions, dtype=np.float64), device=dev)
    na = pos.shape[0]
    # projectors: per atom one s and three p columns, (4 pi / sqrt Omega)(-i)^l Y_lm F_l e^{-i(k+G).tau}
    q = torch.sqrt((kpg * kpg).sum(-1))
    unit = kpg / torch.where(q > 1e-14, q, torch.ones_like(q))[..., None]
    c00 = 0.5 * math.sqrt(1.0 / math.pi)
    c1 = math.sqrt(3.0 / (4.0 * math.pi))
    pref = 4.0 * math.pi / math.sqrt(vol)
    cols, free, atom_idx, dblocks = [], [], [], []
    for a, sp in enumerate(species):
        f0 = kb_form_factor(sp, 0, q)
        f1 = kb_form_factor(sp, 1, q)
        phase = torch.exp(-1j * (kpg @ pos[a]).to(torch.complex128))
        ang = [torch.full_like(q, c00) * f0 + 0j,
               -1j * (c1 * unit[..., 2] * f1),
               -1j * (c1 * unit[..., 0] * f1),
               -1j * (c1 * unit[..., 1] * f1)]
        for col in ang:
            cols.append(pref * col * phase)
            free.append(pref * col)
            atom_idx.append(a)
        dblocks.append(torch.diag(torch.tensor([sp.d_s_ha, sp.d_p_ha, sp.d_p_ha, sp.d_p_ha],
                                               dtype=f64) * HARTREE_EV))
    mk = torch.as_tensor(mask, device=dev)[:, None, :]
    p = torch.stack(cols, dim=1) * mk
    pf = torch.stack(free, dim=1) * mk
    del cols, free
    dij = torch.block_diag(*dblocks)
    # local potential on the full box, separable structure factor per atom
    n1, n2, n3 = shape
    freqs = [torch.as_tensor(np.fft.fftfreq(n, d=1.0 / n), dtype=f64, device=dev) for n in shape]
    m0, m1, m2 = torch.meshgrid(*freqs, indexing="ij")
    bt = torch.as_tensor(b, dtype=f64, device=dev)
    g = m0[..., None] * bt[0] + m1[..., None] * bt[1] + m2[..., None] * bt[2]
    gabs = torch.sqrt((g * g).sum(-1))
    frac = pos @ torch.as_tensor(np.linalg.inv(cell), dtype=f64, device=dev)
    vg = torch.zeros(shape, dtype=torch.complex128, device=dev)
    by_species: dict[Species, list[int]] = {}
    for a, sp in enumerate(species):
        by_species.setdefault(sp, []).append(a)
    tables = []
    for sp, atoms in by_species.items():
        rc = sp.rc_bohr * BOHR_ANG
        gz = torch.where(gabs > 1e-10, gabs, torch.ones_like(gabs))
        v = -4.0 * math.pi * sp.z * E2 * torch.exp(-0.25 * (gz * rc) ** 2) / gz**2
        v = torch.where(gabs > 1e-10, v, torch.full_like(v, math.pi * sp.z * E2 * rc**2))
        tables.append(v)
        s = torch.zeros(shape, dtype=torch.complex128, device=dev)
        for a in atoms:
            e = [torch.exp(-2j * math.pi * freqs[i].to(torch.complex128) * frac[a, i]) for i in range(3)]
            s += e[0][:, None, None] * e[1][None, :, None] * e[2][None, None, :]
        vg += s * v
    vg /= vol
    v_r = (torch.fft.ifftn(vg) * (n1 * n2 * n3)).real.contiguous()
    bk = BatchedK(npw=torch.as_tensor(npw, dtype=torch.int64, device=dev),
                  mask=torch.as_tensor(mask, device=dev),
                  flat_idx=torch.as_tensor(flat, device=dev), kpg=kpg,
                  t=torch.as_tensor(t, dtype=f64, device=dev), dij_full=dij.to(dev),
                  proj_atom_index=torch.as_tensor(atom_idx, dtype=torch.int64, device=dev),
                  proj_phase_free=pf.contiguous())
    order = list(by_species)
    sidx = torch.tensor([order.index(sp) for sp in species], dtype=torch.int64, device=dev)
    return Workload(name, cell, pos, tuple(shape), ecut, kfrac, np.asarray(kweights, dtype=np.float64),
                    bk, p.contiguous(), v_r, nbands, float(sum(sp.z for sp in species)),
                    volume=vol, species_index=sidx, vloc_tables=torch.stack(tables).contiguous(),
                    g_cart=g.contiguous(), ecut_ry=float(ecut_ry), species=tuple(species))


def jitter(pos: np.ndarray, delta: float, seed: int) -> np.ndarray:
    return pos + np.random.default_rng(seed).uniform(-delta, delta, pos.shape)


def si2_fcc(device="cuda", nbands: int = 8) -> Workload:
    cell = fcc_primitive(A_SI)
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) * A_SI
    k, w = reduce_mesh(mp_mesh((4, 4, 4), (1, 1, 1)), cell, full_cubic=True)
    return build("si2_fcc_30Ry_24^3_10k", cell, pos, [SI, SI], 30.0, k, w, nbands, (24, 24, 24), device)


def si64_gamma(device="cuda", nbands: int = 160, seed: int = 0) -> Workload:
    cell, pos = si_conventional((2, 2, 2))
    pos = jitter(pos, 0.02, seed)
    return build("si64_gamma_40Ry_90^3", cell, pos, [SI] * 64, 40.0, np.zeros((1, 3)), [1.0], nbands,
                 (90, 90, 90), device)


def si8_mp(device="cuda", nbands: int = 24, seed: int = 0) -> Workload:
    cell, pos = si_conventional()
    pos = jitter(pos, 0.05, seed)
    k, w = reduce_mesh(mp_mesh((4, 4, 4)), cell, full_cubic=False)
    return build("si8_30Ry_36^3_mp444", cell, pos, [SI] * 8, 30.0, k, w, nbands, (36, 36, 36), device)


def si216_gamma(device="cuda", nbands: int = 480, seed: int = 0) -> Workload:
    cell, pos = si_conventional((3, 3, 3))
    pos = jitter(pos, 0.03, seed)
    return build("si216_gamma_30Ry_120^3", cell, pos, [SI] * 216, 30.0, np.zeros((1, 3)), [1.0],
                 nbands, (120, 120, 120), device)


def start_vectors(wl: Workload, seed: int = 0, dtype=torch.complex128) -> torch.Tensor:
    """Seeded complex N(0,1) start block, masked (test_davidson_fp32_expansion.py:51-52 idiom)."""
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(wl.bk.nk, wl.nbands, wl.bk.npw_max, dtype=torch.complex128, generator=g)
    x = x.to(wl.bk.mask.device) * wl.bk.mask[:, None, :]
    return x.to(dtype)
