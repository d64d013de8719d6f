"""FFT box and wavefunction G-sphere -- restates grids.py:27-184.

The sphere ordering and the flat box index are integer maps that the GPU
pack/unpack must reproduce bit for bit, so the arithmetic here follows the
reference step for step (same numpy matmul/einsum, same 1e-12 relative cutoff
slack, same lexsort keys on round(|k+G|^2, 10)).

Pinned by: golden vectors written from the reference's own ``build_gsphere``
(tests/golden/gsphere_*.npz), brute-force enumeration of |k+G|^2 <= ecut, and the
reference's good_fft_size cases (test_grids.py:13-17).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from oracle.units import HBAR2_2M


def reciprocal_cell(cell: np.ndarray) -> np.ndarray:
    """grids.py:27-29 -- rows b_i with a_i.b_j = 2 pi delta_ij."""
    return 2.0 * np.pi * np.linalg.inv(cell).T


def good_fft_size(n: int) -> int:
    """grids.py:32-41 -- smallest size >= n with prime factors <= 7."""
    while True:
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def gmax_from_ecut(ecut: float) -> float:
    """grids.py:44-45."""
    return float(np.sqrt(ecut / HBAR2_2M))


@dataclass(frozen=True)
class FFTGrid:
    """grids.py:48-65."""

    cell: np.ndarray
    shape: tuple[int, int, int]
    g_cart: torch.Tensor  # (n1,n2,n3,3) A^-1
    g2: torch.Tensor  # (n1,n2,n3)

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.cell)))

    @property
    def n_points(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]


def build_fft_grid(cell, ecut: float, shape_override=None) -> FFTGrid:
    """grids.py:68-126 (equal_dims omitted: no caller on this path uses it)."""
    cell = np.asarray(cell, dtype=np.float64)
    b = reciprocal_cell(cell)
    gmax_dens = 2.0 * gmax_from_ecut(ecut)
    if shape_override is not None:
        shape = tuple(int(n) for n in shape_override)
    else:
        dims = []
        for i in range(3):
            m_i = int(np.floor(gmax_dens * np.linalg.norm(cell[i]) / (2.0 * np.pi)))
            dims.append(good_fft_size(2 * m_i + 1))
        shape = tuple(dims)
    millers = np.meshgrid(
        *[np.fft.fftfreq(n, d=1.0 / n).astype(np.int64) for n in shape], indexing="ij"
    )
    m = np.stack(millers, axis=-1)
    g = m @ b
    g2 = np.einsum("...i,...i->...", g, g)
    return FFTGrid(cell=cell, shape=shape,
                   g_cart=torch.as_tensor(g, dtype=torch.float64),
                   g2=torch.as_tensor(g2, dtype=torch.float64))


@dataclass(frozen=True)
class GSphere:
    """grids.py:129-142."""

    k_frac: np.ndarray
    miller: torch.Tensor  # (npw,3) int64
    kpg: torch.Tensor  # (npw,3) A^-1
    kpg2: torch.Tensor  # (npw,)
    flat_idx: torch.Tensor  # (npw,) int64

    @property
    def npw(self) -> int:
        return int(self.miller.shape[0])


def build_gsphere(grid: FFTGrid, ecut: float, k_frac) -> GSphere:
    """grids.py:145-184 -- all G with HBAR2_2M|k+G|^2 <= ecut, reference order."""
    cell = grid.cell
    b = reciprocal_cell(cell)
    k_frac = np.asarray(k_frac, dtype=np.float64)
    n1, n2, n3 = grid.shape
    gmax = gmax_from_ecut(ecut)
    bound = [int(np.floor(gmax * np.linalg.norm(cell[i]) / (2 * np.pi) + 1)) + 1 for i in range(3)]
    axes = [np.arange(-bound[i], bound[i] + 1) for i in range(3)]
    mm = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    kpg = (mm + k_frac) @ b
    kpg2 = np.einsum("ij,ij->i", kpg, kpg)
    keep = HBAR2_2M * kpg2 <= ecut * (1 + 1e-12)
    mm, kpg, kpg2 = mm[keep], kpg[keep], kpg2[keep]
    for i, n in enumerate((n1, n2, n3)):
        if np.any(np.abs(mm[:, i]) > n // 2):
            raise RuntimeError("wavefunction sphere exceeds FFT box -- grid sizing bug")
    flat = (mm[:, 0] % n1) * (n2 * n3) + (mm[:, 1] % n2) * n3 + (mm[:, 2] % n3)
    order = np.lexsort((mm[:, 2], mm[:, 1], mm[:, 0], np.round(kpg2, 10)))
    mm, kpg, kpg2, flat = mm[order], kpg[order], kpg2[order], flat[order]
    return GSphere(k_frac=k_frac,
                   miller=torch.as_tensor(mm, dtype=torch.int64),
                   kpg=torch.as_tensor(kpg, dtype=torch.float64),
                   kpg2=torch.as_tensor(kpg2, dtype=torch.float64),
                   flat_idx=torch.as_tensor(flat, dtype=torch.int64))
