"""Analytic Gaussian pseudopotential for the synthetic Si/Al workloads.

The reference reads UPF files only; it consumes exactly two tables that this
analytic pseudo produces (core/hamiltonian.py:46-53, core/batch.py:249-258):

* local form factor, the reference's Coulomb-tail term with the short-range part
  set to zero (pseudo/local.py:63):
      v(g) = -4 pi Z e^2 exp(-(g rc)^2 / 4) / g^2       [eV A^3], g > 0
  and the reference's G=0 convention for V_sr = 0 (pseudo/local.py:72):
      v(0) = alpha = pi Z e^2 rc^2
  V_loc(G) = (1/Omega) sum_a S_a(G) v(|G|) with S_a = exp(-i G.tau_a)
  (core/energies/local_pp.py:22-38); V(r) = Re N ifftn V_loc(G) on the full box
  (scf/loop.py:532-544 through core/fftbox.py:46-60).
* KB form factors of a Gaussian radial projector beta_l(r) = N_l r^l e^{-r^2/2w^2},
  normalised to int r^2 beta_l^2 dr = 1:
      F_l(q) = int r^2 beta_l(r) j_l(q r) dr
             = sqrt(pi / Gamma(l + 3/2)) w^(l+3/2) q^l exp(-q^2 w^2 / 2)   [A^(3/2)]
  with one s (D_s) and one p channel (D_p), D block-diagonal per atom
  (core/hamiltonian.py:78-93 m-expansion).

Pinned by: the closed form of F_l checked against numerical quadrature of the
radial integral (scipy spherical_jn), the local form factor checked against the
Fourier transform of a Gaussian-smeared point charge, and V(r) real-valuedness.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from oracle.units import BOHR_ANG, E2, HARTREE_EV


@dataclass(frozen=True)
class AnalyticSpecies:
    name: str
    z: float  # valence charge
    rc_bohr: float  # local Gaussian smearing radius
    w_bohr: float = 1.0  # KB Gaussian width
    d_s_ha: float = 1.0
    d_p_ha: float = -0.5

    @property
    def rc(self) -> float:
        return self.rc_bohr * BOHR_ANG

    @property
    def w(self) -> float:
        return self.w_bohr * BOHR_ANG

    @property
    def dij(self) -> torch.Tensor:
        """(2,2) [eV] over channels (l=0, l=1)."""
        return torch.tensor([[self.d_s_ha * HARTREE_EV, 0.0],
                             [0.0, self.d_p_ha * HARTREE_EV]], dtype=torch.float64)

    @property
    def beta_ls(self) -> list[int]:
        return [0, 1]


SI = AnalyticSpecies("Si", z=4.0, rc_bohr=0.8)
AL = AnalyticSpecies("Al", z=3.0, rc_bohr=0.6)


def vloc_form_factor(sp: AnalyticSpecies, g: np.ndarray) -> np.ndarray:
    """v(|G|) [eV A^3] on an array of |G| (A^-1); G=0 entries get alpha."""
    g = np.asarray(g, dtype=np.float64)
    out = np.empty_like(g)
    nz = g > 1e-10
    gg = g[nz]
    out[nz] = -4.0 * np.pi * sp.z * E2 * np.exp(-0.25 * (gg * sp.rc) ** 2) / gg**2
    out[~nz] = np.pi * sp.z * E2 * sp.rc**2
    return out


def kb_form_factor(sp: AnalyticSpecies, l: int, q: np.ndarray) -> np.ndarray:
    """F_l(q) [A^1.5] of the normalised Gaussian radial projector."""
    q = np.asarray(q, dtype=np.float64)
    w = sp.w
    pref = math.sqrt(math.pi / math.gamma(l + 1.5)) * w ** (l + 1.5)
    return pref * q**l * np.exp(-0.5 * (q * w) ** 2)


def local_potential_g(positions: torch.Tensor, species_of_atom: list[AnalyticSpecies],
                      g_cart: torch.Tensor, volume: float) -> torch.Tensor:
    """V_loc(G) on the full box [eV] -- core/energies/local_pp.py:22-38 with
    per-atom tables; differentiable in positions (A)."""
    g_abs = np.sqrt((g_cart.detach().numpy() ** 2).sum(-1))
    tabs = {}
    for sp in species_of_atom:
        if sp.name not in tabs:
            tabs[sp.name] = torch.as_tensor(vloc_form_factor(sp, g_abs), dtype=torch.float64)
    # per-species sum of structure factors, one atom at a time (memory stays
    # O(N) for the 216-atom cell instead of the (na, n1, n2, n3) stack)
    v = torch.zeros(g_cart.shape[:-1], dtype=torch.complex128)
    for name, tab in tabs.items():
        s_sum = torch.zeros(g_cart.shape[:-1], dtype=torch.complex128)
        for a, sp in enumerate(species_of_atom):
            if sp.name != name:
                continue
            phase = (g_cart * positions[a]).sum(-1)
            s_sum = s_sum + torch.exp(torch.complex(torch.zeros_like(phase), -phase))
        v = v + s_sum * tab.to(torch.complex128)
    return v / volume


def local_potential_r(vloc_g: torch.Tensor) -> torch.Tensor:
    """V(r) = Re N ifftn V_loc(G) -- scf/loop.py:532-544, fftbox.py:46-60."""
    n = vloc_g.shape[-3] * vloc_g.shape[-2] * vloc_g.shape[-1]
    return (torch.fft.ifftn(vloc_g, dim=(-3, -2, -1)) * n).real
