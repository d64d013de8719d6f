"""Shared inputs for the GPU parity tests: oracle systems moved to the device."""

from __future__ import annotations

import sys
from types import SimpleNamespace

import torch

from oracle import systems as S
from oracle.hamiltonian import BatchedHamiltonian as RefH
from oracle.hamiltonian import BatchedK

_CACHE: dict = {}


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev() -> torch.device:
    return torch.device("cuda:0")


def _si8_box120():
    """Si8 cell, 30 Ry, two k-points, in a 120^3 box (exercises the 120-point lines)."""
    import numpy as np

    cell, pos = S.si_conventional()
    pos = S.jitter(pos, 0.05, 3)
    k = np.array([[0.125, -0.25, 0.375], [0.0, 0.0, 0.0]])
    return S.build_system("si8_box120", cell, pos, [S.SI] * 8, 30.0, k, np.full(2, 0.5), 8,
                          shape=(120, 120, 120))


def system(name: str):
    if name not in _CACHE:
        _CACHE[name] = {"si2": S.si2_c1, "si8": S.si8_c4, "si64": S.si64_c2,
                        "si8_120": _si8_box120}[name]()
    return _CACHE[name]


def select_k(s, sel=None):
    """Oracle BatchedK restricted to k rows `sel` (CPU) and its device twin."""
    bk = s.bk
    if sel is None:
        sel = list(range(bk.nk))
    cpu = BatchedK(npw=bk.npw[sel], mask=bk.mask[sel], flat_idx=bk.flat_idx[sel], kpg=bk.kpg[sel],
                   t=bk.t[sel], proj_phase_free=bk.proj_phase_free[sel],
                   proj_atom_index=bk.proj_atom_index, dij_full=bk.dij_full)
    d = dev()
    gpu = SimpleNamespace(npw=cpu.npw.to(d), mask=cpu.mask.to(d), flat_idx=cpu.flat_idx.to(d),
                          t=cpu.t.to(d), dij_full=cpu.dij_full.to(d))
    return cpu, gpu, s.p[sel]


def ref_h(s, sel=None, v_r=None):
    cpu, _, p = select_k(s, sel)
    return RefH(cpu, s.shape, s.v_r if v_r is None else v_r, p)


def rand_block(nk, nb, m, mask, seed=0, dtype=torch.complex128):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(nk, nb, m, dtype=torch.complex128, generator=g) * mask[:, None, :]
    return x.to(dtype)
