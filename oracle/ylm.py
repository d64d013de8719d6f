"""Real spherical harmonics, QE m-order -- restates core/ylm.py:26-118 (l <= 3).

Order within l: m = 0, +1, -1, +2, -2, +3, -3 (ylm.py:6). Zero vectors give
Y_00 = 1/(2 sqrt(pi)) and 0 for l >= 1 (ylm.py:14-17).

Pinned by: orthonormality on a Lebedev-free quadrature (Gauss-Legendre x
uniform phi), the l=1 closed form sqrt(3/4pi)(z,x,y), and golden projector
tables written from the reference's build_projector_data.
"""

from __future__ import annotations

import math

import torch

C00 = 0.5 * math.sqrt(1.0 / math.pi)
C1 = math.sqrt(3.0 / (4.0 * math.pi))
C20 = 0.25 * math.sqrt(5.0 / math.pi)
C21 = 0.5 * math.sqrt(15.0 / math.pi)
C22 = 0.25 * math.sqrt(15.0 / math.pi)
C30 = 0.25 * math.sqrt(7.0 / math.pi)
C31 = 0.25 * math.sqrt(21.0 / (2.0 * math.pi))
C32 = 0.25 * math.sqrt(105.0 / math.pi)
C32M = 0.5 * math.sqrt(105.0 / math.pi)
C33 = 0.25 * math.sqrt(35.0 / (2.0 * math.pi))


def ylm_all(lmax: int, g: torch.Tensor, eps: float = 1e-14) -> torch.Tensor:
    """ylm.py:45-118 (non-solid branch) for lmax <= 3; g (...,3) -> (...,(lmax+1)^2)."""
    if lmax > 3:
        raise ValueError("oracle ylm_all supports lmax <= 3")
    n2 = (g * g).sum(dim=-1, keepdim=True)
    zero = (n2 < eps * eps).squeeze(-1)
    g_safe = torch.where(zero.unsqueeze(-1), torch.ones_like(g), g)
    norm = torch.linalg.norm(g_safe, dim=-1, keepdim=True)
    unit = g / norm
    x, y, z = unit[..., 0], unit[..., 1], unit[..., 2]
    x = torch.where(zero, torch.zeros_like(x), x)
    y = torch.where(zero, torch.zeros_like(y), y)
    z = torch.where(zero, torch.zeros_like(z), z)
    r2 = x * x + y * y + z * z
    out = [torch.full_like(x, C00)]
    if lmax >= 1:
        out += [C1 * z, C1 * x, C1 * y]
    if lmax >= 2:
        out += [C20 * (3.0 * z * z - r2), C21 * x * z, C21 * y * z,
                C22 * (x * x - y * y), C21 * x * y]
    if lmax >= 3:
        out += [C30 * z * (5.0 * z * z - 3.0 * r2), C31 * x * (5.0 * z * z - r2),
                C31 * y * (5.0 * z * z - r2), C32 * z * (x * x - y * y),
                C32M * x * y * z, C33 * x * (x * x - 3.0 * y * y),
                C33 * y * (3.0 * x * x - y * y)]
    return torch.stack(out, dim=-1)
