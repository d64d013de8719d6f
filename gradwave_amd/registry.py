"""Eigensolver registry with the reference's plugin API (solvers/registry.py:79-222).

``solve(apply_H, X0, precond, mask, *, tol, nbands, **kw) -> EigResult``;
``register(name, solver, overwrite=False)`` raises ValueError on a duplicate
name and ``get`` raises ValueError listing what is registered, as the
reference does. The MI355X Davidson is registered as "davidson"; a caller of
the reference registry swaps it in with ``register("davidson",
davidson_adapter, overwrite=True)`` (the trick benchmarks/bench_fp32_expansion.py:41-52
uses).
"""

from __future__ import annotations

from collections.abc import Callable
from typing import Any

import torch

from gradwave_amd.davidson import EigResult, davidson_batched, davidson_batched_ms

_REGISTRY: dict[str, Callable[..., EigResult]] = {}


def register(name: str, solver: Callable[..., EigResult], *, overwrite: bool = False) -> None:
    """registry.py:115-123."""
    if name in _REGISTRY and not overwrite:
        raise ValueError(f"eigensolver {name!r} already registered; pass overwrite=True to replace")
    _REGISTRY[name] = solver


def get(name: str) -> Callable[..., EigResult]:
    """registry.py:126-133."""
    try:
        return _REGISTRY[name]
    except KeyError:
        raise ValueError(f"unknown eigensolver {name!r}; registered: {sorted(_REGISTRY)}") from None


def available() -> list[str]:
    return sorted(_REGISTRY)


def is_registered(name: str) -> bool:
    return name in _REGISTRY


def davidson_adapter(apply_H: Callable[[torch.Tensor], torch.Tensor], X0: torch.Tensor,
                     precond: torch.Tensor, mask: torch.Tensor, *, tol: float,
                     nbands: int | None = None, max_iter: int = 40, max_dim_factor: int = 4,
                     **kw: Any) -> EigResult:
    """registry.py:154-182 on the MI355X davidson_batched."""
    r = davidson_batched(apply_H, X0, precond, mask, tol=tol, max_iter=max_iter,
                         max_dim_factor=max_dim_factor)
    return EigResult(r.eigenvalues, r.eigenvectors, r.n_iter, r.residual_norms,
                     {"solver": "davidson", "max_dim_factor": max_dim_factor,
                      "max_iter": max_iter, "hit_max_iter": r.n_iter >= max_iter,
                      "n_apply_low": r.n_apply_low, "n_apply_full": r.n_apply_full,
                      "backend": "gradwave_amd"})


def davidson_ms_adapter(apply_H, X0, precond, mask, *, tol: float, nbands: int | None = None,
                        max_iter: int = 40, max_dim_factor: int = 4, crossover: float = 1e-5,
                        **kw: Any) -> EigResult:
    """Mixed-precision two-stage solve (complex64 draft, complex128 polish)."""
    r = davidson_batched_ms(apply_H, X0, precond, mask, tol=tol, max_iter=max_iter,
                            max_dim_factor=max_dim_factor, crossover=crossover)
    return EigResult(r.eigenvalues, r.eigenvectors, r.n_iter, r.residual_norms,
                     {"solver": "davidson_ms", "crossover": crossover, "max_iter": max_iter,
                      "n_apply_low": r.n_apply_low, "n_apply_full": r.n_apply_full,
                      "backend": "gradwave_amd"})


register("davidson", davidson_adapter)
register("davidson_ms", davidson_ms_adapter)
