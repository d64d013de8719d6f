"""gradwave_amd -- MI355X (gfx950) H|psi> apply and batched block Davidson.

Drop-in replacements for gradwave's ``core.batch.BatchedHamiltonian`` and
``solvers.davidson.davidson_batched`` (plus the eigensolver registry), with
all arithmetic in the HIP library ``libgradwave_hip.so`` built from ``csrc/``.
"""

from gradwave_amd._lib import GW_FP32, GW_FP64, GWError, lib  # noqa: F401

__all__ = ["GWError", "lib", "GW_FP32", "GW_FP64"]
