"""Write the oracle's converged FP64 Davidson solution of the Si64 Gamma workload.

The oracle (a torch-CPU complex128 restatement of solvers/davidson.py:527-774
and core/batch.py:480-505) solves the 160-band Si64 problem (2x2x2
conventional cell, U(+-0.02 A) jitter with seed 0, 40 Ry, 90^3 box) from the
seeded start block ``torch.randn(1, 160, npw_max, complex128)`` of a
``torch.Generator().manual_seed(0)`` (masked), to a residual-norm tolerance of
1e-10 Ha. The eigenvalues, residual norms and iteration count are stored as
data so the GPU test can compare its FP64 solve from the same start block
without running the oracle (several minutes on the CPU). Usage:

    python tests/golden/make_si64_fp64_eigs.py
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "si64_fp64_eigs.npz")
SEED, NB = 0, 160


def start_block(mask: torch.Tensor, nb: int, seed: int) -> torch.Tensor:
    g = torch.Generator().manual_seed(seed)
    nk, m = mask.shape
    return torch.randn(nk, nb, m, dtype=torch.complex128, generator=g) * mask[:, None, :]


def main() -> None:
    sys.path.insert(0, REPO)
    from oracle import systems as S
    from oracle.davidson import davidson_batched
    from oracle.hamiltonian import BatchedHamiltonian
    from oracle.units import HARTREE_EV

    torch.set_num_threads(min(8, os.cpu_count() or 1))
    s = S.si64_c2(nbands=NB, seed=0)
    h = BatchedHamiltonian(s.bk, s.shape, s.v_r, s.p)
    x0 = start_block(s.bk.mask, NB, SEED)
    tol = 1e-10 * HARTREE_EV
    t0 = time.time()
    hist = []
    r = davidson_batched(h.apply, x0, s.bk.t, s.bk.mask, tol=tol, max_iter=400, history_out=hist)
    dt = time.time() - t0
    print(f"oracle si64 fp64: {r.n_iter} iterations, {dt:.0f} s, max residual "
          f"{float(r.residual_norms.max()):.2e} eV", flush=True)
    assert float(r.residual_norms.max()) < tol, "oracle did not converge"
    np.savez_compressed(OUT, eigenvalues_eV=r.eigenvalues.numpy(), residual_norms_eV=r.residual_norms.numpy(),
                        n_iter=np.int64(r.n_iter), tol_eV=np.float64(tol), seed=np.int64(SEED),
                        nbands=np.int64(NB), npw=s.bk.npw.numpy(), seconds=np.float64(dt))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
