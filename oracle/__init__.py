"""CPU oracle for gradwave's plane-wave H|psi> apply and batched block Davidson.

TEST INFRASTRUCTURE ONLY. This package is a torch-CPU (complex128) restatement,
op for op, of the reference algorithm on the hot path. It is imported only by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py`` -- always as the checker, never as the thing measured or shipped.
The product package ``gradwave_amd`` never imports it.

Modules and the reference files they restate (paths under the reference's
``src/gradwave/``):

* ``units``        -- constants.py:10-20
* ``grids``        -- grids.py:27-184 (FFT box sizing, G-sphere, flat index maps)
* ``ylm``          -- core/ylm.py:45-118 (real Y_lm, QE m-order)
* ``pseudo``       -- analytic Gaussian pseudopotential; the Coulomb-tail form of
                      pseudo/local.py:63 with the short-range part zero, the G=0
                      alpha convention of pseudo/local.py:67-72, and Gaussian KB
                      form factors in closed form
* ``hamiltonian``  -- core/hamiltonian.py:35-191, core/batch.py:126-235,238-505,590-595
* ``davidson``     -- solvers/davidson.py:66-308,406-486,511-812, solvers/precond.py:8-39,
                      solvers/_ms.py:26-52, solvers/registry.py:86-182
* ``systems``      -- the synthetic Si/Al workloads (cells, k-lists, jitter)
* ``dense``        -- brute-force dense H matrix (closed-form check of the apply)
* ``generalized``  -- scf/uspp_batch.py:27-233 (S-metric generalized Davidson)
* ``energies``     -- band energy from eigenvalues (postscf/mae.py:349) and at
                      fixed orbitals, with frozen-potential forces by autograd
                      in the positions (postscf/forces.py:168-213 pattern)
* ``gamma``        -- core/gamma.py:70-290 (Gamma half sphere, real embedding)

Parity status of each function is stated in its module docstring. Each public
function is run by at least one test under ``tests/`` that a plausible mistake
would fail (closed forms, brute force, invariants, finite differences, or the
golden vectors written from the reference's leaf modules under
``tests/golden/``): ``test_oracle_closed_forms.py``, ``test_oracle_golden.py``,
``test_oracle_adjoint.py``, ``test_oracle_gamma.py``,
``test_oracle_generalized.py`` and ``test_oracle_solvers.py`` (two-stage solve,
registry adapter, apply_cols, band energy).
"""
