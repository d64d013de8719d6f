"""Pins the oracle functions the closed-form tests do not reach: the two-stage
solve (solvers/_ms.py:26-52, solvers/davidson.py:777-812), the registry
adapter (solvers/registry.py:154-182), ``apply_cols`` (core/batch.py:507-547)
and the band-energy module.

The two-stage checks follow the reference's own tests
(tests/unit/test_review_dedup.py:129-196): the skip predicate is bit-exact
against the single full-precision solve, and the two-stage result reaches the
dense eigenvalues. A recording operator then fixes the schedule itself: the
draft applies run in complex64, the polish starts from the draft's
eigenvectors cast back up, and nothing runs in complex64 after the polish
starts.
"""

import pytest
import torch

from oracle import systems as S
from oracle.davidson import (
    LOW,
    BatchedDavidsonResult,
    _orthonormalize_b,
    EigResult,
    davidson_adapter,
    davidson_batched,
    davidson_batched_ms,
    mixed_precision_solve,
)
from oracle.dense import dense_h
from oracle.energies import band_energy_eigs, band_energy_psi
from oracle.hamiltonian import BatchedHamiltonian


def _hermitian(m, seed=7):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(m, m, generator=g, dtype=torch.float64) + 1j * torch.randn(m, m, generator=g,
                                                                              dtype=torch.float64)
    return a + a.conj().T + m * torch.eye(m, dtype=torch.complex128)


def _apply_of(a, log=None):
    at = a.transpose(-1, -2)

    def apply(v):
        if log is not None:
            log.append(v.clone())
        return v @ at.to(v.dtype)

    return apply


def _start(nb, m, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.view_as_complex(torch.randn(1, nb, m, 2, dtype=torch.float64, generator=g))


def test_ms_skip_predicate_bitexact():
    m, nb = 14, 3
    h = _apply_of(_hermitian(m))
    x0 = _start(nb, m, 3)
    t = torch.ones(1, m, dtype=torch.float64)
    mask = torch.ones(1, m, dtype=torch.bool)
    base = davidson_batched(h, x0, t, mask, tol=1e-10)
    off = davidson_batched_ms(h, x0, t, mask, tol=1e-10, mixed_precision=False)
    assert torch.equal(base.eigenvalues, off.eigenvalues) and off.n_iter == base.n_iter
    skip = davidson_batched_ms(h, x0, t, mask, tol=1e-10, crossover=1e-3)
    assert torch.equal(base.eigenvalues, skip.eigenvalues)
    # a complex64 start is already low precision: one solve in complex64
    x32 = x0.to(LOW)
    low = davidson_batched(h, x32, t, mask, tol=1e-4)
    low_ms = davidson_batched_ms(h, x32, t, mask, tol=1e-4, crossover=1e-6)
    assert low_ms.eigenvectors.dtype == LOW and torch.equal(low.eigenvalues, low_ms.eigenvalues)


def test_ms_full_precision_result_and_schedule():
    m, nb = 80, 4
    a = _hermitian(m, seed=11)
    ref = torch.linalg.eigvalsh(a)[:nb]
    log = []
    x0 = _start(nb, m, 5)
    t = torch.ones(1, m, dtype=torch.float64)
    mask = torch.ones(1, m, dtype=torch.bool)
    # crossover >= tol skips the draft (the reference's own test runs that case)
    d = davidson_batched_ms(_apply_of(a, log), x0, t, mask, tol=1e-9, crossover=1e-4, max_iter=300)
    assert torch.allclose(d.eigenvalues[0], ref, atol=1e-7)
    assert all(v.dtype == torch.complex128 for v in log)
    # crossover < tol runs the draft: complex64 to `crossover`, then the polish
    log.clear()
    d = davidson_batched_ms(_apply_of(a, log), x0, t, mask, tol=1e-9, crossover=1e-10, max_iter=300)
    assert torch.allclose(d.eigenvalues[0], ref, atol=1e-9)
    assert float(d.residual_norms.max()) < 1e-9 and d.eigenvectors.dtype == torch.complex128
    dts = [v.dtype for v in log]
    first_hi = dts.index(torch.complex128)
    assert first_hi > 0 and all(dt == LOW for dt in dts[:first_hi])
    assert all(dt == torch.complex128 for dt in dts[first_hi:])
    # the polish starts from the draft's eigenvectors: its first apply is their
    # cast, orthonormalised as every solve's start block is (davidson.py:624)
    draft = davidson_batched(_apply_of(a), x0.to(LOW), t.to(torch.float32), mask, tol=1e-10, max_iter=300)
    assert torch.equal(log[first_hi], _orthonormalize_b(draft.eigenvectors.to(torch.complex128), mask))
    # and the polish needs fewer iterations than a cold full-precision solve
    cold = davidson_batched(_apply_of(a), x0, t, mask, tol=1e-9, max_iter=300)
    assert d.n_iter < cold.n_iter


def test_mixed_precision_solve_stage_contract():
    calls = []
    hi = torch.randn(1, 2, 5, dtype=torch.complex128)

    def res(x):
        return BatchedDavidsonResult(torch.zeros(1, 2), x * 2, 1, torch.zeros(1, 2))

    def full():
        calls.append("full")
        return res(hi)

    def make_stages():
        calls.append("make")

        def draft(x):
            calls.append(("draft", x.dtype))
            return res(x)

        def polish(x):
            calls.append(("polish", x.dtype))
            return res(x)

        return draft, polish

    r = mixed_precision_solve(hi, 1e-4, 1e-6, True, full=full, make_stages=make_stages)
    assert calls == ["make", ("draft", LOW), ("polish", torch.complex128)]
    assert torch.equal(r.eigenvectors, (hi.to(LOW) * 2).to(torch.complex128) * 2)
    for args in ((hi, 1e-4, 1e-6, False), (hi.to(LOW), 1e-4, 1e-6, True), (hi, 1e-9, 1e-9, True),
                 (hi, 1e-9, 1e-5, True)):
        calls.clear()
        mixed_precision_solve(*args, full=full, make_stages=make_stages)
        assert calls == ["full"]


def test_davidson_adapter_equals_solver():
    m, nb = 18, 4
    a = _hermitian(m, seed=17)
    h = _apply_of(a)
    x0 = _start(nb, m, 9)
    t = torch.ones(1, m, dtype=torch.float64)
    mask = torch.ones(1, m, dtype=torch.bool)
    base = davidson_batched(h, x0, t, mask, tol=1e-10, max_iter=100)
    r = davidson_adapter(h, x0, t, mask, tol=1e-10, nbands=nb, max_iter=100)
    assert isinstance(r, EigResult)
    assert torch.equal(r.eigenvalues, base.eigenvalues) and torch.equal(r.eigenvectors, base.eigenvectors)
    assert r.n_iter == base.n_iter and torch.equal(r.residual_norms, base.residual_norms)
    assert r.diagnostics["solver"] == "davidson" and r.diagnostics["hit_max_iter"] is False
    short = davidson_adapter(h, x0, t, mask, tol=1e-14, nbands=nb, max_iter=2)
    assert short.n_iter == 2 and short.diagnostics["hit_max_iter"] is True
    assert torch.allclose(r.eigenvalues[0], torch.linalg.eigvalsh(a)[:nb], atol=1e-9)


@pytest.fixture(scope="module")
def si2():
    return S.si2_c1()


def test_apply_cols_matches_apply_rows(si2):
    s = si2
    h = BatchedHamiltonian(s.bk, s.shape, s.v_r, s.p)
    nk, m = s.bk.mask.shape
    g = torch.Generator().manual_seed(2)
    c = torch.randn(nk, 3, m, dtype=torch.complex128, generator=g) * s.bk.mask[:, None, :]
    full = h.apply(c)
    kcol = torch.tensor([3, 0, 3, 7, 9, 0, 5])
    bcol = torch.tensor([1, 2, 0, 2, 1, 0, 2])
    cols = h.apply_cols(c[kcol, bcol], kcol)
    assert float((cols - full[kcol, bcol]).abs().max()) < 1e-12 * float(full.abs().max())
    # the k index matters: the same columns applied at other k differ
    wrong = h.apply_cols(c[kcol, bcol], torch.roll(kcol, 1))
    assert float((wrong - full[kcol, bcol]).abs().max()) > 1e-3 * float(full.abs().max())
    # and the dense matrix agrees column by column
    for j in (0, 3):
        k = int(kcol[j])
        n = int(s.bk.npw[k])
        hd = dense_h(s, k)
        want = hd @ c[k, int(bcol[j]), :n]
        assert float((cols[j, :n] - want).abs().max()) < 1e-10 * float(want.abs().max())
        assert float(cols[j, n:].abs().max()) == 0.0
    cf = h.apply_cols(c[kcol, bcol].to(LOW), kcol)
    assert cf.dtype == LOW
    assert float((cf.to(torch.complex128) - cols).abs().max()) < 2e-5 * float(cols.abs().max())


def test_band_energy_psi_equals_eigenvalue_sum(si2):
    s = si2
    sel = [0, 4]
    h = BatchedHamiltonian(s.bk, s.shape, s.v_r, s.p)
    nk, m = s.bk.mask.shape
    g = torch.Generator().manual_seed(4)
    x0 = torch.randn(nk, 6, m, dtype=torch.complex128, generator=g) * s.bk.mask[:, None, :]
    r = davidson_batched(h.apply, x0, s.bk.t, s.bk.mask, tol=1e-9, max_iter=100)
    occ = torch.zeros(2, 6, dtype=torch.float64)
    occ[:, :4] = 2.0
    kw = torch.as_tensor(s.kweights[sel])
    e_eig = band_energy_eigs(r.eigenvalues[sel], occ, kw)
    e_psi = band_energy_psi(s, r.eigenvectors[sel], kw[:, None] * occ, s.positions, kpoints=sel)
    assert abs(float(e_psi) - float(e_eig)) < 1e-9 * abs(float(e_eig))
    # a dropped occupation or k weight changes the sum
    assert abs(float(band_energy_eigs(r.eigenvalues[sel], occ, torch.ones(2))) - float(e_eig)) > 1e-3
