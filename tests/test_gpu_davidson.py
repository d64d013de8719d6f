"""GPU Davidson parity against the oracle: eigenvalues of the plane-wave H and
of the reference's synthetic Hermitian operator, fp64 to the 1e-10 Ha tolerance,
plus the mixed-precision two-stage solve within 1e-6 Ha."""

import numpy as np
import pytest
import torch

from gpu_cases import dev, rand_block, ref_h, say, select_k, system

pytestmark = pytest.mark.gpu

from gradwave_amd import linalg as LA  # noqa: E402
from gradwave_amd.davidson import davidson_batched, davidson_batched_ms, davidson_mixed  # noqa: E402
from gradwave_amd.hamiltonian import BatchedHamiltonian  # noqa: E402
from oracle.davidson import davidson_batched as ref_davidson  # noqa: E402
from oracle.dense import dense_h  # noqa: E402
from oracle.units import HARTREE_EV  # noqa: E402

TOL = 1e-10 * HARTREE_EV  # 1e-10 Ha residual-norm tolerance, in eV


def test_si2_all_k_fp64_matches_oracle_and_dense():
    say("davidson si2 10k fp64")
    s = system("si2")
    cpu, gpu, p = select_k(s)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    x0 = rand_block(cpu.nk, 8, cpu.npw_max, cpu.mask, seed=0)
    r = davidson_batched(h.apply, x0.to(dev()), gpu.t, gpu.mask, tol=TOL, max_iter=100)
    assert float(r.residual_norms.max()) < TOL
    rr = ref_davidson(ref_h(s).apply, x0, cpu.t, cpu.mask, tol=TOL, max_iter=100)
    assert float((r.eigenvalues.cpu() - rr.eigenvalues).abs().max()) < 1e-10 * HARTREE_EV
    for ik in (0, 8):
        e = torch.linalg.eigvalsh(dense_h(s, ik))[:8]
        assert float((r.eigenvalues[ik].cpu() - e).abs().max()) < 1e-10 * HARTREE_EV
    # returned vectors: orthonormal rows, zero padding, residuals re-measured
    x = r.eigenvectors
    gram = LA.inner(x, x).cpu()
    assert float((gram - torch.eye(8, dtype=gram.dtype)).abs().max()) < 1e-10
    assert float(x.cpu()[~cpu.mask[:, None, :].expand_as(x)].abs().max()) == 0.0
    hx = h.apply(x)
    rn = torch.linalg.norm((hx - r.eigenvalues[..., None] * x).cpu(), dim=-1)
    assert float((rn - r.residual_norms.cpu()).abs().max()) < 1e-12


def test_registry_solvers_on_si2():
    """The SCF's call through the registry (scf/loop.py:953-955:
    get_solver(name)(apply, coeffs, t_solve, mask, tol=, nbands=)) runs the
    MI355X solvers: "davidson" gives davidson_batched's result exactly; the
    two-stage "davidson_ms" skips its draft when crossover >= tol (_ms.py:47) and
    otherwise drafts in complex64 (its applies counted as low precision)."""
    from gradwave_amd import registry as R

    say("registry davidson / davidson_ms si2")
    s = system("si2")
    cpu, gpu, p = select_k(s, [0, 5])
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    x0 = rand_block(2, 8, cpu.npw_max, cpu.mask, seed=4).to(dev())
    want = davidson_batched(h.apply, x0, gpu.t, gpu.mask, tol=TOL, max_iter=100)
    got = R.get("davidson")(h.apply, x0, gpu.t, gpu.mask, tol=TOL, nbands=8, max_iter=100)
    assert torch.equal(got.eigenvalues, want.eigenvalues) and got.n_iter == want.n_iter
    assert torch.equal(got.residual_norms, want.residual_norms)
    assert got.diagnostics["solver"] == "davidson" and not got.diagnostics["hit_max_iter"]
    ms = R.get("davidson_ms")(h.apply, x0, gpu.t, gpu.mask, tol=TOL, nbands=8, max_iter=100)
    assert ms.diagnostics["n_apply_low"] == 0 and torch.equal(ms.eigenvalues, want.eigenvalues)
    loose = 1e-3  # eV
    ms = R.get("davidson_ms")(h.apply, x0, gpu.t, gpu.mask, tol=loose, nbands=8, max_iter=100, crossover=1e-4)
    assert ms.diagnostics["n_apply_low"] > 0 and float(ms.residual_norms.max()) < loose
    assert float((ms.eigenvalues - want.eigenvalues).abs().max()) < 1e-6 * HARTREE_EV


def test_orthonormalize_degenerate_rows_and_breakdown():
    """davidson.py:406-457 on the GPU. A zero row gets the reference's seeded
    jitter, and the result then spans exactly what the oracle's spans. It is
    orthonormal, zero in the padding and orthogonal to the basis it was
    projected against (test_davidson_qr_offload.py:77-87 checks the same for
    QR). A repeated row and a pair equal to 1e-9 break CholQR2, so the
    Householder fallback runs and the result is still orthonormal and masked.
    Orthogonality to the basis is not required there: the reference's
    row-norm test does not catch such rows, and its QR completes the rank
    freely."""
    from gradwave_amd.davidson import _orthonormalize_b
    from oracle.davidson import _orthonormalize_b as ref_orth

    say("orthonormalize degenerate rows / breakdown si2")
    s = system("si2")
    cpu, gpu, p = select_k(s, [0, 3])
    nk, m = cpu.mask.shape
    pad = ~cpu.mask[:, None, :]
    basis = ref_orth(rand_block(nk, 4, m, cpu.mask, seed=21), cpu.mask)
    probe = rand_block(nk, 5, m, cpu.mask, seed=24)

    def check(q, against=True):
        gram = LA.inner(q, q).cpu()
        assert float((gram - torch.eye(q.shape[1], dtype=gram.dtype)).abs().max()) < 1e-10
        qc = q.cpu()
        assert float(qc[pad.expand_as(qc)].abs().max()) == 0.0
        if against:
            assert float((qc.conj() @ basis.mT).abs().max()) < 1e-10
        return qc

    v = rand_block(nk, 6, m, cpu.mask, seed=22)
    v[:, 2] = 0.0
    q = check(_orthonormalize_b(v.to(dev()), gpu.mask, against=basis.to(dev())))
    qr = ref_orth(v, cpu.mask, against=basis)

    def span(x):  # projections of fixed probe rows onto span(x)
        return (probe @ x.conj().mT) @ x

    assert float((span(q) - span(qr)).abs().max()) < 1e-10
    for bad in ("repeat", "near"):
        w = rand_block(nk, 6, m, cpu.mask, seed=23)
        w[:, 4] = w[:, 1] if bad == "repeat" else w[:, 1] + 1e-9 * rand_block(nk, 1, m, cpu.mask, seed=25)[:, 0]
        check(_orthonormalize_b(w.to(dev()), gpu.mask, against=basis.to(dev())), against=False)


def _synthetic(nk, npw, seed):
    """The reference's synthetic operator (test_davidson_fp32_expansion.py:24-42)."""
    torch.manual_seed(seed)
    diag = torch.linspace(1.0, 100.0, npw, dtype=torch.float64)
    h = torch.zeros(nk, npw, npw, dtype=torch.complex128)
    for k in range(nk):
        a = torch.randn(npw, npw, dtype=torch.complex128)
        h[k] = 0.5 * (a + a.conj().T) + torch.diag(diag.to(torch.complex128))
        h[k] = 0.5 * (h[k] + h[k].conj().T)
    return h, diag.expand(nk, npw).contiguous()


@pytest.mark.parametrize("mode", ["off", "on"])
def test_synthetic_operator_and_fp32_expansion(mode, monkeypatch):
    say(f"davidson synthetic fp32_expansion={mode}")
    import os

    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.npz"))
    h = torch.as_tensor(gold["syn_h"])
    x0 = torch.as_tensor(gold["syn_x0"])
    t = torch.linspace(1.0, 100.0, 80, dtype=torch.float64).expand(4, 80).contiguous()
    hT = h.transpose(1, 2).contiguous().to(dev())

    def apply(c):  # out[b][i] = sum_j h[i][j] c[b][j], recomputed in c's precision
        return LA.combine(c.transpose(1, 2).contiguous(), hT.to(c.dtype))

    monkeypatch.setenv("GRADWAVE_FP32_EXPANSION", mode)
    mask = torch.ones(4, 80, dtype=torch.bool, device=dev())
    r = davidson_batched(apply, x0.to(dev()), t.to(dev()), mask, tol=1e-9, max_iter=100)
    assert float(r.residual_norms.max()) < 1e-9
    ref = torch.linalg.eigvalsh(h)[:, :8]
    assert float((r.eigenvalues.cpu() - ref).abs().max()) < 1e-9
    assert float((r.eigenvalues.cpu() - torch.as_tensor(gold[f"syn_{mode}_eig"])).abs().max()) < 1e-9
    if mode == "on":
        assert r.n_apply_low > 0 and r.n_apply_full > 0


def test_si8_mixed_precision_two_stage():
    """Precision form on a multi-k cell: complex64 draft + complex128 polish vs the fp64 solve."""
    say("davidson si8 mixed vs fp64")
    s = system("si8")
    sel = [0, 7, 20]
    cpu, gpu, p = select_k(s, sel)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    x0 = rand_block(3, 24, cpu.npw_max, cpu.mask, seed=1).to(dev())
    r64 = davidson_batched(h.apply, x0, gpu.t, gpu.mask, tol=TOL, max_iter=200)
    rms = davidson_mixed(h.apply, x0, gpu.t, gpu.mask, tol=TOL, draft_tol=1e-3, max_iter=200)
    assert float(r64.residual_norms.max()) < TOL and float(rms.residual_norms.max()) < TOL
    assert float((rms.eigenvalues - r64.eigenvalues).abs().max()) < 1e-6 * HARTREE_EV
    assert rms.n_apply_low > 0
    # the reference's two-stage wrapper skips its draft when crossover >= tol (_ms.py:47)
    r_skip = davidson_batched_ms(h.apply, x0, gpu.t, gpu.mask, tol=TOL, max_iter=200, crossover=1e-4)
    assert r_skip.n_apply_low == 0 and float(r_skip.residual_norms.max()) < TOL


def test_si8_mixed_refinement_sweeps():
    """Draft + n FP64 refinement sweeps: Ritz values are upper bounds of the
    converged FP64 eigenvalues and do not rise from one sweep to the next (the
    basis only grows); one sweep lands within 1e-6 Ha; the returned residual
    norms are those of the FP64 operator."""
    from gradwave_amd.davidson import LOW, _refine_steps, drive

    say("davidson si8 draft + fp64 refinement sweeps")
    s = system("si8")
    cpu, gpu, p = select_k(s, [0, 7, 20])
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    x0 = rand_block(3, 24, cpu.npw_max, cpu.mask, seed=1).to(dev())
    r64 = davidson_batched(h.apply, x0, gpu.t, gpu.mask, tol=TOL, max_iter=200)
    # the draft as davidson_mixed runs it: one projection per round
    draft = davidson_batched(h.apply, x0.to(LOW), gpu.t.to(torch.float32), gpu.mask, tol=1e-3, max_iter=200,
                             proj_passes=1)
    xd = draft.eigenvectors.to(torch.complex128)
    res = [drive(_refine_steps(xd, gpu.t, gpu.mask, n), h.apply) for n in (0, 1, 2)]
    eps = 1e-12 * HARTREE_EV
    for n, r in enumerate(res):
        assert r.n_iter == n
        assert float((r.eigenvalues - r64.eigenvalues).min()) > -eps
        hx = h.apply(r.eigenvectors)
        rn = torch.linalg.norm(hx - r.eigenvalues[..., None] * r.eigenvectors, dim=-1)
        assert float((rn - r.residual_norms).abs().max()) < 1e-10
        gram = LA.inner(r.eigenvectors, r.eigenvectors).cpu()
        assert float((gram - torch.eye(24, dtype=gram.dtype)).abs().max()) < 1e-10
    assert float((res[1].eigenvalues - res[0].eigenvalues).max()) < eps
    assert float((res[2].eigenvalues - res[1].eigenvalues).max()) < eps
    assert float((res[1].eigenvalues - r64.eigenvalues).abs().max()) < 1e-6 * HARTREE_EV
    rms = davidson_mixed(h.apply, x0, gpu.t, gpu.mask, draft_tol=1e-3, max_iter=200, sweeps=1)
    assert rms.n_iter == 1 and rms.n_draft_iter == draft.n_iter and rms.n_apply_low > 0
    assert float((rms.eigenvalues - r64.eigenvalues).abs().max()) < 1e-6 * HARTREE_EV
    # buffer bands: the requested 24 come back, within the bound, as upper bounds
    rb = davidson_mixed(h.apply, x0, gpu.t, gpu.mask, draft_tol=5e-3, max_iter=200, sweeps=1, buffer=4)
    assert rb.eigenvalues.shape == (3, 24) and rb.eigenvectors.shape == x0.shape and rb.residual_norms.shape == (3, 24)
    assert float((rb.eigenvalues - r64.eigenvalues).abs().max()) < 1e-6 * HARTREE_EV
    assert float((rb.eigenvalues - r64.eigenvalues).min()) > -eps
    hx = h.apply(rb.eigenvectors)
    rn = torch.linalg.norm(hx - rb.eigenvalues[..., None] * rb.eigenvectors, dim=-1)
    assert float((rn - rb.residual_norms).abs().max()) < 1e-10
    with pytest.raises(ValueError):
        davidson_mixed(h.apply, x0, gpu.t, gpu.mask, sweeps=1, buffer=-1)
    # nconv: only the first 16 bands gate convergence; they match the full solve
    rc = davidson_batched(h.apply, x0, gpu.t, gpu.mask, tol=TOL, max_iter=200, nconv=16)
    assert float(rc.residual_norms[:, :16].max()) < TOL and rc.n_iter <= r64.n_iter
    assert float((rc.eigenvalues[:, :16] - r64.eigenvalues[:, :16]).abs().max()) < 1e-10 * HARTREE_EV
    with pytest.raises(ValueError):
        davidson_mixed(h.apply, x0, gpu.t, gpu.mask, draft_tol=1e-3, max_iter=200, sweeps=-1)


def test_sternheimer_cg_matches_dense_solve():
    """cg_sternheimer (contract of postscf/_response.py:502-595) on Si2, two
    k-points, four occupied bands: against a dense solve of
    (H - eps_n + s P_occ) x = rhs with the explicit H, through the operator's
    compacted column apply and through full-block applies; warm starts and
    padded garbage in x0; complex64 operands."""
    from types import SimpleNamespace

    from gradwave_amd.response import cg_sternheimer
    from oracle.dense import dense_h

    say("sternheimer cg si2")
    s = system("si2")
    sel = [0, 3]
    cpu, gpu, p = select_k(s, sel)
    nk, m = cpu.mask.shape
    nocc, shift = 4, 5.0
    c_occ = torch.zeros(nk, nocc, m, dtype=torch.complex128)
    eps = torch.zeros(nk, nocc, dtype=torch.float64)
    rhs = torch.zeros(nk, nocc, m, dtype=torch.complex128)
    want = torch.zeros(nk, nocc, m, dtype=torch.complex128)
    g = torch.Generator().manual_seed(3)
    for j, ik in enumerate(sel):
        hd = dense_h(s, ik)
        n = hd.shape[0]
        w, u = torch.linalg.eigh(hd)
        c = u[:, :nocc].T.contiguous()
        c_occ[j, :, :n] = c
        eps[j] = w[:nocc]
        pm = c.T @ c.conj()  # P_occ as an (n, n) matrix acting on columns
        b = torch.randn(nocc, n, dtype=torch.complex128, generator=g)
        b = b - (pm @ b.T).T
        rhs[j, :, :n] = b
        for i in range(nocc):
            a = hd - w[i] * torch.eye(n, dtype=torch.complex128) + shift * pm
            x = torch.linalg.solve(a, b[i])
            want[j, i, :n] = x - pm @ x
    d = dev()
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(d), p.to(d))
    x0 = torch.zeros_like(rhs).to(d)
    for op in (h, SimpleNamespace(apply=h.apply)):  # compacted column applies, then full-block applies
        got = cg_sternheimer(op, gpu, c_occ.to(d), eps.to(d), rhs.to(d), x0, shift, tol=1e-10, max_iter=500).cpu()
        err = float((got - want).abs().max() / want.abs().max())
        assert err < 1e-7, err
    # a warm start is certified, not trusted: the exact answer plus a small error,
    # and garbage in the padded slots (zeroed on entry), both converge to the same
    warm = want.clone()
    warm[0, 1] += 1e-3 * torch.randn(m, dtype=torch.complex128, generator=g) * cpu.mask[0]
    warm[1] += 5.0 * (~cpu.mask[1]).to(torch.complex128)
    got = cg_sternheimer(h, gpu, c_occ.to(d), eps.to(d), rhs.to(d), warm.to(d), shift, tol=1e-10,
                         max_iter=500).cpu()
    assert float((got - want).abs().max() / want.abs().max()) < 1e-7
    # complex64 operands run the FP32 kernels to an FP32-reachable tolerance
    got32 = cg_sternheimer(h, gpu, c_occ.to(torch.complex64).to(d), eps.to(d), rhs.to(torch.complex64).to(d),
                           x0.to(torch.complex64), shift, tol=1e-4, max_iter=500).cpu().to(torch.complex128)
    assert float((got32 - want).abs().max() / want.abs().max()) < 1e-3
    # a zero iteration budget returns the projected start
    z = cg_sternheimer(h, gpu, c_occ.to(d), eps.to(d), rhs.to(d), x0, shift, tol=1e-10, max_iter=0)
    assert float(z.abs().max()) == 0.0


@pytest.mark.parametrize("factor", [1, 2])
def test_small_max_dim_matches_oracle(factor):
    """max_dim_factor below 2 (max_dim < nb + n_add after a restart) is a legal
    argument: the reference concatenates past max_dim (davidson.py:757-759).
    Eigenvalues against the oracle's davidson_batched to 1e-10 Ha."""
    say(f"davidson si2 max_dim_factor={factor} vs oracle")
    s = system("si2")
    sel = [0, 6]
    cpu, gpu, p = select_k(s, sel)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    x0 = rand_block(2, 8, cpu.npw_max, cpu.mask, seed=12)
    r = davidson_batched(h.apply, x0.to(dev()), gpu.t, gpu.mask, tol=TOL, max_iter=400, max_dim_factor=factor)
    rr = ref_davidson(ref_h(s, sel).apply, x0, cpu.t, cpu.mask, tol=TOL, max_iter=400, max_dim_factor=factor)
    assert float(r.residual_norms.max()) < TOL and float(rr.residual_norms.max()) < TOL
    assert float((r.eigenvalues.cpu() - rr.eigenvalues).abs().max()) < 1e-10 * HARTREE_EV


def test_two_stage_matches_oracle_two_stage():
    """GPU davidson_batched_ms against the oracle's (solvers/davidson.py:777-812):
    the skip case (crossover >= tol) and the draft + polish case (crossover < tol),
    eigenvalues to 1e-10 Ha, the draft counted as low-precision applies."""
    from oracle.davidson import davidson_batched_ms as ref_ms

    say("davidson_batched_ms si2 vs oracle")
    s = system("si2")
    sel = [1, 9]
    cpu, gpu, p = select_k(s, sel)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    ref = ref_h(s, sel)
    x0 = rand_block(2, 8, cpu.npw_max, cpu.mask, seed=13)
    for tol, crossover in ((TOL, 1e-4), (1e-6, 1e-7)):
        got = davidson_batched_ms(h.apply, x0.to(dev()), gpu.t, gpu.mask, tol=tol, max_iter=200,
                                  crossover=crossover)
        want = ref_ms(ref.apply, x0, cpu.t, cpu.mask, tol=tol, max_iter=200, crossover=crossover)
        assert float(got.residual_norms.max()) < tol
        assert float((got.eigenvalues.cpu() - want.eigenvalues).abs().max()) < 1e-10 * HARTREE_EV
        assert (got.n_apply_low > 0) == (crossover < tol)


def test_cholqr2_complex64_ill_conditioned_matches_oracle():
    """CholQR2 of an ill-conditioned complex64 block (condition number ~3e3):
    the Gram is formed in FP64 from upcast rows (davidson.py:213-216), so the
    rows come out as orthonormal as the oracle's (complex64 storage bound) and
    span the same space."""
    from gradwave_amd.davidson import _cholqr2_dev
    from oracle.davidson import _cholqr2 as ref_cholqr2

    say("cholqr2 complex64 ill-conditioned vs oracle")
    g = torch.Generator().manual_seed(21)
    n, m = 24, 3000
    a = torch.randn(n, m, dtype=torch.complex128, generator=g)
    u, _, vh = torch.linalg.svd(a, full_matrices=False)
    sv = torch.logspace(0, -3.5, n, dtype=torch.float64)
    v = ((u * sv) @ vh).to(torch.complex64)[None]
    q, infos = _cholqr2_dev(v.to(dev()))
    assert int(infos.abs().max()) == 0
    q = q.cpu()
    qr = ref_cholqr2(v[0].T).T[None]
    eye = torch.eye(n, dtype=torch.complex128)

    def orth(z):
        z = z[0].to(torch.complex128)
        return float((z @ z.mH - eye).abs().max())

    assert orth(q) < max(4 * orth(qr), 1e-5), (orth(q), orth(qr))
    # same span: the oracle's rows lie in the span of the GPU rows
    z, zr = q[0].to(torch.complex128), qr[0].to(torch.complex128)
    resid = zr - (zr @ z.mH) @ z
    assert float(resid.abs().max()) < 1e-4
