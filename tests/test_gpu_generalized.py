"""GPU parity of the S-metric generalized Davidson (scf/uspp_batch.py:27-233): the
HIP S apply against the oracle's, and the generalized eigenpairs against the
oracle solve and the dense generalized eigenproblem (Si2, two k-points, 8 bands)."""

import dataclasses

import numpy as np
import pytest
import scipy.linalg as sl
import torch

from tests.gpu_cases import dev, say, select_k, system
from tests.uspp_cases import dense_s, q_blocks

pytestmark = pytest.mark.gpu

SEL = [0, 3]
HARTREE_EV = 27.211386245988


def _cases():
    from oracle.dense import dense_h
    from oracle.generalized import BatchedHS as RefHS

    s = system("si2")
    cpu, gpu, p = select_k(s, SEL)
    q = q_blocks(cpu.proj_atom_index.numpy(), seed=1)
    ref = RefHS(cpu, s.shape, s.v_r, p, cpu.dij_full, q)
    sub = dataclasses.replace(s, bk=cpu, p=p)
    return s, cpu, gpu, p, q, ref, sub, dense_h


def _start(mask, nb, seed=5):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(mask.shape[0], nb, mask.shape[1], dtype=torch.complex128, generator=g)
    return x * mask[:, None, :]


@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-12), (torch.complex64, 2e-5)])
def test_s_apply_matches_oracle(dtype, tol):
    from gradwave_amd.generalized import BatchedHS

    say(f"S apply {dtype}")
    s, cpu, gpu, p, q, ref, _, _ = _cases()
    hs = BatchedHS(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), cpu.dij_full, q)
    c = _start(cpu.mask, 5, seed=3)
    want = ref.s(c)
    got = hs.s(c.to(dtype).to(dev())).cpu().to(torch.complex128)
    assert float((got - want).abs().max()) < tol * float(want.abs().max())
    got_h = hs.h(c.to(dtype).to(dev())).cpu().to(torch.complex128)
    want_h = ref.h(c)
    assert float((got_h - want_h).abs().max()) < tol * float(want_h.abs().max())


def test_generalized_davidson_matches_oracle_and_dense():
    from gradwave_amd.generalized import BatchedHS, davidson_gen_batched
    from oracle.generalized import davidson_gen_batched as ref_gen

    say("generalized davidson si2 2k nb=8")
    s, cpu, gpu, p, q, ref, sub, dense_h = _cases()
    hs = BatchedHS(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), cpu.dij_full, q)
    nb = 8
    x0 = _start(cpu.mask, nb)
    eig, x = davidson_gen_batched(hs, x0.to(dev()), nb, tol=1e-9, max_iter=200)
    eig = eig.cpu()
    want, _ = ref_gen(ref, x0, nb, tol=1e-9, max_iter=200)
    assert float((eig - want).abs().max()) / HARTREE_EV < 1e-10
    for i in range(len(SEL)):
        n = int(cpu.npw[i])
        dense = sl.eigh(dense_h(sub, i).numpy(), dense_s(p[i], q, n).numpy(), eigvals_only=True,
                        subset_by_index=[0, nb - 1])
        assert np.abs(eig[i].numpy() - dense).max() / HARTREE_EV < 1e-10
    xd = x
    r = hs.h(xd) - eig.to(dev())[..., None].to(xd.dtype) * hs.s(xd)
    assert float(torch.linalg.norm(r, dim=-1).max()) < 1e-8


def test_refuses_unsupported_terms():
    from gradwave_amd.generalized import BatchedHS

    s, cpu, gpu, p, q, *_ = _cases()
    with pytest.raises(NotImplementedError):
        BatchedHS(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), cpu.dij_full, q, metagga_v=torch.zeros(1))
    hq = torch.zeros(len(SEL), 2, cpu.mask.shape[1], dtype=torch.complex128, device=dev())
    with pytest.raises(NotImplementedError):
        BatchedHS(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), cpu.dij_full, q, hub_sphi=hq,
                  hub_d=torch.eye(2, dtype=torch.complex128) * 1j)
    with pytest.raises(ValueError):
        BatchedHS(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), cpu.dij_full, q.to(torch.complex128) * 1j)


def _dual_grid_case():
    """Si2 at two k-points in a 36^3 dense box with the local term on a 24^3
    smooth box (its own sphere map and a smooth V), plus a DFT+U term of five
    seeded unit projectors with a real symmetric D."""
    from oracle import systems as S

    cell = S.fcc_primitive(S.A_SI)
    pos = np.array([[0.0, 0.0, 0.0], [0.25 * S.A_SI] * 3])
    k = np.array([[0.125, 0.25, -0.125], [0.375, 0.0, 0.125]])
    small = S.build_system("s24", cell, pos, [S.SI, S.SI], 30.0, k, np.full(2, 0.5), 8, shape=(24, 24, 24))
    dense = S.build_system("s36", cell, pos, [S.SI, S.SI], 30.0, k, np.full(2, 0.5), 8, shape=(36, 36, 36))
    g = torch.Generator().manual_seed(31)
    hub_q = torch.randn(2, 5, dense.bk.npw_max, dtype=torch.complex128, generator=g) * dense.bk.mask[:, None, :]
    hub_q = hub_q / torch.linalg.norm(hub_q, dim=-1, keepdim=True)  # unit projectors: a term of a few eV
    a = torch.randn(5, 5, dtype=torch.float64, generator=g)
    hub_d = 0.3 * (a + a.T)
    return small, dense, hub_q, hub_d


def test_dual_grid_and_hubbard_term_match_oracle():
    """BatchedHamiltonian(..., hub_q, hub_dij, smooth) (core/batch.py:249-279,
    501-504) against the oracle's, FP64 1e-12 / FP32 2e-5, and the generalized
    Davidson on that H (scf/uspp_batch.py:37-41) against the oracle's solve."""
    from types import SimpleNamespace

    from gradwave_amd.generalized import BatchedHS, davidson_gen_batched
    from gradwave_amd.hamiltonian import BatchedHamiltonian
    from oracle.generalized import BatchedHS as RefHS
    from oracle.generalized import davidson_gen_batched as ref_gen
    from oracle.hamiltonian import BatchedHamiltonian as RefH

    say("dual grid + DFT+U term vs oracle")
    small, dense, hub_q, hub_d = _dual_grid_case()
    d = dev()
    bk = dense.bk
    gpu = SimpleNamespace(npw=bk.npw.to(d), mask=bk.mask.to(d), flat_idx=bk.flat_idx.to(d), t=bk.t.to(d),
                          dij_full=bk.dij_full.to(d))
    smooth_cpu = ((24, 24, 24), small.bk.flat_idx, small.v_r)
    smooth_gpu = ((24, 24, 24), small.bk.flat_idx.to(d), small.v_r.to(d))
    ref = RefH(bk, dense.shape, dense.v_r, dense.p, hub_q=hub_q, hub_dij=hub_d, smooth=smooth_cpu)
    h = BatchedHamiltonian(gpu, dense.shape, dense.v_r.to(d), dense.p.to(d), hub_q=hub_q.to(d),
                           hub_dij=hub_d.to(d), smooth=smooth_gpu)
    assert h.shape == (24, 24, 24)
    c = _start(bk.mask, 4, seed=7)
    want = ref.apply(c)
    assert float((h.apply(c.to(d)).cpu() - want).abs().max()) < 1e-12 * float(want.abs().max())
    got32 = h.apply(c.to(torch.complex64).to(d)).cpu().to(torch.complex128)
    assert float((got32 - want).abs().max()) < 2e-5 * float(want.abs().max())
    kcol = torch.tensor([1, 0, 1])
    cols = h.apply_cols(c[kcol, torch.tensor([0, 2, 3])].to(d), kcol.to(d)).cpu()
    assert float((cols - want[kcol, torch.tensor([0, 2, 3])]).abs().max()) < 1e-12 * float(want.abs().max())
    # the generalized solve on the same H
    q = q_blocks(bk.proj_atom_index.numpy(), seed=2)
    hs = BatchedHS(gpu, dense.shape, dense.v_r.to(d), dense.p.to(d), bk.dij_full, q, hub_sphi=hub_q.to(d),
                   hub_d=hub_d.to(d), smooth=smooth_gpu)
    rhs = RefHS(bk, dense.shape, dense.v_r, dense.p, bk.dij_full, q, hub_sphi=hub_q, hub_d=hub_d, smooth=smooth_cpu)
    x0 = _start(bk.mask, 8)
    eig, _ = davidson_gen_batched(hs, x0.to(d), 8, tol=1e-9, max_iter=200)
    want_e, _ = ref_gen(rhs, x0, 8, tol=1e-9, max_iter=200)
    assert float((eig.cpu() - want_e).abs().max()) / HARTREE_EV < 1e-10


def test_sternheimer_s_metric_matches_dense_solve():
    """S-metric cg_sternheimer (postscf/_response.py:691-745) on Si2, two
    k-points, the four lowest pairs of H c = eps S c (c^H S c = 1): against a
    dense solve of [(1 - S P)(H - eps_n S) + s S P S] x = rhs with P = sum |c><c|
    and rhs in the S-conduction space, returned as (1 - P S) x. With S = 1 and
    s_occ = c_occ (orthonormal) it runs the norm-conserving solve's operations
    (_response.py:706-709)."""
    from types import SimpleNamespace

    from gradwave_amd.generalized import BatchedHS
    from gradwave_amd.response import cg_sternheimer

    say("sternheimer cg, S metric, si2")
    s, cpu, gpu, p, q, _, sub, dense_h = _cases()
    d = dev()
    hs = BatchedHS(gpu, s.shape, s.v_r.to(d), p.to(d), cpu.dij_full, q)
    nk, m = cpu.mask.shape
    nocc, shift = 4, 5.0
    c_occ = torch.zeros(nk, nocc, m, dtype=torch.complex128)
    eps = torch.zeros(nk, nocc, dtype=torch.float64)
    rhs = torch.zeros_like(c_occ)
    want = torch.zeros_like(c_occ)
    rng = np.random.default_rng(7)
    for i in range(nk):
        n = int(cpu.npw[i])
        hd, sd = dense_h(sub, i).numpy(), dense_s(p[i], q, n).numpy()
        w, v = sl.eigh(hd, sd)
        c = v[:, :nocc]
        c_occ[i, :, :n] = torch.as_tensor(c.T)
        eps[i] = torch.as_tensor(w[:nocc])
        pm = c @ c.conj().T
        b = rng.standard_normal((n, nocc)) + 1j * rng.standard_normal((n, nocc))
        b = b - sd @ (pm @ b)
        rhs[i, :, :n] = torch.as_tensor(b.T)
        for j in range(nocc):
            a = (np.eye(n) - sd @ pm) @ (hd - w[j] * sd) + shift * (sd @ pm @ sd)
            x = np.linalg.solve(a, b[:, j])
            want[i, j, :n] = torch.as_tensor(x - pm @ (sd @ x))
    x0 = torch.zeros_like(rhs).to(d)
    c_d = c_occ.to(d)
    got = cg_sternheimer(hs.ham, gpu, c_d, eps.to(d), rhs.to(d), x0, shift, tol=1e-10, max_iter=500,
                         s_apply=hs.s, s_occ=hs.s(c_d)).cpu()
    err = float((got - want).abs().max() / want.abs().max())
    assert err < 1e-7, err
    with pytest.raises(ValueError, match="s_occ"):
        cg_sternheimer(hs.ham, gpu, c_d, eps.to(d), rhs.to(d), x0, shift, s_apply=hs.s)
    # the norm-conserving limit: S = 1 and s_occ = c_occ (H's own orthonormal
    # eigenvectors here) run the plain solve's operations
    c_nc = torch.zeros_like(c_occ)
    eps_nc = torch.zeros_like(eps)
    rhs_nc = torch.zeros_like(rhs)
    for i in range(nk):
        n = int(cpu.npw[i])
        w, u = torch.linalg.eigh(dense_h(sub, i))
        c = u[:, :nocc]
        c_nc[i, :, :n] = c.T
        eps_nc[i] = w[:nocc]
        b = torch.as_tensor(rng.standard_normal((n, nocc)) + 1j * rng.standard_normal((n, nocc)))
        rhs_nc[i, :, :n] = (b - c @ (c.conj().T @ b)).T
    nc = SimpleNamespace(apply=hs.ham.apply)
    args = (gpu, c_nc.to(d), eps_nc.to(d), rhs_nc.to(d), x0, shift)
    plain = cg_sternheimer(nc, *args, tol=1e-10, max_iter=500)
    ident = cg_sternheimer(nc, *args, tol=1e-10, max_iter=500, s_apply=lambda y: y.clone(),
                           s_occ=c_nc.to(d))
    assert bool(torch.isfinite(plain).all())
    assert float((plain - ident).abs().max() / plain.abs().max()) < 1e-13
