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
        BatchedHS(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), cpu.dij_full, q, hub_d=torch.zeros(1))
    with pytest.raises(ValueError):
        BatchedHS(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), cpu.dij_full, q.to(torch.complex128) * 1j)
