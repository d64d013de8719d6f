"""gw_hpsi_apply_noalloc (the apply whose scratch is a fixed device array)
against gw_hpsi_apply_strided, called through the C ABI: two calls in a row on
one GPU, the second with a smaller input (fewer bands, other precision, a
strided band slice), each equal to the workspace entry point's result (the
same kernels run, so the results agree to round-off; fp64 1e-14, fp32 1e-6
relative); a block whose scratch need exceeds the maximum is refused before
any launch."""

from __future__ import annotations

import ctypes as C

import pytest
import torch

from tests.gpu_cases import dev, rand_block, say, select_k, system

pytestmark = pytest.mark.gpu


def _noalloc(h, c, out, terms=7):
    from gradwave_amd import _lib as L

    return L.lib().gw_hpsi_apply_noalloc(h.plan.handle, L.precision_of(c.dtype), terms, c.data_ptr(), c.stride(1),
                                         c.stride(0), out.data_ptr(), out.stride(1), out.stride(0), c.shape[1],
                                         L.stream_ptr(c.device))


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def test_noalloc_apply_matches_workspace_apply_twice():
    from gradwave_amd import _lib as L
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    say("no-allocation apply vs workspace apply, si2 10 k")
    s = system("si2")
    cpu, gpu, p = select_k(s)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    nk, m = cpu.mask.shape
    x = rand_block(nk, 8, m, cpu.mask, seed=31).to(dev())
    out = torch.full_like(x, float("nan"))
    assert _noalloc(h, x, out) == 0, L.lib().gw_last_error()
    want = h.apply_terms(x, L.GW_TERM_ALL)
    torch.cuda.synchronize()
    assert _rel(out, want) < 1e-14
    # second call: 3 bands of a strided band slice, complex64
    x2 = rand_block(nk, 5, m, cpu.mask, seed=32).to(dev()).to(torch.complex64)
    xs = x2[:, 1:4]
    out2 = torch.full((nk, 3, m), float("nan"), dtype=torch.complex64, device=dev())
    assert _noalloc(h, xs, out2) == 0, L.lib().gw_last_error()
    want2 = h.apply_terms(xs.contiguous(), L.GW_TERM_ALL)
    torch.cuda.synchronize()
    assert _rel(out2, want2) < 1e-6
    # the first call's result is unchanged by the second (outputs are the caller's)
    assert _rel(out, want) < 1e-14


def test_noalloc_apply_refuses_oversized_block():
    from gradwave_amd import _lib as L
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    say("no-allocation apply refuses a block above the scratch maximum")
    s = system("si8")
    cpu, gpu, p = select_k(s)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()))
    nk, m = cpu.mask.shape
    cap = int(L.lib().gw_hpsi_noalloc_max_bytes())
    nb = 24
    assert int(L.lib().gw_hpsi_workspace_bytes(h.plan.handle, nb, L.GW_FP64)) > cap
    x = torch.zeros(nk, nb, m, dtype=torch.complex128, device=dev())
    out = torch.zeros_like(x)
    rc = _noalloc(h, x, out)
    assert rc == 4 and b"fixed maximum" in L.lib().gw_last_error()
