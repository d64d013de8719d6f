"""GPU parity of the MFMA complex contractions against CPU complex128 products.

Shapes cover tile edges (not multiples of 16/64), batches, strided views,
the conjugated combine and beta accumulation.
"""

import pytest
import torch

from gpu_cases import dev, say

pytestmark = pytest.mark.gpu

from gradwave_amd import linalg as LA  # noqa: E402


def _rand(*shape, seed=0, dtype=torch.complex128):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(*shape, dtype=torch.complex128, generator=g).to(dtype)


TOL = {torch.complex128: 1e-13, torch.complex64: 3e-6}


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
@pytest.mark.parametrize("nk,M,N,K", [(1, 8, 8, 763), (3, 17, 33, 1000), (2, 70, 5, 4097),
                                      (1, 160, 160, 37073), (1, 256, 96, 2969), (4, 1, 1, 3)])
def test_inner(dtype, nk, M, N, K):
    say(f"inner {dtype} {nk}x{M}x{N}x{K}")
    x, y = _rand(nk, M, K, seed=1, dtype=dtype), _rand(nk, N, K, seed=2, dtype=dtype)
    want = torch.matmul(x.to(torch.complex128).conj(), y.to(torch.complex128).mT)
    got = LA.inner(x.to(dev()), y.to(dev())).cpu().to(torch.complex128)
    scale = float((x.abs() ** 2).sum(-1).max().sqrt() * (y.abs() ** 2).sum(-1).max().sqrt())
    assert float((got - want).abs().max()) < TOL[dtype] * scale


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
@pytest.mark.parametrize("nk,A,G,J", [(1, 8, 763, 8), (3, 33, 1000, 17), (2, 5, 4097, 70),
                                      (1, 160, 37073, 256), (1, 24, 2969, 96), (2, 1, 3, 1)])
@pytest.mark.parametrize("conj_u,beta", [(False, 0.0), (True, 1.0)])
def test_combine(dtype, nk, A, G, J, conj_u, beta):
    say(f"combine {dtype} {nk}x{A}x{G}x{J} conj={conj_u} beta={beta}")
    u, y, z0 = _rand(nk, J, A, seed=3, dtype=dtype), _rand(nk, J, G, seed=4, dtype=dtype), \
        _rand(nk, A, G, seed=5, dtype=dtype)
    uu = u.to(torch.complex128)
    uu = uu.conj() if conj_u else uu
    want = beta * z0.to(torch.complex128) + torch.einsum("kja,kjg->kag", uu, y.to(torch.complex128))
    zd = z0.to(dev())
    got = LA.combine(u.to(dev()), y.to(dev()), out=zd, beta=beta, conj_u=conj_u).cpu()
    scale = float(u.abs().max() * y.abs().max() * J) + beta * float(z0.abs().max())
    assert float((got.to(torch.complex128) - want).abs().max()) < TOL[dtype] * scale


def test_strided_views():
    say("inner/combine on strided views")
    big = _rand(2, 40, 300, seed=6).to(dev())
    x, y = big[:, 3:11], big[:, 20:37]
    want = torch.matmul(x.cpu().conj(), y.cpu().mT)
    assert float((LA.inner(x, y).cpu() - want).abs().max()) < 1e-12
    u = _rand(2, 8, 17, seed=7).to(dev())
    out = torch.zeros(2, 50, 300, dtype=torch.complex128, device=dev())
    LA.combine(u, x, out=out[:, 30:47], beta=0.0)
    want2 = torch.einsum("kja,kjg->kag", u.cpu(), x.cpu())
    assert float((out[:, 30:47].cpu() - want2).abs().max()) < 1e-12
    assert float(out[:, :30].abs().max()) == 0.0 and float(out[:, 47:].abs().max()) == 0.0


def test_residual_and_teter_kernels():
    say("residual + teter kernels")
    from oracle.davidson import teter_b

    nk, nb, m = 3, 6, 500
    x, hx = _rand(nk, nb, m, seed=8), _rand(nk, nb, m, seed=9)
    eig = torch.randn(nk, nb, dtype=torch.float64)
    r, rn = LA.residual(x.to(dev()), hx.to(dev()), eig.to(dev()))
    r_ref = hx - eig[..., None] * x
    assert float((r.cpu() - r_ref).abs().max()) < 1e-13
    assert float((rn.cpu() - torch.linalg.norm(r_ref, dim=-1)).abs().max()) < 1e-12
    t = torch.rand(nk, m, dtype=torch.float64) * 50
    t[:, -20:] = 0.0  # padded slots: K(0) = 1
    sel = torch.argsort(rn.cpu(), dim=1, descending=True)[:, :4]
    d = LA.teter_expand(x.to(dev()), r, t.to(dev()), sel.to(dev())).cpu()
    tb = torch.einsum("kbg,kg,kbg->kb", x.conj(), t.to(x.dtype), x).real
    dref = teter_b(torch.gather(r_ref, 1, sel[..., None].expand(-1, -1, m)), t, torch.gather(tb, 1, sel))
    dref = dref / torch.linalg.norm(dref, dim=-1, keepdim=True)
    assert float((d - dref).abs().max()) < 1e-13


RTOL = {torch.float64: 1e-13, torch.float32: 3e-6}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("nk,M,N,K", [(1, 8, 8, 763), (3, 17, 33, 1000), (1, 640, 160, 37073),
                                      (1, 70, 65, 4097), (2, 1, 1, 3)])
def test_real_inner(dtype, nk, M, N, K):
    say(f"rinner {dtype} {nk}x{M}x{N}x{K}")
    g = torch.Generator().manual_seed(3)
    x = torch.randn(nk, M, K, dtype=torch.float64, generator=g)
    y = torch.randn(nk, N, K, dtype=torch.float64, generator=g)
    want = torch.matmul(x.to(dtype).double(), y.to(dtype).double().mT)
    got = LA.rinner(x.to(dev(), dtype), y.to(dev(), dtype)).cpu().double()
    scale = float(x.norm(dim=-1).max() * y.norm(dim=-1).max())
    assert float((got - want).abs().max()) < RTOL[dtype] * scale
    if nk == 1:  # unbatched views with a row stride
        xs = x[0].to(dev(), dtype)[::2]
        got2 = LA.rinner(xs, y[0].to(dev(), dtype)).cpu().double()
        assert float((got2 - want[0][::2]).abs().max()) < RTOL[dtype] * scale


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("nk,A,G,J", [(1, 8, 763, 8), (3, 33, 1000, 17), (1, 160, 37073, 640),
                                      (1, 65, 4097, 70), (2, 1, 3, 1)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0)])
def test_real_combine(dtype, nk, A, G, J, alpha, beta):
    say(f"rcombine {dtype} {nk}x{A}x{G}x{J} a={alpha} b={beta}")
    g = torch.Generator().manual_seed(4)
    u = torch.randn(nk, J, A, dtype=torch.float64, generator=g).to(dtype)
    y = torch.randn(nk, J, G, dtype=torch.float64, generator=g).to(dtype)
    z0 = torch.randn(nk, A, G, dtype=torch.float64, generator=g).to(dtype)
    want = alpha * torch.matmul(u.double().mT, y.double()) + beta * z0.double()
    z = z0.to(dev())
    LA.rcombine(u.to(dev()), y.to(dev()), out=z, alpha=alpha, beta=beta)
    scale = float(u.double().norm(dim=-2).max() * y.double().norm(dim=-2).max()) + float(z0.abs().max())
    assert float((z.cpu().double() - want).abs().max()) < RTOL[dtype] * scale


@pytest.mark.parametrize("tile", list(range(1, 14)))
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_tile_variants_match(tile, dtype):
    """The autotuner's MFMA tile shapes give the same contractions: 1-5 the
    expanded-operand tiles, 6-9 the complex-accumulator tiles with four real
    products, 10-13 with three (Gauss); edges, batches, K tails, both combine forms."""
    from gradwave_amd import _lib as L

    say(f"gemm tile variant {tile} {dtype}")
    lib = L.lib()
    prec = L.precision_of(dtype)
    try:
        assert lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, prec, tile) == 0
        for (nk, M, N, K) in [(1, 256, 160, 37073), (3, 17, 33, 1000), (2, 70, 5, 4097), (1, 97, 130, 9)]:
            x, y = _rand(nk, M, K, seed=1, dtype=dtype), _rand(nk, N, K, seed=2, dtype=dtype)
            want = torch.matmul(x.to(torch.complex128).conj(), y.to(torch.complex128).mT)
            got = LA.inner(x.to(dev()), y.to(dev())).cpu().to(torch.complex128)
            scale = float((x.abs() ** 2).sum(-1).max().sqrt() * (y.abs() ** 2).sum(-1).max().sqrt())
            assert float((got - want).abs().max()) < TOL[dtype] * scale
        for (nk, A, G, J) in [(1, 160, 37073, 256), (3, 33, 1000, 17), (2, 5, 4097, 70), (1, 97, 130, 9)]:
            u, y = _rand(nk, J, A, seed=3, dtype=dtype), _rand(nk, J, G, seed=4, dtype=dtype)
            z0 = _rand(nk, A, G, seed=5, dtype=dtype)
            scale = float(u.abs().pow(2).sum(-2).max().sqrt() * y.abs().pow(2).sum(-2).max().sqrt()) + 1.0
            for conj_u, beta in ((True, 1.0), (False, 0.0)):
                uu = u.to(torch.complex128)
                uu = uu.conj() if conj_u else uu
                want = torch.matmul(uu.mT, y.to(torch.complex128)) + beta * z0.to(torch.complex128)
                z = z0.to(dev())
                LA.combine(u.to(dev()), y.to(dev()), out=z, beta=beta, conj_u=conj_u)
                assert float((z.cpu().to(torch.complex128) - want).abs().max()) < TOL[dtype] * scale
    finally:
        lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, prec, 0)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_combine_on_rocblas_matches(dtype):
    """The autotuner may move the combine (back-projection, Ritz) to rocBLAS."""
    from gradwave_amd import _lib as L

    say(f"combine on rocBLAS {dtype}")
    lib = L.lib()
    prec = L.precision_of(dtype)
    try:
        assert lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, 1) == 0
        for (nk, A, G, J) in [(1, 160, 37073, 256), (3, 33, 1000, 17), (2, 5, 4097, 70)]:
            for conj_u, beta in ((False, 0.0), (True, 1.0)):
                u, y = _rand(nk, J, A, seed=3, dtype=dtype), _rand(nk, J, G, seed=4, dtype=dtype)
                z0 = _rand(nk, A, G, seed=5, dtype=dtype)
                uu = u.to(torch.complex128)
                want = torch.matmul((uu.conj() if conj_u else uu).mT, y.to(torch.complex128)) + beta * z0.to(
                    torch.complex128)
                z = z0.to(dev())
                LA.combine(u.to(dev()), y.to(dev()), out=z, beta=beta, conj_u=conj_u)
                scale = float(u.abs().pow(2).sum(-2).max().sqrt() * y.abs().pow(2).sum(-2).max().sqrt()) + 1.0
                assert float((z.cpu().to(torch.complex128) - want).abs().max()) < TOL[dtype] * scale
    finally:
        lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, 0)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_inner_on_rocblas_matches(dtype):
    """The autotuner may run the split-K chunks of the inner product on rocBLAS;
    shapes cover a short last chunk, K below one chunk, batches and strided rows."""
    from gradwave_amd import _lib as L

    say(f"inner on rocBLAS {dtype}")
    lib = L.lib()
    prec = L.precision_of(dtype)
    try:
        assert lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, 1) == 0
        for (nk, M, N, K) in [(1, 256, 160, 37073), (3, 17, 33, 1000), (2, 70, 5, 4097), (1, 9, 7, 5)]:
            x, y = _rand(nk, M, K, seed=1, dtype=dtype), _rand(nk, N, K, seed=2, dtype=dtype)
            want = torch.matmul(x.to(torch.complex128).conj(), y.to(torch.complex128).mT)
            got = LA.inner(x.to(dev()), y.to(dev())).cpu().to(torch.complex128)
            scale = float((x.abs() ** 2).sum(-1).max().sqrt() * (y.abs() ** 2).sum(-1).max().sqrt())
            assert float((got - want).abs().max()) < TOL[dtype] * scale
        # strided rows: every other row of a wider block
        xw, yw = _rand(1, 40, 3000, seed=6, dtype=dtype), _rand(1, 24, 3000, seed=7, dtype=dtype)
        xs, ys = xw[:, ::2], yw[:, ::2]
        want = torch.matmul(xs.to(torch.complex128).conj(), ys.to(torch.complex128).mT)
        got = LA.inner(xw.to(dev())[:, ::2], yw.to(dev())[:, ::2]).cpu().to(torch.complex128)
        scale = float((xs.abs() ** 2).sum(-1).max().sqrt() * (ys.abs() ** 2).sum(-1).max().sqrt())
        assert float((got - want).abs().max()) < TOL[dtype] * scale
    finally:
        lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, 0)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_empty_contraction_is_zero(dtype):
    """K = 0: the inner product is an M x N zero matrix (written by a kernel, so
    an output full of garbage must come back zeroed); the combine with J = 0
    leaves beta * Z."""
    say(f"gemm empty contraction {dtype}")
    x = torch.zeros(2, 3, 0, dtype=dtype, device=dev())
    y = torch.zeros(2, 4, 0, dtype=dtype, device=dev())
    out = torch.full((2, 3, 4), 7.0, dtype=dtype, device=dev())
    c = LA.inner(x, y, out=out)
    assert c.shape == (2, 3, 4) and float(c.abs().max().cpu()) == 0.0
    z = torch.full((2, 3, 5), 2.0, dtype=dtype, device=dev())
    u = torch.zeros(2, 0, 3, dtype=dtype, device=dev())
    v = torch.zeros(2, 0, 5, dtype=dtype, device=dev())
    r = LA.combine(u, v, out=z, beta=0.5)
    assert float((r - 1.0).abs().max().cpu()) == 0.0
