"""GPU parity of the batched subspace eigensolver (gw_heev_batched) against
LAPACK (numpy.linalg.eigh, complex128) on the host.

Random Hermitian matrices of ragged sizes share one launch; spectra with
exact multiplicities, tight clusters, a split tridiagonal (diagonal input),
the zero matrix and a strided view cover the inverse-iteration and
reorthogonalisation paths. Eigenvalues are compared with LAPACK's; the
eigenvectors are checked through their residuals and orthonormality (their
phases, and the basis inside a degenerate eigenspace, are not unique).
"""

import numpy as np
import pytest
import torch

from gpu_cases import dev, say

pytestmark = pytest.mark.gpu

from gradwave_amd import linalg as LA  # noqa: E402


def _herm(n, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(n, n, dtype=torch.complex128, generator=g)
    return scale * 0.5 * (a + a.mH)


def _with_spectrum(lam, seed):
    n = len(lam)
    g = torch.Generator().manual_seed(seed)
    q, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.complex128, generator=g))
    a = q @ torch.diag(torch.tensor(lam, dtype=torch.complex128)) @ q.mH
    return 0.5 * (a + a.mH)


def _check(a, w, u, nev, tol=1e-12):
    a = a.to(torch.complex128)
    an = max(float(torch.linalg.matrix_norm(a, ord=2)), 1e-300)
    ref = np.linalg.eigvalsh(a.numpy())[:nev]
    w, u = w.cpu(), u.cpu()
    assert bool(torch.isfinite(w).all()) and bool(torch.isfinite(u).all())
    assert float(np.abs(w.numpy() - ref).max()) <= tol * an
    res = a @ u - u * w[None, :].to(torch.complex128)
    assert float(res.abs().max()) <= 10 * tol * an
    gram = u.mH @ u
    assert float((gram - torch.eye(nev, dtype=torch.complex128)).abs().max()) <= 10 * tol


def test_ragged_random_batch_one_launch():
    sizes = [(1, 1), (2, 2), (3, 1), (5, 5), (17, 4), (64, 16), (130, 130), (257, 64), (512, 128), (96, 32)]
    mats = [_herm(n, seed=n)[None] for n, _ in sizes]
    say(f"heev ragged batch {sizes}")
    out = LA.eigh_lowest([m.to(dev()) for m in mats], [nev for _, nev in sizes])
    for (n, nev), m, (w, u) in zip(sizes, mats, out):
        assert w.shape == (1, nev) and u.shape == (1, n, nev)
        _check(m[0], w[0], u[0], nev)


def test_batched_k_and_strided_view():
    nk, cap, m, nev = 3, 80, 61, 20
    big = torch.zeros(nk, cap, cap, dtype=torch.complex128)
    for k in range(nk):
        big[k, :m, :m] = _herm(m, seed=100 + k, scale=30.0)
    say("heev strided view")
    (w, u), = LA.eigh_lowest([big.to(dev())[:, :m, :m]], [nev])
    for k in range(nk):
        _check(big[k, :m, :m], w[k], u[k], nev)


@pytest.mark.parametrize("case", ["multiplicities", "cluster", "diagonal", "zero", "scaled_identity"])
def test_degenerate_and_split_spectra(case):
    if case == "multiplicities":
        lam = [-3.0] * 5 + [-1.0] * 3 + [0.0] * 4 + list(np.linspace(1.0, 9.0, 36))
        a = _with_spectrum(lam, seed=7)
    elif case == "cluster":
        lam = list(-2.0 + 1e-9 * np.arange(6)) + list(-2.0 + 1e-5 * np.arange(6)) + list(np.linspace(0, 40, 52))
        a = _with_spectrum(lam, seed=8)
    elif case == "diagonal":
        a = torch.diag(torch.tensor(np.r_[np.linspace(5, -5, 30), [0.0, 0.0, 2.0]], dtype=torch.complex128))
    elif case == "zero":
        a = torch.zeros(40, 40, dtype=torch.complex128)
    else:
        a = 2.5 * torch.eye(33, dtype=torch.complex128)
    n = a.shape[0]
    nev = min(n, 24)
    say(f"heev {case} n={n}")
    (w, u), = LA.eigh_lowest([a[None].to(dev())], [nev])
    if case == "zero":
        assert float(w.abs().max()) < 1e-13
        gram = u[0].cpu().mH @ u[0].cpu()
        assert float((gram - torch.eye(nev, dtype=torch.complex128)).abs().max()) < 1e-12
        return
    _check(a, w[0], u[0], nev)


def test_davidson_subspace_matrix():
    """A subspace matrix as davidson_batched forms it: Ritz values of a block in
    eV on the diagonal block, dense coupling to high-kinetic correction
    directions (|H| ~ 1e3 eV, spectrum spread over five decades)."""
    n, nb = 192, 48
    g = torch.Generator().manual_seed(11)
    a = torch.zeros(n, n, dtype=torch.complex128)
    a[:nb, :nb] = torch.diag(torch.linspace(-12.0, 8.0, nb, dtype=torch.float64).to(torch.complex128))
    c = 1e-2 * torch.randn(nb, n - nb, dtype=torch.complex128, generator=g)
    a[:nb, nb:] = c
    a[nb:, :nb] = c.mH
    d = _herm(n - nb, seed=12, scale=5.0) + torch.diag(torch.logspace(0, 3, n - nb).to(torch.complex128))
    a[nb:, nb:] = d
    say("heev Davidson-shaped n=192")
    (w, u), = LA.eigh_lowest([a[None].to(dev())], [nb])
    _check(a, w[0], u[0], nb)


def test_size_limits_raise():
    with pytest.raises(Exception):
        LA.eigh_lowest([torch.zeros(1, LA.heev_max_n() + 1, LA.heev_max_n() + 1, dtype=torch.complex128,
                                    device=dev())], [1])
    with pytest.raises(Exception):
        LA.eigh_lowest([torch.zeros(1, 4, 4, dtype=torch.complex128, device=dev())], [5])


def _hpd(n, seed, cond_scale=1.0):
    g = torch.Generator().manual_seed(seed)
    v = torch.randn(n, 3 * n + 5, dtype=torch.complex128, generator=g)
    a = cond_scale * (v @ v.mH)
    return 0.5 * (a + a.mH)


def test_cholesky_inverse_ragged_batch_matches_torch():
    """gw_chol_inv_batched against torch.linalg.cholesky_ex + a triangular
    solve on the host, ragged sizes and a batched k axis in one launch."""
    sizes = [(1, 1), (1, 2), (2, 8), (1, 17), (3, 32), (1, 64)]
    mats = [torch.stack([_hpd(n, seed=10 * n + k, cond_scale=10.0 ** (k - 1)) for k in range(nk)])
            for nk, n in sizes]
    say(f"cholesky inverse {sizes}")
    out = LA.cholesky_inverse([m.to(dev()) for m in mats])
    for (nk, n), m, (linv, info) in zip(sizes, mats, out):
        assert int(info.abs().max()) == 0
        lo = torch.linalg.cholesky(m)
        want = torch.linalg.solve_triangular(lo, torch.eye(n, dtype=torch.complex128).expand_as(lo), upper=False)
        got = linv.cpu()
        assert float((got - want).abs().max()) <= 1e-11 * float(want.abs().max())
        assert float(torch.triu(got, diagonal=1).abs().max()) == 0.0
        # L^-1 G L^-H = I
        eye = got @ m @ got.mH
        assert float((eye - torch.eye(n, dtype=torch.complex128)).abs().max()) < 1e-10


def test_cholesky_inverse_reports_breakdown():
    n = 12
    ok = _hpd(n, seed=3)
    bad = ok.clone()
    bad[5, 5] = -1.0  # the leading 6 x 6 minor is no longer positive definite
    say("cholesky inverse breakdown")
    (linv, info), = LA.cholesky_inverse([torch.stack([ok, bad]).to(dev())])
    info = info.cpu()
    _, want = torch.linalg.cholesky_ex(torch.stack([ok, bad]))
    assert int(info[0]) == 0 and int(info[1]) > 0
    assert int(info[1]) == int(want[1])
    with pytest.raises(Exception):
        LA.cholesky_inverse([torch.eye(LA.chol_max_n() + 1, dtype=torch.complex128, device=dev())[None]])
