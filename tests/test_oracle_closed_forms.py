"""The oracle against closed forms, brute force and invariants (no reference code).

Mirrors the reference's own hot-path tests: test_grids.py:13-48,88-102,
test_exact_limits.py:39-100, test_toeplitz_local.py:61-113,
test_davidson_qr_offload.py:77-87, plus analytic checks of the Gaussian pseudo.
"""

import math

import numpy as np
import pytest
import scipy.constants as sc
import torch
from scipy.integrate import quad
from scipy.special import erfc, spherical_jn

from oracle import grids as G
from oracle import systems as S
from oracle.davidson import _cholqr2, davidson_batched, teter_b
from oracle.dense import dense_h
from oracle.hamiltonian import BatchedHamiltonian, BatchedK, toeplitz_local
from oracle.pseudo import SI, kb_form_factor, vloc_form_factor
from oracle.units import BOHR_ANG, E2, HARTREE_EV, HBAR2_2M, RY_EV
from oracle.ylm import ylm_all


def test_units_codata2018():
    # CODATA 2018 (the reference's stated source); scipy may ship a later
    # adjustment, so compare to it only at the 1e-9 relative level
    assert HARTREE_EV == 27.211386245988 and BOHR_ANG == 0.529177210903
    assert abs(HARTREE_EV / sc.physical_constants["Hartree energy in eV"][0] - 1) < 1e-9
    assert abs(BOHR_ANG / (sc.physical_constants["Bohr radius"][0] * 1e10) - 1) < 1e-9
    assert abs(HBAR2_2M / (sc.hbar**2 / (2 * sc.m_e) / sc.e * 1e20) - 1) < 1e-8
    assert abs(RY_EV * BOHR_ANG**2 - HBAR2_2M) == 0.0


def test_good_fft_size():
    assert [G.good_fft_size(n) for n in (1, 11, 13, 97)] == [1, 12, 14, 98]


def _bk_no_proj(sph) -> BatchedK:
    n = sph.npw
    return BatchedK(npw=torch.tensor([n]), mask=torch.ones(1, n, dtype=torch.bool),
                    flat_idx=sph.flat_idx[None], kpg=sph.kpg[None],
                    t=(HBAR2_2M * sph.kpg2)[None],
                    proj_phase_free=torch.zeros(1, 0, n, dtype=torch.complex128),
                    proj_atom_index=torch.zeros(0, dtype=torch.int64),
                    dij_full=torch.zeros(0, 0, dtype=torch.float64))


def _solve(cell, ecut, k, v_of_grid, nb, tol=1e-11):
    """test_exact_limits.py:39-56 -- lowest nb eigenvalues at one k via apply + Davidson."""
    grid = G.build_fft_grid(np.asarray(cell, dtype=np.float64), ecut)
    sph = G.build_gsphere(grid, ecut, k)
    bk = _bk_no_proj(sph)
    h = BatchedHamiltonian(bk, grid.shape, v_of_grid(grid), bk.proj_phase_free)
    gen = torch.Generator().manual_seed(0)
    x0 = (torch.randn(1, nb, sph.npw, generator=gen, dtype=torch.float64)
          + 1j * torch.randn(1, nb, sph.npw, generator=gen, dtype=torch.float64))
    x0 = (x0 / (1.0 + HBAR2_2M * sph.kpg2)).to(torch.complex128)
    x0 = x0 / torch.linalg.norm(x0, dim=-1, keepdim=True)
    res = davidson_batched(h.apply, x0, bk.t, bk.mask, tol=tol, max_iter=200)
    return res.eigenvalues[0], sph


def test_empty_lattice_triclinic():
    cell = [[3.1, 0.0, 0.0], [0.4, 2.7, 0.0], [0.2, -0.3, 3.3]]
    eigs, sph = _solve(cell, 30 * RY_EV, (0.31, 0.17, -0.23),
                       lambda g: torch.zeros(g.shape, dtype=torch.float64), 12)
    exact = torch.sort(HBAR2_2M * sph.kpg2).values[:12]
    assert float((eigs - exact).abs().max()) < 1e-9


def test_cosine_potential_mathieu():
    from scipy.special import mathieu_a, mathieu_b

    a, L, q = 3.0, 1.2, 1.5
    unit = HBAR2_2M * (np.pi / a) ** 2
    v0 = 2.0 * q * unit

    def v_cos(grid):
        x = torch.arange(grid.shape[0], dtype=torch.float64) / grid.shape[0]
        return (v0 * torch.cos(2.0 * np.pi * x))[:, None, None].expand(grid.shape).contiguous()

    cases = {(0.0, 0.0, 0.0): [mathieu_a(0, q), mathieu_b(2, q), mathieu_a(2, q),
                               mathieu_b(4, q), mathieu_a(4, q)],
             (0.5, 0.0, 0.0): [mathieu_b(1, q), mathieu_a(1, q), mathieu_b(3, q),
                               mathieu_a(3, q)]}
    for k, chars in cases.items():
        exact = np.array([unit * c for c in chars])
        exact = exact[exact < 80.0]
        eigs, _ = _solve(np.diag([a, L, L]), 60 * RY_EV, k, v_cos, len(exact) + 2)
        assert np.abs(eigs.numpy()[: len(exact)] - exact).max() < 1e-7


@pytest.fixture(scope="module")
def si2():
    return S.si2_c1()


def test_kpoint_reduction_and_npw(si2):
    assert np.allclose(si2.kweights * 64, [2, 6, 6, 6, 6, 12, 12, 6, 2, 6])
    assert si2.bk.npw.tolist() == [743, 748, 753, 757, 751, 748, 747, 747, 763, 744]
    # brute-force sphere count at the first k
    b = G.reciprocal_cell(si2.cell)
    ms = np.stack(np.meshgrid(*[np.arange(-12, 13)] * 3, indexing="ij"), -1).reshape(-1, 3)
    kpg = (ms + si2.kfrac[0]) @ b
    assert int((HBAR2_2M * (kpg**2).sum(1) <= si2.ecut * (1 + 1e-12)).sum()) == 743


def test_toeplitz_identity_random_v(si2):
    """test_toeplitz_local.py:61-78 -- FFT local term == dense convolution, padded k."""
    torch.manual_seed(0)
    v = torch.randn(*si2.shape, dtype=torch.float64)
    bk = si2.bk
    h = BatchedHamiltonian(bk, si2.shape, v, si2.p[:, :0])
    c = torch.randn(bk.nk, 4, bk.npw_max, dtype=torch.complex128)  # random in padding too
    out_fft = h.local_fft(c, v) * bk.mask[:, None, :]
    out_toep = toeplitz_local(bk, si2.shape, v, c)
    assert float((out_fft - out_toep).abs().max()) < 1e-12


def test_apply_hermitian_and_masked(si2):
    h = BatchedHamiltonian(si2.bk, si2.shape, si2.v_r, si2.p)
    g = torch.Generator().manual_seed(1)
    m = si2.bk.mask[:, None, :]
    x = torch.randn(si2.bk.nk, 3, si2.bk.npw_max, dtype=torch.complex128, generator=g) * m
    y = torch.randn(si2.bk.nk, 3, si2.bk.npw_max, dtype=torch.complex128, generator=g) * m
    hx, hy = h.apply(x), h.apply(y)
    lhs = (y.conj() * hx).sum(-1)
    rhs = (hy.conj() * x).sum(-1)
    assert float((lhs - rhs).abs().max() / lhs.abs().max()) < 1e-13
    assert float(hx[~si2.bk.mask[:, None, :].expand_as(hx)].abs().max()) == 0.0


def test_davidson_matches_dense_eigvalsh(si2):
    """Brute force: explicit npw x npw H per k, eigvalsh vs apply + Davidson."""
    h = BatchedHamiltonian(si2.bk, si2.shape, si2.v_r, si2.p)
    g = torch.Generator().manual_seed(0)
    x0 = torch.randn(si2.bk.nk, 8, si2.bk.npw_max, dtype=torch.complex128, generator=g)
    x0 = x0 * si2.bk.mask[:, None, :]
    r = davidson_batched(h.apply, x0, si2.bk.t, si2.bk.mask, tol=1e-10 * HARTREE_EV, max_iter=100)
    assert float(r.residual_norms.max()) < 1e-10 * HARTREE_EV
    for ik in (0, 8):
        e = torch.linalg.eigvalsh(dense_h(si2, ik))[:8]
        assert float((r.eigenvalues[ik] - e).abs().max()) < 1e-10 * HARTREE_EV


def test_ylm_orthonormal():
    xg, wg = np.polynomial.legendre.leggauss(24)
    phi = np.arange(48) * 2 * np.pi / 48
    ct, ph = np.meshgrid(xg, phi, indexing="ij")
    st = np.sqrt(1 - ct**2)
    g = torch.as_tensor(np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3))
    w = torch.as_tensor((wg[:, None] * np.full(48, 2 * np.pi / 48)[None]).reshape(-1))
    y = ylm_all(3, g)
    gram = (y * w[:, None]).T @ y
    assert float((gram - torch.eye(16, dtype=torch.float64)).abs().max()) < 1e-12
    # l=1 closed form sqrt(3/4pi) (z, x, y)
    assert torch.allclose(y[:, 1:4], math.sqrt(3 / (4 * math.pi)) * g[:, [2, 0, 1]], atol=1e-15)


@pytest.mark.parametrize("l", [0, 1])
def test_kb_form_factor_quadrature(l):
    w = SI.w
    norm = math.sqrt(2.0 / (math.gamma(l + 1.5) * w ** (2 * l + 3)))
    assert abs(quad(lambda r: r**2 * (norm * r**l * math.exp(-r * r / (2 * w * w))) ** 2,
                    0, 30 * w)[0] - 1.0) < 1e-12
    for q in (0.3, 1.7, 4.2):
        num = quad(lambda r: r**2 * norm * r**l * math.exp(-r * r / (2 * w * w))
                   * spherical_jn(l, q * r), 0, 30 * w, limit=200)[0]
        assert abs(num - float(kb_form_factor(SI, l, np.array([q]))[0])) < 1e-11


def test_local_form_factor_is_smeared_coulomb():
    """v(g) + 4 pi Z e^2/g^2 = 4 pi Z e^2 int r erfc(r/rc) j0(gr) dr; alpha is its g->0 limit."""
    z, rc = SI.z, SI.rc
    for g in (0.5, 1.3, 3.0):
        short = 4 * math.pi * z * E2 * quad(lambda r: r * erfc(r / rc) * math.sin(g * r) / (g * r),
                                            0, 20 * rc, limit=200)[0]
        v = float(vloc_form_factor(SI, np.array([g]))[0])
        assert abs(v + 4 * math.pi * z * E2 / g**2 - short) < 1e-9 * abs(short)
    alpha = 4 * math.pi * z * E2 * quad(lambda r: r * erfc(r / rc), 0, 20 * rc)[0]
    assert abs(float(vloc_form_factor(SI, np.array([0.0]))[0]) - alpha) < 1e-10 * alpha


def test_local_potential_real_and_mean(si2):
    """V(r) is real (imaginary part dropped only at Nyquist) and its mean is V_loc(G=0)."""
    from oracle.pseudo import local_potential_g

    vg = local_potential_g(si2.positions, si2.species, si2.grid.g_cart, si2.grid.volume)
    expect_mean = 2 * math.pi * SI.z * E2 * SI.rc**2 / si2.grid.volume
    assert abs(float(si2.v_r.mean()) - expect_mean) < 1e-12 * abs(expect_mean) + 1e-14
    assert abs(float(vg[0, 0, 0].real) - expect_mean) < 1e-13


def test_teter_closed_form():
    r = torch.ones(1, 2, 3, dtype=torch.complex128)
    t_g = torch.tensor([[0.0, 1.0, 2.0]], dtype=torch.float64)
    out = teter_b(r, t_g, torch.ones(1, 2, dtype=torch.float64))
    k1 = 65.0 / 81.0
    k2 = (27 + 36 + 48 + 64) / (27 + 36 + 48 + 64 + 256)
    assert torch.allclose(out[0, 0].real, torch.tensor([1.0, k1, k2], dtype=torch.float64))


def test_cholqr2_spans_and_orthonormal():
    g = torch.Generator().manual_seed(3)
    x = torch.randn(2, 200, 6, dtype=torch.complex128, generator=g)
    q = _cholqr2(x)
    eye = torch.eye(6, dtype=torch.complex128)
    assert float((q.mH @ q - eye).abs().max()) < 1e-13
    # span: x = q (q^H x)
    assert float((q @ (q.mH @ x) - x).abs().max()) < 1e-12


def test_density_integrates_to_weighted_norms(si2):
    """Parseval: (volume / N) sum_r rho = sum_k w_k sum_b f_kb ||c_kb||^2 for the
    oracle's density_b (core/batch.py:562-587)."""
    from oracle.hamiltonian import density_b

    g = torch.Generator().manual_seed(4)
    bk = si2.bk
    c = torch.randn(bk.nk, 3, bk.npw_max, dtype=torch.complex128, generator=g) * bk.mask[:, None, :]
    occ = torch.rand(bk.nk, 3, dtype=torch.float64, generator=g)
    kw = torch.as_tensor(si2.kweights)
    rho = density_b(c, occ, kw, bk, si2.shape, si2.grid.volume)
    n = si2.shape[0] * si2.shape[1] * si2.shape[2]
    want = float((kw[:, None] * occ * (c.abs() ** 2).sum(-1)).sum())
    assert abs(float(rho.sum()) * si2.grid.volume / n - want) < 1e-12 * want
    assert float(rho.min()) >= 0.0


def _fourier_interp(v_small, shape_big):
    """Real v on a small box -> the same band-limited field on a bigger box (zero-padded G)."""
    import numpy as np

    vg = np.fft.fftn(v_small.numpy()) / v_small.numel()
    n, m = v_small.shape, shape_big
    big = np.zeros(m, dtype=complex)
    freqs = [np.fft.fftfreq(k, 1.0 / k).astype(int) for k in n]
    for i, fi in enumerate(freqs[0]):
        for j, fj in enumerate(freqs[1]):
            for k, fk in enumerate(freqs[2]):
                big[fi % m[0], fj % m[1], fk % m[2]] = vg[i, j, k]
    return torch.as_tensor(np.fft.ifftn(big).real * (m[0] * m[1] * m[2]))


def test_dual_grid_local_term_equals_dense_box():
    """core/batch.py:258-266: the local term on the smooth box (its own sphere
    map) equals the dense-box term for a V band-limited to the smooth box, since
    V-hat(G_i - G_j) never aliases when the smooth box holds 2m+1 points per axis."""
    cell = S.fcc_primitive(S.A_SI)
    pos = np.array([[0.0, 0.0, 0.0], [0.25 * S.A_SI] * 3])
    k = np.array([[0.125, 0.25, -0.125], [0.375, 0.0, 0.125]])
    small = S.build_system("s24", cell, pos, [SI, SI], 30.0, k, np.full(2, 0.5), 4, shape=(24, 24, 24))
    dense = S.build_system("s36", cell, pos, [SI, SI], 30.0, k, np.full(2, 0.5), 4, shape=(36, 36, 36))
    assert torch.equal(small.bk.npw, dense.bk.npw)
    g = torch.Generator().manual_seed(9)
    v_s = torch.randn(24, 24, 24, dtype=torch.float64, generator=g)
    vg = torch.fft.fftn(v_s)  # drop the Nyquist planes: the field is then band-limited below 12
    vg[12], vg[:, 12], vg[:, :, 12] = 0, 0, 0
    v_s = torch.fft.ifftn(vg).real
    v_d = _fourier_interp(v_s, (36, 36, 36))
    c = torch.randn(2, 4, dense.bk.npw_max, dtype=torch.complex128, generator=g) * dense.bk.mask[:, None, :]
    h_dense = BatchedHamiltonian(dense.bk, dense.shape, v_d, dense.p)
    h_smooth = BatchedHamiltonian(dense.bk, dense.shape, None, dense.p,
                                  smooth=((24, 24, 24), small.bk.flat_idx, v_s))
    want = h_dense.apply(c)
    assert h_smooth.shape == (24, 24, 24)
    assert float((h_smooth.apply(c) - want).abs().max()) < 1e-11 * float(want.abs().max())
    # the smooth box with another k-point's map scrambles the term
    h_bad = BatchedHamiltonian(dense.bk, dense.shape, None, dense.p,
                               smooth=((24, 24, 24), torch.roll(small.bk.flat_idx, 1, dims=0), v_s))
    assert float((h_bad.apply(c) - want).abs().max()) > 1e-5 * float(want.abs().max())


def test_hubbard_term_is_a_second_projector_contraction():
    """core/batch.py:501-504, 543-546: the DFT+U term sum_pq <q_p|c> U_pq q_q equals
    the KB contraction with the projector sets stacked and D block-diagonal, and
    matches the dense matrix q^T U conj(q) added to H."""
    s = S.si2_c1()
    sel = [0, 5]
    bk = BatchedK(npw=s.bk.npw[sel], mask=s.bk.mask[sel], flat_idx=s.bk.flat_idx[sel], kpg=s.bk.kpg[sel],
                  t=s.bk.t[sel], proj_phase_free=s.bk.proj_phase_free[sel], proj_atom_index=s.bk.proj_atom_index,
                  dij_full=s.bk.dij_full)
    g = torch.Generator().manual_seed(12)
    nq = 5
    hub_q = torch.randn(2, nq, bk.npw_max, dtype=torch.complex128, generator=g) * bk.mask[:, None, :]
    a = torch.randn(nq, nq, dtype=torch.float64, generator=g)
    hub_d = a + a.T
    h = BatchedHamiltonian(bk, s.shape, s.v_r, s.p[sel], hub_q=hub_q, hub_dij=hub_d)
    c = torch.randn(2, 3, bk.npw_max, dtype=torch.complex128, generator=g) * bk.mask[:, None, :]
    out = h.apply(c)
    stacked = BatchedK(**{**bk.__dict__, "dij_full": torch.block_diag(bk.dij_full, hub_d)})
    h2 = BatchedHamiltonian(stacked, s.shape, s.v_r, torch.cat([s.p[sel], hub_q], dim=1))
    assert float((out - h2.apply(c)).abs().max()) < 1e-12 * float(out.abs().max())
    for j, ik in enumerate(sel):
        n = int(bk.npw[j])
        hq = hub_q[j, :, :n]
        hd = dense_h(s, ik) + hq.T @ hub_d.to(torch.complex128).T @ hq.conj()
        want = c[j, :, :n] @ hd.T
        assert float((out[j, :, :n] - want).abs().max()) < 1e-10 * float(want.abs().max())
    cols = h.apply_cols(c[[1, 0], [2, 1]], torch.tensor([1, 0]))
    assert float((cols - out[[1, 0], [2, 1]]).abs().max()) < 1e-12 * float(out.abs().max())
