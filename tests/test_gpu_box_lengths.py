"""GPU parity on boxes other than the cubic ones of the workloads.

build_fft_grid (grids.py:44-60) gives a non-cubic box for a non-cubic cell or
an explicit shape, and good_fft_size (grids.py:32-41) returns every 7-smooth
length. Two-codelet lengths run the generic box FFT kernels on any box shape;
a box with a length that has no such split (125, 162, ..., every length above
256) takes its local term through the Toeplitz contraction or, with Toeplitz
off or too large, the library FFT between the plan's scatter and gather. The
apply matches the oracle's (core/batch.py:405-505) to 1e-12 (FP64) and 2e-5
(FP32) relative, and the Davidson eigenvalues the oracle's solve to 1e-10 Ha.
"""

import numpy as np
import pytest
import torch

from gpu_cases import dev, rand_block, ref_h, say, select_k

pytestmark = pytest.mark.gpu

from gradwave_amd import _lib as L  # noqa: E402
from gradwave_amd.hamiltonian import BatchedHamiltonian  # noqa: E402
from oracle import systems as S  # noqa: E402
from oracle.units import HARTREE_EV  # noqa: E402

_KPTS = np.array([[0.125, 0.125, 0.125], [0.375, -0.125, 0.25]])
_SYS: dict = {}


def _si2(shape):
    if shape not in _SYS:
        cell = S.fcc_primitive(S.A_SI)
        pos = np.array([[0.0, 0.0, 0.0], [0.25 * S.A_SI] * 3])
        _SYS[shape] = S.build_system(f"si2_{shape}", cell, pos, [S.SI, S.SI], 30.0, _KPTS, np.full(2, 0.5), 8,
                                     shape=shape)
    return _SYS[shape]


def _gpu_h(s, toeplitz=None, flags=0):
    cpu, gpu, p = select_k(s)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), toeplitz=toeplitz)
    if flags:
        h.set_flags(flags)
    return cpu, gpu, h


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def _check_apply(s, h, cpu, seed):
    c = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=seed)
    c_pad = c + 5.0 * (~cpu.mask[:, None, :])  # padded input must be ignored
    want = ref_h(s).apply(c)
    got = h.apply(c_pad.to(dev())).cpu()
    assert _rel(got, want) < 1e-12
    assert float(got[~cpu.mask[:, None, :].expand_as(got)].abs().max()) == 0.0
    got32 = h.apply(c_pad.to(torch.complex64).to(dev())).cpu().to(torch.complex128)
    assert _rel(got32, want) < 2e-5


@pytest.mark.parametrize("shape,flags", [((24, 30, 36), 0), ((36, 24, 30), 0), ((30, 36, 24), 0),
                                         ((24, 30, 36), L.GW_PLAN_FORCE_UNFUSED | L.GW_PLAN_LEGACY),
                                         ((24, 120, 96), 0)])
def test_noncubic_boxes_match_oracle(shape, flags):
    """Non-cubic two-codelet boxes on the generic kernels; (24, 120, 96) has
    FP64 x-planes above the LDS budget (split stick/column kernels), its FP32
    planes fit (one plane per workgroup)."""
    say(f"non-cubic box {shape} flags {flags}")
    s = _si2(shape)
    cpu, _, h = _gpu_h(s, toeplitz="off", flags=flags)
    info = h.info()
    assert info["nosplit_len"] == 0 and h.local_path == "fft"
    _check_apply(s, h, cpu, seed=11)


@pytest.mark.parametrize("shape", [(125, 24, 30), (24, 30, 125), (27, 200, 21), (270, 24, 25)])
@pytest.mark.parametrize("path", ["toeplitz", "library_fft"])
def test_boxes_without_codelet_split_match_oracle(shape, path):
    say(f"box without codelet split {shape} via {path}")
    s = _si2(shape)
    cpu, _, h = _gpu_h(s, toeplitz="on" if path == "toeplitz" else "off")
    assert h.local_path == path and h.info()["nosplit_len"] in shape
    _check_apply(s, h, cpu, seed=12)
    # every term separately (the kinetic and nonlocal terms still run on the plan)
    c = rand_block(cpu.nk, 2, cpu.npw_max, cpu.mask, seed=13)
    want_all = ref_h(s).apply(c)
    parts = sum(h.apply_terms(c.to(dev()), t).cpu() for t in (L.GW_TERM_KINETIC, L.GW_TERM_LOCAL,
                                                               L.GW_TERM_NONLOCAL))
    assert _rel(parts, want_all) < 1e-12


def test_box_without_codelet_split_real_space_and_refusals():
    """psi(r) (core/batch.py:208-216) and density_b (batch.py:562-587) on a
    125 x 24 x 30 box through the plan's scatter and rocFFT, against the
    oracle's; the V(r) adjoint and ragged batches refuse the box."""
    import ctypes as C

    from gradwave_amd import batching as B
    from gradwave_amd.hamiltonian import density_b
    from oracle.hamiltonian import density_b as ref_density
    from oracle.hamiltonian import g_to_r

    say("psi(r), density and refusals on a 125x24x30 box")
    s = _si2((125, 24, 30))
    cpu, gpu, h = _gpu_h(s)
    c = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=14)
    psi = h.wfc_to_real(c.to(dev())).cpu()
    for k in range(cpu.nk):
        n = int(cpu.npw[k])
        assert _rel(psi[k], g_to_r(c[k, :, :n], cpu.flat_idx[k, :n], s.shape)) < 1e-12
    g = torch.Generator().manual_seed(2)
    occ = torch.rand(cpu.nk, 3, dtype=torch.float64, generator=g)
    kw = torch.rand(cpu.nk, dtype=torch.float64, generator=g)
    want = ref_density(c, occ, kw, cpu, s.shape, s.grid.volume)
    got = density_b(c.to(dev()), occ.to(dev()), kw.to(dev()), gpu, s.shape, s.grid.volume).cpu()
    assert _rel(got, want) < 1e-12
    got32 = density_b(c.to(torch.complex64).to(dev()), occ.to(dev()), kw.to(dev()), gpu, s.shape,
                      s.grid.volume).cpu()
    assert _rel(got32, want) < 2e-5
    with pytest.raises(L.GWError, match="ragged batch"):
        B.RaggedBatch([h])
    cd = c.to(dev())
    gv = torch.zeros(s.shape, dtype=torch.float64, device=dev())
    ws = torch.empty(256, dtype=torch.uint8, device=dev())
    rc = L.lib().gw_hpsi_adjoint_v(h.plan.handle, L.GW_FP64, cd.data_ptr(), cd.stride(1), cd.stride(0),
                                   cd.data_ptr(), cd.stride(1), cd.stride(0), 3, gv.data_ptr(), ws.data_ptr(),
                                   C.c_size_t(ws.numel()), L.stream_ptr(dev()))
    assert rc != 0 and "box FFT kernels" in L.lib().gw_last_error().decode()


@pytest.mark.parametrize("path", ["toeplitz", "library_fft"])
def test_davidson_on_box_without_codelet_split_matches_oracle(path):
    """FP64 Davidson (solvers/davidson.py:527-774) on a 125 x 24 x 30 box, both
    k-points, 8 bands, against the oracle's solve from the same start."""
    from gradwave_amd.davidson import davidson_batched
    from oracle.davidson import davidson_batched as ref_davidson

    say(f"davidson on a 125x24x30 box via {path}")
    tol = 1e-10 * HARTREE_EV
    s = _si2((125, 24, 30))
    cpu, gpu, h = _gpu_h(s, toeplitz="on" if path == "toeplitz" else "off")
    x0 = rand_block(cpu.nk, 8, cpu.npw_max, cpu.mask, seed=15)
    r = davidson_batched(h.apply, x0.to(dev()), gpu.t, gpu.mask, tol=tol, max_iter=100)
    assert float(r.residual_norms.max()) < tol
    rr = ref_davidson(ref_h(s).apply, x0, cpu.t, cpu.mask, tol=tol, max_iter=100)
    assert float((r.eigenvalues.cpu() - rr.eigenvalues).abs().max()) < 1e-10 * HARTREE_EV


def test_noncubic_cells_in_a_ragged_batch():
    """The ragged kernels share the generic bodies: non-cubic boxes applied as
    one batch, against the oracle per cell."""
    from gradwave_amd import batching as B

    say("non-cubic boxes in one ragged batch")
    shapes = [(24, 30, 36), (36, 24, 30), (30, 36, 24)]
    hams, cs, wants = [], [], []
    for i, sh in enumerate(shapes):
        s = _si2(sh)
        cpu, _, h = _gpu_h(s, toeplitz="off")
        hams.append(h)
        c = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=20 + i)
        cs.append(c)
        wants.append(ref_h(s).apply(c))
    batch = B.RaggedBatch(hams)
    for dtype, tol in ((torch.complex128, 1e-12), (torch.complex64, 2e-5)):
        outs = batch.apply([c.to(dtype).to(dev()) for c in cs])
        for o, w in zip(outs, wants):
            assert _rel(o.cpu().to(torch.complex128), w) < tol


@pytest.mark.parametrize("shape", [(24, 24, 24), (24, 30, 36), (24, 120, 96), (120, 120, 120)])
def test_pencil_path_matches_oracle(shape):
    """GW_PLAN_PENCIL: the local term as z-stick transforms around one (y, x)
    slab kernel per (z, band), against the oracle's apply (FP64 1e-12, FP32
    2e-5) and against the plan's default path."""
    say(f"pencil path {shape}")
    s = _si2(shape)
    cpu, _, h = _gpu_h(s, toeplitz="off")
    c = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=31)
    base = h.apply(c.to(dev())).cpu()
    h.set_flags(L.GW_PLAN_PENCIL)
    _check_apply(s, h, cpu, seed=31)
    assert _rel(h.apply(c.to(dev())).cpu(), base) < 1e-12


@pytest.mark.parametrize("name,nb", [("si8", 40), ("si64", 24)])
def test_pencil_path_matches_default_path(name, nb):
    """Many bands (band chunks, every (z, band) work item) on the 36^3 and
    90^3 workloads: the pencil path equals the wave-line path to 1e-12 (FP64)
    and 2e-6 (FP32), and the oracle on three bands."""
    from gpu_cases import system

    say(f"pencil path {name} nb={nb}")
    s = system(name)
    sel = [0, 2] if name == "si8" else None
    cpu, gpu, p = select_k(s, sel)
    h = BatchedHamiltonian(gpu, s.shape, s.v_r.to(dev()), p.to(dev()), toeplitz="off")
    c = rand_block(cpu.npw.shape[0], nb, cpu.mask.shape[1], cpu.mask, seed=32).to(dev())
    base64, base32 = h.apply(c).cpu(), h.apply(c.to(torch.complex64)).cpu()
    h.set_flags(L.GW_PLAN_PENCIL)
    assert _rel(h.apply(c).cpu(), base64) < 1e-12
    assert _rel(h.apply(c.to(torch.complex64)).cpu(), base32) < 2e-6
    want = ref_h(s, sel).apply(c[:, :3].cpu())
    assert _rel(h.apply(c[:, :3].contiguous()).cpu(), want) < 1e-12


@pytest.mark.parametrize("shape", [(24, 30, 36), (125, 24, 30)])
def test_real_space_v_adjoint_matches_kernel_and_oracle_autograd(shape):
    """The V(r) adjoint on boxes without the wave-line adjoint kernel (non-cubic,
    or a length without a codelet split): psi(r) of both blocks and a float64
    sum. It equals the HIP adjoint kernel on a cubic 24^3 box (which pins its
    scaling), and the gradient of Re sum <w|H c> in V(r) equals the oracle's
    autograd to 1e-10 on the given box."""
    from types import SimpleNamespace

    from gradwave_amd import adjoint as A
    from oracle.hamiltonian import HamiltonianK, build_projector_data

    say(f"real-space V(r) adjoint {shape}")
    d = dev()
    s = _si2((24, 24, 24))
    cpu, _, h = _gpu_h(s, toeplitz="off")
    c = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=40).to(d)
    g = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=41).to(d)
    op = SimpleNamespace(fwd=h, shape=s.shape)
    want = A.adjoint_v(op, c, g)
    got = A._adjoint_v_real_space(op, c, g, torch.zeros_like(want))
    assert _rel(got.cpu(), want.cpu()) < 1e-12

    s = _si2(shape)
    cpu, gpu, p = select_k(s)
    gpu.kpg = cpu.kpg.to(d)
    gpu.proj_phase_free = cpu.proj_phase_free.to(d)
    gpu.proj_atom_index = cpu.proj_atom_index.to(d)
    c = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=42)
    w = rand_block(cpu.nk, 3, cpu.npw_max, cpu.mask, seed=43)
    v = s.v_r.to(d).requires_grad_(True)
    hd = A.DifferentiableHamiltonian(gpu, s.shape, v, p.to(d), cpu.dij_full.to(d))
    loss = (w.to(d).conj() * hd.apply(c.to(d))).real.sum()
    (gv,) = torch.autograd.grad(loss, [v])
    vr = s.v_r.clone().requires_grad_(True)
    ref = 0.0
    for j in range(cpu.nk):
        n = int(cpu.npw[j])
        pd = build_projector_data(s.spheres[j], s.species, s.grid.volume)
        pd.dij_full = cpu.dij_full
        hk = HamiltonianK(s.spheres[j], s.shape, vr, pd, p[j, :, :n])
        ref = ref + (w[j, :, :n].conj() * hk.apply(c[j, :, :n])).real.sum()
    (gv_ref,) = torch.autograd.grad(ref, [vr])
    assert abs(float(loss.detach()) - float(ref.detach())) < 1e-11 * abs(float(ref.detach()))
    assert _rel(gv.cpu(), gv_ref) < 1e-10
