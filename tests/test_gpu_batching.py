"""GPU parity of the ragged batch (gradwave_amd.batching): one launch per local
pass over many independent cells against each operator applied alone, the HIP
graph replay against eager applies, and lockstep Davidson against independent
solves. Tolerances: apply 1e-12 relative (fp64), 2e-5 (fp32); eigenvalues 1e-10 Ha.
"""

from __future__ import annotations

import pytest
import torch

from tests.gpu_cases import dev, say

pytestmark = pytest.mark.gpu

HARTREE_EV = 27.211386245988
_CACHE = {}


def _instances(n=8):
    from gradwave_amd import batching as B

    if n not in _CACHE:
        _CACHE[n] = B.ragged_instances(device=dev(), n=n, seed=1)
    return _CACHE[n]


def _hams(wls):
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    return [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in wls]


def _blocks(wls, dtype, seed=0, nbs=None):
    g = torch.Generator().manual_seed(seed)
    out = []
    for i, w in enumerate(wls):
        nb = w.nbands if nbs is None else nbs[i]
        if nb == 0:
            out.append(None)
            continue
        x = torch.randn(w.bk.nk, nb, w.bk.npw_max, dtype=torch.complex128, generator=g)
        out.append((x.to(dev()) * w.bk.mask[:, None, :]).to(dtype))
    return out


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("library", [0, 1])
@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-12), (torch.complex64, 2e-5)])
def test_ragged_apply_matches_single_operators(dtype, tol, library):
    """The batched apply against each operator on its own. With the kernels here
    the cells' projector contractions run as one grouped launch each; with the
    rocBLAS contractions selected, as per-cell chains over four streams."""
    from gradwave_amd import _lib as L
    from gradwave_amd import batching as B

    wls = _instances()
    say(f"ragged apply library={library} " + ", ".join(w.name for w in wls))
    hams = _hams(wls)
    batch = B.RaggedBatch(hams)
    nbs = [w.nbands for w in wls]
    nbs[3] = 0  # one instance sits the call out
    nbs[5] = 3
    cs = _blocks(wls, dtype, nbs=nbs)
    lib = L.lib()
    prec = L.precision_of(dtype)
    try:
        L.check(lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, library), "gw_tune_set")
        L.check(lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, library), "gw_tune_set")
        outs = batch.apply(cs)
    finally:
        L.check(lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, 0), "gw_tune_set")
        L.check(lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, 0), "gw_tune_set")
    assert outs[3] is None
    for i, (h, c, o) in enumerate(zip(hams, cs, outs)):
        if c is None:
            continue
        want = h.apply_terms(c.to(torch.complex128)).to(torch.complex128)
        assert _rel(o.to(torch.complex128), want) < tol, (i, wls[i].name)


def test_ragged_graph_replay_matches_eager():
    """A captured batched apply keeps replaying correctly while eager applies
    with other band counts, and a second capture, run through the same batch in
    between (each graph owns its tables and workspace)."""
    from gradwave_amd import batching as B

    wls = _instances()
    say("ragged graph replay with eager calls through the same batch in between")
    hams = _hams(wls)
    batch = B.RaggedBatch(hams)
    cs = _blocks(wls, torch.complex128, seed=3)
    outs = [torch.empty_like(c) for c in cs]
    graph = batch.capture(cs, outs)
    small = _blocks(wls, torch.complex128, seed=4, nbs=[max(1, w.nbands // 3) for w in wls])
    small_outs = [torch.empty_like(c) for c in small]
    graph2 = batch.capture(small, small_outs)
    for rep in range(3):
        g = torch.Generator().manual_seed(10 + rep)
        for c, w in zip(cs, wls):
            c.copy_((torch.randn(c.shape, dtype=torch.complex128, generator=g).to(dev()) * w.bk.mask[:, None, :]))
        other = batch.apply(_blocks(wls, torch.complex128, seed=20 + rep,
                                    nbs=[(w.nbands + rep) % 5 for w in wls]))  # eager, other shapes
        graph.replay()
        graph2.replay()
        torch.cuda.synchronize()
        for h, c, o, cs2, o2 in zip(hams, cs, outs, small, small_outs):
            assert _rel(o, h.apply(c)) < 1e-13
            assert _rel(o2, h.apply(cs2)) < 1e-13
        assert len(other) == len(wls)


@pytest.mark.parametrize("subspace_eigh", ["kernel", "library"])
def test_lockstep_davidson_matches_independent_solves(subspace_eigh):
    from gradwave_amd import batching as B
    from gradwave_amd.davidson import davidson_batched
    from gradwave_amd.workloads import start_vectors

    wls = _instances()
    say(f"lockstep davidson vs independent ({subspace_eigh} subspace eigensolves)")
    hams = _hams(wls)
    batch = B.RaggedBatch(hams)
    tol = 1e-10 * HARTREE_EV
    x0s = [start_vectors(w, seed=i) for i, w in enumerate(wls)]
    stats = B.RaggedSolveStats()
    res = B.davidson_ragged(batch, x0s, [w.bk.t for w in wls], [w.bk.mask for w in wls], tol=tol, max_iter=200,
                            stats=stats, check_finite=True, subspace_eigh=subspace_eigh)
    say(f"rounds {stats.rounds} applies {stats.batched_applies} eighs {stats.batched_eighs}")
    for i, (w, h, x0, r) in enumerate(zip(wls, hams, x0s, res)):
        ref = davidson_batched(h.apply, x0, w.bk.t, w.bk.mask, tol=tol, max_iter=200)
        assert float(r.residual_norms.max()) < tol, (i, w.name)
        assert float((r.eigenvalues - ref.eigenvalues).abs().max()) / HARTREE_EV < 1e-10, (i, w.name)
        assert abs(r.n_iter - ref.n_iter) <= 2, (i, w.name, r.n_iter, ref.n_iter)
    mixed = B.davidson_ragged_mixed(batch, x0s, [w.bk.t for w in wls], [w.bk.mask for w in wls], tol=tol,
                                    draft_tol=1e-3, max_iter=200, subspace_eigh=subspace_eigh)
    for r, m in zip(res, mixed):
        assert float(m.residual_norms.max()) < tol
        assert float((m.eigenvalues - r.eigenvalues).abs().max()) / HARTREE_EV < 1e-10
    # sweep form in lockstep: each instance as its own draft + one FP64 sweep would give
    swept = B.davidson_ragged_mixed(batch, x0s, [w.bk.t for w in wls], [w.bk.mask for w in wls],
                                    draft_tol=1e-3, max_iter=200, sweeps=1, subspace_eigh=subspace_eigh)
    for i, (w, h, x0, r, sw) in enumerate(zip(wls, hams, x0s, res, swept)):
        assert sw.n_iter == 1
        assert float((sw.eigenvalues - r.eigenvalues).min()) > -1e-12 * HARTREE_EV, (i, w.name)
        assert float((sw.eigenvalues - r.eigenvalues).abs().max()) / HARTREE_EV < 1e-6, (i, w.name)
        hx = h.apply(sw.eigenvectors)
        rn = torch.linalg.norm(hx - sw.eigenvalues[..., None] * sw.eigenvectors, dim=-1)
        assert float((rn - sw.residual_norms).abs().max()) < 1e-10, (i, w.name)


def test_autotuner_choices_are_valid_and_exact(tmp_path, monkeypatch):
    from gradwave_amd import _lib as L
    from gradwave_amd import autotune
    from gradwave_amd import batching as B

    say("autotune radix + split-K on 4 cells")
    wls = _instances()[:4]
    radix = autotune.tune_radix(wls, reps=2)
    tiles, wgs = autotune.tune_gemm(wls, reps=2, allow_library=False)
    assert wgs in autotune.GEMM_TARGETS and tiles["fp64"] in autotune.GEMM_TILES
    assert tiles["fp32"] in autotune.GEMM_TILES and tiles["combine_lib"] == {"fp64": 0, "fp32": 0}
    assert set(tiles["combine"].values()) <= set(autotune.GEMM_TILES)
    assert tiles["inner_lib"] == {"fp64": 0, "fp32": 0}
    assert L.lib().gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP32 + 512) == tiles["combine"]["fp32"]
    assert L.lib().gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP64 + 256) == tiles["fp64"]
    for n, a in radix.items():
        assert a in autotune.candidate_splits(n)
    hams = _hams(wls)
    for h, w in zip(hams, wls):
        info = h.info(torch.complex64)
        for ax, n in enumerate(w.shape):
            if n in radix:
                assert info[("split_x", "split_y", "split_z")[ax]][0] == radix[n]
    cs = _blocks(wls, torch.complex128, seed=9)
    outs = B.RaggedBatch(hams).apply(cs)
    for h, c, o in zip(hams, cs, outs):
        assert _rel(o, h.apply_terms(c)) < 1e-12
    path = str(tmp_path / "tune.json")
    monkeypatch.setenv("GW_TUNE_FILE", path)
    autotune.save()
    L.lib().gw_tune_set(L.GW_TUNE_GEMM_WGS, 0, 0)
    assert autotune.load()["gemm_wgs"] == wgs
    assert L.lib().gw_tune_get(L.GW_TUNE_GEMM_WGS, 0) == wgs
    for n in radix:  # leave the default splits for later tests
        L.lib().gw_tune_set(L.GW_TUNE_RADIX, n, 0)
    L.lib().gw_tune_set(L.GW_TUNE_GEMM_WGS, 0, 0)
    L.lib().gw_tune_set(L.GW_TUNE_GEMM_TILE, L.GW_FP64, 0)
    L.lib().gw_tune_set(L.GW_TUNE_GEMM_TILE, L.GW_FP32, 0)


def test_xpass_tuner_round_trip(tmp_path, monkeypatch):
    """tune_xpass picks shapes from its list for a 90^3 and a 36^3 operator, the
    choice survives save/load, and the applies stay exact."""
    from tests.gpu_cases import select_k, system
    from gradwave_amd import _lib as L
    from gradwave_amd import autotune
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    say("autotune x-pass shapes")
    lib = L.lib()

    class _W:  # the fields tune_xpass reads
        def __init__(self, s):
            _, self.bk, p = select_k(s, [0])
            self.bk.nk, self.bk.npw_max = self.bk.mask.shape
            self.shape, self.v_r, self.p = s.shape, s.v_r.to(dev()), p.to(dev())

    wls = [_W(system("si64")), _W(system("si8"))]
    c = (torch.randn(1, 6, wls[0].bk.npw_max, dtype=torch.complex128, generator=torch.Generator().manual_seed(3))
         .to(dev()) * wls[0].bk.mask[:, None, :])
    h = BatchedHamiltonian(wls[0].bk, wls[0].shape, wls[0].v_r, wls[0].p)
    before = h.apply(c)
    try:
        got = autotune.tune_xpass(wls, nb=8, reps=2)
        assert got["fp64"] in autotune.XPASS_SHAPES and got["fp32"] in autotune.XPASS_SHAPES
        assert lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP32) == got["fp32"]
        assert torch.equal(h.apply(c), before)
        monkeypatch.setenv("GW_TUNE_FILE", str(tmp_path / "tune.json"))
        autotune.save()
        lib.gw_tune_set(L.GW_TUNE_XPASS, 0, 208)
        assert autotune.load()["xpass"] == got
        assert lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP64) == got["fp64"]
    finally:
        lib.gw_tune_set(L.GW_TUNE_XPASS, L.GW_FP64, 408)
        lib.gw_tune_set(L.GW_TUNE_XPASS, L.GW_FP32, 10416)
