"""GPU checks of the graph-captured Davidson rounds (gradwave_amd.graph_davidson):
replayed rounds against the same rounds run eagerly, and the drafted solve
refined by one FP64 sweep against the lockstep FP64 solve of the same cells.
"""

from __future__ import annotations

import pytest
import torch

from tests.gpu_cases import dev, say

pytestmark = pytest.mark.gpu

HARTREE_EV = 27.211386245988
_CACHE = {}


def _setup():
    if "wls" not in _CACHE:
        from gradwave_amd import batching as B
        from gradwave_amd.hamiltonian import BatchedHamiltonian
        from gradwave_amd.workloads import start_vectors

        wls = B.ragged_instances(device=dev(), n=8, seed=1)
        hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in wls]
        x0s = [start_vectors(w, seed=i) for i, w in enumerate(wls)]
        _CACHE["wls"] = (wls, hams, B.RaggedBatch(hams), x0s)
    return _CACHE["wls"]


def test_replayed_rounds_match_eager_rounds():
    """Every replay of the three captured rounds must reproduce the eager
    rounds: same kernels, same inputs, so the drafted Ritz values agree to the
    last bits (and the cells converge in the same round)."""
    from gradwave_amd import graph_davidson as GD

    wls, _, batch, x0s = _setup()
    ts, ms = [w.bk.t for w in wls], [w.bk.mask for w in wls]
    say("graph davidson: eager rounds")
    eager = GD.davidson_ragged_graph(batch, x0s, ts, ms, tol=1e-3, max_iter=60, capture=False)
    st = GD.GraphSolveStats()
    say("graph davidson: captured rounds")
    solver = GD.RaggedGraphDavidson(batch, ts, ms, [int(x.shape[1]) for x in x0s])
    graph = GD.davidson_ragged_graph(batch, x0s, ts, ms, tol=1e-3, max_iter=60, stats=st, solver=solver)
    assert st.captured_graphs == 3 and st.rounds > 4
    # a second solve on the same solver only replays (new start vectors: the seeds shifted)
    from gradwave_amd.workloads import start_vectors

    wls2 = _setup()[0]
    x1s = [start_vectors(w, seed=100 + i) for i, w in enumerate(wls2)]
    eager2 = GD.davidson_ragged_graph(batch, x1s, ts, ms, tol=1e-3, max_iter=60, capture=False)
    st2 = GD.GraphSolveStats()
    graph2 = GD.davidson_ragged_graph(batch, x1s, ts, ms, tol=1e-3, max_iter=60, stats=st2, solver=solver)
    assert st2.captured_graphs == 3
    for e, g in ((eager, graph), (eager2, graph2)):
        for i, (a, b) in enumerate(zip(e, g)):
            assert a.n_iter == b.n_iter, (i, a.n_iter, b.n_iter)
            assert float((a.eigenvalues - b.eigenvalues).abs().max()) < 1e-12, i
            assert float(b.residual_norms.max()) < 1e-3, i


def test_graph_draft_plus_fp64_sweep_matches_lockstep_fp64():
    """Draft (graph rounds) + one FP64 sweep against the converged lockstep FP64
    solve: eigenvalues within the 1e-6 Ha bound and not below it (Ritz values of
    a subspace are upper bounds), FP64 residual norms re-measured."""
    from gradwave_amd import batching as B
    from gradwave_amd import graph_davidson as GD

    wls, hams, batch, x0s = _setup()
    ts, ms = [w.bk.t for w in wls], [w.bk.mask for w in wls]
    tol = 1e-10 * HARTREE_EV
    say("graph davidson: lockstep FP64 reference")
    ref = B.davidson_ragged(batch, x0s, ts, ms, tol=tol, max_iter=200)
    say("graph davidson: draft + sweep")
    res = GD.davidson_ragged_graph_mixed(batch, x0s, ts, ms, draft_tol=1e-3, sweeps=1)
    for i, (w, h, r, g) in enumerate(zip(wls, hams, ref, res)):
        assert g.eigenvectors.dtype == torch.complex128 and g.n_iter == 1
        assert float((g.eigenvalues - r.eigenvalues).min()) > -1e-12 * HARTREE_EV, (i, w.name)
        assert float((g.eigenvalues - r.eigenvalues).abs().max()) / HARTREE_EV < 1e-6, (i, w.name)
        hx = h.apply(g.eigenvectors)
        rn = torch.linalg.norm(hx - g.eigenvalues[..., None] * g.eigenvectors, dim=-1)
        assert float((rn - g.residual_norms).abs().max()) < 1e-10, (i, w.name)


def test_cells_too_small_for_fixed_rounds_raise():
    from gradwave_amd import graph_davidson as GD

    wls, _, batch, x0s = _setup()
    tiny = torch.zeros_like(wls[0].bk.mask)
    tiny[:, :3] = True
    with pytest.raises(ValueError, match="plane waves"):
        GD.davidson_ragged_graph(batch, [x0s[0]], [wls[0].bk.t], [tiny])
