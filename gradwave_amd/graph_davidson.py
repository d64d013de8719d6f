"""Graph-captured Davidson rounds for a ragged batch of cells (the batching form).

The lockstep solver (``batching.davidson_ragged``) advances 32 Python step
generators and launches every cell's small products one by one; its rounds
are shaped by host reads (convergence gate, breakdown flags). Here the draft
runs as fixed-shape rounds captured once as HIP graphs and replayed:

* every round expands the basis with ALL nb preconditioned residuals (the
  reference's sync-free skeleton, solvers/davidson.py:545-554 and 691-716,
  with its first-round expansion count nb kept for good), so the basis of
  every cell cycles through 2nb -> 3nb -> 4nb -> restart, and three graphs
  (one per basis size) hold the whole round of every cell;
* orthonormalisation follows the reference's sync-free path
  (davidson.py:406-457 with ``jitter``): a blanket 1e-10 jitter from a
  pre-drawn block (drawn on the device, seed nb + 7919), two projections,
  CholQR2 -- no degenerate-row test, no host read;
* the subspace eigensolve, the Cholesky inverses and the FP32 apply are one
  launch each for all cells (``gw_heev_batched``, ``gw_chol_inv_batched``, a
  bound ragged apply); each cell's small products run on one of four side
  streams between them, so a captured round has four parallel branches;
* ``RaggedGraphDavidson`` keeps the cells' buffers and the captured graphs:
  its first ``solve`` captures, every later solve of the same batch only
  replays;
* after each replay the host reads one status vector (max residual norm and
  factorisation breakdown per cell); a cell whose residuals are all below the
  tolerance keeps the Ritz block of that round (at least one refinement past
  convergence, as in the reference's delayed gate).

The draft is complex64 with FP64 subspace algebra (the precision form's
draft); ``davidson_ragged_graph_mixed`` refines it with FP64 Davidson sweeps
in lockstep, as ``batching.davidson_ragged_mixed(..., sweeps=n)`` does.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import torch

from gradwave_amd import linalg as LA
from gradwave_amd.batching import RaggedBatch, _lockstep
from gradwave_amd.davidson import BatchedDavidsonResult, _gram, _hermitize, _refine_steps

LOW = torch.complex64
JITTER = 1e-10


@dataclass
class GraphSolveStats:
    rounds: int = 0          # rounds run (eager and replayed)
    captured_graphs: int = 0
    host_reads: int = 0
    seconds: dict | None = None  # wall time per stage (setup, eager, capture, replay), synchronised


def _mark(st: GraphSolveStats, key: str, t0: float) -> float:
    if st.seconds is None:
        return t0
    torch.cuda.synchronize()
    now = time.perf_counter()
    st.seconds[key] = st.seconds.get(key, 0.0) + now - t0
    return now


class _Cell:
    """Static device state of one cell: basis, images, subspace matrix, Ritz block."""

    def __init__(self, nb: int, t: torch.Tensor, mask: torch.Tensor) -> None:
        nk, m = int(mask.shape[0]), int(mask.shape[1])
        if int(mask.sum(dim=1).min()) < 4 * nb:
            raise ValueError(f"graph rounds need >= 4 nb = {4 * nb} plane waves per k, got "
                             f"{int(mask.sum(dim=1).min())}")
        dev = mask.device
        self.nk, self.nb, self.m = nk, nb, m
        self.t = t.to(torch.float64).contiguous()
        self.mask = mask
        self.maskc = mask[:, None, :].to(LOW)
        self.v = torch.zeros(nk, 4 * nb, m, dtype=LOW, device=dev)
        self.hv = torch.zeros_like(self.v)
        self.s = torch.zeros(nk, 4 * nb, 4 * nb, dtype=torch.complex128, device=dev)
        self.x = torch.zeros(nk, nb, m, dtype=LOW, device=dev)
        self.hx = torch.zeros_like(self.x)
        self.eig = torch.zeros(nk, nb, dtype=torch.float64, device=dev)
        self.rn = torch.zeros(nk, nb, dtype=torch.float64, device=dev)
        self.d = torch.zeros(nk, nb, m, dtype=LOW, device=dev)   # expansion block (apply input)
        self.hd = torch.zeros_like(self.d)                        # its image (apply output)
        self.sel = torch.arange(nb, device=dev).expand(nk, nb).contiguous()
        g = torch.Generator(device=dev).manual_seed(nb + 7919)
        noise = torch.randn(nk, nb, m, 2, generator=g, dtype=torch.float32, device=dev)
        self.jit = torch.view_as_complex(noise) * mask[:, None, :]
        self.bad = torch.zeros(nk, dtype=torch.int32, device=dev)
        self.gram = None  # Gram matrix handed to the next batched Cholesky inverse


def _project(cell: _Cell, d: torch.Tensor, n: int) -> None:
    """d -= (d V^H) V twice against the first n basis rows (in place)."""
    a = cell.v[:, :n]
    for _ in range(2):
        coef = LA.inner(a, d)
        LA.combine((-coef).contiguous(), a, out=d, beta=1.0)


class _Lanes:
    """Side streams for the per-cell segments of a round: cell i runs on lane
    i % k between a fork from and a join to the launching stream, so a captured
    round has k parallel branches of small kernels."""

    def __init__(self, k: int, device) -> None:
        self.streams = [torch.cuda.Stream(device=device) for _ in range(k)]  # default priority

    def run(self, cells: list, fn) -> None:
        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        for i, c in enumerate(cells):
            with torch.cuda.stream(self.streams[i % len(self.streams)]):
                fn(i, c)
        for s in self.streams:
            main.wait_stream(s)


def _round(cells: list, binding, lanes: _Lanes, n_of: list, restart: bool, status: torch.Tensor) -> None:
    """One Davidson round of every cell at basis size n_of[i]: Rayleigh-Ritz,
    residuals, expansion of all nb bands (restart on the repaired Ritz block
    first when ``restart``), orthonormalisation, FP32 apply, subspace update.
    Issues GPU work only (no host read), so it can be captured; the batched
    launches (eigensolve, Cholesky inverses, apply) run on the launching
    stream, the per-cell products on the lanes."""
    eighs = LA.eigh_lowest([c.s[:, :n, :n] for c, n in zip(cells, n_of)], [c.nb for c in cells])

    def ritz(i, c):
        n, (w, u) = n_of[i], eighs[i]
        uu = u.to(LOW).contiguous()
        LA.combine(uu, c.v[:, :n], out=c.x)
        LA.combine(uu, c.hv[:, :n], out=c.hx)
        c.eig.copy_(w)
        r, rn = LA.residual(c.x, c.hx, c.eig)
        c.rn.copy_(rn)
        LA.teter_expand(c.x, r, c.t, c.sel, out=c.d)
        status[i, 0] = rn.max()
        if restart:
            c.gram = _gram(c.x)

    lanes.run(cells, ritz)
    if restart:  # basis = Ritz block with its drift repaired (x = L q, hx -> L^-1 hx)
        inv = LA.cholesky_inverse([c.gram for c in cells])

        def repair(i, c):
            linv, info = inv[i]
            ut = linv.transpose(-1, -2).to(LOW).contiguous()
            LA.combine(ut, c.x, out=c.v[:, :c.nb])
            LA.combine(ut, c.hx, out=c.hv[:, :c.nb])
            c.s[:, :c.nb, :c.nb] = _hermitize(LA.inner(c.v[:, :c.nb], c.hv[:, :c.nb]))
            torch.maximum(c.bad, info, out=c.bad)

        lanes.run(cells, repair)
        base = [c.nb for c in cells]
    else:
        base = list(n_of)

    def project(i, c):
        c.d.add_(c.jit, alpha=JITTER)
        c.d.mul_(c.maskc)
        _project(c, c.d, base[i])
        c.gram = _gram(c.d)

    lanes.run(cells, project)
    for rep in range(2):  # CholQR2: two rounds, each with one batched inverse
        inv = LA.cholesky_inverse([c.gram for c in cells])

        def qr(i, c, inv=inv, last=rep == 1):
            linv, info = inv[i]
            LA.combine(linv.transpose(-1, -2).to(LOW).contiguous(), c.d.clone(), out=c.d)
            torch.maximum(c.bad, info, out=c.bad)
            if not last:
                c.gram = _gram(c.d)

        lanes.run(cells, qr)
    binding.apply()  # hd = H d for every cell, one launch per pass

    def append(i, c):
        n = base[i]
        n1 = n + c.nb
        c.v[:, n:n1].copy_(c.d)
        c.hv[:, n:n1].copy_(c.hd)
        col = LA.inner(c.v[:, :n1], c.hd).to(torch.complex128)
        c.s[:, :n, n:n1] = col[:, :n]
        c.s[:, n:n1, :n] = col[:, :n].mH
        c.s[:, n:n1, n:n1] = _hermitize(col[:, n:])
        status[i, 1] = c.bad.max().to(torch.float64)

    lanes.run(cells, append)


class RaggedGraphDavidson:
    """Graph-captured draft solver for one ragged batch (fixed cells and band
    counts): the first ``solve`` runs the cycle's three rounds eagerly once and
    captures them; every later ``solve`` (new start vectors, same shapes) only
    replays them ("decode-style": capture once, replay per solve)."""

    def __init__(self, batch: RaggedBatch, ts: list, masks: list, nbs: list, lanes: int = 4) -> None:
        self.batch = batch
        dev = masks[0].device
        self.cells = [_Cell(int(nb), t, m) for t, m, nb in zip(ts, masks, nbs)]
        self.status = torch.zeros(len(self.cells), 2, dtype=torch.float64, device=dev)
        self.binding = batch.bind([c.d for c in self.cells], [c.hd for c in self.cells])
        self.lanes = _Lanes(lanes, dev)
        self.graphs: list = []
        self.shapes = [([2 * c.nb for c in self.cells], False), ([3 * c.nb for c in self.cells], False),
                       ([4 * c.nb for c in self.cells], True)]

    def _start(self, x0s: list) -> None:
        """V = orthonormalised x0 (jittered), HV = H V, S = V^H H V; flags cleared."""
        cells = self.cells
        for c, x0 in zip(cells, x0s):
            c.d.copy_(x0.to(LOW))
            c.d.add_(c.jit, alpha=JITTER)
            c.d.mul_(c.maskc)
            c.bad.zero_()
        for _ in range(2):
            for c, (linv, info) in zip(cells, LA.cholesky_inverse([_gram(c.d) for c in cells])):
                LA.combine(linv.transpose(-1, -2).to(LOW).contiguous(), c.d.clone(), out=c.d)
                torch.maximum(c.bad, info, out=c.bad)
        self.binding.apply()
        for c in cells:
            c.v[:, :c.nb].copy_(c.d)
            c.hv[:, :c.nb].copy_(c.hd)
            c.s[:, :c.nb, :c.nb] = _hermitize(LA.inner(c.v[:, :c.nb], c.hv[:, :c.nb]))

    def solve(self, x0s: list, tol: float = 1e-4, max_iter: int = 100,
              stats: GraphSolveStats | None = None, capture: bool = True) -> list:
        st = stats if stats is not None else GraphSolveStats()
        cells, status = self.cells, self.status
        t0 = time.perf_counter()
        self._start(x0s)
        results: list = [None] * len(cells)

        def harvest(it: int) -> bool:
            flags = status.cpu()
            st.host_reads += 1
            if bool((flags[:, 1] > 0).any()):
                raise FloatingPointError("CholQR factorisation broke down in a graph round")
            for i, c in enumerate(cells):
                if results[i] is None and float(flags[i, 0]) < tol:
                    results[i] = BatchedDavidsonResult(c.eig.clone(), c.x.clone(), it, c.rn.clone(), 0, 0)
            return all(r is not None for r in results)

        def finish(it: int) -> list:
            for i, c in enumerate(cells):
                if results[i] is None:
                    results[i] = BatchedDavidsonResult(c.eig.clone(), c.x.clone(), it, c.rn.clone(), 0, 0)
            return results

        t0 = _mark(st, "setup", t0)
        # first round (basis nb -> 2nb) runs once per solve, eagerly
        _round(cells, self.binding, self.lanes, [c.nb for c in cells], False, status)
        st.rounds += 1
        it = 1
        if harvest(it) or it >= max_iter:
            return finish(it)
        if not self.graphs:
            # the cycle's rounds eagerly once (warms every workspace), then captured
            for n_of, restart in self.shapes:
                it += 1
                _round(cells, self.binding, self.lanes, n_of, restart, status)
                st.rounds += 1
                if harvest(it) or it >= max_iter:
                    _mark(st, "eager", t0)
                    return finish(it)
            t0 = _mark(st, "eager", t0)
            if not capture:
                while True:
                    for n_of, restart in self.shapes:
                        it += 1
                        _round(cells, self.binding, self.lanes, n_of, restart, status)
                        st.rounds += 1
                        if harvest(it) or it >= max_iter:
                            return finish(it)
            torch.cuda.synchronize()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for n_of, restart in self.shapes:
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=side):
                        _round(cells, self.binding, self.lanes, n_of, restart, status)
                    self.graphs.append(g)
            torch.cuda.current_stream().wait_stream(side)
            # capture records without running: the cycle starts again at the
            # 2nb round, which matches the basis the eager 4nb round left
            t0 = _mark(st, "capture", t0)
        st.captured_graphs = len(self.graphs)
        while it < max_iter:
            for g in self.graphs:
                it += 1
                g.replay()
                st.rounds += 1
                if harvest(it) or it >= max_iter:
                    _mark(st, "replay", t0)
                    return finish(it)
        _mark(st, "replay", t0)
        return finish(it)


def davidson_ragged_graph(batch: RaggedBatch, x0s: list, ts: list, masks: list, tol: float = 1e-4,
                          max_iter: int = 100, stats: GraphSolveStats | None = None,
                          capture: bool = True, solver: RaggedGraphDavidson | None = None) -> list:
    """complex64 draft of every cell to ``tol`` (eV, on every residual norm) in
    graph-captured fixed-shape rounds (module docstring). Returns one
    BatchedDavidsonResult per cell (complex64 vectors, float64 eigenvalues and
    FP32-sourced residual norms); raises FloatingPointError if a CholQR
    factorisation broke down. Cells need at least 4 nb plane waves per k.
    ``capture=False`` runs the same rounds eagerly (the replay's reference);
    pass a ``RaggedGraphDavidson`` as ``solver`` to reuse its captured graphs."""
    if solver is None:
        solver = RaggedGraphDavidson(batch, ts, masks, [int(x.shape[1]) for x in x0s])
    return solver.solve(x0s, tol=tol, max_iter=max_iter, stats=stats, capture=capture)


def davidson_ragged_graph_mixed(batch: RaggedBatch, x0s: list, ts: list, masks: list,
                                draft_tol: float = 1e-3, sweeps: int = 1, max_iter: int = 100,
                                stats: GraphSolveStats | None = None, capture: bool = True,
                                solver: RaggedGraphDavidson | None = None) -> list:
    """The batching form's solve: graph-captured complex64 draft of every cell
    (``davidson_ragged_graph``), then ``sweeps`` FP64 Davidson sweeps per cell in
    lockstep from the drafted vectors (``_refine_steps``: Rayleigh-Ritz with the
    FP64 operator, every band's residual appended, Rayleigh-Ritz again), whose
    FP64 residual norms the results carry."""
    if sweeps < 0:
        raise ValueError(f"sweeps must be >= 0, got {sweeps}")
    d = davidson_ragged_graph(batch, x0s, ts, masks, tol=draft_tol, max_iter=max_iter, stats=stats,
                              capture=capture, solver=solver)
    t0 = time.perf_counter()
    gens = [_refine_steps(r.eigenvectors.to(x0.dtype), t, m, int(sweeps)) for r, x0, t, m in zip(d, x0s, ts, masks)]
    p = _lockstep(batch, gens, None, False, "kernel")
    if stats is not None:
        _mark(stats, "fp64_sweeps", t0)
    for pi, di in zip(p, d):
        pi.n_draft_iter = di.n_iter
    return p
