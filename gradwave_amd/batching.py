"""Ragged batches of independent plane-wave problems on MI355X (the batching form).

The reference solves one system at a time: one ``BatchedHamiltonian`` per
cell (core/batch.py:238-505) and one ``davidson_batched`` call per solve
(solvers/davidson.py:527-774), so many small cells leave a large GPU idle
between tiny launches. Here:

* ``RaggedBatch`` applies many operators -- different boxes, G-spheres, V(r),
  projector counts and band counts -- as ONE launch per local-term pass over the
  concatenated work items of every instance (``gw_batch_apply``), followed by
  each instance's MFMA projector GEMMs; ``capture`` records a fixed-shape
  batched apply in a HIP graph and ``replay`` re-launches it; the graph owns
  its binding (device tables, workspace), so eager calls in between leave it
  valid.
* ``davidson_ragged`` runs one ``davidson_steps`` generator per instance (the
  same solver as ``davidson_batched``) in lockstep, decode-style: each round it
  gathers every pending apply request into one ragged apply, the pending
  subspace eigensolves into a few batched eigensolves (sizes bucketed), and
  every pending host read (convergence gate, breakdown tests) into one copy,
  so the per-solve results are those of ``davidson_batched`` on each instance
  alone.
"""

from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import time

import numpy as np
import torch

from gradwave_amd import _lib as L
from gradwave_amd import linalg as LA
from gradwave_amd import workloads as W
from gradwave_amd.davidson import BatchedDavidsonResult, _cholinv_torch, _refine_steps, davidson_steps
from gradwave_amd.hamiltonian import BatchedHamiltonian


class RaggedBatch:
    """Independent operators applied together; ``hams`` are BatchedHamiltonian
    objects (their plans hold the geometry, V(r), projectors and D)."""

    def __init__(self, hams: list[BatchedHamiltonian]) -> None:
        self.hams = list(hams)
        n = len(self.hams)
        arr = (C.c_void_p * n)(*[h.plan.handle.value for h in self.hams])
        self.handle = C.c_void_p()
        self._lib = L.lib()
        L.check(self._lib.gw_batch_create(C.byref(self.handle), arr, n), "gw_batch_create")
        self._ws: dict[int, torch.Tensor] = {}
        self.device = self.hams[0].device

    def __len__(self) -> int:
        return len(self.hams)

    def __del__(self) -> None:
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.gw_batch_destroy(h)
            except Exception:
                pass
            self.handle = C.c_void_p()

    def _args(self, cs, outs):
        n = len(self.hams)
        cp, op = (C.c_void_p * n)(), (C.c_void_p * n)()
        ldc, sc, ldo, so = (C.c_long * n)(), (C.c_long * n)(), (C.c_long * n)(), (C.c_long * n)()
        nb = (C.c_int * n)()
        prec = None
        for i, (c, o) in enumerate(zip(cs, outs)):
            if c is None or c.shape[1] == 0:
                continue
            h = self.hams[i]
            if c.dim() != 3 or c.shape[0] != h.plan.nk or c.shape[2] != h.plan.npw_max or c.stride(2) != 1:
                raise ValueError(f"instance {i}: block must be (nk={h.plan.nk}, nb, npw_max={h.plan.npw_max}) "
                                 f"with a contiguous last axis, got {tuple(c.shape)}")
            p = L.precision_of(c.dtype)
            if prec is None:
                prec = p
            elif p != prec:
                raise TypeError("all blocks of one batched apply must share a precision")
            cp[i], op[i] = c.data_ptr(), o.data_ptr()
            ldc[i], sc[i], ldo[i], so[i] = c.stride(1), c.stride(0), o.stride(1), o.stride(0)
            nb[i] = c.shape[1]
        return prec, cp, ldc, sc, op, ldo, so, nb

    def _workspace(self, nb, prec: int) -> torch.Tensor:
        need = int(self._lib.gw_batch_workspace_bytes(self.handle, nb, prec))
        ws = self._ws.get(prec)
        if ws is None or ws.numel() < need:
            ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
            self._ws[prec] = ws
        return ws

    def apply(self, cs: list, terms: int = L.GW_TERM_ALL) -> list:
        """H_i c_i for every instance i with a block (None or 0 bands: skipped)."""
        outs = [None if c is None else torch.empty_like(c) for c in cs]
        prec, cp, ldc, sc, op, ldo, so, nb = self._args(cs, outs)
        if prec is None:
            return outs
        ws = self._workspace(nb, prec)
        L.check(self._lib.gw_batch_apply(self.handle, prec, terms, cp, ldc, sc, op, ldo, so, nb, ws.data_ptr(),
                                         ws.numel(), L.stream_ptr(self.device)), "gw_batch_apply")
        return outs

    def bind(self, cs: list, outs: list, terms: int = L.GW_TERM_ALL) -> "RaggedBinding":
        """Fix the blocks of a batched apply: a binding with device tables and a
        workspace of its own (uploaded on the current stream), launched by
        ``RaggedBinding.apply``; later calls through this batch leave it valid."""
        prec, cp, ldc, sc, op, ldo, so, nb = self._args(cs, outs)
        if prec is None:
            raise ValueError("nothing to bind")
        need = int(self._lib.gw_batch_workspace_bytes(self.handle, nb, prec))
        ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        handle = C.c_void_p()
        L.check(self._lib.gw_batch_bind(self.handle, prec, terms, cp, ldc, sc, op, ldo, so, nb, ws.data_ptr(),
                                        ws.numel(), L.stream_ptr(self.device), C.byref(handle)), "gw_batch_bind")
        return RaggedBinding(self, handle, (list(cs), list(outs), ws))

    def capture(self, cs: list, outs: list, terms: int = L.GW_TERM_ALL) -> "RaggedGraph":
        """Bind fixed blocks and record the batched apply in a HIP graph; the
        graph's ``replay()`` re-runs it on whatever the bound inputs hold. The
        graph owns its binding (tables, workspace, references to the blocks), so
        later ``apply``/``bind``/``capture`` calls through the batch do not
        touch what it replays."""
        binding = self.bind(cs, outs, terms)
        binding.apply()  # warm (first-use module loads stay out of the graph)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                binding.apply()
        torch.cuda.current_stream().wait_stream(s)
        return RaggedGraph(g, binding)


class RaggedBinding:
    """One bound batched apply (``RaggedBatch.bind``): owns its device tables and
    workspace and keeps its input and output blocks alive."""

    def __init__(self, batch: RaggedBatch, handle, keep) -> None:
        self.batch, self.handle, self._keep = batch, handle, keep

    def apply(self) -> None:
        L.check(self.batch._lib.gw_batch_apply_bound(self.handle, L.stream_ptr(self.batch.device)),
                "gw_batch_apply_bound")

    def __del__(self) -> None:
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.batch._lib.gw_batch_binding_destroy(h)
            except Exception:
                pass
            self.handle = C.c_void_p()


class RaggedGraph:
    """A HIP graph of one bound batched apply; holds the binding it replays."""

    def __init__(self, graph: torch.cuda.CUDAGraph, binding: RaggedBinding) -> None:
        self.graph, self.binding = graph, binding

    def replay(self) -> None:
        self.graph.replay()


# ---------------------------------------------------------------- lockstep solves


_EIGH_BUCKET = 32
_EIGH_HOST_MAX = int(os.environ.get("GW_EIGH_CPU_MAX", "64"))


def _eigh_exact_groups(mats: list[torch.Tensor]) -> list:
    groups: dict[int, list[int]] = {}
    for j, s in enumerate(mats):
        groups.setdefault(int(s.shape[-1]), []).append(j)
    out: list = [None] * len(mats)
    for idx in groups.values():
        w, u = torch.linalg.eigh(torch.cat([mats[j].to(torch.complex128) for j in idx]))
        row = 0
        for j in idx:
            nk = int(mats[j].shape[0])
            out[j] = (w[row:row + nk], u[row:row + nk])
            row += nk
    return out


_EIGH_STREAMS: list = []


def _eigh_batched(mats: list[torch.Tensor]) -> list[tuple[torch.Tensor, torch.Tensor]]:
    """Eigenpairs of many Hermitian subspace matrices (each (nk, m_i, m_i)).

    Matrices up to GW_EIGH_CPU_MAX go to host LAPACK, one batched call per size
    (the reference offloads small subspace solves to the CPU). Larger ones are
    latency-bound rocSOLVER solves: each runs on one of eight side streams so
    that independent solves overlap on the GPU. (Padding every matrix into one
    common-size batch gave non-finite eigenvectors from the batched GPU solver.)"""
    out: list = [None] * len(mats)
    small = [j for j, s in enumerate(mats) if int(s.shape[-1]) <= _EIGH_HOST_MAX]
    large = [j for j, s in enumerate(mats) if int(s.shape[-1]) > _EIGH_HOST_MAX]
    if small:
        host = [mats[j].to(torch.complex128).cpu() for j in small]
        for j, (w, u) in zip(small, _eigh_exact_groups(host)):
            out[j] = (w.to(mats[j].device), u.to(mats[j].device))
    if large:
        if not _EIGH_STREAMS:
            _EIGH_STREAMS.extend(torch.cuda.Stream() for _ in range(8))  # default priority
        cur = torch.cuda.current_stream()
        used = []
        for n, j in enumerate(large):
            st = _EIGH_STREAMS[n % len(_EIGH_STREAMS)]
            st.wait_stream(cur)
            s = mats[j].to(torch.complex128)
            s.record_stream(st)
            with torch.cuda.stream(st):
                w, u = torch.linalg.eigh(s)
            w.record_stream(cur)
            u.record_stream(cur)
            out[j] = (w, u)
            if st not in used:
                used.append(st)
        for st in used:
            cur.wait_stream(st)
    return out


def _eigh_lowest_batched(reqs: list, mode: str = "kernel") -> list:
    """Lowest nev eigenpairs for each (matrix, nev) request. ``mode="kernel"``:
    every matrix up to ``gw_heev_max_n`` in ONE launch of the batched
    subspace eigensolver (one workgroup per matrix); larger ones, and
    ``mode="library"``, take ``_eigh_batched`` (host LAPACK / rocSOLVER) and keep
    the first nev pairs."""
    out: list = [None] * len(reqs)
    if mode not in ("kernel", "library"):
        raise ValueError(f"subspace eigensolve mode must be 'kernel' or 'library', got {mode!r}")
    small = [j for j, (s, _) in enumerate(reqs) if mode == "kernel" and int(s.shape[-1]) <= LA.heev_max_n()]
    taken = set(small)
    rest = [j for j in range(len(reqs)) if j not in taken]
    if small:
        for j, wu in zip(small, LA.eigh_lowest([reqs[j][0].to(torch.complex128) for j in small],
                                               [reqs[j][1] for j in small])):
            out[j] = wu
    if rest:
        for j, (w, u) in zip(rest, _eigh_batched([reqs[j][0] for j in rest])):
            nev = reqs[j][1]
            out[j] = (w[..., :nev], u[..., :nev])
    return out


def _cholinv_batched(gs: list, mode: str = "kernel") -> list:
    """(L^-1, info) for each Gram matrix G = L L^H. ``mode="kernel"``: every
    matrix up to ``gw_chol_max_n`` in ONE launch of the batched Cholesky
    inverse; larger ones, and ``mode="library"``, take torch (rocSOLVER potrf,
    rocBLAS trsm) one by one as ``drive`` does."""
    if mode not in ("kernel", "library"):
        raise ValueError(f"factorisation mode must be 'kernel' or 'library', got {mode!r}")
    out: list = [None] * len(gs)
    small = [j for j, g in enumerate(gs) if mode == "kernel" and int(g.shape[-1]) <= LA.chol_max_n()]
    if small:
        for j, r in zip(small, LA.cholesky_inverse([gs[j] for j in small])):
            out[j] = r
    for j in range(len(gs)):
        if out[j] is None:
            out[j] = _cholinv_torch(gs[j])
    return out


@dataclass
class RaggedSolveStats:
    rounds: int = 0
    batched_applies: int = 0
    batched_eighs: int = 0
    host_reads: int = 0
    seconds: dict | None = None  # phase wall times when profiling (synchronising)


def _check_finite(kind: str, i: int, req, resp) -> None:
    tensors = resp if isinstance(resp, tuple) else (resp,)
    for t in tensors:
        if not bool(torch.isfinite(t).all()):
            src = req if isinstance(req, torch.Tensor) else None
            raise FloatingPointError(f"non-finite {kind} result for instance {i}: request finite="
                                     f"{None if src is None else bool(torch.isfinite(src).all())}, "
                                     f"shape {tuple(src.shape) if src is not None else None}")


def davidson_ragged(batch: RaggedBatch, x0s: list[torch.Tensor], ts: list[torch.Tensor],
                    masks: list[torch.Tensor], tol: float = 1e-9, max_iter: int = 40,
                    max_dim_factor: int = 4, stats: RaggedSolveStats | None = None,
                    check_finite: bool = False, subspace_eigh: str = "kernel") -> list[BatchedDavidsonResult]:
    """One ``davidson_steps`` solve per instance, advanced in lockstep: each round
    serves all pending apply requests with one ragged apply (one per precision)
    and the pending subspace eigensolves and Cholesky inverses (CholQR2) with
    one launch each of the batched kernels (``subspace_eigh="kernel"``;
    "library" serves them with host LAPACK / rocSOLVER / rocBLAS, as ``drive``
    would).
    ``check_finite`` raises on the first non-finite result (debugging aid)."""
    gens = [davidson_steps(x0, t, m, tol=tol, max_iter=max_iter, max_dim_factor=max_dim_factor)
            for x0, t, m in zip(x0s, ts, masks)]
    return _lockstep(batch, gens, stats, check_finite, subspace_eigh)


def _lockstep(batch: RaggedBatch, gens: list, stats: RaggedSolveStats | None,
              check_finite: bool, subspace_eigh: str = "kernel") -> list:
    """Advance step generators (``davidson_steps``, ``_refine_steps``) together:
    each round serves every pending apply with one ragged apply per precision,
    every pending eigensolve with one batched call and every host read with one
    copy."""
    results: list = [None] * len(gens)
    pending: dict[int, tuple] = {}

    def advance(i, resp=None, first=False):
        try:
            pending[i] = next(gens[i]) if first else gens[i].send(resp)
        except StopIteration as stop:
            results[i] = stop.value
            pending.pop(i, None)

    for i in range(len(gens)):
        advance(i, first=True)
    st = stats if stats is not None else RaggedSolveStats()
    prof = st.seconds

    def mark(key, t0):
        if prof is None:
            return t0
        torch.cuda.synchronize()
        now = time.perf_counter()
        prof[key] = prof.get(key, 0.0) + now - t0
        return now

    while pending:
        st.rounds += 1
        resp: dict[int, object] = {}
        t0 = mark("solver", time.perf_counter())
        applies = {i: a for i, (k, a) in pending.items() if k == "apply"}
        by_dtype: dict[torch.dtype, dict[int, torch.Tensor]] = {}
        for i, z in applies.items():
            by_dtype.setdefault(z.dtype, {})[i] = z if z.stride(-1) == 1 else z.contiguous()
        for group in by_dtype.values():
            cs = [group.get(i) for i in range(len(batch))]
            outs = batch.apply(cs)
            st.batched_applies += 1
            for i in group:
                resp[i] = outs[i]
        t0 = mark("apply", t0)
        eighs = {i: s for i, (k, s) in pending.items() if k == "eigh"}
        if eighs:
            keys = list(eighs)
            for i, wu in zip(keys, _eigh_batched([eighs[i] for i in keys])):
                resp[i] = wu
            st.batched_eighs += 1
        lows = {i: a for i, (k, a) in pending.items() if k == "eigh_low"}
        if lows:
            keys = list(lows)
            for i, wu in zip(keys, _eigh_lowest_batched([lows[i] for i in keys], subspace_eigh)):
                resp[i] = wu
            st.batched_eighs += 1
        chols = {i: g for i, (k, g) in pending.items() if k == "cholinv"}
        if chols:
            keys = list(chols)
            for i, r in zip(keys, _cholinv_batched([chols[i] for i in keys], subspace_eigh)):
                resp[i] = r
        t0 = mark("eigh", t0)
        hosts = {i: t for i, (k, t) in pending.items() if k == "host"}
        if hosts:  # every pending host read in one device-to-host copy
            keys = list(hosts)
            flat = torch.cat([hosts[i].reshape(-1).to(torch.float64) for i in keys]).cpu()
            off = 0
            for i in keys:
                t = hosts[i]
                resp[i] = flat[off:off + t.numel()].reshape(t.shape).to(t.dtype)
                off += t.numel()
            st.host_reads += 1
        t0 = mark("host", t0)
        if check_finite:
            for i, r in resp.items():
                _check_finite(pending[i][0], i, pending[i][1], r)
        for i, r in resp.items():
            advance(i, r)
        mark("solver", t0)
    return results


def davidson_ragged_mixed(batch: RaggedBatch, x0s, ts, masks, tol: float = 1e-9, draft_tol: float = 1e-4,
                          max_iter: int = 100, max_dim_factor: int = 4, sweeps: int | None = None,
                          stats: RaggedSolveStats | None = None,
                          subspace_eigh: str = "kernel") -> list[BatchedDavidsonResult]:
    """Precision form of the batch: all instances draft in complex64 (FP32 FFTs,
    V multiply and MFMA GEMMs) to ``draft_tol`` in lockstep, then polish in
    complex128 to ``tol`` from the drafted Ritz vectors, FP64 subspace algebra
    and residual norms throughout (the scaffold of solvers/_ms.py:26-52). With
    ``sweeps=n`` the polish is n FP64 refinement sweeps per instance in lockstep
    (no gate, ``tol`` unused), as in ``davidson_mixed``."""
    hi = [x.dtype for x in x0s]
    d = davidson_ragged(batch, [x.to(torch.complex64) for x in x0s], [t.to(torch.float32) for t in ts], masks,
                        tol=draft_tol, max_iter=max_iter, max_dim_factor=max_dim_factor, stats=stats,
                        subspace_eigh=subspace_eigh)
    if sweeps is not None:
        if sweeps < 0:
            raise ValueError(f"sweeps must be >= 0, got {sweeps}")
        gens = [_refine_steps(r.eigenvectors.to(h), t, m, int(sweeps)) for r, h, t, m in zip(d, hi, ts, masks)]
        p = _lockstep(batch, gens, stats, False, subspace_eigh)
    else:
        p = davidson_ragged(batch, [r.eigenvectors.to(h) for r, h in zip(d, hi)], ts, masks, tol=tol,
                            max_iter=max_iter, max_dim_factor=max_dim_factor, stats=stats,
                            subspace_eigh=subspace_eigh)
    for pi, di in zip(p, d):
        pi.n_apply_low += di.n_apply_full
        pi.n_draft_iter = di.n_iter
    return p


# ---------------------------------------------------------------- the ragged workload


def fft_shape(cell: np.ndarray, ecut: float) -> tuple[int, int, int]:
    """Box sizing of grids.py:68-126: n_i = good size >= 2 m_i + 1 for the density
    cutoff 4 ecut, 7-smooth."""
    gmax = 2.0 * math.sqrt(ecut / W.HBAR2_2M)
    dims = []
    for i in range(3):
        n = 2 * int(np.floor(gmax * np.linalg.norm(cell[i]) / (2.0 * np.pi))) + 1
        while True:
            m = n
            for p in (2, 3, 5, 7):
                while m % p == 0:
                    m //= p
            if m == 1:
                break
            n += 1
        dims.append(n)
    return tuple(dims)


def ragged_instances(device="cuda", n: int = 32, seed: int = 0) -> list[W.Workload]:
    """32 independent cells: fcc Al (a = 4.05 A, Z = 3, rc = 0.6 bohr), Si2 fcc,
    Si8 (conventional) and Si16 (2x2x2 fcc primitive) with seeded U(+-0.03 A) jitter, each at a
    seeded random k-point, Ecut uniform in [25, 40] Ry, boxes sized by the
    reference's rule, 8 (Al1), 8 (Si2), 32 (Si8) and 64 (Si16) bands."""
    rng = np.random.default_rng(seed)
    out = []
    kinds = ["al1", "si2", "si8", "si16"]
    for i in range(n):
        kind = kinds[i % 4]
        ecut_ry = float(rng.uniform(25.0, 40.0))
        k = rng.uniform(-0.5, 0.5, 3)
        if kind == "al1":
            cell = W.fcc_primitive(W.A_AL)
            pos, species, nb = np.zeros((1, 3)), [W.AL], 8
        elif kind == "si2":
            cell = W.fcc_primitive(W.A_SI)
            pos, species, nb = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) * W.A_SI, [W.SI] * 2, 8
        elif kind == "si8":
            cell, pos = W.si_conventional()
            pos, species, nb = W.jitter(pos, 0.03, seed + i), [W.SI] * 8, 32
        else:
            prim = W.fcc_primitive(W.A_SI)
            cell = 2.0 * prim
            basis = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) * W.A_SI
            shifts = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], dtype=np.float64) @ prim
            pos = (shifts[:, None, :] + basis[None]).reshape(-1, 3)
            pos, species, nb = W.jitter(pos, 0.03, seed + i), [W.SI] * 16, 64
        shape = fft_shape(cell, ecut_ry * W.RY_EV)
        out.append(W.build(f"{kind}_{ecut_ry:.1f}Ry_{shape[0]}x{shape[1]}x{shape[2]}", cell, pos, species, ecut_ry,
                           k[None], [1.0], nb, shape, device))
    return out
