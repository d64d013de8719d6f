"""Autotuner for the FFT radix splits, the MFMA tile shape and the split-K width of
the GEMMs (no reference counterpart: the reference uses torch.fft and BLAS).

* Radix: every box length n is transformed as two codelet passes n = a * b
  (a, b in {1..10, 12, 14, 15, 16}); plans take the most balanced split unless
  ``gw_tune_set(GW_TUNE_RADIX, n, a)`` says otherwise. ``tune_radix`` times
  the generic-size local-term kernels of real operators for each candidate a
  (splits into factors <= 10 when there are any) and keeps the fastest.
* GEMM tiles and split-K: the complex MFMA kernels run 2x2, 4x2, 2x4, 4x4 or
  band-sized (2x5 / 5x2, 4x5 / 5x4) 16x16 blocks per wave (``GW_TUNE_GEMM_TILE``,
  per precision and per contraction), and ``gw_zgemm_inner`` / ``gw_rgemm_inner``
  split the long axis until the launch has about ``GW_TUNE_GEMM_WGS``
  workgroups; ``tune_gemm`` times the projector products of real operators for
  each tile shape and target.
* x pass of the wave-line boxes (24, 36, 90, 120 cubed): lines straight from
  HBM or planes staged through LDS, waves and bands per workgroup
  (``GW_TUNE_XPASS``, per precision); ``tune_xpass`` times the local term.

Choices persist as JSON (``GW_TUNE_FILE``, default
``~/.cache/gradwave_amd/tune.json``) and are loaded with the library.
"""

from __future__ import annotations

import json
import os
import statistics
import time

import torch

from gradwave_amd import _lib as L

CODELETS = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16}
GEMM_TARGETS = (512, 1024, 2048, 4096, 8192)


PLANE_BANDS = (1, 2, 4, 8)


def tune_plane_bands(workloads, nb: int = 64, reps: int = 5) -> dict:
    """For each precision, time P1 and P3 (the plan's stage timer) of the local
    term of ``workloads`` whose box is a wave-line cube with every bands-per-item
    count in PLANE_BANDS and keep the fastest per pass. Returns {"fp64": [p1, p3],
    "fp32": [p1, p3]} (the current values if no workload qualifies)."""
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    lib = L.lib()
    users = [w for w in workloads if len(set(w.shape)) == 1 and w.shape[0] in WAVE_LENGTHS]
    chosen = {name: [lib.gw_tune_get(L.GW_TUNE_PLANE_BANDS, prec + 256), lib.gw_tune_get(L.GW_TUNE_PLANE_BANDS,
                                                                                        prec + 512)]
              for name, prec in (("fp64", L.GW_FP64), ("fp32", L.GW_FP32))}
    if not users:
        return chosen
    hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in users]
    for h in hams:
        h.set_flags(L.GW_PLAN_TIMING)
    g = torch.Generator().manual_seed(6)
    b64 = [(torch.randn(w.bk.nk, nb, w.bk.npw_max, dtype=torch.complex128, generator=g).to(w.v_r.device)
            * w.bk.mask[:, None, :]) for w in users]
    try:
        for name, prec, blocks in (("fp64", L.GW_FP64, b64),
                                   ("fp32", L.GW_FP32, [b.to(torch.complex64) for b in b64])):
            best = {"plane_inv": (0, float("inf")), "plane_fwd": (0, float("inf"))}
            for bpw in PLANE_BANDS:
                L.check(lib.gw_tune_set(L.GW_TUNE_PLANE_BANDS, prec, bpw), "gw_tune_set")
                for h, x in zip(hams, blocks):
                    h.apply_terms(x, L.GW_TERM_LOCAL)
                torch.cuda.synchronize()
                for h in hams:
                    h.stage_times(reset=True)
                for _ in range(reps):
                    for h, x in zip(hams, blocks):
                        h.apply_terms(x, L.GW_TERM_LOCAL)
                torch.cuda.synchronize()
                tot = {"plane_inv": 0.0, "plane_fwd": 0.0}
                for h in hams:
                    st = h.stage_times(reset=True)
                    for k in tot:
                        tot[k] += st[k][0] / max(st[k][1], 1)
                for k in best:
                    if tot[k] < best[k][1]:
                        best[k] = (bpw, tot[k])
            L.check(lib.gw_tune_set(L.GW_TUNE_PLANE_BANDS, prec + 256, best["plane_inv"][0]), "gw_tune_set")
            L.check(lib.gw_tune_set(L.GW_TUNE_PLANE_BANDS, prec + 512, best["plane_fwd"][0]), "gw_tune_set")
            chosen[name] = [best["plane_inv"][0], best["plane_fwd"][0]]
    finally:
        for h in hams:
            h.set_flags(0)
    return chosen


def tune_file() -> str:
    return os.environ.get("GW_TUNE_FILE", os.path.join(os.path.expanduser("~"), ".cache", "gradwave_amd",
                                                       "tune.json"))


def candidate_splits(n: int) -> list[int]:
    """First factors a of the two-codelet splits n = a * b; when splits with both
    factors <= 10 exist only those are offered (the kernels built for codelets up
    to 10 need far fewer registers than the ones carrying 12-16)."""
    all_ = [a for a in range(1, 17) if n % a == 0 and a in CODELETS and n // a in CODELETS]
    small = [a for a in all_ if a <= 10 and n // a <= 10]
    return small or all_


def current() -> dict:
    lib = L.lib()
    radix = {n: lib.gw_tune_get(L.GW_TUNE_RADIX, n) for n in range(1, 1024)}
    return {"radix": {str(n): a for n, a in radix.items() if a > 0},
            "gemm_wgs": lib.gw_tune_get(L.GW_TUNE_GEMM_WGS, 0),
            "gemm_tile": {"fp64": lib.gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP64 + 256),
                          "fp32": lib.gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP32 + 256)},
            "gemm_tile_combine": {"fp64": lib.gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP64 + 512),
                                  "fp32": lib.gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP32 + 512)},
            "combine_lib": {"fp64": lib.gw_tune_get(L.GW_TUNE_COMBINE_LIB, L.GW_FP64),
                            "fp32": lib.gw_tune_get(L.GW_TUNE_COMBINE_LIB, L.GW_FP32)},
            "inner_lib": {"fp64": lib.gw_tune_get(L.GW_TUNE_INNER_LIB, L.GW_FP64),
                          "fp32": lib.gw_tune_get(L.GW_TUNE_INNER_LIB, L.GW_FP32)},
            "xpass": {"fp64": lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP64),
                      "fp32": lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP32)},
            "plane_bands": {name: [lib.gw_tune_get(L.GW_TUNE_PLANE_BANDS, prec + 256),
                                   lib.gw_tune_get(L.GW_TUNE_PLANE_BANDS, prec + 512)]
                            for name, prec in (("fp64", L.GW_FP64), ("fp32", L.GW_FP32))}}


def apply(choices: dict) -> None:
    lib = L.lib()
    for n, a in choices.get("radix", {}).items():
        L.check(lib.gw_tune_set(L.GW_TUNE_RADIX, int(n), int(a)), "gw_tune_set")
    if "gemm_wgs" in choices:
        L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_WGS, 0, int(choices["gemm_wgs"])), "gw_tune_set")
    for name, prec in (("fp64", L.GW_FP64), ("fp32", L.GW_FP32)):
        if name in choices.get("gemm_tile", {}):
            L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, prec + 256, int(choices["gemm_tile"][name])), "gw_tune_set")
        if name in choices.get("gemm_tile_combine", {}):
            L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, prec + 512, int(choices["gemm_tile_combine"][name])),
                    "gw_tune_set")
        if name in choices.get("combine_lib", {}):
            L.check(lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, int(choices["combine_lib"][name])), "gw_tune_set")
        if name in choices.get("inner_lib", {}):
            L.check(lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, int(choices["inner_lib"][name])), "gw_tune_set")
        if name in choices.get("xpass", {}):
            L.check(lib.gw_tune_set(L.GW_TUNE_XPASS, prec, int(choices["xpass"][name])), "gw_tune_set")
        if name in choices.get("plane_bands", {}):
            p1, p3 = choices["plane_bands"][name]
            L.check(lib.gw_tune_set(L.GW_TUNE_PLANE_BANDS, prec + 256, int(p1)), "gw_tune_set")
            L.check(lib.gw_tune_set(L.GW_TUNE_PLANE_BANDS, prec + 512, int(p3)), "gw_tune_set")


def load(path: str | None = None) -> dict | None:
    path = path or tune_file()
    if not os.path.exists(path):
        return None
    with open(path) as f:
        choices = json.load(f)
    apply(choices)
    return choices


def save(path: str | None = None) -> str:
    path = path or tune_file()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(current(), f, indent=1, sort_keys=True)
    return path


def _median_ms(fn, reps: int = 5) -> float:
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts) * 1e3


def tune_radix(workloads, nb: int = 8, reps: int = 5) -> dict[int, int]:
    """For every box length of ``workloads`` (objects with bk, shape, v_r, p),
    time the generic-size local term of the operators that use it under each
    candidate split and keep the fastest (lengths tuned one at a time)."""
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    lib = L.lib()
    lengths = sorted({n for w in workloads for n in w.shape})
    chosen = {}
    for n in lengths:
        cands = candidate_splits(n)
        users = [w for w in workloads if n in w.shape]
        if len(cands) < 2 or not users:
            continue
        g = torch.Generator().manual_seed(n)
        blocks = [(torch.randn(w.bk.nk, nb, w.bk.npw_max, dtype=torch.complex128, generator=g).to(w.v_r.device)
                   * w.bk.mask[:, None, :]).to(torch.complex64) for w in users]
        best, best_ms = None, float("inf")
        for a in cands:
            L.check(lib.gw_tune_set(L.GW_TUNE_RADIX, n, a), "gw_tune_set")
            hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in users]
            for h in hams:
                h.set_flags(L.GW_PLAN_LEGACY)
            ms = _median_ms(lambda: [h.apply_terms(x, L.GW_TERM_LOCAL) for h, x in zip(hams, blocks)], reps)
            if ms < best_ms:
                best, best_ms = a, ms
        L.check(lib.gw_tune_set(L.GW_TUNE_RADIX, n, best), "gw_tune_set")
        chosen[n] = best
    return chosen


# 0-5: expanded-operand tiles (MFMA blocks per wave 2x2, 4x2, 2x4, 4x4, band-sized
# 2x5/5x2, 4x5/5x4); 6-9: complex-accumulator tiles, four real products per
# complex one; 10-13: the same with three (Gauss)
GEMM_TILES = tuple(range(14))


def tune_gemm(workloads, nb: int = 32, reps: int = 5, allow_library: bool = True) -> tuple[dict, int]:
    """Tune the projector products (nonlocal term) of ``workloads`` on the plan's
    per-stage HIP-event times: per precision, the inner-product tile on the becp
    stage and the back-projection tile on the backproj stage, each on its own
    (with ``allow_library``, rocBLAS is also tried for both contractions); then
    the split-K workgroup target on the becp stage of both precisions. Returns
    ({"fp64": inner tile, "fp32": inner tile, "combine": {"fp64": tile, "fp32": tile},
    "combine_lib": {"fp64": 0|1, "fp32": 0|1}, "inner_lib": {...}}, target)."""
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    lib = L.lib()
    hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in workloads]
    for h in hams:
        h.set_flags(L.GW_PLAN_TIMING)
    g = torch.Generator().manual_seed(7)
    b64 = [(torch.randn(w.bk.nk, nb, w.bk.npw_max, dtype=torch.complex128, generator=g).to(w.v_r.device)
            * w.bk.mask[:, None, :]) for w in workloads]
    blocks = {"fp64": b64, "fp32": [b.to(torch.complex64) for b in b64]}

    def stage_ms(name):
        def run():
            for h, x in zip(hams, blocks[name]):
                h.apply_terms(x, L.GW_TERM_NONLOCAL)
        run()
        torch.cuda.synchronize()
        for h in hams:
            h.stage_times(reset=True)
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
        tot = {"becp": 0.0, "backproj": 0.0}
        for h in hams:
            st = h.stage_times(reset=True)
            for k in tot:
                tot[k] += st[k][0] / max(st[k][1], 1)
        return tot

    tiles, combine, libs, ilibs = {}, {}, {}, {}
    try:
        for name, prec in (("fp64", L.GW_FP64), ("fp32", L.GW_FP32)):
            L.check(lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, 0), "gw_tune_set")
            L.check(lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, 0), "gw_tune_set")
            best = {"becp": (0, float("inf")), "backproj": (0, float("inf"))}
            for tile in GEMM_TILES:
                L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, prec, tile), "gw_tune_set")
                t = stage_ms(name)
                for k in best:
                    if t[k] < best[k][1]:
                        best[k] = (tile, t[k])
            tiles[name], combine[name] = best["becp"][0], best["backproj"][0]
            L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, prec + 256, tiles[name]), "gw_tune_set")
            L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, prec + 512, combine[name]), "gw_tune_set")
            libs[name] = ilibs[name] = 0
            if allow_library:
                L.check(lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, 1), "gw_tune_set")
                L.check(lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, 1), "gw_tune_set")
                t = stage_ms(name)
                libs[name] = int(t["backproj"] < best["backproj"][1])
                ilibs[name] = int(t["becp"] < best["becp"][1])
                L.check(lib.gw_tune_set(L.GW_TUNE_COMBINE_LIB, prec, libs[name]), "gw_tune_set")
                L.check(lib.gw_tune_set(L.GW_TUNE_INNER_LIB, prec, ilibs[name]), "gw_tune_set")
        best_target, best_ms = None, float("inf")
        for target in GEMM_TARGETS:
            L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_WGS, 0, target), "gw_tune_set")
            ms = stage_ms("fp64")["becp"] + stage_ms("fp32")["becp"]
            if ms < best_ms:
                best_target, best_ms = target, ms
        L.check(lib.gw_tune_set(L.GW_TUNE_GEMM_WGS, 0, best_target), "gw_tune_set")
    finally:
        for h in hams:
            h.set_flags(0)
    tiles["combine"] = combine
    tiles["combine_lib"] = libs
    tiles["inner_lib"] = ilibs
    return tiles, best_target


# x-pass shapes: mode * 10000 + waves * 100 + bands (mode 1: planes staged through LDS)
XPASS_SHAPES = (408, 416, 208, 10408, 10416, 10432, 10208)
WAVE_LENGTHS = (24, 36, 90, 120)


def tune_xpass(workloads, nb: int = 32, reps: int = 5) -> dict:
    """For each precision, time the local term of the ``workloads`` whose box is a
    wave-line cube under each x-pass shape and keep the fastest. Returns
    {"fp64": shape, "fp32": shape} (the current shapes if no workload qualifies)."""
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    lib = L.lib()
    users = [w for w in workloads if len(set(w.shape)) == 1 and w.shape[0] in WAVE_LENGTHS]
    chosen = {"fp64": lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP64),
              "fp32": lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP32)}
    if not users:
        return chosen
    hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in users]
    g = torch.Generator().manual_seed(5)
    b64 = [(torch.randn(w.bk.nk, nb, w.bk.npw_max, dtype=torch.complex128, generator=g).to(w.v_r.device)
            * w.bk.mask[:, None, :]) for w in users]
    for name, prec, blocks in (("fp64", L.GW_FP64, b64), ("fp32", L.GW_FP32, [b.to(torch.complex64) for b in b64])):
        # two interleaved rounds, each shape keeping its best median: a single
        # pass in a fixed order let clock ramp-up and noise pick a slower shape
        times = {}
        for _ in range(2):
            for shape in XPASS_SHAPES:
                L.check(lib.gw_tune_set(L.GW_TUNE_XPASS, prec, shape), "gw_tune_set")
                ms = _median_ms(lambda: [h.apply_terms(x, L.GW_TERM_LOCAL) for h, x in zip(hams, blocks)], reps)
                times[shape] = min(ms, times.get(shape, float("inf")))
        best = min(times, key=times.get)
        L.check(lib.gw_tune_set(L.GW_TUNE_XPASS, prec, best), "gw_tune_set")
        chosen[name] = best
    return chosen
