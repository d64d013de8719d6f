"""Benchmark: H|psi> band-applies/s and Davidson s/k-pt on one MI355X.

Workload (default): Si216 Gamma supercell (3x3x3 conventional, L=16.29 A,
seeded U(+-0.03 A) jitter), 30 Ry, 80,797 plane waves, 120^3 box, 480 bands,
864 KB projectors, analytic Gaussian pseudopotential. One step = one H|psi>
apply of the full (1, 480, npw) block.

Printed JSON line: value = FP64 (complex128) band-applies/s with the block
resident in HBM; the FP32 (complex64) rate and its error against FP64 beside
it; the Davidson solve times per k-point (FP64, the mixed two-stage solve, the
sweep form: FP32 draft + final FP64 refinement sweeps, and the FP64 solve
stopped at the sweep's residual) with eigenvalue errors against FP64; the
adjoint forces; the Gamma real layout; the roofline of the dominant FP64
kernel and of every stage (HIP-event timed on the launch stream); and the CPU
baseline (the oracle's op-for-op restatement, torch CPU complex128: a bounded
apply sample and, for large cells, an extrapolated Davidson estimate).
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# The headline metric pairs the apply rate (`value`, band-applies/s) with the
# Davidson time per k-point (the "davidson" section, FP64 and mixed, with the
# mixed solve's eigenvalue error against FP64).
METRIC = "Hψ band-applies/s; Davidson s/k-pt, mixed vs FP64 GPU @ ≤1e-6 Ha eigval err"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK = {"fp64": 78.6, "fp32": 157.3}  # dense TFLOP/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="si216_gamma",
                    choices=["si64_gamma", "si2_fcc", "si8_mp", "si216_gamma", "ragged32"])
    ap.add_argument("--no-davidson", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-bands", type=int, default=8)
    ap.add_argument("--cpu-davidson-iters", type=int, default=2,
                    help="oracle Davidson iterations timed for the CPU solve estimate on large cells")
    ap.add_argument("--davidson-max-iter", type=int, default=300)
    ap.add_argument("--no-forces", action="store_true")
    ap.add_argument("--no-gamma", action="store_true")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged32 batching-form section")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default GEMM tile and split-K width")
    ap.add_argument("--only", choices=["apply", "gamma", "forces", "copy", "davidson"], default=None,
                    help="profiling runs: time just the block apply (FP64 then FP32), the Gamma-layout "
                         "apply, the adjoint forces, or a 1 GiB device copy (traffic-counter calibration); "
                         "with --workload ragged32, 'davidson' runs only the lockstep FP64 solve and two "
                         "graph-captured draft + sweep solves")
    ap.add_argument("--profile-davidson", action="store_true")
    ap.add_argument("--tol-ha", type=float, default=1e-10)
    ap.add_argument("--draft-tol", type=float, default=1e-3,
                    help="draft tolerance [eV] of the mixed solve that is polished to --tol-ha")
    ap.add_argument("--sweep-draft-tol", type=float, default=5e-3,
                    help="draft tolerance [eV] of the sweep form (FP32 draft + FP64 sweeps), set for the metric's "
                         "1e-6 Ha eigenvalue bound; the eigenvalue and force errors are measured in every run")
    ap.add_argument("--sweep-buffer", type=int, default=32,
                    help="buffer bands the sweep form carries above the requested ones (seeded start vectors; "
                         "only the requested bands are returned and compared)")
    ap.add_argument("--sweeps", type=int, default=1,
                    help="FP64 refinement sweeps after the draft in the sweep form of the mixed solve")
    return ap.parse_args()


def main() -> None:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "gradwave_amd runs on one GPU; --gpus > 1 is not supported "
                          "(k-point/instance sharding is a design note only)"}))
        sys.exit(2)
    import torch

    if args.workload == "ragged32":
        bench_ragged(args)
        return
    from gradwave_amd import _lib as L
    from gradwave_amd import workloads as W
    from gradwave_amd.davidson import davidson_batched, davidson_mixed
    from gradwave_amd.hamiltonian import BatchedHamiltonian

    dev = torch.device("cuda:0")
    t0 = time.time()
    wl = getattr(W, args.workload)(device=dev)
    bk = wl.bk
    nk, nb, m = bk.nk, wl.nbands, bk.npw_max
    h = BatchedHamiltonian(bk, wl.shape, wl.v_r, wl.p)
    print(f"[bench] setup {wl.name}: nk={nk} nb={nb} npw_max={m} nproj={wl.p.shape[1]} "
          f"box={wl.shape} ({time.time() - t0:.1f}s) info={h.info(torch.complex64)}",
          file=sys.stderr, flush=True)
    x64 = W.start_vectors(wl, seed=0)
    x32 = x64.to(torch.complex64)
    tuned = None
    if not args.no_autotune:
        from gradwave_amd import autotune

        t1 = time.perf_counter()
        tiles, wgs = autotune.tune_gemm([wl], nb=nb)
        plane_bands = autotune.tune_plane_bands([wl], nb=nb)
        xpass = autotune.tune_xpass([wl], nb=nb)
        tuned = {"gemm_tile_variant": tiles, "gemm_split_k_target_wgs": wgs,
                 "plane_bands_per_item": plane_bands, "xpass_shape": xpass, "seconds": time.perf_counter() - t1,
                 "saved_to": autotune.save()}  # later runs with --no-autotune load these choices
        print(f"[bench] autotune GEMM: tiles {tiles}, split-K target {wgs}; plane bands {plane_bands}, "
              f"x-pass {xpass} "
              f"({tuned['seconds']:.1f}s)",
              file=sys.stderr, flush=True)
    if args.only is not None:
        profile_only(args, wl, x64)
        return

    def timed_applies(x, steps, warmup):
        out = torch.empty_like(x)
        for _ in range(warmup):
            h.apply_terms(x, L.GW_TERM_ALL, out=out)
        torch.cuda.synchronize()
        ts = []
        t_all = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            h.apply_terms(x, L.GW_TERM_ALL, out=out)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        total = time.perf_counter() - t_all
        return total / steps, statistics.median(ts)

    ms64, med64 = timed_applies(x64, args.steps, args.warmup)
    ms32, med32 = timed_applies(x32, args.steps, args.warmup)
    print(f"[bench] apply fp64 {ms64 * 1e3:.3f} ms/step, fp32 {ms32 * 1e3:.3f} ms/step",
          file=sys.stderr, flush=True)
    h64 = h.apply_terms(x64, L.GW_TERM_ALL)
    h32 = h.apply_terms(x32, L.GW_TERM_ALL).to(torch.complex128)
    err32 = float((h32 - h64).abs().max() / h64.abs().max())
    del h64, h32

    # per-stage HIP-event timing (separate loop: events sit between kernels)
    def stages(x, reps=10):
        h.set_flags(L.GW_PLAN_TIMING)
        h.stage_times(reset=True)
        out = torch.empty_like(x)
        for _ in range(reps):
            h.apply_terms(x, L.GW_TERM_ALL, out=out)
        torch.cuda.synchronize()
        st = h.stage_times(reset=True)
        h.set_flags(0)
        return {k: v[0] / max(v[1], 1) for k, v in st.items()}

    st64, st32 = stages(x64), stages(x32)
    n23 = wl.shape[1] * wl.shape[2]
    info = h.info(torch.complex128)
    na = info["active_planes"]
    nbt = nk * nb
    nproj = int(wl.p.shape[1])
    npw_tot = int(bk.npw.sum())

    def stage_model(e):
        """Algorithmic HBM bytes / flops per launch of each stage (whole block)."""
        N = wl.shape[0] * n23
        box = nbt * na * n23 * e
        return {
            "plane_inv": ("hbm", nb * npw_tot * e + box),
            "xpass": ("hbm", 2 * box + N * (e // 2)),
            "plane_fwd": ("hbm", box + 2 * nb * npw_tot * e),
            "becp": ("mfma", 8.0 * nproj * nb * npw_tot),
            "dmul": ("hbm", 2 * nk * nproj * nb * e),
            "backproj": ("mfma", 8.0 * nproj * nb * npw_tot),
        }

    def roofline_of(stage, st, e, prec):
        kind, work = stage_model(e)[stage]
        t = st[stage] * 1e-3
        if kind == "hbm":
            r = {"kernel": stage, "bound": "hbm", "achieved": work / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
        else:
            # a three-product (Gauss) kernel issues 6 real MFMA flops per complex
            # multiply-add instead of 8: its MFMA rate is measured on those, and
            # the 8-flop rate is kept beside it for comparison with 4-product GEMMs
            frac_mfma = gemm_mfma_fraction(stage, "double" if prec == "fp64" else "float")
            work = work * frac_mfma
            r = {"kernel": stage, "bound": "mfma", "achieved": work / t / 1e12, "peak": MFMA_PEAK[prec],
                 "unit": "TFLOP/s", "real_products_per_complex": int(round(4 * frac_mfma)),
                 "complex_equiv_TFLOPs": work / frac_mfma / t / 1e12}
        r["frac"] = r["achieved"] / r["peak"]
        r["avg_ms"] = st[stage]
        r["algorithmic_per_launch"] = work
        return r

    real64 = "double"
    dom = max(st64, key=lambda k: st64[k])
    roof = roofline_of(dom, st64, 16, "fp64")
    roof["kernel_name"] = stage_kernel(dom, real64, wl.name)
    roof["traffic"], roof["traffic_source"] = profiled_traffic(args.workload, dom, real64)
    roof["stage_ms_fp64"] = st64
    roof["stage_ms_fp32"] = st32
    roof["stages_fp64"] = {k: {f: v for f, v in roofline_of(k, st64, 16, "fp64").items() if f != "kernel"}
                           for k in st64 if k != "dmul"}
    roof["stages_fp32"] = {k: {f: v for f, v in roofline_of(k, st32, 8, "fp32").items() if f != "kernel"}
                           for k in st32 if k != "dmul"}
    roof["copy_GBps"] = copy_bandwidth(dev)

    result = {
        "metric": METRIC,
        "value": nbt / ms64,
        "unit": "band-applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64",
        "data": "synthetic (analytic Gaussian pseudopotential, seeded start vectors)",
        "config": {"workload": wl.name, "nk": nk, "nb": nb, "npw_max": m, "box": list(wl.shape),
                   "nproj": nproj, "apply": "complex128 H|psi> of the full block (value); complex64 below"},
        "fp32_band_applies_per_s": nbt / ms32,
        "fp32_ms_per_step": ms32 * 1e3,
        "fp32_apply_max_rel_err_vs_fp64": err32,
        "roofline": roof,
        "autotune": tuned,
    }
    if nk * m * m * 16 <= (4 << 30):
        result["toeplitz"] = toeplitz_rates(args, wl, x32, x64)

    if not args.no_davidson:
        from gradwave_amd.workloads import HARTREE_EV

        tol = args.tol_ha * HARTREE_EV
        torch.cuda.synchronize()
        a = time.perf_counter()
        prof64, profms = {}, {}
        r64 = davidson_batched(h.apply, x64, bk.t, bk.mask, tol=tol, max_iter=args.davidson_max_iter,
                               profile_out=prof64 if args.profile_davidson else None)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - a
        print(f"[bench] davidson fp64 {t64:.2f}s iters {r64.n_iter} rn {float(r64.residual_norms.max()):.2e}",
              file=sys.stderr, flush=True)
        a = time.perf_counter()
        rms = davidson_mixed(h.apply, x64, bk.t, bk.mask, tol=tol, draft_tol=args.draft_tol,
                             max_iter=args.davidson_max_iter,
                             profile_out=profms if args.profile_davidson else None)
        torch.cuda.synchronize()
        tms = time.perf_counter() - a
        de = float((rms.eigenvalues - r64.eigenvalues).abs().max()) / HARTREE_EV
        print(f"[bench] davidson mixed {tms:.2f}s iters {rms.n_iter} max|de| {de:.2e} Ha",
              file=sys.stderr, flush=True)
        sweep, rsw = sweep_form(args, lambda x: davidson_mixed(
            h.apply, x, bk.t, bk.mask, draft_tol=args.sweep_draft_tol, max_iter=args.davidson_max_iter,
            sweeps=args.sweeps, buffer=args.sweep_buffer),
                           lambda x, tl: davidson_batched(h.apply, x, bk.t, bk.mask, tol=tl,
                                                          max_iter=args.davidson_max_iter),
                           x64, r64.eigenvalues, t64, nk)
        result["davidson"] = {
            "fp64_s_per_kpt": t64 / nk, "mixed_s_per_kpt": tms / nk, "speedup_mixed_vs_fp64": t64 / tms,
            "fp64_iters": r64.n_iter, "mixed_polish_iters": rms.n_iter,
            "fp64_max_residual_eV": float(r64.residual_norms.max()),
            "mixed_max_residual_eV": float(rms.residual_norms.max()),
            "max_eig_diff_mixed_vs_fp64_Ha": de, "tol_Ha": args.tol_ha,
            "mixed_draft_iters": getattr(rms, "n_draft_iter", None),
            "phase_s_fp64": prof64 or None, "phase_s_mixed": profms or None, "sweep_form": sweep,
            "band_energy_fp64_Ha": float((r64.eigenvalues[:, : int(wl.n_electrons // 2)].sum(-1)
                                          * torch.as_tensor(wl.kweights, device=dev)).sum() * 2
                                         / HARTREE_EV)}
        if not args.no_forces:
            result["forces"] = forces_check(wl, r64.eigenvectors, rms.eigenvectors, rsw.eigenvectors)
        if not args.no_gamma and nk == 1 and not np.any(wl.kfrac):
            result["gamma"] = gamma_form(args, wl, x64, r64)

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args, wl, result.get("davidson", {}).get("fp64_iters"))
    if not args.no_ragged and not args.no_davidson:
        result["batching_form"] = ragged_section(args)

    print(json.dumps(result), flush=True)


def ragged_section(args) -> dict:
    """The batching form next to the headline line: 32 independent cells
    (Al1/Si2/Si8/Si16, 25-40 Ry, 8-64 bands) solved in lockstep FP64 (every
    round's eigensolves and Cholesky inverses in one launch each), one cell at
    a time (davidson_batched each), and as graph-captured draft rounds + one
    FP64 sweep (captured once, then replayed). ``bench.py --workload ragged32``
    has the full set (apply rates, library eigensolves, mixed forms)."""
    import torch

    from gradwave_amd import autotune
    from gradwave_amd import batching as B
    from gradwave_amd import graph_davidson as GD
    from gradwave_amd.davidson import davidson_batched
    from gradwave_amd.hamiltonian import BatchedHamiltonian
    from gradwave_amd.workloads import HARTREE_EV, start_vectors

    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    wls = B.ragged_instances(device=dev, n=32, seed=0)
    autotune.tune_radix(wls)
    hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in wls]
    batch = B.RaggedBatch(hams)
    x64 = [start_vectors(w, seed=i) for i, w in enumerate(wls)]
    ts_, ms_ = [w.bk.t for w in wls], [w.bk.mask for w in wls]
    tol = args.tol_ha * HARTREE_EV
    t_setup = time.perf_counter() - t0

    def timed(fn):
        torch.cuda.synchronize()
        a = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - a

    lock, t_lock = timed(lambda: B.davidson_ragged(batch, x64, ts_, ms_, tol=tol, max_iter=args.davidson_max_iter))
    seq, t_seq = timed(lambda: [davidson_batched(h.apply, x, w.bk.t, w.bk.mask, tol=tol,
                                                 max_iter=args.davidson_max_iter)
                                for h, x, w in zip(hams, x64, wls)])
    solver = GD.RaggedGraphDavidson(batch, ts_, ms_, [int(x.shape[1]) for x in x64])

    def graph_solve():
        return GD.davidson_ragged_graph_mixed(batch, x64, ts_, ms_, draft_tol=args.draft_tol, sweeps=args.sweeps,
                                              max_iter=args.davidson_max_iter, solver=solver)

    _, t_graph_first = timed(graph_solve)
    gr, t_graph = timed(graph_solve)
    de_lock = max(float((a.eigenvalues - b.eigenvalues).abs().max()) for a, b in zip(lock, seq)) / HARTREE_EV
    de_graph = max(float((a.eigenvalues - b.eigenvalues).abs().max()) for a, b in zip(gr, seq)) / HARTREE_EV
    out = {"workload": "ragged32: Al1/Si2/Si8/Si16 x 8, 25-40 Ry, seeded k, 8-64 bands", "setup_s": t_setup,
           "lockstep_fp64_s": t_lock, "one_by_one_fp64_s": t_seq, "lockstep_speedup_vs_one_by_one": t_seq / t_lock,
           "max_eig_diff_lockstep_vs_one_by_one_Ha": de_lock,
           "graph_draft_plus_fp64_sweep_s": t_graph, "graph_first_solve_with_capture_s": t_graph_first,
           "graph_speedup_vs_one_by_one_fp64": t_seq / t_graph,
           "max_eig_diff_graph_vs_one_by_one_Ha": de_graph, "graph_draft_tol_eV": args.draft_tol,
           "tol_Ha": args.tol_ha}
    print(f"[bench] batching form: lockstep FP64 {t_lock:.2f}s, one-by-one {t_seq:.2f}s "
          f"({t_seq / t_lock:.1f}x), graph draft + sweep {t_graph:.2f}s replayed ({t_graph_first:.2f}s first), "
          f"max|de| {de_lock:.1e} / {de_graph:.1e} Ha", file=sys.stderr, flush=True)
    return out


def profile_only(args, wl, x64) -> None:
    """Short runs for rocprofv3: ``--only gamma`` times the Gamma-layout apply
    (FP32 and FP64 embeddings), ``--only forces`` the adjoint forces on the
    seeded start block; one JSON line with the medians."""
    import torch

    from gradwave_amd.adjoint import band_energy_forces

    out = {"workload": wl.name, "only": args.only}
    if args.only == "copy":
        out["copy_GBps"] = copy_bandwidth(x64.device, nbytes=1 << 30, reps=5)
        out["bytes_read_per_copy"] = 1 << 30
    elif args.only == "apply":
        from gradwave_amd import _lib as L
        from gradwave_amd.hamiltonian import BatchedHamiltonian

        h = BatchedHamiltonian(wl.bk, wl.shape, wl.v_r, wl.p)
        for name, x in (("fp64", x64), ("fp32", x64.to(torch.complex64))):
            o = torch.empty_like(x)
            ts = []
            for i in range(args.warmup + args.steps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                h.apply_terms(x, L.GW_TERM_ALL, out=o)
                torch.cuda.synchronize()
                if i >= args.warmup:
                    ts.append(time.perf_counter() - t0)
            out[f"{name}_ms"] = statistics.median(ts) * 1e3
    elif args.only == "gamma":
        from gradwave_amd import gamma as G

        sph = type("Sphere", (), {})()
        sph.flat_idx = wl.bk.flat_idx[0]
        sph.kpg2 = (wl.bk.kpg[0] ** 2).sum(-1)
        gb = G.build_gamma_basis(sph, wl.shape)
        gh = G.GammaHamiltonian(gb, wl.v_r, wl.p[0], wl.bk.dij_full)
        xh = G.full_to_half(gb, x64[0])
        for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
            phi = G.embed_real(gb, xh).to(dt)
            ts = []
            for i in range(args.warmup + args.steps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                gh.apply_real(phi)
                torch.cuda.synchronize()
                if i >= args.warmup:
                    ts.append(time.perf_counter() - t0)
            out[f"{name}_ms"] = statistics.median(ts) * 1e3
    else:
        nocc = int(wl.n_electrons // 2)
        w = torch.zeros(wl.bk.nk, x64.shape[1], dtype=torch.float64, device=x64.device)
        w[:, :nocc] = 2.0 * torch.as_tensor(wl.kweights, device=x64.device)[:, None]
        fargs = (wl.bk, wl.shape, wl.positions, wl.species_index, wl.vloc_tables, wl.g_cart, wl.volume)
        ts = []
        for i in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            band_energy_forces(*fargs, x64, w)
            torch.cuda.synchronize()
            if i >= args.warmup:
                ts.append(time.perf_counter() - t0)
        out["forces_ms"] = statistics.median(ts) * 1e3
    print(json.dumps(out), flush=True)


def bench_ragged(args) -> None:
    """Batching form: 32 independent cells (Al1, Si2, Si8, Si16 at seeded random
    k, 25-40 Ry, 8-64 bands) applied as one ragged batch and solved by lockstep
    Davidson, next to the same work done one cell at a time."""
    import torch

    from gradwave_amd import batching as B
    from gradwave_amd.davidson import davidson_batched
    from gradwave_amd.hamiltonian import BatchedHamiltonian
    from gradwave_amd.workloads import HARTREE_EV, start_vectors

    dev = torch.device("cuda", 0)
    a = time.perf_counter()
    wls = B.ragged_instances(device=dev, n=32, seed=0)
    hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in wls]
    batch = B.RaggedBatch(hams)
    nbt = sum(w.bk.nk * w.nbands for w in wls)
    print(f"[bench] ragged32 setup {time.perf_counter() - a:.1f}s: "
          + ", ".join(w.name for w in wls), file=sys.stderr, flush=True)
    from gradwave_amd import autotune

    t0 = time.perf_counter()
    radix = autotune.tune_radix(wls)
    # no rocBLAS back-projection here: the batched apply is captured in a HIP graph below
    gemm_tile, gemm_wgs = autotune.tune_gemm(wls, allow_library=False)
    t_tune = time.perf_counter() - t0
    hams = [BatchedHamiltonian(w.bk, w.shape, w.v_r, w.p) for w in wls]  # plans with the tuned splits
    batch = B.RaggedBatch(hams)
    print(f"[bench] autotune {t_tune:.1f}s: radix {radix}, GEMM tile {gemm_tile}, split-K target {gemm_wgs} "
          "workgroups",
          file=sys.stderr, flush=True)
    x64 = [start_vectors(w, seed=i) for i, w in enumerate(wls)]
    if args.only == "davidson":
        ragged_davidson_only(args, batch, wls, x64)
        return
    x32 = [x.to(torch.complex64) for x in x64]

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ts = []
        t0 = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t1)
        return (time.perf_counter() - t0) / steps, statistics.median(ts)

    rates = {}
    for name, xs in (("fp32", x32), ("fp64", x64)):
        _, med_b = timed(lambda: batch.apply(xs), args.steps, args.warmup)
        _, med_s = timed(lambda: [h.apply(x) for h, x in zip(hams, xs)], max(5, args.steps // 2), 2)
        outs = [torch.empty_like(x) for x in xs]
        graph = batch.capture(xs, outs)
        _, med_g = timed(graph.replay, args.steps, args.warmup)
        rates[name] = {"batched_ms": med_b * 1e3, "one_by_one_ms": med_s * 1e3, "graph_ms": med_g * 1e3,
                       "batched_band_applies_per_s": nbt / med_b, "one_by_one_band_applies_per_s": nbt / med_s,
                       "graph_band_applies_per_s": nbt / med_g}
        print(f"[bench] ragged apply {name}: batched {med_b * 1e3:.3f} ms, graph {med_g * 1e3:.3f} ms, "
              f"one-by-one {med_s * 1e3:.3f} ms", file=sys.stderr, flush=True)
    if args.no_davidson:
        print(json.dumps({"workload": "ragged32", "apply": rates}), flush=True)
        return
    tol = args.tol_ha * HARTREE_EV
    ts_, ms_ = [w.bk.t for w in wls], [w.bk.mask for w in wls]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = B.RaggedSolveStats(seconds={} if args.profile_davidson else None)
    r64 = B.davidson_ragged(batch, x64, ts_, ms_, tol=tol, max_iter=args.davidson_max_iter, stats=st)
    torch.cuda.synchronize()
    t_lock = time.perf_counter() - t0
    t0 = time.perf_counter()
    st_lib = B.RaggedSolveStats(seconds={} if args.profile_davidson else None)
    rlib = B.davidson_ragged(batch, x64, ts_, ms_, tol=tol, max_iter=args.davidson_max_iter, stats=st_lib,
                             subspace_eigh="library")
    torch.cuda.synchronize()
    t_lock_lib = time.perf_counter() - t0
    t0 = time.perf_counter()
    rms = B.davidson_ragged_mixed(batch, x64, ts_, ms_, tol=tol, draft_tol=args.draft_tol,
                                  max_iter=args.davidson_max_iter)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    t0 = time.perf_counter()
    seq = [davidson_batched(h.apply, x, w.bk.t, w.bk.mask, tol=tol, max_iter=args.davidson_max_iter)
           for h, x, w in zip(hams, x64, wls)]
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    de = max(float((a.eigenvalues - b.eigenvalues).abs().max()) for a, b in zip(r64, seq)) / HARTREE_EV
    de_lib = max(float((a.eigenvalues - b.eigenvalues).abs().max()) for a, b in zip(rlib, seq)) / HARTREE_EV
    dm = max(float((a.eigenvalues - b.eigenvalues).abs().max()) for a, b in zip(rms, seq)) / HARTREE_EV
    t0 = time.perf_counter()
    rsw = B.davidson_ragged_mixed(batch, x64, ts_, ms_, draft_tol=args.draft_tol,
                                  max_iter=args.davidson_max_iter,
                                  sweeps=args.sweeps)
    torch.cuda.synchronize()
    t_sw = time.perf_counter() - t0
    rn_sw = max(float(r.residual_norms.max()) for r in rsw)
    t0 = time.perf_counter()
    req = B.davidson_ragged(batch, x64, ts_, ms_, tol=rn_sw, max_iter=args.davidson_max_iter)
    torch.cuda.synchronize()
    t_eq = time.perf_counter() - t0
    from gradwave_amd import graph_davidson as GD

    gst = GD.GraphSolveStats(seconds={})
    t0 = time.perf_counter()
    solver = GD.RaggedGraphDavidson(batch, ts_, ms_, [int(x.shape[1]) for x in x64])
    rgr = GD.davidson_ragged_graph_mixed(batch, x64, ts_, ms_, draft_tol=args.draft_tol, sweeps=args.sweeps,
                                         max_iter=args.davidson_max_iter, stats=gst, solver=solver)
    torch.cuda.synchronize()
    t_graph_first = time.perf_counter() - t0
    # the same batch solved again: the captured rounds are only replayed
    gst2 = GD.GraphSolveStats(seconds={})
    t0 = time.perf_counter()
    rgr = GD.davidson_ragged_graph_mixed(batch, x64, ts_, ms_, draft_tol=args.draft_tol, sweeps=args.sweeps,
                                         max_iter=args.davidson_max_iter, stats=gst2, solver=solver)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    graph = {"s_per_kpt": t_graph / sum(w.bk.nk for w in wls), "draft_tol_eV": args.draft_tol,
             "sweeps": args.sweeps, "rounds": gst2.rounds, "captured_graphs": gst2.captured_graphs,
             "host_reads": gst2.host_reads, "stage_s": gst2.seconds,
             "first_solve_with_capture_s": t_graph_first, "first_solve_stage_s": gst.seconds,
             "max_residual_eV": max(float(r.residual_norms.max()) for r in rgr),
             "max_eig_diff_vs_one_by_one_Ha": max(float((a.eigenvalues - b.eigenvalues).abs().max())
                                                  for a, b in zip(rgr, seq)) / HARTREE_EV,
             "speedup_vs_one_by_one_fp64": t_seq / t_graph,
             "speedup_vs_lockstep_sweep_form": t_sw / t_graph}
    print(f"[bench] ragged graph-captured draft + {args.sweeps} FP64 sweep(s) {t_graph:.2f}s replayed "
          f"({t_graph_first:.2f}s with capture; {gst2.rounds} rounds, {gst2.captured_graphs} graphs), "
          f"max|de| {graph['max_eig_diff_vs_one_by_one_Ha']:.1e} Ha",
          file=sys.stderr, flush=True)
    sweep = {"sweeps": args.sweeps, "draft_tol_eV": args.draft_tol, "s_per_kpt": t_sw / sum(w.bk.nk for w in wls),
             "max_residual_eV": rn_sw,
             "max_eig_diff_vs_one_by_one_Ha": max(float((a.eigenvalues - b.eigenvalues).abs().max())
                                                  for a, b in zip(rsw, seq)) / HARTREE_EV,
             "speedup_vs_lockstep_fp64": t_lock / t_sw,
             "lockstep_fp64_same_residual_s_per_kpt": t_eq / sum(w.bk.nk for w in wls),
             "speedup_vs_lockstep_fp64_same_residual": t_eq / t_sw}
    print(f"[bench] ragged draft + {args.sweeps} FP64 sweep(s) {t_sw:.2f}s, max|de| "
          f"{sweep['max_eig_diff_vs_one_by_one_Ha']:.1e} Ha; lockstep FP64 to residual {rn_sw:.1e} eV "
          f"{t_eq:.2f}s", file=sys.stderr, flush=True)
    rn = max(float(r.residual_norms.max()) for r in r64 + rms)
    nkp = sum(w.bk.nk for w in wls)
    print(f"[bench] ragged davidson: lockstep fp64 {t_lock:.2f}s ({st.rounds} rounds; library eigensolves "
          f"{t_lock_lib:.2f}s), mixed {t_mixed:.2f}s, one-by-one fp64 {t_seq:.2f}s; max|de| {de:.1e} / {dm:.1e} Ha",
          file=sys.stderr, flush=True)
    result = {
        "metric": METRIC,
        "value": rates["fp32"]["batched_band_applies_per_s"], "unit": "band-applies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": rates["fp32"]["batched_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (analytic Gaussian pseudopotential, seeded cells, k-points and start vectors)",
        "config": {"workload": "ragged32: Al1/Si2/Si8/Si16 x 8, 25-40 Ry, seeded k, 8-64 bands",
                   "instances": len(wls), "band_rows": nbt, "shapes": [list(w.shape) for w in wls]},
        "apply": rates,
        "autotune": {"seconds": t_tune, "radix_first_factor": {str(k): v for k, v in radix.items()},
                     "gemm_tile_variant": gemm_tile, "gemm_split_k_target_wgs": gemm_wgs},
        "davidson": {"lockstep_fp64_s_per_kpt": t_lock / nkp, "lockstep_mixed_s_per_kpt": t_mixed / nkp,
                     "one_by_one_fp64_s_per_kpt": t_seq / nkp, "rounds": st.rounds,
                     "lockstep_speedup_vs_one_by_one": t_seq / t_lock,
                     "lockstep_phase_s": st.seconds,
                     "subspace_eigensolve": "gw_heev_batched: one launch per round for every instance's matrix",
                     "lockstep_library_eigh_s_per_kpt": t_lock_lib / nkp,
                     "lockstep_library_eigh_phase_s": st_lib.seconds,
                     "max_eig_diff_lockstep_library_eigh_vs_one_by_one_Ha": de_lib,
                     "max_eig_diff_lockstep_vs_one_by_one_Ha": de, "max_eig_diff_mixed_vs_one_by_one_Ha": dm,
                     "max_residual_eV": rn, "tol_Ha": args.tol_ha, "sweep_form": sweep,
                     "graph_captured_rounds": graph},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def ragged_davidson_only(args, batch, wls, x64) -> None:
    """Profiling run of the batching form's solves: the lockstep FP64 solve,
    then the graph-captured draft + FP64 sweep twice on one solver (the second
    only replays)."""
    import torch

    from gradwave_amd import batching as B
    from gradwave_amd import graph_davidson as GD
    from gradwave_amd.workloads import HARTREE_EV

    ts_, ms_ = [w.bk.t for w in wls], [w.bk.mask for w in wls]
    out = {"workload": "ragged32", "only": "davidson"}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = B.RaggedSolveStats()
    B.davidson_ragged(batch, x64, ts_, ms_, tol=args.tol_ha * HARTREE_EV, max_iter=args.davidson_max_iter, stats=st)
    torch.cuda.synchronize()
    out["lockstep_fp64_s"], out["lockstep_rounds"] = time.perf_counter() - t0, st.rounds
    solver = GD.RaggedGraphDavidson(batch, ts_, ms_, [int(x.shape[1]) for x in x64])
    for key in ("graph_first_s", "graph_replayed_s"):
        t0 = time.perf_counter()
        GD.davidson_ragged_graph_mixed(batch, x64, ts_, ms_, draft_tol=args.draft_tol, sweeps=args.sweeps,
                                       max_iter=args.davidson_max_iter, solver=solver)
        torch.cuda.synchronize()
        out[key] = time.perf_counter() - t0
    print(f"[bench] ragged davidson only: {out}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def copy_bandwidth(dev, nbytes: int = 2 << 30, reps: int = 10) -> float:
    """Bytes read + written per second by a plain device-to-device copy of
    ``nbytes`` (torch copy_), timed with HIP events: the bandwidth a streaming
    kernel can reach on this box, beside the published 8 TB/s."""
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * nbytes / t / 1e9


def toeplitz_rates(args, wl, x32, x64) -> dict:
    """Small-cell local term as a matrix (core/batch.py:333-464): full apply
    band-applies/s with the path forced on, next to the FFT path."""
    import torch

    from gradwave_amd.hamiltonian import BatchedHamiltonian

    ht = BatchedHamiltonian(wl.bk, wl.shape, wl.v_r, wl.p, toeplitz="on")
    out = {}
    for name, x in (("fp32", x32), ("fp64", x64)):
        for _ in range(3):
            ht.apply(x)
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(5, args.steps // 2)):
            a = time.perf_counter()
            ht.apply(x)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        ms = statistics.median(ts)
        out[f"{name}_band_applies_per_s"] = x.shape[0] * x.shape[1] / ms
        out[f"{name}_ms_per_step"] = ms * 1e3
    print(f"[bench] toeplitz apply fp32 {out['fp32_ms_per_step']:.3f} ms fp64 {out['fp64_ms_per_step']:.3f} ms",
          file=sys.stderr, flush=True)
    return out


def sweep_form(args, mixed_sweep, fp64_to, x64, eig64, t64: float, nk: int):
    """The precision form as a draft plus ``args.sweeps`` final FP64 refinement
    sweeps (no tolerance gate): time, eigenvalue error against the converged FP64
    solve, the FP64 residual it reached, and the FP64 solve stopped at that same
    residual, so the speedup is stated both against the converged FP64 solve and
    at equal residual."""
    import torch

    from gradwave_amd.workloads import HARTREE_EV

    torch.cuda.synchronize()
    a = time.perf_counter()
    rsw = mixed_sweep(x64)
    torch.cuda.synchronize()
    tsw = time.perf_counter() - a
    rmax = float(rsw.residual_norms.max())
    a = time.perf_counter()
    req = fp64_to(x64, rmax)
    torch.cuda.synchronize()
    teq = time.perf_counter() - a
    out = {"sweeps": args.sweeps, "draft_tol_eV": args.sweep_draft_tol, "buffer_bands": args.sweep_buffer,
           "s_per_kpt": tsw / nk,
           "draft_iters": getattr(rsw, "n_draft_iter", None), "max_residual_eV": rmax,
           "max_eig_diff_vs_fp64_Ha": float((rsw.eigenvalues - eig64).abs().max()) / HARTREE_EV,
           "speedup_vs_fp64": t64 / tsw, "fp64_same_residual_s_per_kpt": teq / nk,
           "fp64_same_residual_iters": req.n_iter,
           "max_eig_diff_fp64_same_residual_vs_fp64_Ha": float((req.eigenvalues - eig64).abs().max()) / HARTREE_EV,
           "speedup_vs_fp64_same_residual": teq / tsw}
    print(f"[bench] davidson draft + {args.sweeps} FP64 sweep(s) {tsw:.2f}s ({out['draft_iters']} draft it), "
          f"max|de| {out['max_eig_diff_vs_fp64_Ha']:.1e} Ha, residual {rmax:.1e} eV; FP64 to that residual "
          f"{teq:.2f}s ({req.n_iter} it)", file=sys.stderr, flush=True)
    return out, rsw


def gamma_form(args, wl, x64, r64) -> dict:
    """Gamma real layout (core/gamma.py): half-sphere apply at two bands per
    complex apply, real-symmetric Davidson (FP64 and mixed) from the same start
    vectors; eigenvalues against the complex FP64 solve."""
    import torch

    from gradwave_amd import gamma as G
    from gradwave_amd.workloads import HARTREE_EV

    sph = type("Sphere", (), {})()
    sph.flat_idx = wl.bk.flat_idx[0]
    sph.kpg2 = (wl.bk.kpg[0] ** 2).sum(-1)
    gb = G.build_gamma_basis(sph, wl.shape)
    gh = G.GammaHamiltonian(gb, wl.v_r, wl.p[0], wl.bk.dij_full)
    def half(x):
        v = G.full_to_half(gb, x)
        v[:, 0] = v[:, 0].real.to(v.dtype)
        return v

    xh = half(x64[0])
    nb = xh.shape[0]
    out = {}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        phi = G.embed_real(gb, xh).to(dt)
        for _ in range(3):
            gh.apply_real(phi)
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(5, args.steps // 2)):
            a = time.perf_counter()
            gh.apply_real(phi)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        ms = statistics.median(ts)
        out[f"{name}_band_applies_per_s"] = nb / ms
        out[f"{name}_ms_per_step"] = ms * 1e3
    tol = args.tol_ha * HARTREE_EV
    prof64, profms = {}, {}
    torch.cuda.synchronize()
    a = time.perf_counter()
    g64 = G.davidson_gamma(gh, xh, tol=tol, max_iter=args.davidson_max_iter,
                           profile_out=prof64 if args.profile_davidson else None)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - a
    a = time.perf_counter()
    gms = G.davidson_gamma_mixed(gh, xh, tol=tol, draft_tol=args.draft_tol, max_iter=args.davidson_max_iter,
                                 profile_out=profms if args.profile_davidson else None)
    torch.cuda.synchronize()
    tms = time.perf_counter() - a
    ref = r64.eigenvalues[0]
    out["sweep_form"], _ = sweep_form(
        args, lambda x: G.davidson_gamma_mixed(gh, x, draft_tol=args.sweep_draft_tol, max_iter=args.davidson_max_iter,
                                               sweeps=args.sweeps, buffer=args.sweep_buffer),
        lambda x, tl: G.davidson_gamma(gh, x, tol=tl, max_iter=args.davidson_max_iter), xh, g64.eigenvalues, t64, 1)
    out.update({
        "fp64_s_per_kpt": t64, "mixed_s_per_kpt": tms, "fp64_iters": g64.n_iter,
        "mixed_draft_iters": gms.n_draft_iter, "mixed_polish_iters": gms.n_iter,
        "fp64_max_residual_eV": float(g64.residual_norms.max()),
        "mixed_max_residual_eV": float(gms.residual_norms.max()),
        "max_eig_diff_fp64_vs_complex_Ha": float((g64.eigenvalues - ref).abs().max()) / HARTREE_EV,
        "max_eig_diff_mixed_vs_complex_Ha": float((gms.eigenvalues - ref).abs().max()) / HARTREE_EV,
        "phase_s_fp64": prof64 or None, "phase_s_mixed": profms or None})
    print(f"[bench] gamma apply fp32 {out['fp32_ms_per_step']:.3f} ms fp64 {out['fp64_ms_per_step']:.3f} ms; "
          f"davidson fp64 {t64:.2f}s ({g64.n_iter} it) mixed {tms:.2f}s; max|de| "
          f"{out['max_eig_diff_fp64_vs_complex_Ha']:.1e} / {out['max_eig_diff_mixed_vs_complex_Ha']:.1e} Ha",
          file=sys.stderr, flush=True)
    return out


def forces_check(wl, psi64, psims, psisw=None) -> dict:
    """Band-energy forces at fixed orbitals (adjoint kernels) from the FP64, the
    mixed (polished) and the sweep-form solves' orbitals: time and max |dF| in
    Ha/bohr."""
    import torch

    from gradwave_amd.adjoint import band_energy_forces
    from gradwave_amd.workloads import BOHR_ANG, HARTREE_EV

    nocc = int(wl.n_electrons // 2)
    w = torch.zeros(wl.bk.nk, psi64.shape[1], dtype=torch.float64, device=psi64.device)
    w[:, :nocc] = 2.0 * torch.as_tensor(wl.kweights, device=psi64.device)[:, None]
    args = (wl.bk, wl.shape, wl.positions, wl.species_index, wl.vloc_tables, wl.g_cart, wl.volume)
    band_energy_forces(*args, psi64, w)  # warm-up (plans, workspaces)
    torch.cuda.synchronize()
    a = time.perf_counter()
    e64, f64 = band_energy_forces(*args, psi64, w)
    torch.cuda.synchronize()
    t = time.perf_counter() - a
    _, fms = band_energy_forces(*args, psims, w)
    to_ha_bohr = BOHR_ANG / HARTREE_EV
    df = float((fms - f64).abs().max()) * to_ha_bohr
    dsw = None
    if psisw is not None:
        _, fsw = band_energy_forces(*args, psisw, w)
        dsw = float((fsw - f64).abs().max()) * to_ha_bohr
    print(f"[bench] forces {3 * f64.shape[0]} coords {t * 1e3:.1f} ms, max|dF| mixed vs fp64 {df:.2e} Ha/bohr, "
          f"sweep form vs fp64 {dsw} Ha/bohr", file=sys.stderr, flush=True)
    return {"coords": 3 * int(f64.shape[0]), "s": t, "max_abs_Ha_per_bohr": float(f64.abs().max()) * to_ha_bohr,
            "max_diff_mixed_vs_fp64_Ha_per_bohr": df, "max_diff_sweep_form_vs_fp64_Ha_per_bohr": dsw,
            "band_energy_fixed_psi_eV": float(e64)}


STAGE_KERNEL = {"plane_inv": "k_plane_inv_w", "xpass": "k_xpass_wave", "plane_fwd": "k_plane_fwd_w",
                "becp": "k_gemm_inner_lds", "dmul": "k_dmul", "backproj": "k_gemm_combine_lds"}
PROFILE_OF = {"si64_gamma": "hpsi_apply_si64_gamma_90cubed.txt",
              "si216_gamma": "hpsi_apply_si216_gamma_120cubed.txt"}


# complex-accumulator GEMM variants: (inner BM, BN), (combine BM, BN), Gauss form
CTILE = {6: ((2, 2), (2, 2), False), 7: ((3, 3), (3, 2), False), 8: ((2, 5), (5, 2), False),
         9: ((4, 3), (3, 4), False), 10: ((2, 2), (2, 2), True), 11: ((3, 2), (3, 2), True),
         12: ((2, 3), (2, 3), True), 13: ((3, 3), (5, 2), True)}


def gemm_mfma_fraction(stage: str, real: str) -> float:
    """Real MFMA flops per 8 algorithmic flops of the projector contraction that
    runs ``stage``: 0.75 for the three-product tiles, 1 otherwise."""
    from gradwave_amd import _lib as L

    prec = L.GW_FP32 if real == "float" else L.GW_FP64
    lib_key, op = (L.GW_TUNE_INNER_LIB, 256) if stage == "becp" else (L.GW_TUNE_COMBINE_LIB, 512)
    if L.lib().gw_tune_get(lib_key, prec):
        return 1.0
    tile = L.lib().gw_tune_get(L.GW_TUNE_GEMM_TILE, prec + op)
    return 0.75 if tile in CTILE and CTILE[tile][2] else 1.0


def stage_kernel(stage: str, real: str, workload: str = "") -> str:
    """Kernel that runs a stage: the x pass is k_xpass_stage when the launch shape
    of that precision stages the planes through LDS (mode digit of GW_TUNE_XPASS);
    a projector contraction routed to rocBLAS by the autotuner runs its GEMM."""
    from gradwave_amd import _lib as L

    prec = L.GW_FP32 if real == "float" else L.GW_FP64
    if stage == "xpass" and L.lib().gw_tune_get(L.GW_TUNE_XPASS, prec) // 10000 == 1:
        return "k_xpass_stage"
    if stage == "becp" and L.lib().gw_tune_get(L.GW_TUNE_INNER_LIB, prec):
        return "rocBLAS " + ("zgemm" if prec == L.GW_FP64 else "cgemm") + " strided batched (Cijk_*)"
    if stage == "backproj" and L.lib().gw_tune_get(L.GW_TUNE_COMBINE_LIB, prec):
        return "rocBLAS " + ("zgemm" if prec == L.GW_FP64 else "cgemm") + " strided batched (Cijk_*)"
    if stage in ("becp", "backproj"):
        tile = L.lib().gw_tune_get(L.GW_TUNE_GEMM_TILE, prec + (256 if stage == "becp" else 512))
        if tile in CTILE:
            inner, comb, g3 = CTILE[tile]
            bm, bn = inner if stage == "becp" else comb
            return (f"k_cgemm_{'inner' if stage == 'becp' else 'combine'}<{real}, {bm}, {bn}, "
                    f"{'true' if g3 else 'false'}{'' if stage == 'becp' else ', ...'}>")
    return STAGE_KERNEL[stage]


def profiled_traffic(workload: str, stage: str, real: str):
    """(HBM bytes per launch, source) of the stage's kernel: calibrated
    FETCH_SIZE + WRITE_SIZE from a rocprofv3 --pmc run of the same workload,
    read from the summary committed under profiles/ (a line
    ``traffic <stage> <fp64|fp32> <MB read> <MB written>``), or (None, None)."""
    name = PROFILE_OF.get(workload)
    if name is None:
        return None, None
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None, None
    prec = "fp64" if real == "double" else "fp32"
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0] == "traffic" and f[1] == stage and f[2] == prec:
            try:
                return (float(f[3]) + float(f[4])) * 1e6, f"profiles/{name} (committed rocprofv3 --pmc figure)"
            except ValueError:
                return None, None
    return None, None


def cpu_baseline(args, wl, gpu_fp64_iters: int | None = None) -> dict:
    """The oracle's BatchedHamiltonian.apply (core/batch.py:480-505 restated, torch CPU
    complex128) on a bounded band sample of the same workload."""
    import torch

    from oracle.hamiltonian import BatchedHamiltonian as RefH
    from oracle.hamiltonian import BatchedK

    cores = min(os.cpu_count() or 1, 8)
    torch.set_num_threads(cores)
    bk = wl.bk
    cpu = BatchedK(npw=bk.npw.cpu(), mask=bk.mask.cpu(), flat_idx=bk.flat_idx.cpu(), kpg=bk.kpg.cpu(),
                   t=bk.t.cpu(), proj_phase_free=wl.p.cpu(), proj_atom_index=bk.proj_atom_index.cpu(),
                   dij_full=bk.dij_full.cpu())
    h = RefH(cpu, wl.shape, wl.v_r.cpu(), wl.p.cpu())
    nbs = min(args.cpu_bands, wl.nbands)
    g = torch.Generator().manual_seed(1)
    c = torch.randn(bk.nk, nbs, bk.npw_max, dtype=torch.complex128, generator=g) * cpu.mask[:, None, :]
    def median_apply(seconds):
        h.apply(c)  # warm-up
        ts = []
        budget = time.perf_counter() + seconds
        while len(ts) < 5 or (len(ts) < 400 and time.perf_counter() < budget):
            a = time.perf_counter()
            h.apply(c)
            ts.append(time.perf_counter() - a)
        return statistics.median(ts), len(ts)

    med, n = median_apply(12.0)
    out = {"value": bk.nk * nbs / med, "unit": "band-applies/s", "cores": cores, "kind": "port",
           "sample": f"{n} applies of a ({bk.nk}, {nbs}, {bk.npw_max}) complex128 block, median",
           "cpu_model": _cpu_model()}
    all_cores = os.cpu_count() or 1
    if all_cores > cores:
        torch.set_num_threads(all_cores)
        med_all, _ = median_apply(6.0)
        out["value_all_cores"] = bk.nk * nbs / med_all
        out["cores_all"] = all_cores
        torch.set_num_threads(cores)
    from oracle.davidson import davidson_batched as ref_davidson
    from gradwave_amd.workloads import HARTREE_EV, start_vectors

    x0 = start_vectors(wl, seed=0).cpu()
    if bk.nk * wl.nbands * bk.npw_max <= 200_000:  # small cells: the oracle's full Davidson solve
        a = time.perf_counter()
        r = ref_davidson(h.apply, x0, cpu.t, cpu.mask, tol=args.tol_ha * HARTREE_EV,
                         max_iter=args.davidson_max_iter)
        t = time.perf_counter() - a
        out["davidson_fp64_s_per_kpt"] = t / bk.nk
        out["davidson_iters"] = r.n_iter
    elif gpu_fp64_iters and args.cpu_davidson_iters > 0:
        # large cells: time the oracle's first iterations from the same start block and
        # extrapolate to the GPU FP64 solve's iteration count (an ESTIMATE: later
        # iterations carry a larger basis, so it undercounts the CPU time)
        marks = _Marks()
        a = time.perf_counter()
        ref_davidson(h.apply, x0, cpu.t, cpu.mask, tol=args.tol_ha * HARTREE_EV,
                     max_iter=args.cpu_davidson_iters, history_out=marks)
        total = time.perf_counter() - a
        steps = [t1 - t0 for t0, t1 in zip([a] + marks.t[:-1], marks.t)]
        first, rest = steps[0], steps[1:] or steps
        per_iter = statistics.mean(rest)
        est = first + per_iter * (gpu_fp64_iters - 1)
        out["davidson_fp64_s_per_kpt_estimate"] = est / bk.nk
        out["davidson_estimate"] = (f"oracle davidson_batched from the seeded start block, {len(steps)} iterations "
                                    f"timed ({total:.1f} s: first {first:.1f} s incl. setup, then "
                                    f"{per_iter:.1f} s each) and extrapolated to the GPU FP64 solve's "
                                    f"{gpu_fp64_iters} iterations")
    return out


class _Marks(list):
    """history_out sink that stamps the wall time of every finished iteration."""

    def __init__(self):
        super().__init__()
        self.t = []

    def append(self, item):
        self.t.append(time.perf_counter())
        super().append(item)


def _cpu_model() -> str:
    """/proc/cpuinfo model name with its family/model numbers (the name alone can be generic)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = line.split(":", 1)
                info.setdefault(k.strip(), v.strip())
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    return f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')})"


if __name__ == "__main__":
    main()
