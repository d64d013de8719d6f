"""Drop-in ``BatchedHamiltonian`` for gradwave's batched H|psi> on MI355X.

Mirrors core/batch.py:238-559 (constructor signature, ``apply``, ``apply_cols``,
attributes ``bk``, ``shape``, ``v_eff_r``, dtype polymorphism over complex64 /
complex128 per core/batch.py:306-331) and the band*k apply tally of
core/batch.py:111-123. All arithmetic runs in libgradwave_hip (pruned LDS box
FFTs with fused kinetic/local terms, MFMA projector contractions); this module
only owns the plan and the device workspace.
"""

from __future__ import annotations

import ctypes as C
import os
import time

import torch

from gradwave_amd import _lib as L

_HAPPLY_TALLY = {"on": False, "count": 0}

# Small-cell local term as one dense GEMM per k (core/batch.py:48-107, 333-464):
# out(G_i) = sum_j Vhat(G_i - G_j) c(G_j). GRADWAVE_TOEPLITZ in {auto, on, off}
# as in the reference; like the reference on CUDA, "auto" keeps the FFT path on
# the GPU unless GW_TOEPLITZ_ON_GPU=1 lets the timing trial decide (adopt when
# >= 30 % faster). The matrix cap is larger than the reference's 256 MiB because
# an MI355X has 288 GB of HBM.
_TOEPLITZ_MODE = os.environ.get("GRADWAVE_TOEPLITZ", "auto").strip().lower()
_TOEPLITZ_ON_GPU = os.environ.get("GW_TOEPLITZ_ON_GPU", "0") == "1"
_TOEPLITZ_M_BUDGET_BYTES = int(os.environ.get("GW_TOEPLITZ_BUDGET", str(4 << 30)))
_TOEP_TRIAL_MARGIN = 0.7
_TOEP_TRIAL: dict[tuple, bool] = {}
# Device bytes one band chunk of the library-FFT local term may hold (box,
# psi(r) and the transformed box).
_LIB_FFT_BUDGET_BYTES = int(os.environ.get("GW_LIB_FFT_BUDGET", str(2 << 30)))


def reset_happly_tally() -> None:
    """core/batch.py:111-118."""
    _HAPPLY_TALLY["on"] = True
    _HAPPLY_TALLY["count"] = 0


def happly_tally() -> int:
    """core/batch.py:120-123."""
    return _HAPPLY_TALLY["count"]


class Plan:
    """Owner of a gw_plan handle: geometry from BatchedK maps plus the tables."""

    def __init__(self, shape, flat_idx: torch.Tensor, npw: torch.Tensor) -> None:
        n1, n2, n3 = (int(s) for s in shape)
        flat = flat_idx.detach().to("cpu", torch.int64).contiguous()
        npw_h = npw.detach().to("cpu", torch.int64).contiguous()
        nk, npw_max = flat.shape
        self.handle = C.c_void_p()
        self._lib = L.lib()
        L.check(self._lib.gw_plan_create(C.byref(self.handle), n1, n2, n3, nk, npw_max,
                                         flat.data_ptr(), npw_h.data_ptr()), "gw_plan_create")
        self.shape = (n1, n2, n3)
        self.nk, self.npw_max = nk, npw_max
        info = (C.c_int * 15)()
        L.check(self._lib.gw_plan_info(self.handle, L.GW_FP64, info, 15), "gw_plan_info")
        # 0, or a box length with no codelet split: no box FFT kernels on this plan
        self.nosplit_len = int(info[14])

    def set_kinetic(self, t: torch.Tensor) -> None:
        t = t.detach().to(torch.float64).contiguous()
        L.require_gpu(t, "t")
        L.check(self._lib.gw_plan_set_kinetic(self.handle, t.data_ptr(), L.stream_ptr(t.device)),
                "gw_plan_set_kinetic")
        self._t = t  # keep alive until the async copy has run

    def set_potential(self, v_r: torch.Tensor) -> None:
        v = v_r.detach().to(torch.float64).contiguous()
        L.require_gpu(v, "v_eff_r")
        L.check(self._lib.gw_plan_set_potential(self.handle, v.data_ptr(), L.stream_ptr(v.device)),
                "gw_plan_set_potential")
        self._v = v

    def set_projectors(self, p: torch.Tensor, dij: torch.Tensor) -> None:
        nproj = int(p.shape[1])
        if nproj:
            pp = p.detach().to(torch.complex128).contiguous()
            L.require_gpu(pp, "p")
            d = dij.detach().to("cpu", torch.float64).contiguous()
            L.check(self._lib.gw_plan_set_projectors(self.handle, nproj, pp.data_ptr(), d.data_ptr(),
                                                     L.stream_ptr(pp.device)),
                    "gw_plan_set_projectors")
            self._p = pp
        else:
            L.check(self._lib.gw_plan_set_projectors(self.handle, 0, None, None, None),
                    "gw_plan_set_projectors")
        self.nproj = nproj

    def __del__(self) -> None:
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.gw_plan_destroy(h)
            except Exception:
                pass
            self.handle = C.c_void_p()


class BatchedHamiltonian:
    """H apply for all k at once (core/batch.py:238-505), fixed V_eff(r) and projectors.

    ``bk`` needs the BatchedK fields npw, mask, flat_idx, t and dij_full
    (core/batch.py:126-171); ``v_eff_r`` (n1,n2,n3) real and ``p``
    (nk, nproj, npw_max) complex live on the GPU.

    ``smooth = (shape_s, flat_idx_s, v_s)`` runs the local term on that smooth
    box with its own sphere map (the USPP/PAW dual grid, batch.py:258-266; the
    kinetic and nonlocal terms are sphere-based and unchanged), and then
    ``shape`` / ``v_eff_r`` are the smooth box's, as in the reference.
    ``hub_q`` (nk, nproj_U, npw_max) with a real ``hub_dij`` (nproj_U, nproj_U)
    adds the DFT+U term (batch.py:272-274, 501-504): the plan carries the KB and
    Hubbard projectors stacked, with D block-diagonal, so both run in the same
    MFMA contractions. A complex hub_dij is refused.

    Boxes with a length that has no split into two codelets <= 16 (125, 162,
    ..., every length above 256; good_fft_size returns them, grids.py:32-41)
    have no box FFT kernels here: their local term is the Toeplitz contraction
    (M built by gw_toeplitz_build, applied by the MFMA inner product) when the
    matrix fits the memory cap, else scatter -> rocFFT inverse -> x V ->
    rocFFT forward -> gather (torch.fft on the GPU; the scatter and gather are
    the plan's kernels). psi(r), the density and the V(r) adjoint of the
    forces take the scatter and rocFFT there too; ragged batches refuse such
    boxes with GWError.
    """

    def __init__(self, bk, shape, v_eff_r: torch.Tensor, p: torch.Tensor, hub_q=None,
                 hub_dij=None, smooth=None, toeplitz: str | None = None) -> None:
        mode = (toeplitz or _TOEPLITZ_MODE).strip().lower()
        if mode not in ("auto", "on", "off"):
            raise ValueError(f"toeplitz mode must be auto, on or off, got {mode!r}")
        L.require_gpu(bk.mask, "bk.mask")
        self.bk = bk
        flat_idx = bk.flat_idx
        if smooth is not None:
            shape, flat_idx, v_eff_r = smooth
        self.shape = tuple(int(s) for s in shape)
        self.n = self.shape[0] * self.shape[1] * self.shape[2]
        self.v_eff_r = v_eff_r
        self.p = p
        self.hub_q, self.hub_dij = hub_q, hub_dij
        proj, dij = p, bk.dij_full
        if hub_q is not None and hub_dij is not None:
            hd = hub_dij
            if torch.is_complex(hd):
                if float(hd.imag.abs().max()) != 0.0:
                    raise NotImplementedError("a complex DFT+U D matrix is not supported (real D only)")
                hd = hd.real
            proj = torch.cat([p, hub_q.to(p.dtype)], dim=1)
            dij = torch.block_diag(bk.dij_full.to(torch.float64), hd.to(bk.dij_full.device, torch.float64))
        self.plan = Plan(self.shape, flat_idx, bk.npw)
        self.plan.set_kinetic(bk.t)
        self.plan.set_potential(v_eff_r)
        self.plan.set_projectors(proj, dij)
        self._ws: dict[int, torch.Tensor] = {}
        self.device = bk.mask.device
        nk, npw_max = self.plan.nk, self.plan.npw_max
        self._toep_mode = mode
        # the Toeplitz path reads the dense sphere map (ineligible with a dual grid, batch.py:288-296)
        toep_fits = smooth is None and nk * npw_max * npw_max * 16 <= _TOEPLITZ_M_BUDGET_BYTES
        self._toep_eligible = mode != "off" and toep_fits
        self._toep_M: dict[torch.dtype, torch.Tensor] = {}
        # local term of a box without FFT kernels: Toeplitz unless switched off or too large
        self.local_path = "fft" if not self.plan.nosplit_len else ("toeplitz" if self._toep_eligible
                                                                   else "library_fft")
        self._v_lib: dict[torch.dtype, torch.Tensor] = {}

    def _workspace(self, nb: int, prec: int) -> torch.Tensor:
        need = int(L.lib().gw_hpsi_workspace_bytes(self.plan.handle, nb, prec))
        ws = self._ws.get(prec)
        if ws is None or ws.numel() < need:
            ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
            self._ws[prec] = ws
        return ws

    def apply_terms(self, c: torch.Tensor, terms: int = L.GW_TERM_ALL,
                    out: torch.Tensor | None = None) -> torch.Tensor:
        if c.dim() != 3 or c.shape[0] != self.plan.nk or c.shape[2] != self.plan.npw_max:
            raise ValueError(f"c must be (nk={self.plan.nk}, nb, npw_max={self.plan.npw_max}), "
                             f"got {tuple(c.shape)}")
        L.require_gpu(c, "c")
        prec = L.precision_of(c.dtype)
        if c.stride(2) != 1:
            c = c.contiguous()
        nb = int(c.shape[1])
        if out is None:
            out = torch.empty(c.shape, dtype=c.dtype, device=c.device)
        if nb == 0:
            return out
        if (terms & L.GW_TERM_LOCAL) and self.plan.nosplit_len:
            rest = terms & ~L.GW_TERM_LOCAL
            if rest:
                self.apply_terms(c, rest, out=out)
            else:
                out.zero_()
            out += self._local_without_box_fft(c)
            return out
        ws = self._workspace(nb, prec)
        L.check(L.lib().gw_hpsi_apply_strided(
            self.plan.handle, prec, terms, c.data_ptr(), c.stride(1), c.stride(0),
            out.data_ptr(), out.stride(1), out.stride(0), nb, ws.data_ptr(), ws.numel(),
            L.stream_ptr(c.device)), "gw_hpsi_apply")
        return out

    def apply(self, c: torch.Tensor) -> torch.Tensor:
        """(nk, nb, npw_max) -> H c, mask preserved (core/batch.py:480-505)."""
        if _HAPPLY_TALLY["on"]:
            _HAPPLY_TALLY["count"] += int(c.shape[0] * c.shape[1])
        if self._use_toeplitz(c):
            out = self.apply_terms(c, L.GW_TERM_KINETIC | L.GW_TERM_NONLOCAL)
            out += self.local_toeplitz(c)
            return out
        return self.apply_terms(c, L.GW_TERM_ALL)

    def _local_without_box_fft(self, c: torch.Tensor) -> torch.Tensor:
        if self.local_path == "toeplitz":
            return self.local_toeplitz(c)
        return self._local_library_fft(c)

    def _local_library_fft(self, c: torch.Tensor) -> torch.Tensor:
        """V psi on a box the FFT kernels cannot split (batch.py:405-428): the
        plan's scatter, rocFFT ifftn, x V(r), rocFFT fftn, the plan's gather
        (padding 0), in band chunks of at most GW_LIB_FFT_BUDGET bytes."""
        nk, nb, m = (int(x) for x in c.shape)
        prec = L.precision_of(c.dtype)
        rdt = torch.float64 if c.dtype == torch.complex128 else torch.float32
        v = self._v_lib.get(rdt)
        if v is None:
            v = self.v_eff_r.to(self.device, rdt).contiguous()
            self._v_lib[rdt] = v
        lib, st = L.lib(), L.stream_ptr(c.device)
        res = torch.empty(nk, nb, m, dtype=c.dtype, device=c.device)
        step = max(1, min(nb, _LIB_FFT_BUDGET_BYTES // (3 * nk * self.n * c.element_size())))
        for b0 in range(0, nb, step):
            nbs = min(step, nb - b0)
            cb = c[:, b0:b0 + nbs].contiguous()
            box = torch.empty(nk, nbs, *self.shape, dtype=c.dtype, device=c.device)
            L.check(lib.gw_sphere_to_box(self.plan.handle, prec, cb.data_ptr(), box.data_ptr(), nbs, st),
                    "gw_sphere_to_box")
            psi = torch.fft.ifftn(box, dim=(-3, -2, -1))
            del box
            psi *= v
            hb = torch.fft.fftn(psi, dim=(-3, -2, -1)).contiguous()
            del psi
            ob = res[:, b0:b0 + nbs] if nbs == nb else torch.empty(nk, nbs, m, dtype=c.dtype, device=c.device)
            L.check(lib.gw_box_to_sphere(self.plan.handle, prec, hb.data_ptr(), ob.data_ptr(), nbs, st),
                    "gw_box_to_sphere")
            if nbs != nb:
                res[:, b0:b0 + nbs] = ob
        return res

    # -- Toeplitz local term (core/batch.py:333-464) --

    def toeplitz_matrix(self, cdtype: torch.dtype) -> torch.Tensor:
        """M[k][i][j] = Vhat(G_i - G_j), Vhat = fftn(v_eff) / N, zero in padding;
        cached per precision for the operator's lifetime (batch.py:380-393)."""
        M = self._toep_M.get(cdtype)
        if M is None:
            n1, n2, n3 = self.shape
            vhat = (torch.fft.fftn(self.v_eff_r.to(self.device, torch.float64)) / self.n).to(cdtype).contiguous()
            nk, m = self.plan.nk, self.plan.npw_max
            M = torch.empty(nk, m, m, dtype=cdtype, device=self.device)
            npw32 = self.bk.npw.to(self.device, torch.int32).contiguous()
            flat32 = self.bk.flat_idx.to(self.device, torch.int32).contiguous()
            L.check(L.lib().gw_toeplitz_build(L.precision_of(cdtype), nk, m, n1, n2, n3, npw32.data_ptr(),
                                              flat32.data_ptr(), vhat.data_ptr(), M.data_ptr(),
                                              L.stream_ptr(self.device)), "gw_toeplitz_build")
            self._toep_M[cdtype] = M
        return M

    def local_toeplitz(self, c: torch.Tensor) -> torch.Tensor:
        """Local V psi as one MFMA contraction per k: out[b][i] = sum_j M[i][j] c[b][j]
        (batch.py:395-403), computed as inner(conj(c), M); padded input ignored."""
        from gradwave_amd import linalg as LA

        M = self.toeplitz_matrix(c.dtype)
        cm = (c * self.bk.mask[:, None, :]).conj().resolve_conj()
        return LA.inner(cm, M)

    def _use_toeplitz(self, c: torch.Tensor) -> bool:
        """The reference's gate (batch.py:430-464): "on" forces the path within the
        memory cap; "auto" times both local terms once per (nk, npw_max, shape,
        dtype) signature and adopts Toeplitz only when >= 30 % faster -- on the
        GPU only with GW_TOEPLITZ_ON_GPU=1, as the reference keeps it off on CUDA."""
        if not self._toep_eligible:
            return False
        if self._toep_mode == "on":
            return True
        if not _TOEPLITZ_ON_GPU:
            return False
        key = (self.plan.nk, self.plan.npw_max, self.shape, c.dtype)
        use_toep = _TOEP_TRIAL.get(key)
        if use_toep is None:
            def timed(f):
                f()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                f()
                torch.cuda.synchronize()
                return time.perf_counter() - t0

            t_fft = timed(lambda: self.apply_terms(c, L.GW_TERM_LOCAL))
            t_toep = timed(lambda: self.local_toeplitz(c))
            use_toep = t_toep < _TOEP_TRIAL_MARGIN * t_fft
            _TOEP_TRIAL[key] = use_toep
            if not use_toep:
                self._toep_M.clear()
        return use_toep

    __call__ = apply

    def apply_cols(self, c: torch.Tensor, kcol: torch.Tensor) -> torch.Tensor:
        """H on a compacted column batch (ncol, npw_max) with k index kcol (core/batch.py:507-547).

        The columns are grouped by k into a padded (nk, ncol_k_max, npw_max)
        block, applied, and scattered back."""
        if _HAPPLY_TALLY["on"]:
            _HAPPLY_TALLY["count"] += int(c.shape[0])
        ncol = int(c.shape[0])
        out = torch.zeros_like(c)
        if ncol == 0:
            return out
        kc = kcol.to(c.device)
        counts = torch.bincount(kc, minlength=self.plan.nk)
        width = int(counts.max())
        order = torch.argsort(kc, stable=True)
        starts = torch.cumsum(counts, 0) - counts
        slot = torch.arange(ncol, device=c.device) - starts[kc[order]]
        blk = torch.zeros(self.plan.nk, width, c.shape[1], dtype=c.dtype, device=c.device)
        blk[kc[order], slot] = c[order]
        hb = self.apply_terms(blk, L.GW_TERM_ALL)
        out[order] = hb[kc[order], slot]
        return out

    def set_flags(self, flags: int) -> None:
        L.check(L.lib().gw_plan_set_flags(self.plan.handle, int(flags)), "gw_plan_set_flags")

    def stage_times(self, reset: bool = True) -> dict[str, tuple[float, int]]:
        """Per-stage (summed ms, calls) from HIP events, when GW_PLAN_TIMING is set."""
        ms = (C.c_double * 6)()
        calls = (C.c_long * 6)()
        L.check(L.lib().gw_plan_stage_times(self.plan.handle, ms, calls, 6, int(reset)),
                "gw_plan_stage_times")
        return {n: (ms[i], calls[i]) for i, n in enumerate(L.STAGES)}

    def info(self, dtype: torch.dtype = torch.complex128) -> dict[str, int]:
        v = (C.c_int * 15)()
        L.check(L.lib().gw_plan_info(self.plan.handle, L.precision_of(dtype), v, 15), "gw_plan_info")
        split = [tuple(f for f in (v[5 + 2 * a], v[6 + 2 * a], v[11 + a]) if f != 1) or (1,) for a in range(3)]
        return {"active_planes": v[0], "max_sticks": v[1], "fused_planes": v[2], "lds_bytes": v[3],
                "nproj": v[4], "split_x": split[0], "split_y": split[1], "split_z": split[2],
                "nosplit_len": v[14], "local_path": self.local_path}

    def wfc_to_real(self, c: torch.Tensor) -> torch.Tensor:
        """psi(r) = N ifftn(scatter c) on the natural box (nk, nb, n1, n2, n3)
        -- core/batch.py:208-216 (g_to_r_b)."""
        L.require_gpu(c, "c")
        if self.plan.nosplit_len:
            return _to_real_library(self.plan, c)
        prec = L.precision_of(c.dtype)
        c = c.contiguous()
        nk, nb, _ = c.shape
        psi = torch.empty(nk, nb, *self.shape, dtype=c.dtype, device=c.device)
        need = int(L.lib().gw_wfc_to_real_workspace_bytes(self.plan.handle, nb, prec))
        ws = torch.empty(max(need, 256), dtype=torch.uint8, device=c.device)
        L.check(L.lib().gw_wfc_to_real(self.plan.handle, prec, c.data_ptr(), psi.data_ptr(), nb,
                                       ws.data_ptr(), ws.numel(), L.stream_ptr(c.device)),
                "gw_wfc_to_real")
        return psi


def _to_real_library(plan: Plan, c: torch.Tensor) -> torch.Tensor:
    """psi(r) = N ifftn(scatter c) for a plan without box FFT kernels: the plan's
    scatter, then rocFFT (torch.fft on the GPU)."""
    c = c.contiguous()
    nk, nb, _ = (int(x) for x in c.shape)
    box = torch.empty(nk, nb, *plan.shape, dtype=c.dtype, device=c.device)
    if nb:
        L.check(L.lib().gw_sphere_to_box(plan.handle, L.precision_of(c.dtype), c.data_ptr(), box.data_ptr(), nb,
                                         L.stream_ptr(c.device)), "gw_sphere_to_box")
    return torch.fft.ifftn(box, dim=(-3, -2, -1), norm="forward")


_DENSITY_PLANS: dict[tuple, tuple] = {}


def _density_plan(bk, shape) -> Plan:
    flat = bk.flat_idx.detach().to("cpu", torch.int64)
    key = (tuple(int(s) for s in shape), tuple(flat.shape))
    hit = _DENSITY_PLANS.get(key)
    if hit is not None and torch.equal(hit[0], flat):
        return hit[1]
    plan = Plan(shape, flat, bk.npw)
    _DENSITY_PLANS[key] = (flat, plan)
    return plan


def density_b(coeffs: torch.Tensor, occ: torch.Tensor, kweights: torch.Tensor, bk, shape, volume: float) -> torch.Tensor:
    """rho(r) on the dense grid [e/A^3] (core/batch.py:562-587):
    sum_k w_k sum_b f_kb |psi_kb(r)|^2 / volume, psi = N ifftn(scatter c).

    Cubic 24/36/90/120 boxes run the fused wave-line density kernel (one inverse
    transform per band, FP64 accumulation); other boxes transform with
    ``gw_wfc_to_real`` and reduce on the GPU; boxes with a length that has no
    codelet split transform with the plan's scatter and rocFFT, in band chunks."""
    L.require_gpu(coeffs, "coeffs")
    plan = _density_plan(bk, shape)
    prec = L.precision_of(coeffs.dtype)
    c = coeffs if coeffs.stride(2) == 1 else coeffs.contiguous()
    nk, nb, _ = c.shape
    w = (kweights.to(c.device, torch.float64)[:, None] * occ.to(c.device, torch.float64)).contiguous()
    rho = torch.zeros(tuple(int(s) for s in shape), dtype=torch.float64, device=c.device)
    if plan.nosplit_len:
        step = max(1, min(nb, _LIB_FFT_BUDGET_BYTES // (2 * nk * rho.numel() * c.element_size())))
        for b0 in range(0, nb, step):
            psi = _to_real_library(plan, c[:, b0:b0 + step])
            rho += torch.einsum("kb,kbxyz->xyz", w[:, b0:b0 + step],
                                psi.real.double() ** 2 + psi.imag.double() ** 2)
        return rho / volume
    info = (C.c_int * 11)()
    L.check(L.lib().gw_plan_info(plan.handle, prec, info, 11), "gw_plan_info")
    if info[2] == 2:
        need = int(L.lib().gw_density_workspace_bytes(plan.handle, nb, prec))
        ws = torch.empty(max(need, 256), dtype=torch.uint8, device=c.device)
        L.check(L.lib().gw_density(plan.handle, prec, c.data_ptr(), c.stride(1), c.stride(0), nb, w.data_ptr(),
                                   rho.data_ptr(), ws.data_ptr(), ws.numel(), L.stream_ptr(c.device)), "gw_density")
    else:
        need = int(L.lib().gw_wfc_to_real_workspace_bytes(plan.handle, nb, prec))
        ws = torch.empty(max(need, 256), dtype=torch.uint8, device=c.device)
        psi = torch.empty(nk, nb, *rho.shape, dtype=c.dtype, device=c.device)
        cc = c.contiguous()
        L.check(L.lib().gw_wfc_to_real(plan.handle, prec, cc.data_ptr(), psi.data_ptr(), nb, ws.data_ptr(),
                                       ws.numel(), L.stream_ptr(c.device)), "gw_wfc_to_real")
        rho += torch.einsum("kb,kbxyz->xyz", w, (psi.real.double() ** 2 + psi.imag.double() ** 2))
    return rho / volume
