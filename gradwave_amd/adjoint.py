"""Differentiable H|psi> on MI355X: hand-written vector-Jacobian products.

The reference gets its gradients by torch autograd through plain tensor code:
``HamiltonianK.apply`` (core/hamiltonian.py:179-191) for the operator,
``projectors``/``projectors_b`` (core/hamiltonian.py:108-112, core/batch.py:229-235)
for the projector phases e^{-i(k+G).tau}, and ``local_potential_g``
(core/energies/local_pp.py:22-38) for the structure factor of the local
potential; ``postscf/forces.py:168-215`` chains them into forces. Here each of
those is a ``torch.autograd.Function`` whose forward and backward are
libgradwave_hip kernels:

* ``hpsi(op, c, v_eff_r, p, dij)`` -- H c with gradients for all four inputs:
    grad c   = H^dagger g                          (the apply kernels, D transposed)
    grad V   = (1/N) sum Re[conj(N ifftn g) (N ifftn c)]   (gw_hpsi_adjoint_v)
    grad D   = Re sum_k B_k Gamma_k^H,  B = <p|c>, Gamma = <p|g>   (MFMA inner)
    grad p   = sum_b g_b conj(w_b) + c_b u_b,  w = B^T D, u = conj(Gamma)^T D^T  (MFMA combine)
  FFT intermediates are recomputed in the backward, nothing box-sized is saved.
* ``projectors_b(bk, positions)`` -- phases and their position adjoint (gw_proj_phase*).
* ``local_potential_g(positions, species_index, vloc_tables, g_cart, volume)``
  -- structure-factor sum and its position adjoint (gw_vloc_g*).

Gradients follow torch's complex convention (grad z = dL/dRe z + i dL/dIm z),
so they compose with autograd through the rest of a caller's graph. The one
step between V(G) and V(r), ``local_potential_r``, is a single full-box
inverse FFT per geometry and uses torch.fft (rocFFT); it is not on the
per-band path.
"""

from __future__ import annotations

import ctypes as C

from types import SimpleNamespace

import torch

from gradwave_amd import _lib as L
from gradwave_amd import linalg
from gradwave_amd.hamiltonian import BatchedHamiltonian

HARTREE_EV = 27.211386245988
BOHR_ANG = 0.529177210903
HBAR2_2M = HARTREE_EV / 2.0 * BOHR_ANG**2

_WS: dict[tuple, torch.Tensor] = {}


def _scratch(device: torch.device, nbytes: int, tag: str) -> torch.Tensor:
    key = (tag, device.index)
    ws = _WS.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def _symmetric(d: torch.Tensor) -> bool:
    return d.numel() == 0 or bool(torch.equal(d, d.transpose(0, 1)))


class HOperator:
    """Owns the forward plan (and, for a non-symmetric D, the adjoint plan) of
    one differentiable apply: bk, shape and the detached values of V, p, D."""

    def __init__(self, bk, shape, v_eff_r: torch.Tensor, p: torch.Tensor, dij: torch.Tensor) -> None:
        dij = dij.detach().to(torch.float64)
        self.bk = bk
        self.dij = dij
        fwd_bk = SimpleNamespace(npw=bk.npw, mask=bk.mask, flat_idx=bk.flat_idx, t=bk.t, dij_full=dij)
        self.fwd = BatchedHamiltonian(fwd_bk, shape, v_eff_r.detach(), p.detach())
        if _symmetric(dij):
            self.adj = self.fwd
        else:
            adj_bk = SimpleNamespace(npw=bk.npw, mask=bk.mask, flat_idx=bk.flat_idx, t=bk.t,
                                     dij_full=dij.transpose(0, 1).contiguous())
            self.adj = BatchedHamiltonian(adj_bk, shape, v_eff_r.detach(), p.detach())
        self.p = self.fwd.plan._p if int(p.shape[1]) else None  # complex128 device copy
        self.shape = self.fwd.shape


def _adjoint_v_real_space(op: HOperator, c: torch.Tensor, g: torch.Tensor, gv: torch.Tensor) -> torch.Tensor:
    """The V(r) adjoint on boxes without the wave-line adjoint kernel: psi(r) of
    both blocks from the operator's own transform (the generic box FFT kernels,
    or the plan's scatter and rocFFT when a length has no codelet split), in
    band chunks, the sum in float64."""
    from gradwave_amd.hamiltonian import _LIB_FFT_BUDGET_BYTES

    nk, nb = int(c.shape[0]), int(c.shape[1])
    n = gv.numel()
    step = max(1, min(nb, _LIB_FFT_BUDGET_BYTES // (4 * nk * n * c.element_size())))
    for b0 in range(0, nb, step):
        pc = op.fwd.wfc_to_real(c[:, b0:b0 + step].contiguous()).to(torch.complex128)
        pg = op.fwd.wfc_to_real(g[:, b0:b0 + step].contiguous()).to(torch.complex128)
        gv += (pg.conj() * pc).real.sum(dim=(0, 1))
    return gv.div_(n)


def adjoint_v(op: HOperator, c: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """(1/N) sum_{k,b} Re[conj(N ifftn g) (N ifftn c)] on the natural box, float64."""
    plan = op.fwd.plan
    prec = L.precision_of(c.dtype)
    c = c if c.stride(2) == 1 else c.contiguous()
    g = g.to(c.dtype)
    g = g if g.stride(2) == 1 else g.contiguous()
    nb = int(c.shape[1])
    gv = torch.zeros(op.shape, dtype=torch.float64, device=c.device)
    if nb == 0:
        return gv
    info = (C.c_int * 3)()
    L.check(L.lib().gw_plan_info(plan.handle, prec, info, 3), "gw_plan_info")
    if info[2] != 2:  # no wave-line path on this box: transform to r and sum
        return _adjoint_v_real_space(op, c, g, gv)
    need = int(L.lib().gw_hpsi_adjoint_v_workspace_bytes(plan.handle, nb, prec))
    ws = _scratch(c.device, need, "adjv")
    L.check(L.lib().gw_hpsi_adjoint_v(plan.handle, prec, c.data_ptr(), c.stride(1), c.stride(0),
                                      g.data_ptr(), g.stride(1), g.stride(0), nb, gv.data_ptr(),
                                      ws.data_ptr(), ws.numel(), L.stream_ptr(c.device)),
            "gw_hpsi_adjoint_v")
    return gv


def _nonlocal_grads(op: HOperator, c: torch.Tensor, g: torch.Tensor, want_d: bool, want_p: bool):
    p = op.p
    c = c.to(torch.complex128)
    g = g.to(torch.complex128)
    bmat = linalg.inner(p, c)  # (nk, nproj, nb): B[k][p][b] = <p|c_b>
    gam = linalg.inner(p, g)   # (nk, nproj, nb): Gamma[k][p][b] = <p|g_b>
    grad_d = grad_p = None
    if want_d:
        # grad D[p][q] = Re sum_{k,b} B[p][b] conj(Gamma[q][b]);  inner(Gamma, B)[q][p] is that sum per k
        grad_d = linalg.inner(gam, bmat).sum(0).real.transpose(0, 1).contiguous()
    if want_p:
        d = op.dij.to(device=c.device, dtype=torch.complex128)
        w = torch.matmul(bmat.transpose(1, 2).contiguous(), d)                        # (nk, nb, nproj)
        u = torch.matmul(gam.conj().transpose(1, 2).contiguous(), d.transpose(0, 1).contiguous())
        grad_p = linalg.combine(w, g, conj_u=True)                                   # sum_b g_b conj(w_b)
        grad_p = linalg.combine(u.contiguous(), c, out=grad_p, beta=1.0)             # + sum_b c_b u_b
    return grad_d, grad_p


class _HpsiFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, op: HOperator, c, v_eff_r, p, dij):
        ctx.op = op
        ctx.v_dtype = v_eff_r.dtype
        ctx.save_for_backward(c)
        return op.fwd.apply_terms(c, L.GW_TERM_ALL)

    @staticmethod
    def backward(ctx, g):
        op: HOperator = ctx.op
        (c,) = ctx.saved_tensors
        g = (g * op.bk.mask[:, None, :]).to(c.dtype)
        need_c, need_v, need_p, need_d = ctx.needs_input_grad[1:5]
        grad_c = op.adj.apply_terms(g, L.GW_TERM_ALL) if need_c else None
        grad_v = adjoint_v(op, c, g).to(ctx.v_dtype) if need_v else None
        grad_p = grad_d = None
        if (need_p or need_d) and op.p is not None:
            grad_d, grad_p = _nonlocal_grads(op, c, g, need_d, need_p)
            if grad_d is not None:
                grad_d = grad_d.to(op.dij.device)
        return None, grad_c, grad_v, grad_p, grad_d


def hpsi(op: HOperator, c: torch.Tensor, v_eff_r: torch.Tensor, p: torch.Tensor,
         dij: torch.Tensor) -> torch.Tensor:
    """H c (core/batch.py:480-505) differentiable in c, v_eff_r, p and dij.

    ``op`` must have been built from the values of v_eff_r, p and dij; the
    tensors are passed again so autograd can route their gradients."""
    return _HpsiFn.apply(op, c, v_eff_r, p, dij)


class DifferentiableHamiltonian:
    """``BatchedHamiltonian(bk, shape, v_eff_r, p)`` whose ``apply`` carries
    gradients to c, v_eff_r, p and ``dij`` (default ``bk.dij_full``)."""

    def __init__(self, bk, shape, v_eff_r: torch.Tensor, p: torch.Tensor,
                 dij: torch.Tensor | None = None) -> None:
        self.bk = bk
        self.v_eff_r = v_eff_r
        self.p = p
        self.dij = bk.dij_full if dij is None else dij
        self.op = HOperator(bk, shape, v_eff_r, p, self.dij)
        self.shape = self.op.shape

    def apply(self, c: torch.Tensor) -> torch.Tensor:
        return hpsi(self.op, c, self.v_eff_r, self.p, self.dij)

    __call__ = apply


class HamiltonianK:
    """Single-k differentiable apply with the reference's signature
    ``HamiltonianK(sphere, shape, v_eff_r, pd, p).apply(c (nb, npw))``
    (core/hamiltonian.py:120-191); gradients reach c, v_eff_r, p and
    ``pd.dij_full``. ``sphere`` needs flat_idx and kpg2, ``pd`` dij_full."""

    def __init__(self, sphere, shape, v_eff_r: torch.Tensor, pd, p: torch.Tensor) -> None:
        dev = v_eff_r.device
        npw = int(sphere.flat_idx.shape[0])
        flat = sphere.flat_idx.to(dev, torch.int64)[None]
        t = (HBAR2_2M * sphere.kpg2.to(dev, torch.float64))[None]
        bk = SimpleNamespace(npw=torch.tensor([npw], dtype=torch.int64, device=dev),
                             mask=torch.ones(1, npw, dtype=torch.bool, device=dev), flat_idx=flat, t=t,
                             dij_full=pd.dij_full)
        self.sphere, self.shape, self.v_eff_r, self.pd, self.p = sphere, tuple(shape), v_eff_r, pd, p
        self._h = DifferentiableHamiltonian(bk, shape, v_eff_r, p[None], pd.dij_full)

    def apply(self, c: torch.Tensor) -> torch.Tensor:
        h = self._h
        return hpsi(h.op, c[None], h.v_eff_r, self.p[None], h.dij)[0]


# ---------------------------------------------------------------- positions


class _ProjPhaseFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, pf, kpg, atom_index, npw32):
        nk, nproj, npw_max = pf.shape
        p = torch.empty_like(pf)
        pos = positions.detach().to(torch.float64).contiguous()
        L.check(L.lib().gw_proj_phase(nk, nproj, npw_max, npw32.data_ptr(), pf.data_ptr(),
                                      kpg.data_ptr(), pos.data_ptr(), atom_index.data_ptr(),
                                      p.data_ptr(), L.stream_ptr(pf.device)), "gw_proj_phase")
        ctx.save_for_backward(p, kpg, atom_index)
        ctx.na = int(positions.shape[0])
        return p

    @staticmethod
    def backward(ctx, grad_p):
        p, kpg, atom_index = ctx.saved_tensors
        nk, nproj, npw_max = p.shape
        gp = grad_p.to(torch.complex128).contiguous()
        gtau = torch.zeros(ctx.na, 3, dtype=torch.float64, device=p.device)
        need = int(L.lib().gw_proj_phase_grad_workspace_bytes(nk, nproj))
        ws = _scratch(p.device, need, "phase")
        L.check(L.lib().gw_proj_phase_grad(nk, nproj, npw_max, ctx.na, gp.data_ptr(), p.data_ptr(),
                                           kpg.data_ptr(), atom_index.data_ptr(), gtau.data_ptr(),
                                           ws.data_ptr(), ws.numel(), L.stream_ptr(p.device)),
                "gw_proj_phase_grad")
        return gtau, None, None, None, None


def projectors_b(bk, positions: torch.Tensor) -> torch.Tensor:
    """(nk, nproj, npw_max) projectors, differentiable in positions (core/batch.py:229-235).

    ``bk`` needs proj_phase_free (complex), kpg (nk, npw_max, 3), proj_atom_index
    and npw on the GPU."""
    pf = bk.proj_phase_free
    if pf.shape[1] == 0:
        return pf
    L.require_gpu(pf, "proj_phase_free")
    pf = pf.to(torch.complex128).contiguous()
    kpg = bk.kpg.to(torch.float64).contiguous()
    ai = bk.proj_atom_index.to(pf.device, torch.int64).contiguous()
    npw32 = bk.npw.to(pf.device, torch.int32).contiguous()
    return _ProjPhaseFn.apply(positions, pf, kpg, ai, npw32)


def projectors(pd, positions: torch.Tensor) -> torch.Tensor:
    """Single-k (nproj, npw) projectors (core/hamiltonian.py:108-112); ``pd`` needs
    f_ylm_phase_free, kpg and atom_index on the GPU."""
    bk = SimpleNamespace(proj_phase_free=pd.f_ylm_phase_free[None], kpg=pd.kpg[None],
                         proj_atom_index=pd.atom_index,
                         npw=torch.tensor([pd.kpg.shape[0]], device=pd.kpg.device))
    return projectors_b(bk, positions)[0]


class _VlocGFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, species_index, tables, g_cart, volume):
        na = int(positions.shape[0])
        npts = int(g_cart.shape[0])
        pos = positions.detach().to(torch.float64).contiguous()
        vg = torch.empty(npts, dtype=torch.complex128, device=g_cart.device)
        L.check(L.lib().gw_vloc_g(na, npts, pos.data_ptr(), species_index.data_ptr(), tables.data_ptr(),
                                  g_cart.data_ptr(), float(volume), vg.data_ptr(),
                                  L.stream_ptr(g_cart.device)), "gw_vloc_g")
        ctx.save_for_backward(species_index, tables, g_cart)
        ctx.pos = pos
        ctx.volume = float(volume)
        return vg

    @staticmethod
    def backward(ctx, grad_vg):
        sidx, tables, g_cart = ctx.saved_tensors
        pos = ctx.pos
        na, npts = int(pos.shape[0]), int(g_cart.shape[0])
        gv = grad_vg.to(torch.complex128).contiguous()
        gtau = torch.zeros(na, 3, dtype=torch.float64, device=pos.device)
        need = int(L.lib().gw_vloc_g_grad_workspace_bytes(na))
        ws = _scratch(pos.device, need, "vloc")
        L.check(L.lib().gw_vloc_g_grad(na, npts, pos.data_ptr(), sidx.data_ptr(), tables.data_ptr(),
                                       g_cart.data_ptr(), ctx.volume, gv.data_ptr(), gtau.data_ptr(),
                                       ws.data_ptr(), ws.numel(), L.stream_ptr(pos.device)),
                "gw_vloc_g_grad")
        return gtau, None, None, None, None


def local_potential_g(positions: torch.Tensor, species_index: torch.Tensor, vloc_tables: torch.Tensor,
                      g_cart: torch.Tensor, volume: float) -> torch.Tensor:
    """V_loc(G) on the full box [eV], complex128 (core/energies/local_pp.py:22-38),
    differentiable in positions (A). vloc_tables (nspecies, n1, n2, n3) eV A^3,
    g_cart (n1, n2, n3, 3) A^-1."""
    L.require_gpu(g_cart, "g_cart")
    shape = tuple(g_cart.shape[:3])
    dev = g_cart.device
    tab = vloc_tables.to(dev, torch.float64).reshape(vloc_tables.shape[0], -1).contiguous()
    gc = g_cart.to(torch.float64).reshape(-1, 3).contiguous()
    si = species_index.to(dev, torch.int64).contiguous()
    return _VlocGFn.apply(positions, si, tab, gc, volume).reshape(shape)


def local_potential_r(vloc_g: torch.Tensor) -> torch.Tensor:
    """V(r) = Re N ifftn V_loc(G) (scf/loop.py:532-544, one full-box FFT per geometry)."""
    n = vloc_g.shape[-3] * vloc_g.shape[-2] * vloc_g.shape[-1]
    return (torch.fft.ifftn(vloc_g, dim=(-3, -2, -1)) * n).real


# ---------------------------------------------------------------- forces


def band_energy_forces(bk, shape, positions: torch.Tensor, species_index: torch.Tensor,
                       vloc_tables: torch.Tensor, g_cart: torch.Tensor, volume: float,
                       psi: torch.Tensor, weights: torch.Tensor, v_frozen: torch.Tensor | None = None):
    """Band energy at fixed orbitals and its position gradient.

    E = sum_{k,b} weights[k,b] Re <psi_kb| H(tau) |psi_kb>   [eV]
    F = -dE/dtau                                           [eV/A], (na, 3)

    H(tau) carries V_loc(r; tau) (+ ``v_frozen``, a position-independent
    potential such as Hartree + xc) and the projectors p(tau): the
    Hellmann-Feynman pattern of postscf/forces.py:168-215 (orbitals detached,
    gradient through the projector phases and the structure factor).
    ``bk`` needs npw, mask, flat_idx, t, kpg, dij_full, proj_atom_index and
    proj_phase_free on the GPU.
    """
    with torch.enable_grad():
        pos = positions.detach().to(torch.float64).clone().requires_grad_(True)
        v = local_potential_r(local_potential_g(pos, species_index, vloc_tables, g_cart, volume))
        if v_frozen is not None:
            v = v + v_frozen.detach()
        p = projectors_b(bk, pos)
        h = DifferentiableHamiltonian(bk, shape, v, p, bk.dij_full)
        x = psi.detach()
        hx = h.apply(x)
        e = (weights.to(torch.float64)[..., None] * (x.conj() * hx).real).sum()
        (g,) = torch.autograd.grad(e, pos)
    return e.detach(), -g
