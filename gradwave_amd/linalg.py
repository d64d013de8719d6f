"""Thin tensor wrappers over the MFMA contractions and Davidson vector kernels.

Every function takes GPU tensors whose last axis is contiguous and passes raw
pointers plus strides (in complex elements) to libgradwave_hip.
"""

from __future__ import annotations

import torch

from gradwave_amd import _lib as L

_WS: dict[tuple, torch.Tensor] = {}


def _scratch(device: torch.device, nbytes: int) -> torch.Tensor:
    """Split-K workspace of the contractions, one per (device, stream): work on
    different streams never shares it."""
    key = (device.type, device.index, torch.cuda.current_stream(device).cuda_stream)
    ws = _WS.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def _rows(t: torch.Tensor) -> torch.Tensor:
    if t.stride(-1) != 1:
        t = t.contiguous()
    return t


def inner(x: torch.Tensor, y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """C[k] = X[k]^* Y[k]^T, i.e. C[k][i][j] = sum_g conj(x[k][i][g]) y[k][j][g].

    The reference's subspace matrix torch.matmul(v.conj(), hv.mT)
    (solvers/davidson.py:648) and projection coefficients (davidson.py:431)."""
    L.require_gpu(x, "x")
    x, y = _rows(x), _rows(y)
    nk, M, K = x.shape
    N = y.shape[1]
    assert y.shape[0] == nk and y.shape[2] == K and x.dtype == y.dtype
    prec = L.precision_of(x.dtype)
    if out is None:
        out = torch.empty(nk, M, N, dtype=x.dtype, device=x.device)
    need = int(L.lib().gw_zgemm_inner_workspace_bytes(prec, nk, M, N, K))
    ws = _scratch(x.device, need)
    L.check(L.lib().gw_zgemm_inner(prec, nk, M, N, K, x.data_ptr(), x.stride(1), x.stride(0),
                                   y.data_ptr(), y.stride(1), y.stride(0), out.data_ptr(),
                                   out.stride(1), out.stride(0), ws.data_ptr(), ws.numel(),
                                   L.stream_ptr(x.device)), "gw_zgemm_inner")
    return out


def combine(u: torch.Tensor, y: torch.Tensor, out: torch.Tensor | None = None,
            beta: float = 0.0, conj_u: bool = False) -> torch.Tensor:
    """Z[k][a][g] = beta Z + sum_j u[k][j][a] y[k][j][g] -- the Ritz combination
    einsum("kja,kjg->kag", u, v) of solvers/davidson.py:653-654."""
    L.require_gpu(y, "y")
    u, y = _rows(u), _rows(y)
    nk, J, A = u.shape
    G = y.shape[2]
    assert y.shape[0] == nk and y.shape[1] == J and u.dtype == y.dtype
    prec = L.precision_of(y.dtype)
    if out is None:
        out = torch.empty(nk, A, G, dtype=y.dtype, device=y.device)
        beta = 0.0
    L.check(L.lib().gw_zgemm_combine(prec, nk, A, G, J, u.data_ptr(), u.stride(1), u.stride(0),
                                     int(conj_u), y.data_ptr(), y.stride(1), y.stride(0),
                                     out.data_ptr(), out.stride(1), out.stride(0), float(beta),
                                     L.stream_ptr(y.device)), "gw_zgemm_combine")
    return out


def residual(x: torch.Tensor, hx: torch.Tensor, eig: torch.Tensor):
    """r = hx - eig x and rn = ||r|| in fp64 (solvers/davidson.py:656-657)."""
    assert x.stride() == hx.stride() and x.stride(-1) == 1
    nk, nb, m = x.shape
    r = torch.empty_like(x)
    assert r.stride() == x.stride()
    rn = torch.empty(nk, nb, dtype=torch.float64, device=x.device)
    e = eig.to(torch.float64).contiguous()
    L.check(L.lib().gw_residual(L.precision_of(x.dtype), nk, nb, m, x.data_ptr(), hx.data_ptr(),
                                x.stride(1), x.stride(0), e.data_ptr(), r.data_ptr(), rn.data_ptr(),
                                L.stream_ptr(x.device)), "gw_residual")
    return r, rn


def teter_expand(x: torch.Tensor, r: torch.Tensor, t: torch.Tensor, sel: torch.Tensor,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """Teter-preconditioned, unit-normalised residuals of the selected bands
    (solvers/davidson.py:729-743, solvers/precond.py:20-39)."""
    assert x.stride() == r.stride() and x.stride(-1) == 1
    nk, nb, m = x.shape
    nsel = int(sel.shape[1])
    if out is None:
        out = torch.empty(nk, nsel, m, dtype=x.dtype, device=x.device)
    tt = t.to(torch.float64).contiguous()
    s = sel.to(torch.int64).contiguous()
    L.check(L.lib().gw_teter_expand(L.precision_of(x.dtype), nk, nb, nsel, m, x.data_ptr(),
                                    r.data_ptr(), x.stride(1), x.stride(0), tt.data_ptr(),
                                    tt.stride(0), s.data_ptr(), out.data_ptr(), out.stride(1),
                                    out.stride(0), L.stream_ptr(x.device)), "gw_teter_expand")
    return out


def row_norms(v: torch.Tensor) -> torch.Tensor:
    v = _rows(v)
    nk, mm, g = v.shape
    out = torch.empty(nk, mm, dtype=torch.float64, device=v.device)
    L.check(L.lib().gw_row_norms(L.precision_of(v.dtype), nk, mm, g, v.data_ptr(), v.stride(1),
                                 v.stride(0), out.data_ptr(), L.stream_ptr(v.device)),
            "gw_row_norms")
    return out


def rinner(x: torch.Tensor, y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Real C = X Y^T over the long last axis (split-K MFMA, fp64 partial sums):
    x (M, K) or (nk, M, K), y (N, K) or (nk, N, K) -> (M, N) / (nk, M, N)."""
    L.require_gpu(x, "x")
    x, y = _rows(x), _rows(y)
    batched = x.dim() == 3
    xb = x if batched else x[None]
    yb = y if batched else y[None]
    nk, M, K = xb.shape
    N = yb.shape[1]
    assert yb.shape[0] == nk and yb.shape[2] == K and x.dtype == y.dtype
    prec = L.precision_of(x.dtype)
    if out is None:
        out = torch.empty(nk, M, N, dtype=x.dtype, device=x.device) if batched else \
            torch.empty(M, N, dtype=x.dtype, device=x.device)
    ob = out if batched else out[None]
    if M == 0 or N == 0:
        return out
    need = int(L.lib().gw_rgemm_inner_workspace_bytes(prec, nk, M, N, K))
    ws = _scratch(x.device, need)
    L.check(L.lib().gw_rgemm_inner(prec, nk, M, N, K, xb.data_ptr(), xb.stride(1), xb.stride(0), yb.data_ptr(),
                                   yb.stride(1), yb.stride(0), ob.data_ptr(), ob.stride(1), ob.stride(0),
                                   ws.data_ptr(), ws.numel(), L.stream_ptr(x.device)), "gw_rgemm_inner")
    return out


def rcombine(u: torch.Tensor, y: torch.Tensor, out: torch.Tensor | None = None, alpha: float = 1.0,
             beta: float = 0.0) -> torch.Tensor:
    """Real Z = beta Z + alpha U^T Y: u (J, A), y (J, G) -> (A, G) (batched with a leading nk)."""
    L.require_gpu(y, "y")
    u, y = _rows(u), _rows(y)
    batched = y.dim() == 3
    ub = u if batched else u[None]
    yb = y if batched else y[None]
    nk, J, A = ub.shape
    G = yb.shape[2]
    assert yb.shape[0] == nk and yb.shape[1] == J and u.dtype == y.dtype
    prec = L.precision_of(y.dtype)
    if out is None:
        out = torch.empty(nk, A, G, dtype=y.dtype, device=y.device) if batched else \
            torch.empty(A, G, dtype=y.dtype, device=y.device)
        beta = 0.0
    ob = out if batched else out[None]
    L.check(L.lib().gw_rgemm_combine(prec, nk, A, G, J, ub.data_ptr(), ub.stride(1), ub.stride(0), yb.data_ptr(),
                                     yb.stride(1), yb.stride(0), ob.data_ptr(), ob.stride(1), ob.stride(0),
                                     float(alpha), float(beta), L.stream_ptr(y.device)), "gw_rgemm_combine")
    return out


_HEEV_WS: dict[tuple, torch.Tensor] = {}


def heev_max_n() -> int:
    return int(L.lib().gw_heev_max_n())


def eigh_lowest(mats: list, nevs: list) -> list:
    """Lowest ``nev`` eigenpairs of Hermitian complex128 matrices, all in one
    batched kernel launch (``gw_heev_batched``): the subspace eigensolve of
    solvers/davidson.py:640-652, of which davidson.py:651-652 keeps the first nb
    Ritz pairs. ``mats[i]`` is (nk, m_i, m_i) on the GPU with a contiguous last
    axis (views such as ``s[:, :m, :m]`` are read in place); returns
    [(w (nk, nev_i) float64 ascending, u (nk, m_i, nev_i) complex128)]. Sizes
    above ``heev_max_n()`` raise ``GWError``."""
    if not mats:
        return []
    lib = L.lib()
    dev = mats[0].device
    out, jobs, sizes = [], [], []
    for s, nev in zip(mats, nevs):
        L.require_gpu(s, "s")
        if s.dtype != torch.complex128 or s.dim() != 3 or s.shape[1] != s.shape[2]:
            raise TypeError(f"expected (nk, m, m) complex128, got {tuple(s.shape)} {s.dtype}")
        if s.stride(-1) != 1:
            s = s.contiguous()
        nk, m = int(s.shape[0]), int(s.shape[1])
        nev = int(nev)
        need = int(lib.gw_heev_workspace_bytes(m, nev))
        if need == 0:
            raise L.GWError(f"eigh_lowest: unsupported size m={m}, nev={nev} (max m {heev_max_n()})")
        w = torch.empty(nk, nev, dtype=torch.float64, device=dev)
        u = torch.empty(nk, m, nev, dtype=torch.complex128, device=dev)
        out.append((w, u))
        for k in range(nk):
            jobs.append((s, k, w, u, m, nev))
            sizes.append(need)
    total = sum(sizes)
    key = (dev.type, dev.index)
    ws = _HEEV_WS.get(key)
    if ws is None or ws.numel() < total:
        ws = torch.empty(max(total, 1 << 20), dtype=torch.uint8, device=dev)
        _HEEV_WS[key] = ws
    arr = (L.HeevJob * len(jobs))()
    off = 0
    for j, ((s, k, w, u, m, nev), need) in enumerate(zip(jobs, sizes)):
        arr[j].a = s.data_ptr() + k * s.stride(0) * 16
        arr[j].w = w.data_ptr() + k * nev * 8
        arr[j].z = u.data_ptr() + k * m * nev * 16
        arr[j].ws = ws.data_ptr() + off
        arr[j].n, arr[j].lda, arr[j].nev, arr[j].ldz = m, int(s.stride(1)), nev, nev
        off += need
    L.check(lib.gw_heev_batched(len(jobs), arr, L.stream_ptr(dev)), "gw_heev_batched")
    return out


def chol_max_n() -> int:
    return int(L.lib().gw_chol_max_n())


def cholesky_inverse(mats: list) -> list:
    """L^-1 of the Cholesky factor G = L L^H of Hermitian positive definite
    complex128 matrices, all in one launch (``gw_chol_inv_batched``): the CholQR
    step of the orthonormalisation (solvers/davidson.py:195-222). ``mats[i]`` is
    (nk, n_i, n_i) on the GPU (lower triangle read); returns [(linv (nk, n_i,
    n_i), info (nk,) int32)], info > 0 where the factorisation broke down (as
    torch.linalg.cholesky_ex reports). Sizes above ``chol_max_n()`` raise."""
    if not mats:
        return []
    lib = L.lib()
    dev = mats[0].device
    out, jobs = [], []
    for g in mats:
        L.require_gpu(g, "g")
        if g.dtype != torch.complex128 or g.dim() != 3 or g.shape[1] != g.shape[2]:
            raise TypeError(f"expected (nk, n, n) complex128, got {tuple(g.shape)} {g.dtype}")
        if g.stride(-1) != 1:
            g = g.contiguous()
        nk, n = int(g.shape[0]), int(g.shape[1])
        if n > chol_max_n():
            raise L.GWError(f"cholesky_inverse: n={n} above {chol_max_n()}")
        linv = torch.empty(nk, n, n, dtype=torch.complex128, device=dev)
        info = torch.empty(nk, dtype=torch.int32, device=dev)
        out.append((linv, info))
        for k in range(nk):
            jobs.append((g, k, linv, info, n))
    arr = (L.CholJob * len(jobs))()
    for j, (g, k, linv, info, n) in enumerate(jobs):
        arr[j].a = g.data_ptr() + k * g.stride(0) * 16
        arr[j].linv = linv.data_ptr() + k * n * n * 16
        arr[j].info = info.data_ptr() + k * 4
        arr[j].n, arr[j].lda, arr[j].ldl = n, int(g.stride(1)), n
    L.check(lib.gw_chol_inv_batched(len(jobs), arr, L.stream_ptr(dev)), "gw_chol_inv_batched")
    return out
