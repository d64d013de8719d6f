"""The C-ABI library loads here (no GPU) and exports every symbol the header declares."""

import os
import re

import pytest

from gradwave_amd import _lib as L

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "gradwave_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gw_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gw_hpsi_apply" in names and "gw_zgemm_inner" in names and len(names) >= 15


@pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="library not built")
def test_library_exports_all_declared_symbols():
    lib = L.lib()
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} declared in gradwave_hip.h but not exported"
        assert name in L.SIGNATURES, f"{name} has no ctypes signature in _lib.py"
    assert lib.gw_version() >= 1


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(L.GWError, match="not built"):
        L.lib()


@pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="library not built")
def test_argument_errors_are_reported_without_a_gpu():
    """Entry points validate on the host before any launch and report through
    gw_last_error (the reference raises ValueError/RuntimeError, e.g.
    solvers/registry.py:87-94, grids.py:168-170)."""
    import ctypes as C

    lib = L.lib()

    def err(rc):
        assert rc != 0
        return lib.gw_last_error().decode()

    assert "unknown tuning key" in err(lib.gw_tune_set(99, 0, 0))
    assert "does not split" in err(lib.gw_tune_set(L.GW_TUNE_RADIX, 24, 5))
    assert "bad tile" in err(lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, L.GW_FP32, 14))
    assert "contraction selector" in err(lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, L.GW_FP32 + 768, 0))
    assert "inner-product choice" in err(lib.gw_tune_set(L.GW_TUNE_INNER_LIB, L.GW_FP32, 2))
    assert lib.gw_tune_get(L.GW_TUNE_INNER_LIB, L.GW_FP64) == 0
    assert lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, L.GW_FP32 + 512, 4) == 0  # back-projection only
    assert lib.gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP32 + 512) == 4
    assert lib.gw_tune_get(L.GW_TUNE_GEMM_TILE, L.GW_FP32) == 0
    assert lib.gw_tune_set(L.GW_TUNE_GEMM_TILE, L.GW_FP32, 0) == 0
    assert "x-pass shape" in err(lib.gw_tune_set(L.GW_TUNE_XPASS, 0, 305))
    assert lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP64) == 408
    assert lib.gw_tune_get(L.GW_TUNE_XPASS, L.GW_FP32) == 10416
    assert err(lib.gw_batch_create(C.byref(C.c_void_p()), None, 0))
    assert err(lib.gw_toeplitz_build(0, 1, 0, 4, 4, 4, None, None, None, None, None))
    assert err(lib.gw_gamma_pack(0, 0, 5, 5, None, 5, None, None, None, 5, None))
    assert err(lib.gw_vloc_g(1, 8, None, None, None, None, 0.0, None, None))
    assert err(lib.gw_proj_phase(1, 1, 1, None, None, None, None, None, None, None))
    # batched subspace eigensolve: sizes and pointers are checked before the launch
    big = lib.gw_heev_max_n() + 1
    assert f"outside [1, {big - 1}]" in err(lib.gw_heev_batched(1, C.byref(L.HeevJob(None, None, None, None,
                                                                                     big, big, 1, 1)), None))
    assert "nev = 5" in err(lib.gw_heev_batched(1, C.byref(L.HeevJob(None, None, None, None, 4, 4, 5, 5)), None))
    assert "lda 3 < n" in err(lib.gw_heev_batched(1, C.byref(L.HeevJob(None, None, None, None, 4, 3, 2, 2)), None))
    assert "null pointer" in err(lib.gw_heev_batched(1, C.byref(L.HeevJob(None, None, None, None, 4, 4, 2, 2)),
                                                     None))
    assert lib.gw_heev_workspace_bytes(big, 1) == 0 and lib.gw_heev_workspace_bytes(4, 5) == 0
    assert lib.gw_heev_workspace_bytes(4, 2) % 256 == 0 and lib.gw_heev_workspace_bytes(4, 2) > 0
    assert lib.gw_heev_batched(0, None, None) == 0
    assert "n = 65 outside [1, 64]" in err(lib.gw_chol_inv_batched(1, C.byref(L.CholJob(None, None, None, 65, 65, 65)),
                                                                   None))
    assert "null pointer" in err(lib.gw_chol_inv_batched(1, C.byref(L.CholJob(None, None, None, 4, 4, 4)), None))
    assert "ldl 3 < n" in err(lib.gw_chol_inv_batched(1, C.byref(L.CholJob(None, None, None, 4, 4, 3)), None))
    assert lib.gw_chol_inv_batched(0, None, None) == 0 and lib.gw_chol_max_n() == 64
    # a box length that no two codelets <= 16 can split (grids.py sizes are 7-smooth, 11 is not)
    f = (C.c_int * 3)()
    assert "box length 11 has no split" in err(lib.gw_box_split(11, f)) and tuple(f) == (0, 0, 0)


@pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="library not built")
def test_box_splits_of_good_fft_sizes():
    """Box lengths the reference's good_fft_size returns (grids.py:32-41): every
    one with a two-codelet split plans with the most balanced such split; the
    others (125, 162, 200, 216, 243, 250, above 256) have none, because the
    three-factor splits are not enabled, and plans on such boxes run no box
    FFT kernels (their local term takes the Toeplitz or library-FFT path)."""
    import ctypes as C

    from oracle.grids import good_fft_size

    codelets = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16}
    lib = L.lib()
    f = (C.c_int * 3)()
    lengths = sorted({good_fft_size(m) for m in range(1, 625)})
    refused = []
    for n in lengths:
        two = [(x, n // x) for x in sorted(codelets) if n % x == 0 and n // x in codelets]
        rc = lib.gw_box_split(n, f)
        if two:
            assert rc == 0 and f[2] == 1 and (f[0], f[1]) in two, n
            assert abs(f[0] - f[1]) == min(abs(a - b) for a, b in two), n
        else:
            assert rc == 3, n  # GW_ERR_UNSUPPORTED
            refused.append(n)
    assert {125, 162, 200, 216, 243, 250} <= set(refused) and min(x for x in refused if x > 256) > 256
    assert refused[:9] == [125, 147, 162, 175, 189, 200, 216, 243, 245], refused
