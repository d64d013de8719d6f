"""The in-register DFT codelets and the wave-line index maps (Cooley-Tukey and
Good-Thomas splits) against a direct DFT, on the host: tests/fft_codelets_host.hip
includes the same headers the box kernels use and is built here with hipcc."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_codelets_and_waveline_maps(tmp_path):
    cc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "fft_codelets")
    r = subprocess.run([cc, "-O1", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", os.path.join(HERE, "fft_codelets_host.hip"), "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "failures 0" in r.stdout
    # every codelet length of the box plans and every wave-line split was exercised
    for n in (2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16):
        assert f"dft inverse                      n={n:<4d}" in r.stdout
    assert "90 =  9 x 10 prime-factor" in r.stdout
