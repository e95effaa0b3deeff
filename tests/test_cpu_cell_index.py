"""The march's FP32 cell index equals the integer cell index for every legal
volume shape (host build of drr_internal.h's cell_index_biased and
vol_index_params; no GPU needed)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_fp32_cell_index_matches_integer_index(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "cell_index_check"
    src = os.path.join(HERE, "native", "cell_index_check.hip")
    r = subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT}/include", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
