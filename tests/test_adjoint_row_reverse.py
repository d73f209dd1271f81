"""The adjoint's hand-written reverse of the PGS row (``ElPgs::vjpRev`` in
``peclet-dem_amd/csrc/pd_adjoint.hpp``) against the Jacobian of the forward-mode evaluation of the
same row, on the CPU: ``tests/native/row_reverse_check.cpp`` is built with hipcc (host code only
runs) and compares J^T yb on 60,000 random rows covering every branch of the row."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_pgs_row_reverse_equals_forward_mode_jacobian(tmp_path):
    exe = tmp_path / "row_reverse_check"
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "--offload-arch=gfx950",
           "-I", os.path.join(ROOT, "peclet-dem_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "native", "row_reverse_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "0 above 1e-10" in out.stdout
