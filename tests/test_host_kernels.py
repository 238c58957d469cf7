"""CPU checks of device helper functions compiled for the host with hipcc."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_leaf_distance_equals_box_distance(tmp_path):
    src = os.path.join(ROOT, "tests", "host", "leaf_distance_check.cpp")
    exe = str(tmp_path / "leaf_distance_check")
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "--cuda-host-only",
                           "-x", "hip", src, "-o", exe], stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert "total mismatches: 0" in out.stdout
