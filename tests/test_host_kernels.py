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


HEADER_CONFIGS = [
    ["-DFASTTREE_INTEGER_COORDS", "-DPOSITIONS_IN_64BIT", "-DPERIODIC_BC"],
    ["-DFASTTREE_INTEGER_COORDS", "-DPOSITIONS_IN_32BIT", "-DSFC_TYPE_MORTON"],
    ["-DCOORDS_TYPE_DOUBLE", "-DSFC_TYPE_MORTON", "-DBITS_PER_DIMENSION=10"],
    [],
]


@pytest.mark.parametrize("defs", HEADER_CONFIGS, ids=["u64-ph-periodic", "u32-morton",
                                                      "f64-morton10", "f32-default"])
def test_cxx_dropin_header_builds_and_maps_errors(tmp_path, defs):
    """include/fasttree/hlbvh.hpp compiles under the reference's macro sets,
    links against the library and maps k > _MAX_K_ to std::invalid_argument
    with the reference's message."""
    lib = os.path.join(ROOT, "fasttree_amd", "libfasttree_amd.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    src = os.path.join(ROOT, "tests", "host", "hlbvh_header_check.cpp")
    exe = str(tmp_path / "hlbvh_header_check")
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include")] + defs +
                          [src, "-L" + os.path.dirname(lib), "-lfasttree_amd",
                           "-Wl,-rpath," + os.path.dirname(lib), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert ("Instantiate self_knn_query with larger _MAX_K_." in out.stdout)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_float_prefilter_never_rejects_a_leaf_within_the_bound(tmp_path):
    """The packet walk's float32 filters for integer periodic metrics
    (wrap_filter_near; wrap_sq_pre and wrap_box_gap_pre on pre-shifted
    coordinates) pass every leaf and box whose exact d2 is at most the bound,
    over bounds equal to, just below and just above d2, B up to 31, and box
    coordinates with bits above B; wrap_gap_pre equals wrap_abs32."""
    src = os.path.join(ROOT, "tests", "host", "wrap_filter_check.cpp")
    exe = str(tmp_path / "wrap_filter_check")
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "--cuda-host-only",
                           "-x", "hip", src, "-o", exe], stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "false rejections: 0," in out.stdout
    assert "gap mismatches: 0," in out.stdout and "box false rejections: 0\n" in out.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_double_coordinate_float_filter_never_rejects(tmp_path):
    """The packet walk's float32 pre-filter for double coordinates (fpf_*):
    for periodic and open metrics, box sizes 1 to 12345, gaps down to 1e-7 of
    the box and pairs at the wrap, every leaf with exact d2 <= w passes and no
    leaf with d2 >= lo falls below the floor."""
    src = os.path.join(ROOT, "tests", "host", "fp_filter_check.cpp")
    exe = str(tmp_path / "fp_filter_check")
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "--cuda-host-only",
                           "-x", "hip", src, "-o", exe], stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "false rejections: 0, false floor rejections: 0," in out.stdout
