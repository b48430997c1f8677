"""The C++ drop-in adapters (ultrahigh_ann_amd/adapters) compiled against the
reference's own public headers and linked with its library sources and
libuhann.so (oracle/Makefile target `adapted`, built by
__graft_entry__.build() where the reference checkout exists). Here, without
a GPU, each of the reference's eight device test programs must take its
"no device" branch: every engine constructor and GPU probability call throws
std::runtime_error (e.g. tests/exact/l2/cuda_exact_l2_index_test.cpp:61-81)."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")
PROGRAMS = ["cuda_exact_l1_index_test", "cuda_exact_l2_index_test", "cuda_sampled_coordinate_l1_ann_index_test",
            "cuda_sampled_coordinate_l2_ann_index_test", "cuda_hierarchical_l1_ann_index_test",
            "cuda_hierarchical_l2_ann_index_test", "l1_importance_sampling_test", "l2_importance_sampling_test"]
# the reference's CTest environment for its probability tests (CMakeLists.txt:582-588)
ENV = dict(os.environ, OMP_NUM_THREADS="4", OMP_DYNAMIC="FALSE")


def _program(name):
    path = os.path.join(REF_BIN, name)
    if not os.path.exists(path):
        pytest.skip("adapted reference tests not built (needs the reference checkout)")
    return path


def _has_gpu():
    import ultrahigh_ann_amd as uh

    return uh.device_available()


@pytest.mark.parametrize("name", PROGRAMS)
def test_reference_device_tests_without_device(name):
    path = _program(name)
    if _has_gpu():
        pytest.skip("a device is visible: test_gpu_adapters.py runs these programs")
    res = subprocess.run([path], capture_output=True, text=True, timeout=120, env=ENV)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "FAILED" not in res.stderr
