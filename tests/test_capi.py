"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls: those need the GPU)."""
import ctypes
import glob
import os
import re

import pytest

import toposzp_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b((?:szp_cuda|tszp)_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_drop_in_surface():
    names = declared_functions()
    for required in ("szp_cuda_find_critical_points", "szp_cuda_sort_critical_points_by_original_data",
                     "szp_cuda_compress_sort_positions",
                     "szp_cuda_float_compress_randomaccess_topology_preserved",
                     "szp_cuda_float_decompress_randomaccess_topology_preserved",
                     "szp_cuda_decompress_sort_positions", "szp_cuda_find_critical_points_3d",
                     "szp_cuda_sort_critical_points_3d", "szp_cuda_compress_sort_positions_3d",
                     "szp_cuda_float_compress_topology_3d", "szp_cuda_float_compress_randomaccess",
                     "szp_cuda_float_decompress_randomaccess"):
        assert required in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    L = toposzp_hip.lib()
    assert hasattr(L, name), f"{name} declared in include/ but not exported"


def test_version_without_gpu():
    assert toposzp_hip.lib().tszp_version() >= 1
    L = toposzp_hip.lib()
    # pure host arithmetic: bounds of the worst-case stream sizes
    assert L.tszp_blocks(130, 64) == 3
    worst64 = 5 + 8 + 4 * 63 + 16
    assert L.tszp_topo_bound(128, 64, 1) == 8 + 2 * worst64 + 8


def test_missing_library_raises(monkeypatch):
    monkeypatch.setattr(toposzp_hip, "_lib", None)
    monkeypatch.setattr(toposzp_hip, "LIB_PATH", "/nonexistent/libtoposzp_hip.so")
    with pytest.raises(toposzp_hip.TszpError):
        toposzp_hip.lib()
