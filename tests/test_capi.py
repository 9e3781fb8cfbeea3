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
                     "szp_cuda_float_decompress_randomaccess",
                     # the rest of SZp/include/szp_cuda_compress.cuh, szp_cuda_decompress.cuh and
                     # szp_cuda_topology_3d.cuh
                     "szp_cuda_double_compress_randomaccess", "szp_cuda_double_compress_randomaccess_arg",
                     "szp_cuda_double_decompress_randomaccess", "szp_cuda_double_decompress_randomaccess_arg",
                     "szp_cuda_float_compress_threadblock", "szp_cuda_float_compress_threadblock_arg",
                     "szp_cuda_double_compress_threadblock", "szp_cuda_double_compress_threadblock_arg",
                     "szp_cuda_float_decompress_threadblock", "szp_cuda_float_decompress_threadblock_arg",
                     "szp_cuda_double_decompress_threadblock", "szp_cuda_double_decompress_threadblock_arg",
                     "szp_cuda_float_direct_predict_quantization",
                     "szp_cuda_float_threadblock_predict_quantization",
                     "szp_cuda_double_compress_topology_3d"):
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


@pytest.mark.parametrize("n,bs,T,ra", [(1, 8, 1, 0), (5, 64, 8, 0), (17, 1, 4, 0), (65, 64, 1, 1), (200, 7, 3, 1),
                                       (4097, 1025, 5, 0), (50000, 2, 4096, 0), (1000, 64, 3, 1)])
def test_elem_unit_counts_without_gpu(n, bs, T, ra):
    """Host-only sizing of the element-partitioned streams: one unit per SZp block of
    each chunk (random access: ceil(len / bs); threadblock: the anchor plus
    ceil((len - 1) / bs) blocks, at least one unit for a one-element chunk), and a
    bound no smaller than the oracle's stream on incompressible data."""
    import numpy as np
    import oracle
    L = toposzp_hip.lib()
    ce = n // T
    want = 0
    for t in range(T):
        length = (n if t == T - 1 else (t + 1) * ce) - t * ce
        if length == 0:
            continue
        want += -(-length // bs) if ra else max(1, -(-(length - 1) // bs))
    assert L.tszp_elem_units(n, bs, T, ra) == want
    for ds, dt in ((4, np.float32), (8, np.float64)):
        if ra and ds == 4:
            continue
        x = np.random.default_rng(n).normal(0, 1e6, n).astype(dt)  # every bit count up to 32
        s = oracle.elem_compress(x, 1e-3, bs, T=T, ra=bool(ra))
        assert L.tszp_elem_bound(n, bs, T, ds, ra) >= s.size + 4
