"""Python front end of libtoposzp_hip.so (ctypes).

Two layers, mirroring the C-ABI:
  * the reference-compatible drop-in (szp_cuda_*; numpy arrays in host memory),
  * the native device API (tszp_*; torch tensors resident on the GPU).

The HIP library is the only implementation: if it is missing or fails to load,
every call raises.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtoposzp_hip.so")

CP2 = np.dtype([("x", "<i4"), ("y", "<i4"), ("type", "<i4"), ("quantized_bin", "<i4"),
                ("sort_position", "<i4")])
CP3 = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("type", "<i4"),
                ("quantized_bin", "<i4"), ("sort_position", "<i4")])


class TszpError(RuntimeError):
    pass


class TopoInfo(C.Structure):
    _fields_ = [("topo_bytes", C.c_uint64), ("sortpos_bytes", C.c_uint64),
                ("cp_count", C.c_uint64), ("ext_count", C.c_uint64)]


class RestoreInfo(C.Structure):
    _fields_ = [("missing", C.c_uint64), ("spurious", C.c_uint64), ("changed", C.c_uint64),
                ("moves", C.c_uint64), ("windows", C.c_uint64), ("windows_reordered", C.c_uint64), ("rounds", C.c_uint32),
                ("max_dev_over_eb", C.c_float), ("solve_tiles", C.c_uint32), ("solve_chain", C.c_uint32)]

    def mismatches(self):
        return self.missing + self.spurious + self.changed


class BatchField(C.Structure):
    """tszp_batch_field (include/toposzp_hip.h)."""
    _fields_ = [("d_field", C.c_void_p), ("n", C.c_uint64), ("eb", C.c_float), ("block_size", C.c_int),
                ("d_stream", C.c_void_p), ("stream_cap", C.c_uint64), ("d_blk", C.c_void_p),
                ("stream_bytes", C.c_uint64)]


class TopoBatchField(C.Structure):
    _fields_ = [("d_field", C.c_void_p), ("ndim", C.c_int), ("dims", C.c_uint32 * 3), ("eb", C.c_float),
                ("block_size", C.c_int), ("d_topo", C.c_void_p), ("topo_cap", C.c_uint64),
                ("d_topo_blk", C.c_void_p), ("d_sortpos", C.c_void_p), ("sortpos_cap", C.c_uint64),
                ("d_sortpos_blk", C.c_void_p), ("info", TopoInfo), ("restore", RestoreInfo)]


class TicketInfo(C.Structure):
    """tszp_ticket_info (include/toposzp_hip.h)."""
    _fields_ = [("stream_bytes", C.c_uint64), ("queue_us", C.c_double), ("latency_us", C.c_double)]


class ServerStats(C.Structure):
    """tszp_server_stats_t (include/toposzp_hip.h)."""
    _fields_ = [(k, C.c_uint64) for k in ("batches", "requests", "queued", "hbm_in_use", "hbm_peak", "hbm_budget")]


_lib = None
_libc = C.CDLL("libc.so.6")
_libc.free.argtypes = [C.c_void_p]


class Quality(C.Structure):
    """tszp_quality (include/toposzp_hip.h)."""
    _fields_ = [("orig_cp", C.c_uint64 * 4), ("dec_cp", C.c_uint64 * 4), ("preserved", C.c_uint64),
                ("new_cps", C.c_uint64), ("lost", C.c_uint64 * 4), ("changed", C.c_uint64 * 4),
                ("over_1eb", C.c_uint64), ("over_2eb", C.c_uint64), ("saddles", C.c_uint64),
                ("saddles_above", C.c_uint64 * 6), ("saddles_kept_above", C.c_uint64 * 6),
                ("pairs", C.c_uint64 * 4), ("violations", C.c_uint64 * 4), ("max_err", C.c_double),
                ("pers_min", C.c_double), ("pers_max", C.c_double), ("pers_sum", C.c_double)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TszpError(f"{LIB_PATH} is not built; run toposzp-cuda_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, sz, i32, u32, f32 = C.c_void_p, C.c_size_t, C.c_int, C.c_uint, C.c_float
    u64p = C.POINTER(C.c_uint64)
    L.tszp_last_error.restype = C.c_char_p
    L.tszp_version.restype = i32
    for f in ("tszp_topo_bound", "tszp_sortpos_bound", "tszp_plain_bound"):
        getattr(L, f).argtypes = [sz, i32, u32]
        getattr(L, f).restype = sz
    L.tszp_blocks.argtypes = [sz, i32]; L.tszp_blocks.restype = sz
    L.tszp_rel_to_abs.argtypes = [vp, sz, f32, C.POINTER(f32), C.POINTER(f32), vp]
    L.tszp_topo_compress.argtypes = [vp, i32, vp, f32, i32, u32, vp, sz, vp, vp, sz, vp, vp,
                                     C.POINTER(TopoInfo), vp]
    L.tszp_topo_decompress.argtypes = [vp, vp, u32, sz, f32, i32, vp, vp, vp]
    L.tszp_sortpos_decompress.argtypes = [vp, vp, u32, sz, i32, vp, vp]
    L.tszp_plain_compress.argtypes = [vp, sz, f32, i32, u32, vp, sz, vp, u64p, vp]
    L.tszp_plain_decompress.argtypes = [vp, vp, u32, sz, f32, i32, vp, vp]
    L.tszp_server_last_error.restype = C.c_char_p
    L.tszp_server_create.argtypes = [C.c_uint64, u32, C.POINTER(vp)]
    L.tszp_server_destroy.argtypes = [vp]
    L.tszp_server_compress.argtypes = [vp, vp, C.c_uint64, f32, i32, u64p]
    L.tszp_server_decompress.argtypes = [vp, C.c_uint64, vp, u64p]
    L.tszp_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(TicketInfo)]
    L.tszp_server_stream.argtypes = [vp, C.c_uint64, C.POINTER(vp), C.POINTER(vp), u64p, C.POINTER(i32)]
    L.tszp_server_release.argtypes = [vp, C.c_uint64]
    L.tszp_server_copy_stream.argtypes = [vp, C.c_uint64, vp, C.c_uint64, vp]
    L.tszp_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
    for f in ("tszp_server_create", "tszp_server_destroy", "tszp_server_compress", "tszp_server_decompress",
              "tszp_server_wait", "tszp_server_stream", "tszp_server_release", "tszp_server_stats",
              "tszp_server_copy_stream"):
        getattr(L, f).restype = i32
    L.tszp_topo_server_last_error.restype = C.c_char_p
    L.tszp_topo_server_create.argtypes = [C.c_uint64, u32, C.c_uint64, u32, C.POINTER(vp)]
    L.tszp_topo_server_destroy.argtypes = [vp]
    L.tszp_topo_server_compress.argtypes = [vp, vp, i32, vp, f32, i32, u64p]
    L.tszp_topo_server_decompress.argtypes = [vp, C.c_uint64, vp, vp, i32, u64p]
    L.tszp_topo_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(TicketInfo)]
    L.tszp_topo_server_result.argtypes = [vp, C.c_uint64, C.POINTER(TopoInfo), C.POINTER(RestoreInfo),
                                          C.POINTER(i32)]
    L.tszp_topo_server_streams.argtypes = [vp, C.c_uint64, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                           C.POINTER(vp)]
    L.tszp_topo_server_release.argtypes = [vp, C.c_uint64]
    L.tszp_topo_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
    for f in ("tszp_topo_server_create", "tszp_topo_server_destroy", "tszp_topo_server_compress",
              "tszp_topo_server_decompress", "tszp_topo_server_wait", "tszp_topo_server_result",
              "tszp_topo_server_streams", "tszp_topo_server_release", "tszp_topo_server_stats"):
        getattr(L, f).restype = i32
    L.tszp_topo_batch_mem_bound.argtypes = [C.c_uint64, u32]
    L.tszp_topo_batch_mem_bound.restype = C.c_uint64
    L.tszp_batch_plan_create.argtypes = [C.POINTER(BatchField), u32, vp, C.POINTER(vp), vp]
    L.tszp_batch_plan_destroy.argtypes = [vp]
    L.tszp_batch_plan_compress.argtypes = [vp, vp]
    L.tszp_batch_plan_decompress.argtypes = [vp, vp]
    L.tszp_batch_plan_results.argtypes = [vp, u64p, vp]
    for f in ("tszp_batch_plan_create", "tszp_batch_plan_destroy", "tszp_batch_plan_compress",
              "tszp_batch_plan_decompress", "tszp_batch_plan_results"):
        getattr(L, f).restype = i32
    for f in ("tszp_plain_compress_batch", "tszp_plain_decompress_batch", "tszp_batch_tune"):
        getattr(L, f).argtypes = [C.POINTER(BatchField), u32, vp]
        getattr(L, f).restype = i32
    L.tszp_topo_batch_create.argtypes = [C.POINTER(TopoBatchField), u32, C.POINTER(vp), vp]
    L.tszp_topo_batch_destroy.argtypes = [vp]
    L.tszp_topo_batch_compress.argtypes = [vp, vp]
    L.tszp_topo_batch_decompress.argtypes = [vp, vp]
    L.tszp_topo_batch_restore.argtypes = [vp, i32, vp]
    L.tszp_topo_batch_results.argtypes = [vp, C.POINTER(TopoBatchField), vp]
    L.tszp_topo_batch_copy.argtypes = [vp, u32, i32, vp, C.c_uint64, vp]
    for f in ("tszp_topo_batch_create", "tszp_topo_batch_destroy", "tszp_topo_batch_compress",
              "tszp_topo_batch_decompress", "tszp_topo_batch_restore", "tszp_topo_batch_results",
              "tszp_topo_batch_copy"):
        getattr(L, f).restype = i32
    L.tszp_classify.argtypes = [vp, i32, vp, vp, u64p, vp]
    L.tszp_gen_blobs.argtypes = [vp, i32, vp, C.c_uint64, i32, f32, f32, f32, i32, vp]
    L.tszp_gen_fbm.argtypes = [vp, vp, C.c_uint64, i32, f32, vp]
    for f in ("tszp_rel_to_abs", "tszp_topo_compress", "tszp_topo_decompress",
              "tszp_sortpos_decompress", "tszp_plain_compress", "tszp_plain_decompress",
              "tszp_classify", "tszp_gen_blobs", "tszp_gen_fbm"):
        getattr(L, f).restype = i32
    # drop-in
    L.szp_cuda_find_critical_points.argtypes = [vp, C.POINTER(sz), i32, i32, f32]
    L.szp_cuda_find_critical_points.restype = vp
    L.szp_cuda_find_critical_points_3d.argtypes = [vp, C.POINTER(sz), i32, i32, i32, f32]
    L.szp_cuda_find_critical_points_3d.restype = vp
    L.szp_cuda_sort_critical_points_by_original_data.argtypes = [vp, sz, vp, i32]
    L.szp_cuda_sort_critical_points_by_original_data.restype = None
    L.szp_cuda_sort_critical_points_3d.argtypes = [vp, sz, vp, i32, i32]
    L.szp_cuda_sort_critical_points_3d.restype = None
    L.szp_cuda_compress_sort_positions.argtypes = [vp, sz, C.POINTER(sz), i32]
    L.szp_cuda_compress_sort_positions.restype = vp
    L.szp_cuda_compress_sort_positions_3d.argtypes = [vp, sz, C.POINTER(sz), i32]
    L.szp_cuda_compress_sort_positions_3d.restype = vp
    L.szp_cuda_float_compress_randomaccess_topology_preserved.argtypes = [
        vp, C.POINTER(sz), f32, sz, i32, vp, i32, i32, i32]
    L.szp_cuda_float_compress_randomaccess_topology_preserved.restype = vp
    L.szp_cuda_float_compress_topology_3d.argtypes = [vp, C.POINTER(sz), f32, sz, i32, vp, i32,
                                                      i32, i32, i32]
    L.szp_cuda_float_compress_topology_3d.restype = vp
    L.szp_cuda_float_decompress_randomaccess_topology_preserved.argtypes = [
        C.POINTER(vp), sz, f32, i32, vp, C.POINTER(vp)]
    L.szp_cuda_float_decompress_randomaccess_topology_preserved.restype = None
    L.szp_cuda_decompress_sort_positions.argtypes = [vp, sz, i32]
    L.szp_cuda_decompress_sort_positions.restype = vp
    L.szp_cuda_float_compress_randomaccess.argtypes = [vp, C.POINTER(sz), f32, sz, i32]
    L.szp_cuda_float_compress_randomaccess.restype = vp
    L.szp_cuda_float_compress_randomaccess_arg.argtypes = [vp, vp, C.POINTER(sz), f32, sz, i32]
    L.szp_cuda_float_compress_randomaccess_arg.restype = None
    L.szp_cuda_float_compress_randomaccess_bound.argtypes = [sz, i32]
    L.szp_cuda_float_compress_randomaccess_bound.restype = sz
    L.szp_cuda_float_decompress_randomaccess.argtypes = [sz, f32, i32, vp]
    L.szp_cuda_float_decompress_randomaccess.restype = vp
    L.szp_cuda_float_decompress_randomaccess_arg.argtypes = [vp, sz, f32, i32, vp]
    L.szp_cuda_float_decompress_randomaccess_arg.restype = None
    L.tszp_topo_restore.argtypes = [vp, vp, vp, vp, C.c_uint64, i32, vp, f32, i32, C.POINTER(RestoreInfo), vp]
    L.tszp_topo_restore.restype = i32
    L.tszp_topo_verify.argtypes = [vp, vp, i32, vp, u64p, vp]
    L.tszp_topo_verify.restype = i32
    L.tszp_quality_metrics.argtypes = [vp, vp, i32, vp, f32, C.POINTER(Quality), vp]
    L.tszp_quality_metrics.restype = i32
    L.tszp_parse_blocks.argtypes = [vp, sz, sz, i32, i32, vp, C.POINTER(u32), vp]
    L.tszp_parse_blocks.restype = i32
    L.tszp_elem_parse.argtypes = [vp, sz, sz, i32, i32, i32, vp, C.POINTER(u32), vp]
    L.tszp_elem_parse.restype = i32
    L.tszp_szp_pack.argtypes = [vp, sz, vp, sz, vp, sz, C.POINTER(sz), vp]
    L.tszp_szp_pack.restype = i32
    L.tszp_szp_unpack.argtypes = [vp, sz, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz), u64p]
    L.tszp_szp_unpack.restype = i32
    f64 = C.c_double
    L.tszp_elem_units.argtypes = [sz, i32, u32, i32]; L.tszp_elem_units.restype = sz
    L.tszp_elem_bound.argtypes = [sz, i32, u32, i32, i32]; L.tszp_elem_bound.restype = sz
    L.tszp_threadblock_compress.argtypes = [vp, i32, sz, f64, i32, u32, vp, sz, vp, u64p, vp]
    L.tszp_threadblock_decompress.argtypes = [vp, vp, u32, sz, f64, i32, i32, vp, vp]
    L.tszp_plain_compress_f64.argtypes = [vp, sz, f64, i32, u32, vp, sz, vp, u64p, vp]
    L.tszp_plain_decompress_f64.argtypes = [vp, vp, u32, sz, f64, i32, vp, vp]
    L.tszp_topo_compress_f64_3d.argtypes = [vp, vp, f64, i32, vp, C.c_uint64, vp, sz, vp, u64p, vp,
                                            u64p, vp]
    for f in ("tszp_threadblock_compress", "tszp_threadblock_decompress", "tszp_plain_compress_f64",
              "tszp_plain_decompress_f64", "tszp_topo_compress_f64_3d"):
        getattr(L, f).restype = i32
    L.tszp_prof_enable.argtypes = [i32]; L.tszp_prof_enable.restype = None
    L.tszp_prof_reset.argtypes = []; L.tszp_prof_reset.restype = None
    L.tszp_prof_read.argtypes = [i32, C.POINTER(C.c_double), u64p]; L.tszp_prof_read.restype = i32
    _lib = L
    return L


PROF_KINDS = {"codec_topo": 0, "codec_int": 1, "codec_plain": 2, "decode": 3, "sort": 4, "rank": 5,
              "restore": 6, "scatter": 7, "solve": 8, "query": 9, "elem": 10, "elem_decode": 11,
              "saddle": 12, "batch": 13, "batch_decode": 14, "topo_batch": 15, "topo_batch_decode": 16,
              "topo_batch_restore": 17, "topo_batch_sort": 18}


def prof_enable(on=True):
    lib().tszp_prof_enable(int(on))


def prof_reset():
    lib().tszp_prof_reset()


def prof_read():
    """{kind: (total_ms, launches)} of the event-timed kernels since prof_reset()."""
    out = {}
    for k, v in PROF_KINDS.items():
        ms, n = C.c_double(), C.c_uint64()
        _check(lib().tszp_prof_read(v, C.byref(ms), C.byref(n)))
        out[k] = (ms.value, n.value)
    return out


def last_error():
    return lib().tszp_last_error().decode()


def _check(rc):
    if rc != 0:
        raise TszpError(f"toposzp_hip error {rc}: {last_error()}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _take(ptr, nbytes, dtype):
    if not ptr:
        raise TszpError(f"NULL result: {last_error()}")
    buf = (C.c_ubyte * nbytes).from_address(ptr)
    out = np.frombuffer(bytes(buf), dtype=dtype).copy()
    _libc.free(ptr)
    return out


# ------------------------------------------------------------------------------------------
# drop-in (host pointers), reference names
# ------------------------------------------------------------------------------------------
def find_critical_points(data, eb):
    d = np.ascontiguousarray(data, dtype=np.float32)
    cnt = C.c_size_t()
    if d.ndim == 2:
        p = lib().szp_cuda_find_critical_points(_p(d), C.byref(cnt), d.shape[0], d.shape[1], eb)
        dt = CP2
    else:
        p = lib().szp_cuda_find_critical_points_3d(_p(d), C.byref(cnt), *d.shape, eb)
        dt = CP3
    if not p:
        if last_error():
            raise TszpError(last_error())
        return np.zeros(0, dtype=dt)
    return _take(p, cnt.value * dt.itemsize, dt)


def sort_critical_points(cps, data):
    d = np.ascontiguousarray(data, dtype=np.float32)
    out = np.ascontiguousarray(cps).copy()
    if out.size == 0:
        return out
    if out.dtype == CP2:
        lib().szp_cuda_sort_critical_points_by_original_data(_p(out), out.size, _p(d), d.shape[1])
    else:
        lib().szp_cuda_sort_critical_points_3d(_p(out), out.size, _p(d), d.shape[1], d.shape[2])
    return out


def compress_sort_positions(cps, bs):
    c = np.ascontiguousarray(cps)
    n = C.c_size_t()
    f = (lib().szp_cuda_compress_sort_positions if c.dtype == CP2
         else lib().szp_cuda_compress_sort_positions_3d)
    p = f(_p(c), c.size, C.byref(n), bs)
    if not p:
        return np.zeros(0, np.uint8)
    return _take(p, n.value, np.uint8)


def compress_topology(data, eb, bs, cps):
    d = np.ascontiguousarray(data, dtype=np.float32)
    c = np.ascontiguousarray(cps)
    n = C.c_size_t()
    if d.ndim == 2:
        p = lib().szp_cuda_float_compress_randomaccess_topology_preserved(
            _p(d), C.byref(n), eb, d.size, bs, _p(c), c.size, d.shape[0], d.shape[1])
    else:
        p = lib().szp_cuda_float_compress_topology_3d(_p(d), C.byref(n), eb, d.size, bs, _p(c),
                                                      c.size, *d.shape)
    return _take(p, n.value, np.uint8)


def decompress_topology(cmp, n, eb, bs):
    c = np.ascontiguousarray(cmp, dtype=np.uint8)
    dp, fp = C.c_void_p(), C.c_void_p()
    lib().szp_cuda_float_decompress_randomaccess_topology_preserved(C.byref(dp), n, eb, bs, _p(c),
                                                                    C.byref(fp))
    if not dp.value:
        raise TszpError(last_error())
    return _take(dp.value, 4 * n, np.float32), _take(fp.value, 4 * n, np.int32)


def decompress_sort_positions(cmp_at_offsets, count, bs):
    c = np.ascontiguousarray(cmp_at_offsets, dtype=np.uint8)
    p = lib().szp_cuda_decompress_sort_positions(_p(c), count, bs)
    return _take(p, 4 * count, np.int32)


def compress_plain(data, eb, bs):
    d = np.ascontiguousarray(data, dtype=np.float32).ravel()
    n = C.c_size_t()
    p = lib().szp_cuda_float_compress_randomaccess(_p(d), C.byref(n), eb, d.size, bs)
    return _take(p, n.value, np.uint8)


def decompress_plain(cmp_after_header, n, eb, bs):
    c = np.ascontiguousarray(cmp_after_header, dtype=np.uint8)
    p = lib().szp_cuda_float_decompress_randomaccess(n, eb, bs, _p(c))
    return _take(p, 4 * n, np.float32)


# ------------------------------------------------------------------------------------------
# native device API (torch tensors on the GPU)
# ------------------------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _dp(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dims(shape):
    return (C.c_uint32 * len(shape))(*shape)


def bounds(n, bs, T=1):
    L = lib()
    return L.tszp_topo_bound(n, bs, T), L.tszp_sortpos_bound(n, bs, T), L.tszp_plain_bound(n, bs, T)


def rel_to_abs(field, rel):
    eb, rng = C.c_float(), C.c_float()
    _check(lib().tszp_rel_to_abs(_dp(field), field.numel(), rel, C.byref(eb), C.byref(rng), _stream()))
    return eb.value, rng.value


class TopoBuffers:
    """Device buffers sized for one field shape / block size / chunk count."""

    def __init__(self, n, bs, T=1, device="cuda"):
        torch = _torch()
        tb, sb, _ = bounds(n, bs, T)
        nb = lib().tszp_blocks(n, bs)
        self.n, self.bs, self.T = n, bs, T
        self.topo = torch.empty(tb, dtype=torch.uint8, device=device)
        self.topo_blk = torch.empty(nb + 1, dtype=torch.int32, device=device)
        self.sp = torch.empty(sb, dtype=torch.uint8, device=device)
        self.sp_blk = torch.empty(nb + 1, dtype=torch.int32, device=device)
        self.sp_values = torch.empty(n, dtype=torch.int32, device=device)
        self.out = torch.empty(n, dtype=torch.float32, device=device)
        self.types = torch.empty(n, dtype=torch.uint8, device=device)
        self.sp_out = torch.empty(n, dtype=torch.int32, device=device)


def topo_compress(field, eb, bs, T=1, buf=None):
    """field: float32 CUDA tensor (2D or 3D).  Returns (info, buffers)."""
    if buf is None:
        buf = TopoBuffers(field.numel(), bs, T, field.device)
    info = TopoInfo()
    _check(lib().tszp_topo_compress(
        _dp(field), field.dim(), _dims(field.shape), eb, bs, T, _dp(buf.topo), buf.topo.numel(),
        _dp(buf.topo_blk), _dp(buf.sp), buf.sp.numel(), _dp(buf.sp_blk), _dp(buf.sp_values),
        C.byref(info), _stream()))
    return info, buf


def topo_decompress(buf, n, eb, bs, T=1):
    _check(lib().tszp_topo_decompress(_dp(buf.topo), _dp(buf.topo_blk), T, n, eb, bs,
                                      _dp(buf.out), _dp(buf.types), _stream()))
    return buf.out, buf.types


def sortpos_decompress(buf, e, bs, T=1):
    _check(lib().tszp_sortpos_decompress(_dp(buf.sp), _dp(buf.sp_blk), T, e, bs, _dp(buf.sp_out),
                                         _stream()))
    return buf.sp_out[:e]


def plain_compress(field, eb, bs, T=1, out=None, blk=None):
    torch = _torch()
    n = field.numel()
    if out is None:
        out = torch.empty(lib().tszp_plain_bound(n, bs, T), dtype=torch.uint8, device=field.device)
    if blk is None:
        blk = torch.empty(lib().tszp_blocks(n, bs) + 1, dtype=torch.int32, device=field.device)
    nbytes = C.c_uint64()
    _check(lib().tszp_plain_compress(_dp(field), n, eb, bs, T, _dp(out), out.numel(), _dp(blk),
                                     C.byref(nbytes), _stream()))
    return nbytes.value, out, blk


def plain_decompress(stream, blk, n, eb, bs, T=1, out=None):
    torch = _torch()
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=stream.device)
    _check(lib().tszp_plain_decompress(_dp(stream), _dp(blk), T, n, eb, bs, _dp(out), _stream()))
    return out


def _batch_table(fields, ebs, bss, streams=None, blks=None):
    n = len(fields)
    if not (len(ebs) == n and len(bss) == n):
        raise ValueError("one error bound and one block size per field")
    tab = (BatchField * n)()
    for i, f in enumerate(fields):
        t = tab[i]
        t.d_field = f.data_ptr()
        t.n = f.numel()
        t.eb = float(ebs[i])
        t.block_size = int(bss[i])
        if streams is not None:
            t.d_stream = streams[i].data_ptr()
            t.stream_cap = streams[i].numel()
            t.d_blk = blks[i].data_ptr()
    return tab


def batch_tune(fields, ebs):
    """Block size per field chosen by the library's autotuner (tszp_batch_tune)."""
    tab = _batch_table(fields, ebs, [0] * len(fields))
    _check(lib().tszp_batch_tune(tab, len(fields), _stream()))
    return [tab[i].block_size for i in range(len(fields))]


def plain_compress_batch(fields, ebs, bss, streams=None, blks=None):
    """Plain one-chunk streams of many float32 CUDA tensors in one launch per block-size
    class: (bytes per field, streams, side-cars)."""
    torch = _torch()
    L = lib()
    if streams is None:
        streams = [torch.empty(L.tszp_plain_bound(f.numel(), b, 1), dtype=torch.uint8, device=f.device)
                   for f, b in zip(fields, bss)]
        blks = [torch.empty(L.tszp_blocks(f.numel(), b) + 1, dtype=torch.int32, device=f.device)
                for f, b in zip(fields, bss)]
    tab = _batch_table(fields, ebs, bss, streams, blks)
    _check(L.tszp_plain_compress_batch(tab, len(fields), _stream()))
    return [tab[i].stream_bytes for i in range(len(fields))], streams, blks


def plain_decompress_batch(streams, blks, ns, ebs, bss, outs=None):
    """Decodes many plain one-chunk streams in one launch per block-size class."""
    torch = _torch()
    if outs is None:
        outs = [torch.empty(n, dtype=torch.float32, device=s.device) for s, n in zip(streams, ns)]
    tab = _batch_table(outs, ebs, bss, streams, blks)
    _check(lib().tszp_plain_decompress_batch(tab, len(outs), _stream()))
    return outs


class BatchPlan:
    """tszp_batch_plan_*: the layout of a batch kept on the device.  compress() and
    decompress() only launch kernels on the current stream (capturable in a CUDA/HIP
    graph); results() synchronises and returns the stream sizes.  fields, streams, side-
    cars and outputs must stay allocated while the plan lives."""

    def __init__(self, fields, ebs, bss, streams, blks, outs):
        self._keep = (fields, streams, blks, outs)
        self.n = len(fields)
        self._ctab = _batch_table(fields, ebs, bss, streams, blks)
        if len(outs) != self.n or any(o.numel() != f.numel() for o, f in zip(outs, fields)):
            raise ValueError("one output per field, of the field's size")
        self._outs = (C.c_void_p * self.n)(*[o.data_ptr() for o in outs])
        self._h = C.c_void_p()
        _check(lib().tszp_batch_plan_create(self._ctab, self.n, self._outs, C.byref(self._h), _stream()))

    def compress(self):
        _check(lib().tszp_batch_plan_compress(self._h, _stream()))

    def decompress(self):
        _check(lib().tszp_batch_plan_decompress(self._h, _stream()))

    def results(self):
        out = (C.c_uint64 * self.n)()
        _check(lib().tszp_batch_plan_results(self._h, out, _stream()))
        return list(out)

    def close(self):
        if self._h:
            _check(lib().tszp_batch_plan_destroy(self._h))
            self._h = C.c_void_p()


class TopoBatch:
    """Batched topology pipeline over many 2D / 3D float32 CUDA tensors (tszp_topo_batch_*):
    compress / decompress are launches only (capturable in a graph), restore runs the
    restoration of every field, results() reads sizes and counts back.  bss: block size
    per field, 0 for the library's choice."""

    def __init__(self, fields, ebs, bss=None, streams=None):
        torch = _torch()
        L = lib()
        n = len(fields)
        bss = list(bss) if bss is not None else [0] * n
        self.fields = fields
        tab = (TopoBatchField * n)()
        for i, f in enumerate(fields):
            t = tab[i]
            t.d_field = f.data_ptr() if f is not None else None
            t.ndim = f.dim()
            for a in range(f.dim()):
                t.dims[a] = f.shape[a]
            t.eb = float(ebs[i])
            t.block_size = int(bss[i])
        if any(b == 0 for b in bss):  # the tuner needs no buffers: ask it for the block sizes first
            tmp = _batch_table([f.view(-1) for f in fields], ebs, [0] * n)
            _check(L.tszp_batch_tune(tmp, n, _stream()))
            for i in range(n):
                if bss[i] == 0:
                    bss[i] = tmp[i].block_size
                    tab[i].block_size = bss[i]
        self.bss = bss
        if streams is None:
            streams = []
            for i, f in enumerate(fields):
                m, b = f.numel(), bss[i]
                streams.append((torch.empty(L.tszp_topo_bound(m, b, 1), dtype=torch.uint8, device=f.device),
                                torch.empty(L.tszp_blocks(m, b) + 1, dtype=torch.int32, device=f.device),
                                torch.empty(L.tszp_sortpos_bound(m, b, 1), dtype=torch.uint8, device=f.device),
                                torch.empty(L.tszp_blocks(m, b) + 1, dtype=torch.int32, device=f.device)))
        self.streams = streams
        for i, (tp, tb, sp, sb) in enumerate(streams):
            t = tab[i]
            t.d_topo, t.topo_cap, t.d_topo_blk = tp.data_ptr(), tp.numel(), tb.data_ptr()
            t.d_sortpos, t.sortpos_cap, t.d_sortpos_blk = sp.data_ptr(), sp.numel(), sb.data_ptr()
        self.tab = tab
        self._h = C.c_void_p()
        _check(L.tszp_topo_batch_create(tab, n, C.byref(self._h), _stream()))

    def compress(self):
        _check(lib().tszp_topo_batch_compress(self._h, _stream()))

    def decompress(self):
        _check(lib().tszp_topo_batch_decompress(self._h, _stream()))

    def restore(self, max_rounds=16):
        _check(lib().tszp_topo_batch_restore(self._h, max_rounds, _stream()))

    def results(self):
        """[(TopoInfo, RestoreInfo)] per field, in the caller's order."""
        _check(lib().tszp_topo_batch_results(self._h, self.tab, _stream()))
        return [(self.tab[i].info, self.tab[i].restore) for i in range(len(self.fields))]

    def copy(self, i, which, count=None):
        """Field i's decoded values ('decoded'), restored values ('restored'), types
        ('types') or sort positions ('sortpos', count = its extremum count) as a new tensor."""
        torch = _torch()
        f = self.fields[i]
        k = {"decoded": 0, "restored": 1, "types": 2, "sortpos": 3}[which]
        m = f.numel() if count is None else count
        dt = {0: torch.float32, 1: torch.float32, 2: torch.uint8, 3: torch.int32}[k]
        out = torch.empty(max(m, 1), dtype=dt, device=f.device)
        _check(lib().tszp_topo_batch_copy(self._h, i, k, out.data_ptr(), m, _stream()))
        return out[:m]

    def close(self):
        if self._h:
            _check(lib().tszp_topo_batch_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop over the batch codec (tszp_server_*): submit from any thread, wait
    for tickets, release them when done.  Inputs must be complete on the device when
    submitted (the caller synchronises its stream first)."""

    def __init__(self, hbm_budget, max_batch=256):
        self._h = C.c_void_p()
        self._chk(lib().tszp_server_create(int(hbm_budget), int(max_batch), C.byref(self._h)))

    @staticmethod
    def _chk(rc):
        if rc != 0:
            raise TszpError(lib().tszp_server_last_error().decode())

    def compress(self, field, eb, bs):
        t = C.c_uint64()
        self._chk(lib().tszp_server_compress(self._h, _dp(field), field.numel(), float(eb), int(bs), C.byref(t)))
        return t.value

    def decompress(self, ticket, out):
        t = C.c_uint64()
        self._chk(lib().tszp_server_decompress(self._h, int(ticket), _dp(out), C.byref(t)))
        return t.value

    def wait(self, ticket):
        info = TicketInfo()
        self._chk(lib().tszp_server_wait(self._h, int(ticket), C.byref(info)))
        return info

    def stream(self, ticket):
        """(device address of the stream, of its side-car, stream bytes, block size)."""
        ds, db, nb, bs = C.c_void_p(), C.c_void_p(), C.c_uint64(), C.c_int()
        self._chk(lib().tszp_server_stream(self._h, int(ticket), C.byref(ds), C.byref(db), C.byref(nb),
                                           C.byref(bs)))
        return ds.value, db.value, nb.value, bs.value

    def stream_copy(self, ticket):
        """The stream of a finished compress ticket copied into a new uint8 CUDA tensor."""
        torch = _torch()
        _, _, nb, _ = self.stream(ticket)
        out = torch.empty(nb, dtype=torch.uint8, device="cuda")
        self._chk(lib().tszp_server_copy_stream(self._h, int(ticket), _dp(out), nb, _stream()))
        return out

    def release(self, ticket):
        self._chk(lib().tszp_server_release(self._h, int(ticket)))

    def stats(self):
        st = ServerStats()
        self._chk(lib().tszp_server_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            self._chk(lib().tszp_server_destroy(self._h))
            self._h = C.c_void_p()


_hip = None


def _memcpy_d2d(dst, src, nbytes):
    """hipMemcpy device to device between raw addresses (the HIP runtime's own call)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    rc = _hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), int(nbytes), 3)  # hipMemcpyDeviceToDevice
    if rc != 0:
        raise TszpError(f"hipMemcpy failed ({rc})")


class TopoServer:
    """Request loop of the topology pipeline (tszp_topo_server_*): compress requests of
    2D / 3D CUDA tensors and decompress (+ restore) requests of finished compress
    tickets, from any thread; `lanes` batches run on the device at a time.  Inputs must
    be complete on the device when submitted (the caller synchronises its stream
    first)."""

    def __init__(self, hbm_budget, max_batch=256, max_batch_elements=1 << 28, lanes=2):
        self._h = C.c_void_p()
        self._chk(lib().tszp_topo_server_create(int(hbm_budget), int(max_batch), int(max_batch_elements),
                                                int(lanes), C.byref(self._h)))

    @staticmethod
    def _chk(rc):
        if rc != 0:
            raise TszpError(lib().tszp_topo_server_last_error().decode())

    def compress(self, field, eb, bs=0):
        t = C.c_uint64()
        dims = (C.c_uint32 * 3)(*(list(field.shape) + [1] * (3 - field.dim())))
        self._chk(lib().tszp_topo_server_compress(self._h, _dp(field), field.dim(), dims, float(eb), int(bs),
                                                  C.byref(t)))
        return t.value

    def decompress(self, ticket, out, types=None, restore=True):
        t = C.c_uint64()
        self._chk(lib().tszp_topo_server_decompress(self._h, int(ticket), _dp(out),
                                                    _dp(types) if types is not None else None, int(bool(restore)),
                                                    C.byref(t)))
        return t.value

    def wait(self, ticket):
        info = TicketInfo()
        self._chk(lib().tszp_topo_server_wait(self._h, int(ticket), C.byref(info)))
        return info

    def result(self, ticket):
        """(TopoInfo, RestoreInfo, block size) of a finished ticket."""
        info, rinfo, bs = TopoInfo(), RestoreInfo(), C.c_int()
        self._chk(lib().tszp_topo_server_result(self._h, int(ticket), C.byref(info), C.byref(rinfo), C.byref(bs)))
        return info, rinfo, bs.value

    def streams(self, ticket):
        """(topology stream, sort-position stream) of a finished compress ticket, copied
        into new uint8 CUDA tensors."""
        torch = _torch()
        info, _, _ = self.result(ticket)
        dt, dtb, dsp, dspb = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        self._chk(lib().tszp_topo_server_streams(self._h, int(ticket), C.byref(dt), C.byref(dtb), C.byref(dsp),
                                                 C.byref(dspb)))
        out = []
        for addr, nb in ((dt.value, info.topo_bytes), (dsp.value, info.sortpos_bytes)):
            o = torch.empty(nb, dtype=torch.uint8, device="cuda")
            if nb:
                _memcpy_d2d(o.data_ptr(), addr, nb)
            out.append(o)
        return out[0], out[1]

    def release(self, ticket):
        self._chk(lib().tszp_topo_server_release(self._h, int(ticket)))

    def stats(self):
        st = ServerStats()
        self._chk(lib().tszp_topo_server_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            self._chk(lib().tszp_topo_server_destroy(self._h))
            self._h = C.c_void_p()


def elem_buffers(n, bs, T, elem_bytes, ra=False, device="cuda"):
    """(stream, side-car) device buffers for an element-partitioned stream."""
    torch = _torch()
    L = lib()
    out = torch.empty(L.tszp_elem_bound(n, bs, T, elem_bytes, int(ra)), dtype=torch.uint8, device=device)
    blk = torch.empty(L.tszp_elem_units(n, bs, T, int(ra)) + 1, dtype=torch.int32, device=device)
    return out, blk


def threadblock_compress(field, eb, bs, T=1, out=None, blk=None):
    """Threadblock stream of a float32 or float64 CUDA tensor: (bytes, stream, side-car)."""
    n, ds = field.numel(), field.element_size()
    if out is None or blk is None:
        out, blk = elem_buffers(n, bs, T, ds, False, field.device)
    nbytes = C.c_uint64()
    _check(lib().tszp_threadblock_compress(_dp(field), ds, n, float(eb), bs, T, _dp(out), out.numel(),
                                           _dp(blk), C.byref(nbytes), _stream()))
    return nbytes.value, out, blk


def threadblock_decompress(stream, blk, n, eb, bs, T=1, dtype=None, out=None):
    torch = _torch()
    if out is None:
        out = torch.empty(n, dtype=dtype or torch.float32, device=stream.device)
    _check(lib().tszp_threadblock_decompress(_dp(stream), _dp(blk), T, n, float(eb), bs, out.element_size(),
                                             _dp(out), _stream()))
    return out


def plain_compress_f64(field, eb, bs, T=1, out=None, blk=None):
    """Double random-access stream of a float64 CUDA tensor: (bytes, stream, side-car)."""
    n = field.numel()
    if out is None or blk is None:
        out, blk = elem_buffers(n, bs, T, 8, True, field.device)
    nbytes = C.c_uint64()
    _check(lib().tszp_plain_compress_f64(_dp(field), n, float(eb), bs, T, _dp(out), out.numel(), _dp(blk),
                                         C.byref(nbytes), _stream()))
    return nbytes.value, out, blk


def plain_decompress_f64(stream, blk, n, eb, bs, T=1, out=None):
    torch = _torch()
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=stream.device)
    _check(lib().tszp_plain_decompress_f64(_dp(stream), _dp(blk), T, n, float(eb), bs, _dp(out), _stream()))
    return out


def topo_compress_f64_3d(field, eb, bs, cps, want_q=False):
    """Double 3D topology stream with the saddle fix.  field: float64 CUDA tensor
    (d1, d2, d3); cps: int32 CUDA tensor of CriticalPoint3D records (k x 6).
    Returns (bytes, stream, side-car, fixed q or None, saddles moved)."""
    torch = _torch()
    n = field.numel()
    L = lib()
    out = torch.empty(L.tszp_topo_bound(n, bs, 1), dtype=torch.uint8, device=field.device)
    blk = torch.empty(L.tszp_blocks(n, bs) + 1, dtype=torch.int32, device=field.device)
    q = torch.empty(field.shape, dtype=torch.int32, device=field.device) if want_q else None
    nbytes, moved = C.c_uint64(), C.c_uint64()
    count = cps.shape[0] if cps is not None else 0
    _check(L.tszp_topo_compress_f64_3d(_dp(field), _dims(field.shape), float(eb), bs,
                                       _dp(cps) if count else None, count, _dp(out), out.numel(), _dp(blk),
                                       C.byref(nbytes), _dp(q) if want_q else None, C.byref(moved), _stream()))
    return nbytes.value, out, blk, q, moved.value


def classify(field, types=None):
    torch = _torch()
    if types is None:
        types = torch.empty(field.numel(), dtype=torch.uint8, device=field.device)
    cnt = C.c_uint64()
    _check(lib().tszp_classify(_dp(field), field.dim(), _dims(field.shape), _dp(types),
                               C.byref(cnt), _stream()))
    return types, cnt.value


def gen_blobs(shape, seed, nblobs, smin, smax, noise, waves=False, device="cuda"):
    torch = _torch()
    f = torch.empty(tuple(shape), dtype=torch.float32, device=device)
    _check(lib().tszp_gen_blobs(_dp(f), len(shape), _dims(shape), seed, nblobs, smin, smax, noise,
                                int(waves), _stream()))
    return f


def gen_fbm(shape, seed, octaves=6, noise=0.05, device="cuda"):
    torch = _torch()
    f = torch.empty(tuple(shape), dtype=torch.float32, device=device)
    _check(lib().tszp_gen_fbm(_dp(f), _dims(shape), seed, octaves, noise, _stream()))
    return f


def topo_restore(field, types, shape, eb, sp=None, n_ext=0, max_rounds=16, out=None):
    """Restores the critical points of a decoded field (float32 CUDA tensor holding
    (float)q * eb; types: uint8 CUDA tensor).  The result goes to `out` (a float32
    tensor like `field`), or in place into `field` when out is None."""
    info = RestoreInfo()
    dst = field if out is None else out
    _check(lib().tszp_topo_restore(_dp(field), _dp(dst), _dp(types), _dp(sp) if sp is not None else None, n_ext,
                                   len(shape), _dims(shape), eb, max_rounds, C.byref(info), _stream()))
    return info


def topo_verify(field, types, shape):
    out = (C.c_uint64 * 3)()
    _check(lib().tszp_topo_verify(_dp(field), _dp(types), len(shape), _dims(shape), out, _stream()))
    return tuple(out)


# ------------------------------------------------------------------------------------------
# .SZp container (example/testfloat_compress_fastmode1_topology.cc:105-139)
# ------------------------------------------------------------------------------------------
def _addr(a):
    """Address of a numpy array (host) or a torch tensor (device or host)."""
    return _p(a) if isinstance(a, np.ndarray) else _dp(a)


def szp_pack(topo, topo_bytes, sp, sp_bytes, out=None):
    """[topology stream][sort-position stream][u64 topology bytes][u64 sort-position bytes].
    topo / sp: numpy arrays or torch tensors (host or device); out defaults to a numpy
    array when topo is one, else a uint8 tensor on topo's device."""
    total = topo_bytes + sp_bytes + 16
    if out is None:
        if isinstance(topo, np.ndarray):
            out = np.zeros(total, np.uint8)
        else:
            out = _torch().empty(total, dtype=_torch().uint8, device=topo.device)
    nb = C.c_size_t()
    _check(lib().tszp_szp_pack(_addr(topo), topo_bytes, _addr(sp) if sp_bytes else None, sp_bytes, _addr(out),
                               out.nbytes if isinstance(out, np.ndarray) else out.numel(), C.byref(nb), _stream()))
    return out, nb.value


def szp_unpack(buf, length=None):
    """(topology bytes, sort-position offset, sort-position bytes, extrema count) of a
    container, read as the reference's decompress driver reads the footer."""
    n = length if length is not None else (buf.nbytes if isinstance(buf, np.ndarray) else buf.numel())
    tb, so, sb, e = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_uint64()
    _check(lib().tszp_szp_unpack(_addr(buf), n, C.byref(tb), C.byref(so), C.byref(sb), C.byref(e)))
    return tb.value, so.value, sb.value, e.value


# ------------------------------------------------------------------------------------------
# streams without a side-car, parsed on the device
# ------------------------------------------------------------------------------------------
PARSE_KINDS = {"topo": 0, "sortpos": 1, "plain": 2}


def parse_blocks(stream, n, bs, kind, length=None):
    """Block offsets (int32 tensor, tszp_blocks(n, bs) + 1) and chunk count of a device
    stream written without a side-car (e.g. read from a file)."""
    torch = _torch()
    blk = torch.empty(lib().tszp_blocks(n, bs) + 1, dtype=torch.int32, device=stream.device)
    T = C.c_uint32()
    _check(lib().tszp_parse_blocks(_dp(stream), length if length is not None else stream.numel(), n, bs,
                                   PARSE_KINDS[kind], _dp(blk), C.byref(T), _stream()))
    return blk, T.value


def elem_parse(stream, n, bs, elem_bytes, ra=False, length=None):
    """Unit offsets (int32 tensor, tszp_elem_units(n, bs, T, ra) + 1 of them) and chunk
    count of a device element-partitioned stream (eb header included) written without
    its side-car."""
    torch = _torch()
    L = lib()
    cap = L.tszp_elem_units(n, bs, 1, int(ra)) + 130  # enough for any chunk count up to 128
    blk = torch.empty(cap, dtype=torch.int32, device=stream.device)
    T = C.c_uint32()
    _check(L.tszp_elem_parse(_dp(stream), length if length is not None else stream.numel(), n, bs, elem_bytes,
                             int(ra), _dp(blk), C.byref(T), _stream()))
    return blk[:L.tszp_elem_units(n, bs, T.value, int(ra)) + 1], T.value


def quality_metrics(x, y, shape, eb):
    """Topology quality of y against x (device tensors of `shape`), as a dict."""
    q = Quality()
    _check(lib().tszp_quality_metrics(_dp(x), _dp(y), len(shape), _dims(shape), eb, C.byref(q), _stream()))
    return q.as_dict()
