"""numpy front end of the CPU oracle (oracle.c) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker.  The product library never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libszp_ref.so")

CP2 = np.dtype([("x", "<i4"), ("y", "<i4"), ("type", "<i4"), ("quantized_bin", "<i4"),
                ("sort_position", "<i4")])
CP3 = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("type", "<i4"),
                ("quantized_bin", "<i4"), ("sort_position", "<i4")])

_lib = None


def build():
    """Compile oracle.c (and the reference CPU sources when /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
        L = C.CDLL(LIB_PATH)
        vp, sz, i32, f32, u32 = C.c_void_p, C.c_size_t, C.c_int, C.c_float, C.c_uint
        L.or_bitcount.argtypes = [u32]; L.or_bitcount.restype = u32
        L.or_check_bitcount_table.argtypes = [vp, C.c_uint64, C.c_uint64]
        L.or_check_bitcount_table.restype = C.c_uint64
        L.or_cvt_i32.argtypes = [C.c_double]; L.or_cvt_i32.restype = i32
        L.or_rel_to_abs.argtypes = [vp, sz, f32, vp, vp]; L.or_rel_to_abs.restype = i32
        L.or_find_cps_2d.argtypes = [vp, i32, i32, f32, vp]; L.or_find_cps_2d.restype = sz
        L.or_find_cps_3d.argtypes = [vp, i32, i32, i32, f32, vp]; L.or_find_cps_3d.restype = sz
        L.or_sort_positions_2d.argtypes = [vp, sz, vp, i32]; L.or_sort_positions_2d.restype = None
        L.or_sort_positions_3d.argtypes = [vp, sz, vp, i32, i32]
        L.or_sort_positions_3d.restype = None
        for f in ("or_topo_bound", "or_plain_bound", "or_sortpos_bound"):
            getattr(L, f).argtypes = [sz, i32, u32]; getattr(L, f).restype = sz
        L.or_topo_compress.argtypes = [vp, sz, f32, i32, vp, u32, i32, vp]
        L.or_topo_compress.restype = sz
        L.or_plain_compress.argtypes = [vp, sz, f32, i32, u32, i32, vp]
        L.or_plain_compress.restype = sz
        L.or_compress_sortpos.argtypes = [vp, sz, i32, u32, i32, vp]
        L.or_compress_sortpos.restype = sz
        L.or_extrema_sortpos_2d.argtypes = [vp, sz, vp]; L.or_extrema_sortpos_2d.restype = sz
        L.or_extrema_sortpos_3d.argtypes = [vp, sz, vp]; L.or_extrema_sortpos_3d.restype = sz
        L.or_mark_types_2d.argtypes = [vp, sz, i32, i32, vp]; L.or_mark_types_2d.restype = None
        L.or_mark_types_3d.argtypes = [vp, sz, i32, i32, i32, vp]
        L.or_mark_types_3d.restype = None
        L.or_detect_T_topo.argtypes = [vp, sz, i32]; L.or_detect_T_topo.restype = u32
        L.or_detect_T_ra.argtypes = [vp, sz, i32, sz]; L.or_detect_T_ra.restype = u32
        L.or_topo_decompress.argtypes = [vp, sz, f32, i32, vp, vp]
        L.or_topo_decompress.restype = None
        L.or_sortpos_decompress.argtypes = [vp, sz, i32, vp]; L.or_sortpos_decompress.restype = None
        L.or_plain_decompress.argtypes = [vp, sz, f32, i32, vp]; L.or_plain_decompress.restype = None
        L.or_plain_decompress_T.argtypes = [vp, sz, f32, i32, C.c_uint, vp]; L.or_plain_decompress_T.restype = None
        L.or_block_offsets.argtypes = [vp, sz, i32, i32, i32, vp]; L.or_block_offsets.restype = i32
        L.or_classify_2d.argtypes = [vp, i32, i32, vp]; L.or_classify_2d.restype = None
        L.or_classify_3d.argtypes = [vp, i32, i32, i32, vp]; L.or_classify_3d.restype = None
        L.or_chunk_range.argtypes = [sz, u32, i32, u32, vp, vp]; L.or_chunk_range.restype = None
        f64 = C.c_double
        L.or_elem_bound.argtypes = [sz, i32, u32, i32, i32]; L.or_elem_bound.restype = sz
        L.or_elem_compress.argtypes = [vp, i32, sz, f64, i32, u32, i32, vp]
        L.or_elem_compress.restype = sz
        L.or_elem_decompress.argtypes = [vp, sz, f64, i32, u32, i32, i32, vp]
        L.or_elem_decompress.restype = i32
        L.or_fix_saddles_3d_pass.argtypes = [vp, vp, vp, i32, i32, i32]
        L.or_fix_saddles_3d_pass.restype = sz
        L.or_topo3d_f64_compress.argtypes = [vp, i32, i32, i32, f64, i32, vp, vp, vp]
        L.or_topo3d_f64_compress.restype = sz
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def bitcount(m):
    return int(lib().or_bitcount(int(m)))


def rel_to_abs(data, rel):
    d = _f32(data).ravel()
    eb = C.c_float(); rng = C.c_float()
    if lib().or_rel_to_abs(_p(d), d.size, float(rel), C.byref(eb), C.byref(rng)) != 0:
        raise ValueError("invalid relative bound")
    return eb.value, rng.value


def find_cps_2d(data, eb):
    d = _f32(data)
    rows, cols = d.shape
    out = np.zeros(max(0, (rows - 2) * (cols - 2)), dtype=CP2)
    n = lib().or_find_cps_2d(_p(d), rows, cols, float(eb), _p(out))
    return out[:n].copy()


def find_cps_3d(data, eb):
    d = _f32(data)
    d1, d2, d3 = d.shape
    out = np.zeros(max(0, (d1 - 2) * (d2 - 2) * (d3 - 2)), dtype=CP3)
    n = lib().or_find_cps_3d(_p(d), d1, d2, d3, float(eb), _p(out))
    return out[:n].copy()


def sort_positions(cps, data):
    """Fills cps['sort_position'] in place (2D or 3D record arrays)."""
    d = _f32(data)
    if cps.size == 0:
        return cps
    if cps.dtype == CP2:
        lib().or_sort_positions_2d(_p(cps), cps.size, _p(d), d.shape[1])
    else:
        lib().or_sort_positions_3d(_p(cps), cps.size, _p(d), d.shape[1], d.shape[2])
    return cps


def extrema_sort_positions(cps):
    out = np.zeros(cps.size, dtype=np.int32)
    f = lib().or_extrema_sortpos_2d if cps.dtype == CP2 else lib().or_extrema_sortpos_3d
    n = f(_p(cps), cps.size, _p(out))
    return out[:n].copy()


def mark_types(cps, shape):
    types = np.zeros(int(np.prod(shape)), dtype=np.uint8)
    if len(shape) == 2:
        lib().or_mark_types_2d(_p(cps), cps.size, shape[0], shape[1], _p(types))
    else:
        lib().or_mark_types_3d(_p(cps), cps.size, shape[0], shape[1], shape[2], _p(types))
    return types


def topo_compress(data, eb, bs, types, T=1, adaptive=True):
    d = _f32(data).ravel()
    t = np.ascontiguousarray(types, dtype=np.uint8).ravel()
    out = np.zeros(lib().or_topo_bound(d.size, bs, T), dtype=np.uint8)
    n = lib().or_topo_compress(_p(d), d.size, float(eb), bs, _p(t), T, int(adaptive), _p(out))
    return out[:n].copy()


def plain_compress(data, eb, bs, T=1, adaptive=True):
    d = _f32(data).ravel()
    out = np.zeros(lib().or_plain_bound(d.size, bs, T), dtype=np.uint8)
    n = lib().or_plain_compress(_p(d), d.size, float(eb), bs, T, int(adaptive), _p(out))
    return out[:n].copy()


def sortpos_compress(sp, bs, T=1, adaptive=True):
    s = np.ascontiguousarray(sp, dtype=np.int32)
    out = np.zeros(lib().or_sortpos_bound(s.size, bs, T), dtype=np.uint8)
    n = lib().or_compress_sortpos(_p(s), s.size, bs, T, int(adaptive), _p(out))
    return out[:n].copy()


def topo_decompress(cmp, n, eb, bs):
    c = np.ascontiguousarray(cmp, dtype=np.uint8)
    out = np.zeros(n, dtype=np.float32)
    ty = np.zeros(n, dtype=np.int32)
    lib().or_topo_decompress(_p(c), n, float(eb), bs, _p(out), _p(ty))
    return out, ty


def sortpos_decompress(cmp_after_count, e, bs):
    c = np.ascontiguousarray(cmp_after_count, dtype=np.uint8)
    out = np.zeros(e, dtype=np.int32)
    lib().or_sortpos_decompress(_p(c), e, bs, _p(out))
    return out


def plain_decompress(cmp_after_header, n, eb, bs, T=None):
    """T None: the chunk count from the reference's detection, as its decoder does."""
    c = np.ascontiguousarray(cmp_after_header, dtype=np.uint8)
    out = np.zeros(n, dtype=np.float32)
    if T is None:
        lib().or_plain_decompress(_p(c), n, float(eb), bs, _p(out))
    else:
        lib().or_plain_decompress_T(_p(c), n, float(eb), bs, int(T), _p(out))
    return out


def elem_compress(data, eb, bs, T=1, ra=False):
    """Threadblock (ra=False, float or double data) or double random-access (ra=True)
    stream: [eb big-endian][u64 off[T]][chunks]."""
    d = np.ascontiguousarray(data).ravel()
    if d.dtype not in (np.float32, np.float64):
        raise TypeError("float32 or float64 data")
    ds = d.dtype.itemsize
    out = np.zeros(lib().or_elem_bound(d.size, bs, T, ds, int(ra)), dtype=np.uint8)
    n = lib().or_elem_compress(_p(d), ds, d.size, float(eb), bs, T, int(ra), _p(out))
    return out[:n].copy()


def elem_decompress(cmp_after_eb, n, eb, bs, T, ra=False, dtype=np.float32):
    c = np.ascontiguousarray(cmp_after_eb, dtype=np.uint8)
    out = np.zeros(n, dtype=dtype)
    if lib().or_elem_decompress(_p(c), n, float(eb), bs, T, int(ra), out.itemsize, _p(out)) != 0:
        raise ValueError("bad arguments")
    return out


def fix_saddles_pass(q, types):
    """One Jacobi pass of the double 3D saddle fix; returns (new q, points changed)."""
    src = np.ascontiguousarray(q, dtype=np.int32)
    t = np.ascontiguousarray(types, dtype=np.uint8)
    dst = src.copy()
    k = lib().or_fix_saddles_3d_pass(_p(src), _p(dst), _p(t), *src.shape)
    return dst, int(k)


def topo3d_f64_compress(data, eb, bs, types):
    """Double 3D topology stream with the saddle fix; returns (stream, fixed q)."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    t = np.ascontiguousarray(types, dtype=np.uint8)
    out = np.zeros(lib().or_topo_bound(d.size, bs, 1), dtype=np.uint8)
    q = np.zeros(d.shape, dtype=np.int32)
    n = lib().or_topo3d_f64_compress(_p(d), *d.shape, float(eb), bs, _p(t), _p(out), _p(q))
    return out[:n].copy(), q


def block_offsets(payload, n, bs, topo, bc_always=True):
    c = np.ascontiguousarray(payload, dtype=np.uint8)
    nb = (n + bs - 1) // bs
    offs = np.zeros(nb + 1, dtype=np.uint64)
    if lib().or_block_offsets(_p(c), n, bs, int(topo), int(bc_always), _p(offs)) != 0:
        raise ValueError("corrupt stream")
    return offs


def classify(data):
    d = _f32(data)
    t = np.zeros(d.size, dtype=np.uint8)
    if d.ndim == 2:
        lib().or_classify_2d(_p(d), d.shape[0], d.shape[1], _p(t))
    else:
        lib().or_classify_3d(_p(d), d.shape[0], d.shape[1], d.shape[2], _p(t))
    return t.reshape(d.shape)


def detect_T_topo(cmp, n, bs):
    c = np.ascontiguousarray(cmp, dtype=np.uint8)
    return int(lib().or_detect_T_topo(_p(c), n, bs))


def check_bitcount_table(thr, lo, hi):
    t = np.ascontiguousarray(thr, dtype=np.uint32)
    return int(lib().or_check_bitcount_table(_p(t), lo, hi))


def quality_metrics(x, y, eb):
    """Topology quality of y against x, restating the CUDA driver's report
    (example/test_cuda_topology.cu:540-868: steps 5, 6, 8, 9 and 10), written for 2D and
    3D: critical points by the codec rule (classify above), preserved / lost / changed /
    new; pointwise error (double) against 1.01 eb and 2 eb; the saddle persistence proxy
    min |x_c - x_nb| over the stencil (:687-703) with thresholds 0, .5, 1, 2, 5, 10 eb
    (:716-735); local order of each interior point with its forward neighbour along
    every axis when that neighbour is interior too (:778-826)."""
    x = _f32(x)
    y = _f32(y)
    tx = classify(x).astype(np.int64)
    ty = classify(y).astype(np.int64)
    q = {}
    q["orig_cp"] = [int((tx != 0).sum())] + [int((tx == k).sum()) for k in (1, 2, 3)]
    q["dec_cp"] = [int((ty != 0).sum())] + [int((ty == k).sum()) for k in (1, 2, 3)]
    cp = tx != 0
    q["preserved"] = int((cp & (ty == tx)).sum())
    q["new_cps"] = int(((tx == 0) & (ty != 0)).sum())
    lost = cp & (ty == 0)
    chg = cp & (ty != 0) & (ty != tx)
    q["lost"] = [int(lost.sum())] + [int((lost & (tx == k)).sum()) for k in (1, 2, 3)]
    q["changed"] = [int(chg.sum())] + [int((chg & (tx == k)).sum()) for k in (1, 2, 3)]
    err = np.abs(x.astype(np.float64) - y.astype(np.float64))
    q["over_1eb"] = int((err > float(eb) * 1.01).sum())
    q["over_2eb"] = int((err > float(eb) * 2.0).sum())
    q["max_err"] = float(err.max())
    # persistence proxy of the saddles (all interior by the rule)
    sad = np.argwhere(tx == 3)
    xd = x.astype(np.float64)
    pers = np.full(len(sad), np.inf)
    for a in range(x.ndim):
        for s in (-1, 1):
            nb = sad.copy()
            nb[:, a] += s
            pers = np.minimum(pers, np.abs(xd[tuple(sad.T)] - xd[tuple(nb.T)]))
    kept = ty[tuple(sad.T)] == 3 if len(sad) else np.zeros(0, bool)
    q["saddles"] = int(len(sad))
    thr = [0.0, 0.5, 1.0, 2.0, 5.0, 10.0]
    q["saddles_above"] = [int((pers > t * float(eb)).sum()) for t in thr]
    q["saddles_kept_above"] = [int(((pers > t * float(eb)) & kept).sum()) for t in thr]
    q["pers_min"] = float(pers.min()) if len(sad) else 1e300
    q["pers_max"] = float(pers.max()) if len(sad) else 0.0
    q["pers_sum"] = float(pers.sum()) if len(sad) else 0.0
    # local order: interior a with forward neighbour b along each axis, b interior
    pairs = [0, 0, 0, 0]
    viol = [0, 0, 0, 0]
    inner = tuple(slice(1, d - 1) for d in x.shape)
    for a in range(x.ndim):
        sa = list(inner)
        sb = list(inner)
        sa[a] = slice(1, x.shape[a] - 2)
        sb[a] = slice(2, x.shape[a] - 1)
        xa, xb, ya, yb = x[tuple(sa)], x[tuple(sb)], y[tuple(sa)], y[tuple(sb)]
        ta, tb = tx[tuple(sa)], tx[tuple(sb)]
        strict = xa != xb
        oo = np.where(xa < xb, -1, 1)
        od = np.where(ya < yb, -1, np.where(ya > yb, 1, 0))
        v = strict & (oo != od)
        groups = [np.ones_like(strict), (ta != 0) | (tb != 0), (ta != 0) & (tb != 0),
                  (ta == 1) | (ta == 2) | (tb == 1) | (tb == 2)]
        for k, gmask in enumerate(groups):
            pairs[k] += int((strict & gmask).sum())
            viol[k] += int((v & gmask).sum())
    q["pairs"] = pairs
    q["violations"] = viol
    return q
