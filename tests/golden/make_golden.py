"""Writes the golden fixtures under tests/golden/ from the reference's OWN CPU code.

The reference ships no golden vectors, so this script runs its OpenMP CPU sources
(compiled from /root/reference by oracle/Makefile into oracle/_ref/libszp_ref.so) on
small deterministic inputs and stores inputs + outputs as .npz files.  It runs in the
build container only; the GPU tests just read the committed .npz files.

    python tests/golden/make_golden.py [name-prefix ...]

Reference entry points used (SZp/include/szp_float.h:26-60, szpd_float.h:21-33,
szp_topology_3d.h:36-61): szp_find_critical_points[_3d],
szp_sort_critical_points_by_original_data / szp_sort_critical_points_3d,
szp_compress_sort_positions[_3d], szp_float_openmp_threadblock_randomaccess_topology_preserved,
szp_float_compress_topology_3d, szp_float_decompress_openmp_threadblock_randomaccess_topology_preserved,
szp_decompress_sort_positions, szp_float_openmp_threadblock_randomaccess,
szp_float_decompress_openmp_threadblock_randomaccess; and (SZp/include/szp_double.h:30-47,
szpd_double.h:21-29) szp_float_openmp_threadblock, szp_float_decompress_openmp_threadblock,
szp_double_openmp_threadblock[_randomaccess], szp_double_decompress_openmp_threadblock[_randomaccess].
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import fields  # noqa: E402
import oracle  # noqa: E402

vp, sz, i32, f32 = C.c_void_p, C.c_size_t, C.c_int, C.c_float


def load_ref():
    oracle.build()
    if not os.path.exists(oracle.REF_PATH):
        raise SystemExit("reference library not built (needs /root/reference)")
    R = C.CDLL(oracle.REF_PATH)
    R.szp_find_critical_points.argtypes = [vp, C.POINTER(sz), i32, i32, f32]
    R.szp_find_critical_points.restype = vp
    R.szp_find_critical_points_3d.argtypes = [vp, C.POINTER(sz), i32, i32, i32, f32]
    R.szp_find_critical_points_3d.restype = vp
    R.szp_sort_critical_points_by_original_data.argtypes = [vp, sz, vp, i32]
    R.szp_sort_critical_points_3d.argtypes = [vp, sz, vp, i32, i32]
    R.szp_compress_sort_positions.argtypes = [vp, sz, C.POINTER(sz), i32]
    R.szp_compress_sort_positions.restype = vp
    R.szp_compress_sort_positions_3d.argtypes = [vp, sz, C.POINTER(sz), i32]
    R.szp_compress_sort_positions_3d.restype = vp
    R.szp_float_openmp_threadblock_randomaccess_topology_preserved.argtypes = [
        vp, C.POINTER(sz), f32, sz, i32, vp, i32, i32, i32]
    R.szp_float_openmp_threadblock_randomaccess_topology_preserved.restype = vp
    R.szp_float_compress_topology_3d.argtypes = [vp, C.POINTER(sz), f32, sz, i32, vp, i32,
                                                 i32, i32, i32]
    R.szp_float_compress_topology_3d.restype = vp
    R.szp_float_decompress_openmp_threadblock_randomaccess_topology_preserved.argtypes = [
        C.POINTER(vp), sz, f32, i32, vp, C.POINTER(vp)]
    R.szp_decompress_sort_positions.argtypes = [vp, sz, i32]
    R.szp_decompress_sort_positions.restype = vp
    R.szp_float_openmp_threadblock_randomaccess.argtypes = [vp, C.POINTER(sz), f32, sz, i32]
    R.szp_float_openmp_threadblock_randomaccess.restype = vp
    R.szp_float_decompress_openmp_threadblock_randomaccess.argtypes = [sz, f32, i32, vp]
    R.szp_float_decompress_openmp_threadblock_randomaccess.restype = vp
    f64 = C.c_double
    R.szp_float_openmp_threadblock.argtypes = [vp, C.POINTER(sz), f32, sz, i32]
    R.szp_float_openmp_threadblock.restype = vp
    R.szp_float_decompress_openmp_threadblock.argtypes = [sz, f32, i32, vp]
    R.szp_float_decompress_openmp_threadblock.restype = vp
    R.szp_double_openmp_threadblock.argtypes = [vp, C.POINTER(sz), f64, sz, i32]
    R.szp_double_openmp_threadblock.restype = vp
    R.szp_double_decompress_openmp_threadblock.argtypes = [sz, f64, i32, vp]
    R.szp_double_decompress_openmp_threadblock.restype = vp
    R.szp_double_openmp_threadblock_randomaccess.argtypes = [vp, C.POINTER(sz), f64, sz, i32]
    R.szp_double_openmp_threadblock_randomaccess.restype = vp
    R.szp_double_decompress_openmp_threadblock_randomaccess.argtypes = [sz, f64, i32, vp]
    R.szp_double_decompress_openmp_threadblock_randomaccess.restype = vp
    return R


libc = C.CDLL("libc.so.6")
libc.free.argtypes = [vp]
gomp = C.CDLL("libgomp.so.1")


def threads(t):
    gomp.omp_set_num_threads(int(t))


def take(ptr, nbytes, dtype=np.uint8):
    if not ptr or nbytes == 0:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_ubyte * nbytes).from_address(ptr)
    out = np.frombuffer(bytes(buf), dtype=dtype).copy()
    libc.free(ptr)
    return out


def P(a):
    return a.ctypes.data_as(vp)


def decode(R, out, topo, sp_stream, cps, n, eb, bs):
    """Reference decoders on the single-chunk streams.  They infer the chunk count
    from the offset table (SZp/include/szp_detect_threads.h), which can mistake a
    multi-chunk stream for a single chunk, so multi-chunk streams are kept as
    compressor outputs only.  Each decode is checked before it is stored."""
    dptr, fptr = vp(), vp()
    R.szp_float_decompress_openmp_threadblock_randomaccess_topology_preserved(
        C.byref(dptr), n, f32(eb), bs, P(topo), C.byref(fptr))
    dec = take(dptr.value, 4 * n, np.float32)
    fn = take(fptr.value, 4 * n, np.int32)
    x = out["data"].ravel().astype(np.float64)
    # |x - (float)q * eb| < eb up to the rounding of (float)q for |q| > 2^24
    assert np.all(np.abs(dec - x) <= 1.0001 * eb + np.abs(x) * 2.0 ** -22)
    out["topo_dec_T1"], out["topo_fn_T1"] = dec, fn
    ext_sp = cps["sort_position"][(cps["type"] == 1) | (cps["type"] == 2)]
    if ext_sp.size:
        sp_in = np.ascontiguousarray(sp_stream[8:])
        p = R.szp_decompress_sort_positions(P(sp_in), ext_sp.size, bs)
        got = take(p, 4 * ext_sp.size, np.int32)
        if np.array_equal(got, ext_sp):
            out["sortpos_dec_T1"] = got
        else:
            print("  reference sort-position decoder misread this stream (chunk-count "
                  "detection); not stored")


def case_2d(R, name, data, eb, bs, Ts=(1, 4)):
    data = np.ascontiguousarray(data, dtype=np.float32)
    rows, cols = data.shape
    n = data.size
    out = {"data": data, "eb": np.float32(eb), "bs": np.int32(bs), "dims": np.array(data.shape)}
    threads(1)  # canonical row-major CP order (the 2D finder races on `tid` when T > 1)
    cnt = sz()
    p = R.szp_find_critical_points(P(data), C.byref(cnt), rows, cols, f32(eb))
    cps = take(p, cnt.value * oracle.CP2.itemsize, oracle.CP2)
    cps["sort_position"] = 0  # the 2D finder leaves it uninitialised
    out["cps_found"] = cps.copy()
    for T in Ts:
        threads(T)
        cps_s = cps.copy()
        R.szp_sort_critical_points_by_original_data(P(cps_s), cps_s.size, P(data), cols)
        if T == 1:
            out["cps_sorted"] = cps_s.copy()
        osz = sz()
        p = R.szp_compress_sort_positions(P(cps_s), cps_s.size, C.byref(osz), bs)
        sp_stream = take(p, osz.value)
        out[f"sortpos_T{T}"] = sp_stream
        osz = sz()
        p = R.szp_float_openmp_threadblock_randomaccess_topology_preserved(
            P(data), C.byref(osz), f32(eb), n, bs, P(cps_s), cps_s.size, rows, cols)
        topo = take(p, osz.value)
        out[f"topo_T{T}"] = topo
        if T == 1:
            decode(R, out, topo, sp_stream, cps_s, n, eb, bs)
    threads(1)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: n={n} cps={cps.size} topo={out['topo_T1'].size}B sp={out['sortpos_T1'].size}B")


def case_3d(R, name, data, eb, bs, Ts=(1, 4)):
    data = np.ascontiguousarray(data, dtype=np.float32)
    d1, d2, d3 = data.shape
    n = data.size
    out = {"data": data, "eb": np.float32(eb), "bs": np.int32(bs), "dims": np.array(data.shape)}
    threads(1)
    cnt = sz()
    p = R.szp_find_critical_points_3d(P(data), C.byref(cnt), d1, d2, d3, f32(eb))
    cps = take(p, cnt.value * oracle.CP3.itemsize, oracle.CP3)
    out["cps_found"] = cps.copy()
    for T in Ts:
        threads(T)
        cps_s = cps.copy()
        R.szp_sort_critical_points_3d(P(cps_s), cps_s.size, P(data), d2, d3)
        if T == 1:
            out["cps_sorted"] = cps_s.copy()
        osz = sz()
        p = R.szp_compress_sort_positions_3d(P(cps_s), cps_s.size, C.byref(osz), bs)
        sp_stream = take(p, osz.value)
        out[f"sortpos_T{T}"] = sp_stream
        osz = sz()
        p = R.szp_float_compress_topology_3d(P(data), C.byref(osz), f32(eb), n, bs, P(cps_s),
                                             cps_s.size, d1, d2, d3)
        topo = take(p, osz.value)
        out[f"topo_T{T}"] = topo
        if T == 1:
            decode(R, out, topo, sp_stream, cps_s, n, eb, bs)
    threads(1)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: n={n} cps={cps.size} topo={out['topo_T1'].size}B sp={out['sortpos_T1'].size}B")


def case_plain(R, name, data, eb, bs, Ts=(1, 3)):
    data = np.ascontiguousarray(data, dtype=np.float32).ravel()
    n = data.size
    out = {"data": data, "eb": np.float32(eb), "bs": np.int32(bs)}
    for T in Ts:
        threads(T)
        osz = sz()
        p = R.szp_float_openmp_threadblock_randomaccess(P(data), C.byref(osz), f32(eb), n, bs)
        s = take(p, osz.value)
        out[f"plain_T{T}"] = s
        body = np.ascontiguousarray(s[4:])
        p = R.szp_float_decompress_openmp_threadblock_randomaccess(n, f32(eb), bs, P(body))
        out[f"plain_dec_T{T}"] = take(p, 4 * n, np.float32)
    threads(1)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: n={n} plain={out['plain_T1'].size}B")


def case_elem(R, name, data, eb, bs, Ts=(1, 3), ra=False):
    """Threadblock writers (szp_float_openmp_threadblock, szp_double_openmp_threadblock)
    or, with ra, the double random-access writer; the decoders on T = 1 only (they
    infer the chunk count, and the double random-access decoder partitions blocks
    differently from its writer when T > 1), each checked before it is stored."""
    dt = data.dtype
    data = np.ascontiguousarray(data).ravel()
    n = data.size
    ds = dt.itemsize
    out = {"data": data, "eb": np.array(eb, dtype=dt), "bs": np.int32(bs), "ra": np.int32(ra)}
    if dt == np.float32:
        comp, dec, ebc = R.szp_float_openmp_threadblock, R.szp_float_decompress_openmp_threadblock, f32
    elif ra:
        comp = R.szp_double_openmp_threadblock_randomaccess
        dec = R.szp_double_decompress_openmp_threadblock_randomaccess
        ebc = C.c_double
    else:
        comp, dec, ebc = R.szp_double_openmp_threadblock, R.szp_double_decompress_openmp_threadblock, C.c_double
    for T in Ts:
        threads(T)
        osz = sz()
        p = comp(P(data), C.byref(osz), ebc(eb), n, bs)
        s = take(p, osz.value)
        out[f"elem_T{T}"] = s
        if T == 1:
            body = np.ascontiguousarray(s[ds:])
            got = take(dec(n, ebc(eb), bs, P(body)), ds * n, dt)
            if np.all(np.abs(got.astype(np.float64) - data) <= 1.0001 * eb + np.abs(data) * 2.0 ** -22):
                out["elem_dec_T1"] = got
            else:
                print("  reference decoder misread this stream; not stored")
    threads(1)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: n={n} {dt} ra={int(ra)} stream={out['elem_T1'].size}B")


CASES = [
    ("g2d_blobs_bs64", lambda R: case_2d(R, "g2d_blobs_bs64", fields.blobs2d(48, 72, seed=1), 1e-3, 64)),
    ("g2d_blobs_bs8", lambda R: case_2d(R, "g2d_blobs_bs8", fields.blobs2d(37, 53, seed=2), 2e-3, 8)),
    # N % bs == 1: a one-element trailing block (5 * 13 = 65 = 64 + 1)
    # (the reference decoder faults on multi-chunk streams with fewer blocks than chunks)
    ("g2d_ragged_bs64", lambda R: case_2d(R, "g2d_ragged_bs64", fields.blobs2d(5, 13, seed=3, noise=0.05), 1e-3, 64, Ts=(1, 2))),
    ("g2d_ties_bs16", lambda R: case_2d(R, "g2d_ties_bs16", fields.ties2d(31, 29, seed=4), 1e-3, 16)),
    ("g2d_huge_bs32", lambda R: case_2d(R, "g2d_huge_bs32", fields.huge2d(19, 23, seed=5), 1e-3, 32)),
    ("g2d_bs128", lambda R: case_2d(R, "g2d_bs128", fields.blobs2d(40, 70, seed=6, noise=0.02), 5e-4, 128)),
    ("g2d_bs256", lambda R: case_2d(R, "g2d_bs256", fields.blobs2d(33, 90, seed=7, noise=0.02), 5e-4, 256, Ts=(1, 2))),
    ("g3d_blobs_bs16", lambda R: case_3d(R, "g3d_blobs_bs16", fields.blobs3d(9, 11, 13, seed=8, noise=0.01), 1e-3, 16)),
    ("g3d_fbm_bs64", lambda R: case_3d(R, "g3d_fbm_bs64", fields.fbm3d(12, 14, 17, seed=9), 2e-3, 64)),
    ("g3d_ragged_bs64", lambda R: case_3d(R, "g3d_ragged_bs64", fields.blobs3d(4, 5, 13, seed=10, noise=0.05), 1e-3, 64, Ts=(1, 2))),
    ("gplain_bs64", lambda R: case_plain(R, "gplain_bs64", fields.blobs2d(30, 41, seed=11).ravel(), 1e-3, 64)),
    # (the reference writer overruns its per-thread buffer on incompressible data when
    # chunks are uneven, so this incompressible case uses two equal chunks)
    ("gplain_bs128", lambda R: case_plain(R, "gplain_bs128", fields.huge2d(16, 32, seed=12).ravel(), 0.7, 128, Ts=(1, 2))),
    ("gplain_bs32", lambda R: case_plain(R, "gplain_bs32", fields.ties2d(20, 20, seed=13).ravel(), 3e-3, 32)),
    # a one-chunk stream that the reference's chunk-count detection misreads
    ("gplain_misdetect", lambda R: case_plain(R, "gplain_misdetect", fields.blobs3d(12, 35, 5, seed=1000).ravel(),
                                              0.007894032212920654, 2, Ts=(1,))),
    # threadblock (chained) streams: float and double, ragged chunks and blocks
    ("gtb_f32_bs64", lambda R: case_elem(R, "gtb_f32_bs64", fields.blobs2d(30, 41, seed=14).ravel(), 1e-3, 64, Ts=(1, 3, 8))),
    ("gtb_f32_bs7", lambda R: case_elem(R, "gtb_f32_bs7", fields.ties2d(21, 19, seed=15).ravel(), 3e-3, 7, Ts=(1, 4))),
    ("gtb_f64_bs32", lambda R: case_elem(R, "gtb_f64_bs32", fields.blobs2d(26, 33, seed=16).ravel().astype(np.float64) * 1.37,
              2e-4, 32, Ts=(1, 5))),
    ("gtb_f64_bs128", lambda R: case_elem(R, "gtb_f64_bs128", fields.fbm3d(9, 10, 11, seed=17).ravel().astype(np.float64),
              1e-3, 128, Ts=(1, 2))),
    # double random access: n % bs == 1 gives a one-element block
    ("gra_f64_bs64", lambda R: case_elem(R, "gra_f64_bs64", fields.blobs2d(5, 13, seed=18).ravel().astype(np.float64), 1e-3, 64,
              Ts=(1, 3), ra=True)),
    ("gra_f64_bs16", lambda R: case_elem(R, "gra_f64_bs16", fields.blobs2d(20, 24, seed=19).ravel().astype(np.float64) * 0.3,
              5e-4, 16, Ts=(1, 4), ra=True)),
]


def main(prefixes):
    """Writes every case, or only those whose name starts with one of `prefixes`."""
    R = load_ref()
    for name, make in CASES:
        if not prefixes or name.startswith(tuple(prefixes)):
            make(R)


if __name__ == "__main__":
    main(sys.argv[1:])
