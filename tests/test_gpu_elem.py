"""GPU: streams partitioned by elements, against the oracle and the reference's fixtures.

* threadblock codec, float and double data (SZp/src/szp_float.cc:1445-1590,
  SZp/src/szp_double.cc:148-318, decoders SZp/src/szpd_float.cc:28-135,
  szpd_double.cc:6-110);
* double random-access codec (SZp/src/szp_double.cc:759-920);
* double-precision 3D topology stream with the saddle fix
  (SZp/src/szp_cuda_topology_3d.cu:736-826, :1065-1188).

Streams are compared byte for byte with the oracle (whose writers are pinned against the
reference's own outputs in tests/golden), decoded values bit for bit."""
import glob
import os

import numpy as np
import pytest

import fields
import oracle
import toposzp_hip as tz

pytestmark = pytest.mark.gpu

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "g*.npz")))
ELEM = [g for g in GOLD if os.path.basename(g).startswith(("gtb", "gra"))]
G3D = [g for g in GOLD if os.path.basename(g).startswith("g3d")]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _compress(torch, data, eb, bs, T, ra):
    d = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    if ra:
        nbytes, out, blk = tz.plain_compress_f64(d, eb, bs, T)
    else:
        nbytes, out, blk = tz.threadblock_compress(d, eb, bs, T)
    torch.cuda.synchronize()
    return nbytes, out, blk


def _decompress(torch, out, blk, n, eb, bs, T, ra, dtype):
    if ra:
        dec = tz.plain_decompress_f64(out, blk, n, eb, bs, T)
    else:
        dec = tz.threadblock_decompress(out, blk, n, eb, bs, T,
                                        dtype=torch.float64 if dtype == np.float64 else torch.float32)
    torch.cuda.synchronize()
    return dec.cpu().numpy()


def _check(torch, data, eb, bs, T, ra, want=None):
    print(f"elem n={data.size} {data.dtype} bs={bs} T={T} ra={ra}", flush=True)
    ref = oracle.elem_compress(data, eb, bs, T=T, ra=ra) if want is None else want
    nbytes, out, blk = _compress(torch, data, eb, bs, T, ra)
    got = out[:nbytes].cpu().numpy()
    assert nbytes == ref.size, (nbytes, ref.size)
    assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} bytes differ"
    ds = data.dtype.itemsize
    dec = _decompress(torch, out, blk, data.size, eb, bs, T, ra, data.dtype)
    want_dec = oracle.elem_decompress(ref[ds:], data.size, eb, bs, T, ra=ra, dtype=data.dtype)
    assert np.array_equal(dec.view(np.uint8), want_dec.view(np.uint8))
    return dec


@pytest.mark.parametrize("path", ELEM, ids=os.path.basename)
def test_elem_matches_reference_fixtures(path):
    torch = _torch()
    g = _load(path)
    data, eb, bs, ra = g["data"], float(g["eb"]), int(g["bs"]), bool(g["ra"])
    for key in g:
        if key.startswith("elem_T"):
            dec = _check(torch, data, eb, bs, int(key[6:]), ra, want=g[key])
            if key == "elem_T1" and "elem_dec_T1" in g:
                assert np.array_equal(dec.view(np.uint8), g["elem_dec_T1"].view(np.uint8))


SHAPES = [(1, 8, 1), (2, 1, 1), (5, 64, 8), (63, 16, 2), (65, 64, 1), (129, 64, 2), (1000, 7, 3),
          (4097, 1025, 5), (20011, 32, 64), (100003, 128, 7), (50000, 2, 4096)]


@pytest.mark.parametrize("n,bs,T", SHAPES, ids=[f"n{a}_bs{b}_T{c}" for a, b, c in SHAPES])
@pytest.mark.parametrize("dt,ra", [(np.float32, False), (np.float64, False), (np.float64, True)],
                         ids=["tb_f32", "tb_f64", "ra_f64"])
def test_elem_matches_oracle(n, bs, T, dt, ra):
    torch = _torch()
    rng = np.random.default_rng(n * 31 + bs + T)
    x = np.cumsum(rng.normal(0, 1, n)) * 0.01 + 3.0 * np.sin(np.arange(n) * 0.003)
    _check(torch, x.astype(dt), 1e-3, bs, T, ra)


@pytest.mark.parametrize("dt,ra", [(np.float32, False), (np.float64, False), (np.float64, True)],
                         ids=["tb_f32", "tb_f64", "ra_f64"])
def test_elem_large_and_wide(dt, ra):
    """Many tiles chained across one chunk (the segmented look-back), and deltas of
    every width up to 32 bits (huge values saturate the quantizer to INT_MIN)."""
    torch = _torch()
    n = 3_000_017
    rng = np.random.default_rng(7)
    x = np.cumsum(rng.normal(0, 1, n)) * 0.05
    x[::9973] *= 1e4
    x[5::100003] = 1e12
    _check(torch, x.astype(dt), 1e-3, 64, 1, ra)
    _check(torch, x.astype(dt), 1e-3, 256, 3, ra)


def test_elem_rejects_bad_arguments():
    torch = _torch()
    d = torch.zeros(100, dtype=torch.float64, device="cuda")
    with pytest.raises(tz.TszpError):
        tz.threadblock_compress(d, 0.0, 64, 1)
    with pytest.raises(tz.TszpError):
        tz.threadblock_compress(d, 1e-3, 2000, 1)
    with pytest.raises(tz.TszpError):
        tz.plain_compress_f64(d, 1e-3, 64, 5000)


# ------------------------------------------------------------------------------------------
# double 3D topology with the saddle fix
# ------------------------------------------------------------------------------------------
def _cp_records(data, eb):
    cps = oracle.find_cps_3d(np.ascontiguousarray(data, dtype=np.float32), eb)
    rec = np.stack([cps["x"], cps["y"], cps["z"], cps["type"], cps["quantized_bin"],
                    cps["sort_position"]], axis=1).astype(np.int32)
    types = oracle.mark_types(cps, data.shape)
    return np.ascontiguousarray(rec), types


def _check_topo64(torch, d64, eb, bs, rec, types):
    print(f"topo64 shape={d64.shape} bs={bs} cps={rec.shape[0]}", flush=True)
    ref, q_ref = oracle.topo3d_f64_compress(d64, eb, bs, types.reshape(d64.shape))
    field = torch.from_numpy(np.ascontiguousarray(d64)).cuda()
    cps = torch.from_numpy(rec).cuda() if rec.shape[0] else None
    nbytes, out, blk, q, moved = tz.topo_compress_f64_3d(field, eb, bs, cps, want_q=True)
    torch.cuda.synchronize()
    assert np.array_equal(q.cpu().numpy(), q_ref)
    q0 = np.array([oracle.lib().or_cvt_i32(float(v) * (1.0 / eb)) for v in d64.ravel()],
                  dtype=np.int32).reshape(d64.shape)
    assert moved == int((q_ref != q0).sum())
    assert nbytes == ref.size
    got = out[:nbytes].cpu().numpy()
    assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} bytes differ"
    return moved


@pytest.mark.parametrize("path", G3D, ids=os.path.basename)
def test_topo64_fixtures(path):
    torch = _torch()
    g = _load(path)
    data, eb, bs = g["data"], float(g["eb"]), int(g["bs"])
    rec, types = _cp_records(data, eb)
    _check_topo64(torch, data.astype(np.float64), eb, bs, rec, types)


@pytest.mark.parametrize("shape,eb,bs", [((40, 48, 56), 2e-2, 64), ((33, 27, 45), 5e-2, 16),
                                         ((64, 64, 64), 1e-1, 128), ((9, 10, 11), 3e-2, 7)])
def test_topo64_saddles_moved(shape, eb, bs):
    """Error bounds large enough that many saddles break under quantization; the CP list
    comes from the float field, the double field carries extra low bits."""
    torch = _torch()
    data = fields.fbm3d(*shape, seed=sum(shape))
    rec, types = _cp_records(data, eb)
    d64 = data.astype(np.float64) * (1.0 + 2.0 ** -30)
    moved = _check_topo64(torch, d64, eb, bs, rec, types)
    assert moved > 0


def test_topo64_without_critical_points():
    torch = _torch()
    d64 = np.linspace(0, 1, 7 * 8 * 9).reshape(7, 8, 9)
    rec = np.zeros((0, 6), np.int32)
    _check_topo64(torch, d64, 1e-3, 64, rec, np.zeros(d64.shape, np.uint8))


# ------------------------------------------------------------------------------------------
# drop-in entry points (host buffers): szp_cuda_{float,double}_{compress,decompress}_
# threadblock, szp_cuda_double_{compress,decompress}_randomaccess,
# szp_cuda_double_compress_topology_3d, szp_cuda_float_*_predict_quantization
# ------------------------------------------------------------------------------------------
import ctypes as C  # noqa: E402

_libc = C.CDLL("libc.so.6")
_libc.free.argtypes = [C.c_void_p]


def _dropin():
    L = tz.lib()
    vp, sz, i32 = C.c_void_p, C.c_size_t, C.c_int
    for name, f in (("float", C.c_float), ("double", C.c_double)):
        comp = getattr(L, f"szp_cuda_{name}_compress_threadblock")
        comp.argtypes, comp.restype = [vp, C.POINTER(sz), f, sz, i32], vp
        dec = getattr(L, f"szp_cuda_{name}_decompress_threadblock")
        dec.argtypes, dec.restype = [sz, f, i32, vp], vp
    L.szp_cuda_double_compress_randomaccess.argtypes = [vp, C.POINTER(sz), C.c_double, sz, i32]
    L.szp_cuda_double_compress_randomaccess.restype = vp
    L.szp_cuda_double_decompress_randomaccess.argtypes = [sz, C.c_double, i32, vp]
    L.szp_cuda_double_decompress_randomaccess.restype = vp
    L.szp_cuda_double_compress_topology_3d.argtypes = [vp, C.POINTER(sz), C.c_double, sz, i32, vp, i32, i32, i32,
                                                       i32]
    L.szp_cuda_double_compress_topology_3d.restype = vp
    for name in ("direct", "threadblock"):
        f = getattr(L, f"szp_cuda_float_{name}_predict_quantization")
        f.argtypes, f.restype = [vp, C.POINTER(sz), C.c_float, sz, i32], vp
    return L


def _take(ptr, nbytes, dtype):
    assert ptr, tz.last_error()
    out = np.frombuffer(bytes((C.c_ubyte * nbytes).from_address(ptr)), dtype=dtype).copy()
    _libc.free(ptr)
    return out


def _dropin_compress(L, data, eb, bs, ra):
    name = "float" if data.dtype == np.float32 else "double"
    fn = L.szp_cuda_double_compress_randomaccess if ra else getattr(L, f"szp_cuda_{name}_compress_threadblock")
    osz = C.c_size_t()
    p = fn(data.ctypes.data_as(C.c_void_p), C.byref(osz), eb, data.size, bs)
    return _take(p, osz.value, np.uint8)


def _dropin_decompress(L, stream_after_eb, n, eb, bs, ra, dtype):
    name = "float" if dtype == np.float32 else "double"
    fn = L.szp_cuda_double_decompress_randomaccess if ra else getattr(L, f"szp_cuda_{name}_decompress_threadblock")
    buf = np.ascontiguousarray(stream_after_eb)
    p = fn(n, eb, bs, buf.ctypes.data_as(C.c_void_p))
    return _take(p, np.dtype(dtype).itemsize * n, dtype)


@pytest.mark.parametrize("path", ELEM, ids=os.path.basename)
def test_dropin_elem_fixtures(path):
    """The drop-in writes the one-chunk fixture byte for byte and reads every fixture
    (chunk count detected from the offset table)."""
    _torch()
    L = _dropin()
    g = _load(path)
    data, eb, bs, ra = g["data"], float(g["eb"]), int(g["bs"]), bool(g["ra"])
    ds = data.dtype.itemsize
    got = _dropin_compress(L, data, eb, bs, ra)
    assert np.array_equal(got, g["elem_T1"])
    for key in g:
        if key.startswith("elem_T"):
            T = int(key[6:])
            dec = _dropin_decompress(L, g[key][ds:], data.size, eb, bs, ra, data.dtype)
            want = oracle.elem_decompress(g[key][ds:], data.size, eb, bs, T, ra=ra, dtype=data.dtype)
            assert np.array_equal(dec.view(np.uint8), want.view(np.uint8)), key


# (with more chunks than elements, the leading chunks are empty and their zero offsets
# read as a valid stream of zeros under a smaller chunk count -- for the reference's
# detection as for this one -- so detection is tested with every chunk non-empty)
@pytest.mark.parametrize("n,bs,T", [(1000, 64, 1), (1000, 64, 3), (4097, 1025, 5), (20011, 32, 64), (77, 7, 16),
                                    (300, 2, 128)])
@pytest.mark.parametrize("dt,ra", [(np.float32, False), (np.float64, False), (np.float64, True)],
                         ids=["tb_f32", "tb_f64", "ra_f64"])
def test_dropin_elem_detects_chunks(n, bs, T, dt, ra):
    _torch()
    L = _dropin()
    rng = np.random.default_rng(n + bs + T)
    x = (np.cumsum(rng.normal(0, 1, n)) * 0.01).astype(dt)
    s = oracle.elem_compress(x, 1e-3, bs, T=T, ra=ra)
    ds = x.dtype.itemsize
    dec = _dropin_decompress(L, s[ds:], n, 1e-3, bs, ra, dt)
    want = oracle.elem_decompress(s[ds:], n, 1e-3, bs, T, ra=ra, dtype=dt)
    assert np.array_equal(dec.view(np.uint8), want.view(np.uint8))
    if T == 1:
        assert np.array_equal(_dropin_compress(L, x, 1e-3, bs, ra), s)


def test_dropin_topo64():
    _torch()
    L = _dropin()
    data = fields.fbm3d(20, 24, 28, seed=5)
    eb, bs = 3e-2, 64
    cps = oracle.find_cps_3d(data, eb)
    types = oracle.mark_types(cps, data.shape)
    d64 = np.ascontiguousarray(data.astype(np.float64) * (1.0 + 2.0 ** -30))
    ref, _ = oracle.topo3d_f64_compress(d64, eb, bs, types.reshape(data.shape))
    osz = C.c_size_t()
    p = L.szp_cuda_double_compress_topology_3d(d64.ctypes.data_as(C.c_void_p), C.byref(osz), eb, d64.size, bs,
                                               cps.ctypes.data_as(C.c_void_p), cps.size, *data.shape)
    got = _take(p, osz.value, np.uint8)
    assert np.array_equal(got, ref)


def test_dropin_predict_quantization():
    """q[i] - q[i - 1] with q the C cast of x * (1 / (double)eb): IEEE double multiply and
    truncation, which numpy computes identically for |x / eb| < 2^31."""
    _torch()
    L = _dropin()
    rng = np.random.default_rng(3)
    x = (rng.normal(0, 50, 100003)).astype(np.float32)
    eb = np.float32(1e-3)
    q = np.trunc(x.astype(np.float64) * (1.0 / np.float64(eb))).astype(np.int64)
    want = np.diff(q, prepend=0).astype(np.int32)
    for name in ("direct", "threadblock"):
        fn = getattr(L, f"szp_cuda_float_{name}_predict_quantization")
        p = fn(x.ctypes.data_as(C.c_void_p), None, float(eb), x.size, 64)
        assert np.array_equal(_take(p, 4 * x.size, np.int32), want)


@pytest.mark.parametrize("shape", [(3, 3, 3), (3, 4, 5), (5, 3, 4)])
def test_topo64_smallest_fields(shape):
    """Fields with a single interior point or plane: the saddle list, the passes and the
    codec at their smallest sizes."""
    torch = _torch()
    rng = np.random.default_rng(sum(shape))
    data = rng.normal(0, 1, shape).astype(np.float32)
    eb = 0.5
    rec, types = _cp_records(data, eb)
    _check_topo64(torch, data.astype(np.float64), eb, 7, rec, types)


def test_dropin_rejects_corrupt_threadblock_stream():
    """A bit count above 32 in the first block: no chunk count walks, NULL and an error."""
    _torch()
    L = _dropin()
    x = (np.arange(1000) * 0.01).astype(np.float32)
    s = oracle.elem_compress(x, 1e-3, 64, T=1)
    bad = s.copy()
    bad[4 + 8 + 4] = 200  # the first bc byte after the anchor
    buf = np.ascontiguousarray(bad[4:])
    p = L.szp_cuda_float_decompress_threadblock(x.size, 1e-3, 64, buf.ctypes.data_as(C.c_void_p))
    assert not p
    assert "corrupt" in tz.last_error()


# ---- streams without their side-car: unit offsets and chunk count parsed on the device ----
def _parse_and_decode(torch, s, data, eb, bs, ra):
    """Parse a host stream (eb header included) uploaded without its side-car, then
    decode it with the parsed offsets."""
    d = torch.from_numpy(np.ascontiguousarray(s)).cuda()
    blk, T = tz.elem_parse(d, data.size, bs, data.dtype.itemsize, ra)
    dec = _decompress(torch, d, blk, data.size, eb, bs, T, ra, data.dtype)
    return blk.cpu().numpy(), T, dec


@pytest.mark.parametrize("path", ELEM, ids=os.path.basename)
def test_elem_parse_reference_fixtures(path):
    """Every chunk count of the reference's own streams: the parse finds it and decodes
    the stream to the oracle's values."""
    torch = _torch()
    g = _load(path)
    data, eb, bs, ra = g["data"], float(g["eb"]), int(g["bs"]), bool(g["ra"])
    ds = data.dtype.itemsize
    for key in g:
        if not key.startswith("elem_T"):
            continue
        T = int(key[6:])
        print(f"elem parse {os.path.basename(path)} T={T}", flush=True)
        _, got_T, dec = _parse_and_decode(torch, g[key], data, eb, bs, ra)
        assert got_T == T, (key, got_T)
        want = oracle.elem_decompress(g[key][ds:], data.size, eb, bs, T, ra=ra, dtype=data.dtype)
        assert np.array_equal(dec.view(np.uint8), want.view(np.uint8)), key


@pytest.mark.parametrize("n,bs,T", [(1000, 64, 1), (1000, 64, 3), (4097, 400, 5), (20011, 32, 64), (77, 7, 16),
                                    (300, 2, 128), (100003, 128, 7), (2, 8, 1), (65, 64, 2), (1 << 20, 64, 16)])
@pytest.mark.parametrize("dt,ra", [(np.float32, False), (np.float64, False), (np.float64, True)],
                         ids=["tb_f32", "tb_f64", "ra_f64"])
def test_elem_parse_matches_side_car(n, bs, T, dt, ra):
    """The parsed offsets equal the side-car the compressor wrote, the chunk count is
    found (every chunk non-empty), and the decoded values are the side-car decode's."""
    torch = _torch()
    rng = np.random.default_rng(n + 3 * bs + T)
    x = (np.cumsum(rng.normal(0, 1, n)) * 0.01 + np.sin(np.arange(n) * 0.01)).astype(dt)
    print(f"elem parse n={n} bs={bs} T={T} {np.dtype(dt).name} ra={ra}", flush=True)
    nbytes, out, blk = _compress(torch, x, 1e-3, bs, T, ra)
    s = out[:nbytes].cpu().numpy()
    pblk, got_T, dec = _parse_and_decode(torch, s, x, 1e-3, bs, ra)
    assert got_T == T
    assert np.array_equal(pblk, blk.cpu().numpy())
    want = _decompress(torch, out, blk, n, 1e-3, bs, T, ra, dt)
    assert np.array_equal(dec.view(np.uint8), want.view(np.uint8))


def test_elem_parse_rejects_corrupt_streams():
    torch = _torch()
    x = (np.cumsum(np.random.default_rng(5).normal(0, 1, 5000)) * 0.01).astype(np.float32)
    nbytes, out, _ = _compress(torch, x, 1e-3, 64, 4, False)
    s = out[:nbytes].cpu().numpy().copy()
    bad = s.copy()
    bad[4:12] = 0xff  # first chunk offset not zero
    with pytest.raises(tz.TszpError):
        tz.elem_parse(torch.from_numpy(bad).cuda(), x.size, 64, 4, False)
    cut = s[: nbytes // 2].copy()  # truncated: the chunks do not walk inside it
    with pytest.raises(tz.TszpError):
        tz.elem_parse(torch.from_numpy(cut).cuda(), x.size, 64, 4, False)
    with pytest.raises(tz.TszpError):
        tz.elem_parse(torch.from_numpy(s).cuda(), x.size, 64, 4, True)  # random access needs doubles
    with pytest.raises(tz.TszpError, match="too large"):
        tz.elem_parse(torch.from_numpy(s).cuda(), x.size, 1025, 4, False)  # units beyond 2048 bytes
