"""GPU parity: the HIP library against the reference's golden vectors and the CPU oracle.

Bit-exact comparisons of critical-point lists, sort positions, streams, quantized
reconstructions and type arrays; every call goes through the C-ABI of
libtoposzp_hip.so (drop-in host API and native device API).
"""
import glob
import os

import numpy as np
import pytest

import fields
import oracle
import toposzp_hip as tz

pytestmark = pytest.mark.gpu

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "g*.npz")))
TOPO = [g for g in GOLD if os.path.basename(g).startswith(("g2d", "g3d"))]
PLAIN = [g for g in GOLD if os.path.basename(g).startswith("gplain")]


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _eq(a, b, msg=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{msg}: shape {a.shape} vs {b.shape}"
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a.ravel() != b.ravel())
        raise AssertionError(f"{msg}: {bad.size} mismatches, first at {bad[:8]}: "
                             f"{a.ravel()[bad[:4]]} vs {b.ravel()[bad[:4]]}")


# ------------------------------------------------------------------------------------------
# drop-in (host pointers) against the reference's own outputs
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("path", TOPO, ids=os.path.basename)
def test_dropin_matches_reference_golden(path):
    _torch()
    g = _load(path)
    data, eb, bs = g["data"], float(g["eb"]), int(g["bs"])
    n = data.size
    cps = tz.find_critical_points(data, eb)
    _eq(cps.view(np.int32), g["cps_found"].view(np.int32), "find_critical_points")
    srt = tz.sort_critical_points(cps, data)
    _eq(srt.view(np.int32), g["cps_sorted"].view(np.int32), "sort positions")
    sp_stream = tz.compress_sort_positions(srt, bs)
    _eq(sp_stream, g["sortpos_T1"][8:], "sort-position stream")  # no u64 count in the CUDA form
    topo = tz.compress_topology(data, eb, bs, srt)
    _eq(topo, g["topo_T1"], "topology stream")
    dec, fn = tz.decompress_topology(g["topo_T1"], n, eb, bs)
    _eq(dec, g["topo_dec_T1"], "decompressed field")
    _eq(fn, g["topo_fn_T1"], "FN types")
    ext = srt["sort_position"][(srt["type"] == 1) | (srt["type"] == 2)]
    if ext.size:
        _eq(tz.decompress_sort_positions(g["sortpos_T1"][8:], ext.size, bs), ext, "sort-pos decode")
    for key in g:
        if key.startswith("topo_T") and key != "topo_T1":
            d2, f2 = tz.decompress_topology(g[key], n, eb, bs)
            _eq(d2, g["topo_dec_T1"], f"{key} decode")
            _eq(f2, g["topo_fn_T1"], f"{key} types")


@pytest.mark.parametrize("path", PLAIN, ids=os.path.basename)
def test_dropin_plain_matches_reference_golden(path):
    _torch()
    g = _load(path)
    data, eb, bs = g["data"], float(g["eb"]), int(g["bs"])
    s = tz.compress_plain(data, eb, bs)
    _eq(s, g["plain_T1"], "plain stream")
    for key in g:
        if key.startswith("plain_T"):
            dec = tz.decompress_plain(g[key][4:], data.size, eb, bs)
            if "misdetect" in os.path.basename(path):
                # the reference's chunk-count detection misreads this one-chunk stream
                # (its output is 185 eb off); the drop-in walks it as one chunk
                _eq(dec, oracle.plain_decompress(g[key][4:], data.size, eb, bs, T=1), f"{key} decode")
                assert float(np.abs(dec - data).max()) <= eb * 1.0001
            else:
                _eq(dec, g["plain_dec_T1"], f"{key} decode")


# ------------------------------------------------------------------------------------------
# native device API against the oracle
# ------------------------------------------------------------------------------------------
def _oracle_topo(data, eb, bs, T):
    cps = oracle.find_cps_3d(data, eb) if data.ndim == 3 else oracle.find_cps_2d(data, eb)
    oracle.sort_positions(cps, data)
    types = oracle.mark_types(cps, data.shape)
    sp = oracle.extrema_sort_positions(cps)
    adaptive = data.ndim == 2
    topo = oracle.topo_compress(data, eb, bs, types, T=T, adaptive=adaptive)
    sps = oracle.sortpos_compress(sp, bs, T=T, adaptive=adaptive) if sp.size else np.zeros(0, np.uint8)
    return cps, types, sp, topo, sps


CASES = [
    ("blobs2d", lambda: fields.blobs2d(300, 517, nblobs=24, seed=21), 1e-3, 64, 1),
    ("blobs2d_T16", lambda: fields.blobs2d(211, 333, nblobs=24, seed=22), 2e-3, 64, 16),
    ("blobs2d_bs8_T4", lambda: fields.blobs2d(97, 129, seed=23), 1e-3, 8, 4),
    ("ties2d_bs16", lambda: fields.ties2d(150, 151, seed=24), 1e-3, 16, 1),
    ("huge2d_bs128", lambda: fields.huge2d(77, 91, seed=25), 1e-3, 128, 2),
    ("ragged2d_bs64", lambda: fields.blobs2d(5, 13, seed=3, noise=0.05), 1e-3, 64, 1),
    ("bs1_2d", lambda: fields.blobs2d(31, 33, seed=26), 1e-3, 1, 1),
    ("bs7_2d", lambda: fields.blobs2d(41, 43, seed=27), 1e-3, 7, 3),
    ("bs257_2d", lambda: fields.blobs2d(60, 70, seed=28, noise=0.02), 5e-4, 257, 1),
    ("fbm3d", lambda: fields.fbm3d(40, 52, 61, seed=29), 3e-3, 64, 1),
    ("fbm3d_bs128_T8", lambda: fields.fbm3d(33, 40, 45, seed=30), 3e-3, 128, 8),
    ("blobs3d_bs256", lambda: fields.blobs3d(30, 31, 37, seed=31, noise=1e-3), 1e-4, 256, 1),
    ("tiny3d", lambda: fields.blobs3d(3, 3, 3, seed=32, noise=0.1), 1e-3, 64, 1),
    ("const2d", lambda: np.full((40, 50), 1.25, np.float32), 1e-3, 64, 1),
    # fastest axis a multiple of 4 and block size 16 to 256 (powers of 2): 16-byte path
    ("vec2d_bs64", lambda: fields.blobs2d(96, 260, nblobs=24, seed=33), 1e-3, 64, 1),
    ("vec2d_bs32_T5", lambda: fields.blobs2d(77, 132, seed=34, noise=0.03), 1e-3, 32, 5),
    ("vec3d_bs64", lambda: fields.fbm3d(24, 36, 40, seed=35), 3e-3, 64, 1),
    ("vec3d_bs32_T3", lambda: fields.fbm3d(17, 20, 28, seed=36), 2e-3, 32, 3),
    ("vec3d_ragged_tail", lambda: fields.fbm3d(5, 7, 12, seed=37), 2e-3, 64, 1),
    ("vec2d_ties", lambda: fields.ties2d(64, 128, seed=38), 1e-3, 64, 2),
    ("vec3d_bs128", lambda: fields.fbm3d(20, 24, 64, seed=39), 3e-3, 128, 1),
    ("vec2d_bs128_T3", lambda: fields.blobs2d(90, 256, seed=40, noise=0.02), 1e-3, 128, 3),
    ("vec3d_bs128_tail", lambda: fields.fbm3d(7, 9, 20, seed=41), 2e-3, 128, 1),
    ("vec3d_bs16", lambda: fields.fbm3d(18, 22, 36, seed=42), 3e-3, 16, 1),
    ("vec2d_bs16_T7", lambda: fields.blobs2d(70, 96, seed=43, noise=0.02), 1e-3, 16, 7),
    ("vec3d_bs256", lambda: fields.fbm3d(16, 20, 64, seed=44), 3e-3, 256, 1),
    ("vec2d_bs256_T2", lambda: fields.blobs2d(50, 512, seed=45, noise=0.02), 1e-3, 256, 2),
]


@pytest.mark.parametrize("name,make,eb,bs,T", CASES, ids=[c[0] for c in CASES])
def test_native_topology_matches_oracle(name, make, eb, bs, T):
    torch = _torch()
    data = make()
    n = data.size
    cps, types, sp, topo, sps = _oracle_topo(data, eb, bs, T)
    f = torch.from_numpy(data).cuda()
    info, buf = tz.topo_compress(f, eb, bs, T=T)
    torch.cuda.synchronize()
    assert info.cp_count == cps.size
    assert info.ext_count == sp.size
    _eq(buf.topo[:info.topo_bytes].cpu().numpy(), topo, "topology stream")
    if sp.size:
        _eq(buf.sp_values[:sp.size].cpu().numpy(), sp, "sort positions")
        _eq(buf.sp[:info.sortpos_bytes].cpu().numpy(), sps, "sort-position stream")
    else:
        assert info.sortpos_bytes == 0
    # decode with the side-car
    out, ty = tz.topo_decompress(buf, n, eb, bs, T=T)
    dec, fn = oracle.topo_decompress(topo[:], n, eb, bs) if T == 1 else oracle.topo_decompress(
        oracle.topo_compress(data, eb, bs, types, T=1), n, eb, bs)
    _eq(out.cpu().numpy(), dec, "decoded field")
    _eq(ty.cpu().numpy().astype(np.int32), fn, "decoded types")
    _eq(ty.cpu().numpy(), types, "decoded types vs classification")
    if sp.size:
        _eq(tz.sortpos_decompress(buf, sp.size, bs, T=T).cpu().numpy(), sp, "sort-position decode")
    # the GPU classifier agrees with the oracle's
    t2, cnt = tz.classify(f)
    assert cnt == cps.size
    _eq(t2.cpu().numpy(), oracle.classify(data).ravel(), "classify")


@pytest.mark.parametrize("bs,T", [(64, 1), (128, 1), (32, 5), (1025, 1), (2, 2), (100, 16), (16, 3), (256, 1)])
def test_native_plain_matches_oracle(bs, T):
    torch = _torch()
    data = fields.blobs2d(123, 457, seed=bs).ravel()
    eb = 1e-3
    ref = oracle.plain_compress(data, eb, bs, T=T, adaptive=True)
    f = torch.from_numpy(data).cuda()
    nbytes, out, blk = tz.plain_compress(f, eb, bs, T=T)
    torch.cuda.synchronize()
    _eq(out[:nbytes].cpu().numpy(), ref, "plain stream")
    dec = tz.plain_decompress(out, blk, data.size, eb, bs, T=T)
    ref1 = oracle.plain_compress(data, eb, bs, T=1)
    _eq(dec.cpu().numpy(), oracle.plain_decompress(ref1[4:], data.size, eb, bs), "plain decode")


def test_rel_error_bound_matches_reference_rule():
    torch = _torch()
    data = fields.fbm3d(20, 30, 40, seed=5)
    want = oracle.rel_to_abs(data, 1e-3)
    got = tz.rel_to_abs(torch.from_numpy(data).cuda(), 1e-3)
    assert got == want


def test_dropin_plain_round_trip_matches_oracle():
    _torch()
    data = fields.huge2d(64, 97, seed=7).ravel()
    for bs in (16, 64, 1024):
        s = tz.compress_plain(data, 0.5, bs)
        _eq(s, oracle.plain_compress(data, 0.5, bs), f"plain bs={bs}")
        _eq(tz.decompress_plain(s[4:], data.size, 0.5, bs), oracle.plain_decompress(s[4:], data.size, 0.5, bs))


def test_dropin_3d_pipeline_matches_oracle():
    _torch()
    data = fields.fbm3d(24, 30, 36, seed=11)
    eb, bs = 4e-3, 64
    cps = tz.find_critical_points(data, eb)
    ocps = oracle.find_cps_3d(data, eb)
    _eq(cps.view(np.int32), ocps.view(np.int32), "3D critical points")
    srt = tz.sort_critical_points(cps, data)
    oracle.sort_positions(ocps, data)
    _eq(srt.view(np.int32), ocps.view(np.int32), "3D sort positions")
    types = oracle.mark_types(ocps, data.shape)
    _eq(tz.compress_topology(data, eb, bs, srt), oracle.topo_compress(data, eb, bs, types, adaptive=False))
    sp = oracle.extrema_sort_positions(ocps)
    _eq(tz.compress_sort_positions(srt, bs), oracle.sortpos_compress(sp, bs, adaptive=False)[8:])


def test_large_field_matches_oracle():
    """A 3D field of 8.4 M points (the oracle takes a few seconds)."""
    torch = _torch()
    f = tz.gen_fbm((128, 256, 256), seed=2024)
    data = f.cpu().numpy()
    eb, _ = tz.rel_to_abs(f, 1e-3)
    bs = 64
    info, buf = tz.topo_compress(f, eb, bs)
    cps, types, sp, topo, sps = _oracle_topo(data, eb, bs, 1)
    assert info.cp_count == cps.size and info.ext_count == sp.size
    _eq(buf.topo[:info.topo_bytes].cpu().numpy(), topo, "topology stream")
    _eq(buf.sp[:info.sortpos_bytes].cpu().numpy(), sps, "sort-position stream")
    out, ty = tz.topo_decompress(buf, data.size, eb, bs)
    dec, fn = oracle.topo_decompress(topo, data.size, eb, bs)
    _eq(out.cpu().numpy(), dec)
    _eq(ty.cpu().numpy(), types)


# ------------------------------------------------------------------------------------------
# sort positions: per-bin bitmap path against the radix path and the oracle
# ------------------------------------------------------------------------------------------
def _centred(data, scale=1.0):
    return ((data - np.float32(np.median(data))) * np.float32(scale)).astype(np.float32)


SP_CASES = [
    # values around zero: the bins next to zero are wide (radix on their buckets) and
    # their neighbours need several bitmap chunks
    ("zero_centred_3d", lambda: _centred(fields.fbm3d(40, 44, 48, seed=41)), 1e-2),
    ("zero_centred_2d", lambda: _centred(fields.blobs2d(257, 263, seed=42, noise=0.05)), 5e-3),
    # every value within a few error bounds of zero: all bins wide
    ("near_zero_2d", lambda: _centred(fields.blobs2d(120, 130, seed=43, noise=0.05), 1e-3), 1e-3),
    # more bins than the bitmap path takes: radix path
    ("many_bins_2d", lambda: fields.blobs2d(200, 210, seed=44, noise=0.05), 2e-5),
    # one-signed fields, far from zero
    ("negative_3d", lambda: fields.fbm3d(30, 32, 34, seed=45) - np.float32(5.0), 2e-3),
    ("offset_2d", lambda: fields.blobs2d(150, 160, seed=46) + np.float32(300.0), 1e-3),
]


@pytest.mark.parametrize("name,make,eb", SP_CASES, ids=[c[0] for c in SP_CASES])
def test_sort_positions_bin_path(name, make, eb, monkeypatch):
    torch = _torch()
    data = np.ascontiguousarray(make(), dtype=np.float32)
    f = torch.from_numpy(data).cuda()
    bs = 64
    info, buf = tz.topo_compress(f, eb, bs)
    torch.cuda.synchronize()
    sp_bins = buf.sp_values[:info.ext_count].cpu().numpy().copy()
    stream_bins = buf.sp[:info.sortpos_bytes].cpu().numpy().copy()
    monkeypatch.setenv("TSZP_SORTPOS", "radix")
    info2, buf2 = tz.topo_compress(f, eb, bs)
    torch.cuda.synchronize()
    monkeypatch.delenv("TSZP_SORTPOS")
    assert info2.ext_count == info.ext_count and info.ext_count > 0
    _eq(sp_bins, buf2.sp_values[:info2.ext_count].cpu().numpy(), "bin path vs radix path")
    _eq(stream_bins, buf2.sp[:info2.sortpos_bytes].cpu().numpy(), "sort-position streams")
    cps, types, sp, topo, sps = _oracle_topo(data, eb, bs, 1)
    _eq(sp_bins, sp, "sort positions vs oracle")


def test_sort_positions_bin_path_large():
    """8.4 M points, field centred on zero (wide, multi-chunk and narrow bins together);
    the radix path is the reference."""
    torch = _torch()
    f = tz.gen_fbm((128, 256, 256), seed=77)
    f -= f.median()
    eb, _ = tz.rel_to_abs(f, 1e-3)
    info, buf = tz.topo_compress(f, eb, 64)
    torch.cuda.synchronize()
    a = buf.sp_values[:info.ext_count].clone()
    os.environ["TSZP_SORTPOS"] = "radix"
    try:
        info2, buf2 = tz.topo_compress(f, eb, 64)
        torch.cuda.synchronize()
    finally:
        del os.environ["TSZP_SORTPOS"]
    assert info2.ext_count == info.ext_count
    assert torch.equal(a, buf2.sp_values[:info.ext_count])


@pytest.mark.parametrize("shape,bs", [((128, 256, 256), 64), ((40, 96, 100), 32), ((1024, 1028), 64),
                                      ((128, 256, 256), 128), ((33, 70, 92), 128), ((515, 1004), 128),
                                      ((64, 128, 128), 16), ((31, 40, 52), 16), ((64, 128, 256), 256),
                                      ((300, 1024), 256)])
def test_codec_fast_path_matches_generic(shape, bs):
    """The 16-byte-stencil codec against the generic kernel on the same field: identical
    topology, sort-position and plain streams."""
    torch = _torch()
    f = tz.gen_fbm(shape, seed=91) if len(shape) == 3 else torch.from_numpy(
        fields.blobs2d(*shape, nblobs=40, seed=92, noise=0.05)).cuda()
    eb, _ = tz.rel_to_abs(f, 1e-3)
    outs = []
    for generic in (False, True):
        if generic:
            os.environ["TSZP_CODEC_GENERIC"] = "1"
        try:
            info, buf = tz.topo_compress(f, eb, bs)
            nbytes, pout, _ = tz.plain_compress(f, eb, bs)
            torch.cuda.synchronize()
            outs.append((info.topo_bytes, info.sortpos_bytes, buf.topo[:info.topo_bytes].cpu().numpy().copy(),
                         buf.sp[:info.sortpos_bytes].cpu().numpy().copy(), pout[:nbytes].cpu().numpy().copy()))
        finally:
            os.environ.pop("TSZP_CODEC_GENERIC", None)
    (a, b) = outs
    assert a[0] == b[0] and a[1] == b[1]
    _eq(a[2], b[2], "topology stream")
    _eq(a[3], b[3], "sort-position stream")
    _eq(a[4], b[4], "plain stream")


@pytest.mark.parametrize("shape,bs", [((128, 256, 256), 64), ((40, 96, 100), 32), ((33, 70, 92), 128),
                                      ((64, 128, 128), 16), ((515, 1004), 128), ((300, 1024), 256),
                                      ((97, 1000), 64), ((31, 40, 52), 16)])
def test_codec_lds_rows_match_global_rows(shape, bs):
    """Classification with the tile's rows staged in LDS (the tile's points and one row on
    either side, TSZP_CODEC_LDS_ROWS) against the default global-load stencil: identical
    topology and sort-position streams, on grids whose rows start anywhere in a tile and
    whose last tile is partial."""
    torch = _torch()
    f = tz.gen_fbm(shape, seed=93) if len(shape) == 3 else torch.from_numpy(
        fields.blobs2d(*shape, nblobs=40, seed=94, noise=0.05)).cuda()
    eb, _ = tz.rel_to_abs(f, 1e-3)
    outs = []
    for lds in (False, True):
        if lds:
            os.environ["TSZP_CODEC_LDS_ROWS"] = "1"
        try:
            info, buf = tz.topo_compress(f, eb, bs)
            torch.cuda.synchronize()
            outs.append((info.cp_count, info.ext_count, buf.topo[:info.topo_bytes].cpu().numpy().copy(),
                         buf.sp[:info.sortpos_bytes].cpu().numpy().copy()))
        finally:
            os.environ.pop("TSZP_CODEC_LDS_ROWS", None)
    (a, b) = outs
    assert a[0] == b[0] and a[1] == b[1] and a[0] > 0
    _eq(a[2], b[2], "topology stream")
    _eq(a[3], b[3], "sort-position stream")
