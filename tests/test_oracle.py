"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

1. golden vectors produced by the reference's own CPU sources (tests/golden/*.npz),
2. closed forms and brute force (sort positions, error bound, bit_count table),
3. encode -> decode round trips on adversarial inputs.
"""
import glob
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest

import fields
import oracle

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "g*.npz")))
TOPO = [g for g in GOLD if os.path.basename(g).startswith(("g2d", "g3d"))]
PLAIN = [g for g in GOLD if os.path.basename(g).startswith("gplain")]

# Threshold table of the bit-count rule (glibc log2f): bc(x) = 33 - clz(x) iff x >= THR[k].
# The GPU kernels carry the same table; the exhaustive test below re-measures it on this host.
BITCOUNT_THR = np.array([1 << (k + 1) for k in range(20)] + [
    2097151, 4194302, 8388603, 16777205, 33554410, 67108819, 134217637, 268435273,
    536870545, 1073741089, 2147482177, 4294964353], dtype=np.uint64)


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _is3d(g):
    return g["dims"].size == 3


def _fields_equal(a, b, names):
    for n in names:
        np.testing.assert_array_equal(a[n], b[n], err_msg=n)


@pytest.mark.parametrize("path", TOPO, ids=os.path.basename)
def test_find_and_sort_match_reference(path):
    g = _load(path)
    data, eb = g["data"], float(g["eb"])
    cps = oracle.find_cps_3d(data, eb) if _is3d(g) else oracle.find_cps_2d(data, eb)
    names = ["x", "y", "z", "type", "quantized_bin"] if _is3d(g) else ["x", "y", "type",
                                                                         "quantized_bin"]
    assert cps.size == g["cps_found"].size
    _fields_equal(cps, g["cps_found"], names)
    oracle.sort_positions(cps, data)
    _fields_equal(cps, g["cps_sorted"], names + ["sort_position"])


@pytest.mark.parametrize("path", TOPO, ids=os.path.basename)
def test_streams_match_reference(path):
    g = _load(path)
    data, eb, bs = g["data"], float(g["eb"]), int(g["bs"])
    cps = g["cps_sorted"]
    types = oracle.mark_types(cps, tuple(g["dims"]))
    sp = oracle.extrema_sort_positions(cps)
    adaptive = not _is3d(g)  # only the 2D writers use the small-problem partition
    for key in g:
        if key.startswith("topo_T"):
            T = int(key[6:])
            got = oracle.topo_compress(data, eb, bs, types, T=T, adaptive=adaptive)
            np.testing.assert_array_equal(got, g[key], err_msg=key)
        if key.startswith("sortpos_T"):
            T = int(key[9:])
            got = oracle.sortpos_compress(sp, bs, T=T, adaptive=adaptive)
            np.testing.assert_array_equal(got, g[key], err_msg=key)


@pytest.mark.parametrize("path", TOPO, ids=os.path.basename)
def test_decoders_match_reference(path):
    g = _load(path)
    n, eb, bs = g["data"].size, float(g["eb"]), int(g["bs"])
    dec, fn = oracle.topo_decompress(g["topo_T1"], n, eb, bs)
    np.testing.assert_array_equal(dec.view(np.uint32), g["topo_dec_T1"].view(np.uint32))
    np.testing.assert_array_equal(fn, g["topo_fn_T1"])
    if "sortpos_dec_T1" in g:
        e = g["sortpos_dec_T1"].size
        got = oracle.sortpos_decompress(g["sortpos_T1"][8:], e, bs)
        np.testing.assert_array_equal(got, g["sortpos_dec_T1"])


@pytest.mark.parametrize("path", PLAIN, ids=os.path.basename)
def test_plain_matches_reference(path):
    g = _load(path)
    data, eb, bs = g["data"], float(g["eb"]), int(g["bs"])
    for key in g:
        if key.startswith("plain_T"):
            T = int(key[7:])
            got = oracle.plain_compress(data, eb, bs, T=T, adaptive=True)
            np.testing.assert_array_equal(got, g[key], err_msg=key)
            dec = oracle.plain_decompress(got[4:], data.size, eb, bs)
            np.testing.assert_array_equal(dec.view(np.uint32),
                                          g["plain_dec_T" + key[7:]].view(np.uint32))


def _brute_sort_positions(cps, data):
    """sort_position(c) = number of distinct CP values in c's bin below v(c)."""
    flat = data.ravel()
    if cps.dtype == oracle.CP2:
        idx = cps["x"].astype(np.int64) * data.shape[1] + cps["y"]
    else:
        idx = (cps["x"].astype(np.int64) * data.shape[1] + cps["y"]) * data.shape[2] + cps["z"]
    v = flat[idx]
    out = np.zeros(cps.size, dtype=np.int32)
    for k in range(cps.size):
        same = v[cps["quantized_bin"] == cps["quantized_bin"][k]]
        out[k] = np.unique(same[same < v[k]] + np.float32(0)).size  # +0 folds -0.0 into 0.0
    return out


@pytest.mark.parametrize("seed", range(4))
def test_sort_positions_brute_force(seed):
    data = fields.ties2d(23, 31, seed=seed, step=0.002)
    cps = oracle.find_cps_2d(data, 1e-2)
    oracle.sort_positions(cps, data)
    np.testing.assert_array_equal(cps["sort_position"], _brute_sort_positions(cps, data))
    d3 = fields.fbm3d(7, 9, 11, seed=seed)
    c3 = oracle.find_cps_3d(d3, 5e-2)
    oracle.sort_positions(c3, d3)
    np.testing.assert_array_equal(c3["sort_position"], _brute_sort_positions(c3, d3))


def test_classify_brute_force():
    data = fields.ties2d(13, 17, seed=9, step=0.05)
    t = oracle.classify(data)
    for i in range(1, 12):
        for j in range(1, 16):
            c, nb = data[i, j], [data[i - 1, j], data[i + 1, j], data[i, j - 1], data[i, j + 1]]
            want = 0
            if all(c > v for v in nb):
                want = 1
            elif all(c < v for v in nb):
                want = 2
            elif (c < nb[0] and c < nb[1] and c > nb[2] and c > nb[3]) or \
                 (c > nb[0] and c > nb[1] and c < nb[2] and c < nb[3]):
                want = 3
            assert t[i, j] == want
    cps = oracle.find_cps_2d(data, 1e-3)
    np.testing.assert_array_equal(oracle.mark_types(cps, data.shape).reshape(data.shape), t)


def test_rel_to_abs_closed_form():
    data = fields.blobs2d(40, 50, seed=3)
    eb, rng = oracle.rel_to_abs(data, 1e-4)
    r = np.float32(data.max()) - np.float32(data.min())
    assert rng == r and eb == np.float32(np.float32(1e-4) * r)
    eb0, r0 = oracle.rel_to_abs(np.full(10, 3.0, np.float32), 1e-3)
    assert r0 == np.float32(np.finfo(np.float32).eps)


def test_cvt_matches_x86_semantics():
    L = oracle.lib()
    assert L.or_cvt_i32(2.9) == 2 and L.or_cvt_i32(-2.9) == -2
    assert L.or_cvt_i32(2147483647.5) == 2147483647
    assert L.or_cvt_i32(2147483648.0) == -2147483648
    assert L.or_cvt_i32(-2147483649.0) == -2147483648
    assert L.or_cvt_i32(float("nan")) == -2147483648


@pytest.mark.parametrize("bs", [1, 7, 64, 128, 256])
@pytest.mark.parametrize("kind", ["blobs", "ties", "huge"])
def test_topo_round_trip(bs, kind):
    data = {"blobs": fields.blobs2d(29, 37, seed=bs),
            "ties": fields.ties2d(29, 37, seed=bs),
            "huge": fields.huge2d(29, 37, seed=bs)}[kind]
    eb = 1e-3
    types = oracle.classify(data).ravel()
    for T in (1, 3):
        s = oracle.topo_compress(data, eb, bs, types, T=T)
        offs = np.frombuffer(s[:8 * T].tobytes(), dtype=np.uint64)
        assert offs[0] == 0 and np.all(np.diff(offs) >= 0)
    s = oracle.topo_compress(data, eb, bs, types, T=1)
    dec, fn = oracle.topo_decompress(s, data.size, eb, bs)
    q = np.array([oracle.lib().or_cvt_i32(float(np.float64(x) * (1.0 / np.float64(np.float32(eb)))))
                  for x in data.ravel()], dtype=np.int64)
    np.testing.assert_array_equal(dec, (q.astype(np.float32) * np.float32(eb)))
    np.testing.assert_array_equal(fn, types)
    offs = oracle.block_offsets(s[8:], data.size, bs, topo=True)
    assert offs[-1] == s.size - 8


def test_widest_block_round_trip():
    """Deltas wrap in 32-bit two's complement, so |delta| <= 2^31 and bc <= 32."""
    q = np.array([-2147483647, 2147483647, 5, -2147483648, 0, 2147483640, 1, -1],
                 dtype=np.int32)
    s = oracle.sortpos_compress(q, 8)
    assert s[8 + 8 + 4] == 32
    np.testing.assert_array_equal(oracle.sortpos_decompress(s[8:], q.size, 8), q)


def _chunk(args):
    lo, hi = args
    return oracle.check_bitcount_table(BITCOUNT_THR.astype(np.uint32), lo, hi)


def test_bitcount_table_exhaustive():
    """Every 32-bit value: the table rule equals floorf(log2f((float)x)) + 1."""
    edges = np.linspace(1, 1 << 32, 17, dtype=np.uint64)
    with ProcessPoolExecutor(8) as ex:
        bad = sum(ex.map(_chunk, [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:])]))
    assert bad == 0


# ------------------------------------------------------------------------------------------
# quality metrics (example/test_cuda_topology.cu:540-868) on hand-made fields
# ------------------------------------------------------------------------------------------
def test_quality_metrics_known_answers():
    # 5x5: one maximum at (2, 2); the decompressed copy moves it one step right
    x = np.zeros((5, 5), np.float32)
    x[2, 2] = 1.0
    y = x.copy()
    y[2, 2] = 0.0
    y[2, 3] = 1.0 + 1e-6  # (2, 3) is interior? cols 5 -> interior columns 1..3: yes
    q = oracle.quality_metrics(x, y, 0.5)
    assert q["orig_cp"] == [1, 1, 0, 0]
    assert q["dec_cp"][0] == 1 and q["dec_cp"][1] == 1
    assert q["lost"] == [1, 1, 0, 0] and q["preserved"] == 0 and q["new_cps"] == 1
    # errors 1.0 at (2, 2) and 1.0000010 at (2, 3): both above 1.01 eb, one above 2 eb
    assert q["over_1eb"] == 2 and q["over_2eb"] == 1 and abs(q["max_err"] - (1.0 + 1e-6)) < 1e-6
    # local order: interior 3x3, pairs along rows (2 per row, 3 rows) and columns: 12;
    # strict in x only where the max is involved: (2,1)-(2,2), (2,2)-(2,3), (1,2)-(2,2),
    # (2,2)-(3,2); in y (2,1)-(2,2) and (1,2)-(2,2), (2,2)-(3,2) become equal (violations),
    # (2,2)-(2,3) flips
    assert q["pairs"] == [4, 4, 0, 4] and q["violations"] == [4, 4, 0, 4]


def test_quality_metrics_saddle_persistence():
    # a 2D saddle at (2, 2): higher along rows, lower along columns
    x = np.zeros((5, 5), np.float32)
    x[2, 2] = 0.5
    x[2, 1] = x[2, 3] = 0.2   # row neighbours lower
    x[1, 2] = x[3, 2] = 0.9   # column neighbours higher
    q = oracle.quality_metrics(x, x.copy(), 0.1)
    assert q["saddles"] == 1 and q["orig_cp"][3] == 1
    assert abs(q["pers_min"] - 0.3) < 1e-6  # min(|0.5-0.2|, |0.5-0.9|)
    # 0.3 > 0, 0.05, 0.1, 0.2 but not > 0.5 or 1.0
    assert q["saddles_above"] == [1, 1, 1, 1, 0, 0]
    assert q["saddles_kept_above"] == [1, 1, 1, 1, 0, 0]
    assert q["violations"] == [0, 0, 0, 0] and q["lost"][0] == 0
