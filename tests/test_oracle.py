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
ELEM = [g for g in GOLD if os.path.basename(g).startswith(("gtb", "gra"))]

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


@pytest.mark.parametrize("path", PLAIN, ids=os.path.basename)
def test_plain_decoder_with_known_chunk_count(path):
    """The decoder given the chunk count agrees with the reference's detecting decoder
    wherever the detection is right, and stays within the error bound on the one-chunk
    stream that the reference's detection misreads (gplain_misdetect, where the
    reference's own output is 185 eb off)."""
    g = _load(path)
    data, eb, bs = g["data"], float(g["eb"]), int(g["bs"])
    dec = oracle.plain_decompress(g["plain_T1"][4:], data.size, eb, bs, T=1)
    assert float(np.abs(dec - data).max()) <= eb * 1.0001
    ref = g["plain_dec_T1"]
    if "misdetect" in os.path.basename(path):
        assert float(np.abs(ref - data).max()) > 100 * eb
    else:
        np.testing.assert_array_equal(dec.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("path", ELEM, ids=os.path.basename)
def test_elem_streams_match_reference(path):
    """Threadblock (float, double) and double random-access writers, every chunk count
    the fixture holds; the decoders on the one-chunk stream."""
    g = _load(path)
    data, eb, bs, ra = g["data"], float(g["eb"]), int(g["bs"]), bool(g["ra"])
    ds = data.dtype.itemsize
    for key in g:
        if key.startswith("elem_T"):
            T = int(key[6:])
            got = oracle.elem_compress(data, eb, bs, T=T, ra=ra)
            np.testing.assert_array_equal(got, g[key], err_msg=key)
            dec = oracle.elem_decompress(got[ds:], data.size, eb, bs, T, ra=ra, dtype=data.dtype)
            assert np.all(np.abs(dec.astype(np.float64) - data) <= 1.0001 * eb + np.abs(data) * 2.0 ** -22)
    if "elem_dec_T1" in g:
        dec = oracle.elem_decompress(g["elem_T1"][ds:], data.size, eb, bs, 1, ra=ra, dtype=data.dtype)
        np.testing.assert_array_equal(dec.view(np.uint8), g["elem_dec_T1"].view(np.uint8))


@pytest.mark.parametrize("n,bs,T", [(1, 8, 1), (5, 64, 8), (17, 1, 4), (65, 64, 1), (200, 7, 3),
                                    (4097, 1025, 5)])
@pytest.mark.parametrize("dt,ra", [(np.float32, False), (np.float64, False), (np.float64, True)])
def test_elem_layout_closed_form(n, bs, T, dt, ra):
    """Stream length from the layout alone on a constant field (every bc is 0): per
    non-empty chunk a 4-byte anchor and a bc byte per block (threadblock: blocks of bs
    after the anchor; random access: blocks of bs from the chunk start, each with its
    own anchor), offsets the running sums; the decode is exact."""
    data = np.full(n, 0.25, dtype=dt)
    s = oracle.elem_compress(data, 1e-3, bs, T=T, ra=ra)
    ds = np.dtype(dt).itemsize
    ce = n // T
    want, offs = 0, []
    for t in range(T):
        lo, hi = t * ce, (n if t == T - 1 else (t + 1) * ce)
        offs.append(want)
        if hi > lo:
            if ra:
                nb = -(-(hi - lo) // bs)
                want += 5 * nb
            else:
                want += 4 + -(-(hi - lo - 1) // bs)
    assert s.size == ds + 8 * T + want
    np.testing.assert_array_equal(np.frombuffer(s[ds:ds + 8 * T].tobytes(), np.uint64), offs)
    eb_bytes = np.array(1e-3, dtype=dt).tobytes()[::-1]  # big-endian error bound
    assert s[:ds].tobytes() == eb_bytes
    dec = oracle.elem_decompress(s[ds:], n, 1e-3, bs, T, ra=ra, dtype=dt)
    q = int(0.25 * (1 / np.float64(np.float32(1e-3) if dt == np.float32 else 1e-3)))
    want_v = (np.float32(q) * np.float32(1e-3)) if dt == np.float32 else q * 1e-3
    assert np.all(dec == want_v)


def _saddle_box(c, xs, ys, zs):
    """5x5x5 int field, centre (2,2,2) = c with axis neighbours xs, ys, zs (-, +)."""
    q = np.zeros((5, 5, 5), np.int32)
    q[2, 2, 2] = c
    q[1, 2, 2], q[3, 2, 2] = xs
    q[2, 1, 2], q[2, 3, 2] = ys
    q[2, 2, 1], q[2, 2, 3] = zs
    t = np.zeros((5, 5, 5), np.uint8)
    t[2, 2, 2] = 3
    return q, t


def test_saddle_fix_known_answers():
    # intact saddle: high along x, low along y -> untouched
    q, t = _saddle_box(5, (3, 4), (7, 8), (5, 6))
    assert oracle.fix_saddles_pass(q, t)[1] == 0
    # broken (ties along x): c - 1 = 4 is high nowhere; c + 1 = 6 is high along x
    # (> max(5, 3)) and low along y (< min(7, 8))
    q, t = _saddle_box(5, (5, 3), (7, 8), (6, 6))
    got, k = oracle.fix_saddles_pass(q, t)
    assert k == 1 and got[2, 2, 2] == 6 and (got != q).sum() == 1
    # broken (high along x only): c - 1 is tried first and works -- high along x
    # (4 > max(2, 1)), low along y (4 < 5)
    q, t = _saddle_box(5, (2, 1), (5, 5), (5, 9))
    got, k = oracle.fix_saddles_pass(q, t)
    assert k == 1 and got[2, 2, 2] == 4
    # the working trial 6 would reach the -x neighbour (value 5) marked as a maximum
    q, t = _saddle_box(5, (5, 3), (7, 8), (6, 6))
    t[1, 2, 2] = 1
    got, k = oracle.fix_saddles_pass(q, t)
    assert k == 0 and got[2, 2, 2] == 5
    # a neighbouring minimum must stay below the trial: +y (value 8) marked as a
    # minimum blocks 6 as well
    t[1, 2, 2] = 0
    t[2, 3, 2] = 2
    assert oracle.fix_saddles_pass(q, t)[1] == 0
    # no trial within one step: all six neighbours equal to c
    q, t = _saddle_box(5, (5, 5), (5, 5), (5, 5))
    assert oracle.fix_saddles_pass(q, t)[1] == 0


def test_saddle_fix_reads_the_snapshot():
    """Two adjacent broken saddles in one pass both see the old values (Jacobi)."""
    q = np.zeros((4, 5, 5), np.int32)
    t = np.zeros((4, 5, 5), np.uint8)
    t[1, 2, 2] = t[2, 2, 2] = 3
    q[1, 2, 2] = q[2, 2, 2] = 5
    q[0, 2, 2], q[3, 2, 2] = 4, 4          # x: both saddles sit between 4 and a 5
    for x in (1, 2):
        q[x, 1, 2], q[x, 3, 2] = 7, 7      # y: low
        q[x, 2, 1], q[x, 2, 3] = 5, 5      # z: ties
    got, k = oracle.fix_saddles_pass(q, t)
    # each sees x neighbours (4, 5): c - 1 = 4 is not high; c + 1 = 6 > max(4, 5) and < 7
    assert k == 2 and got[1, 2, 2] == 6 and got[2, 2, 2] == 6
    # a Gauss-Seidel order would have let the second see 6 and fail; brute force check
    # of the pass definition on random small fields
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = rng.integers(-3, 4, size=(5, 6, 7)).astype(np.int32)
        t = rng.integers(0, 4, size=q.shape).astype(np.uint8)
        got, k = oracle.fix_saddles_pass(q, t)
        want = q.copy()
        for x in range(1, 4):
            for y in range(1, 5):
                for z in range(1, 6):
                    if t[x, y, z] != 3:
                        continue
                    nb = [(x - 1, y, z), (x + 1, y, z), (x, y - 1, z), (x, y + 1, z), (x, y, z - 1),
                          (x, y, z + 1)]
                    v = [int(q[p]) for p in nb]
                    c = int(q[x, y, z])
                    hi = sum(c > v[2 * a] and c > v[2 * a + 1] for a in range(3))
                    lo = sum(c < v[2 * a] and c < v[2 * a + 1] for a in range(3))
                    if hi and lo:
                        continue
                    for tr in (c - 1, c + 1):
                        th = sum(tr > max(v[2 * a], v[2 * a + 1]) for a in range(3))
                        tl = sum(tr < min(v[2 * a], v[2 * a + 1]) for a in range(3))
                        ok = th >= 1 and tl >= 1 and all(
                            not (t[p] == 1 and vv <= tr) and not (t[p] == 2 and vv >= tr)
                            for p, vv in zip(nb, v))
                        if ok:
                            want[x, y, z] = tr
                            break
        np.testing.assert_array_equal(got, want)
        assert k == int((got != q).sum())


@pytest.mark.parametrize("path", [g for g in TOPO if os.path.basename(g).startswith("g3d")],
                         ids=os.path.basename)
def test_topo3d_f64_equals_float_writer_without_fixes(path):
    """On float-valued double data the double 3D stream carries q = C cast of x / eb and
    the type map: where no saddle needed fixing it is the reference 3D writer's own
    one-chunk stream (golden); otherwise it decodes to the fixed q, and every fixed
    point is a saddle moved by one step."""
    g = _load(path)
    data, eb, bs = g["data"], float(g["eb"]), int(g["bs"])
    types = oracle.classify(data)
    s64, q = oracle.topo3d_f64_compress(data.astype(np.float64), eb, bs, types)
    inv = 1.0 / float(np.float32(eb))
    q0 = np.array([oracle.lib().or_cvt_i32(float(x) * inv) for x in data.ravel()],
                  dtype=np.int32).reshape(data.shape)
    moved = q != q0
    assert np.all(np.abs(q.astype(np.int64) - q0)[moved] == 1) and np.all(types[moved] == 3)
    if not moved.any():
        np.testing.assert_array_equal(s64, g["topo_T1"])
    dec, fn = oracle.topo_decompress(s64, data.size, eb, bs)
    np.testing.assert_array_equal(dec, q.ravel().astype(np.float32) * np.float32(eb))
    np.testing.assert_array_equal(fn, types.ravel())


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
