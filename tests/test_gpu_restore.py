"""GPU topology restoration: after decompression + restore, every interior point must
have the critical-point type of the original field (checked by the CPU oracle's
classifier), with every value inside its quantization cell."""
import numpy as np
import pytest

import fields
import oracle
import toposzp_hip as tz

pytestmark = pytest.mark.gpu


def _pipeline(data, eb, bs=64):
    import torch
    assert torch.cuda.is_available()
    f = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    info, buf = tz.topo_compress(f, eb, bs)
    out, types = tz.topo_decompress(buf, data.size, eb, bs)
    sp = tz.sortpos_decompress(buf, info.ext_count, bs) if info.ext_count else None
    dec = out.clone()
    rinfo = tz.topo_restore(out, types, data.shape, eb, sp=sp, n_ext=info.ext_count)
    torch.cuda.synchronize()
    return dec.cpu().numpy(), out.cpu().numpy().reshape(data.shape), types.cpu().numpy(), rinfo


def _check_cells(x, dec, y, eb):
    """Every restored value stays in the interval of its cell: [d, d + 0.9 eb] above
    zero, [d - 0.9 eb, d] below, [-eb 2^-16, eb 2^-16] in the zero cell."""
    x, dec, y = x.ravel().astype(np.float64), dec.ravel().astype(np.float64), y.ravel().astype(np.float64)
    slack = np.abs(dec) * 2.0 ** -22
    pos, neg, zero = dec > 0, dec < 0, dec == 0
    assert np.all(y[pos] >= dec[pos]) and np.all(y[pos] <= dec[pos] + 0.9 * eb * (1 + 1e-5) + slack[pos])
    assert np.all(y[neg] <= dec[neg]) and np.all(y[neg] >= dec[neg] - 0.9 * eb * (1 + 1e-5) - slack[neg])
    assert np.all(np.abs(y[zero]) <= np.float32(eb) * np.float32(2.0 ** -16))
    # pointwise error bound: |y - x| <= eb outside the zero cell, up to one float spacing
    # of d (the decoder rounds (float)q * eb, as the reference's does), and
    # |y - x| <= eb (1 + 2^-16) inside it, where |x| < eb
    ebf = float(np.float32(eb))
    err = np.abs(y - x)
    ulp = np.spacing(np.abs(dec[~zero]).astype(np.float32)).astype(np.float64)
    assert np.all(err[~zero] <= ebf + ulp)
    assert np.all(err[zero] <= ebf * (1 + 2.0 ** -16))


def _check_extrema_order(x, dec, y, types, eb, info, slack=1):
    """Extrema of one reference bin (outside the zero cell) keep the order of the
    original values: sorted by x, the restored values never decrease.  The only
    exceptions allowed are around windows that had to move an anchored extremum (an
    extremum on a cell boundary whose sort position belongs to the neighbouring bin, or
    a window cut out of a large plateau), and there must be no more such windows than
    boundary extrema, or `slack`."""
    x, dec, y, types = x.ravel(), dec.ravel(), y.ravel(), types.ravel()
    ext = ((types == 1) | (types == 2)) & (dec != 0)
    xe, de, ye = x[ext], dec[ext], y[ext]
    inv = 1.0 / np.float64(np.float32(eb))
    binv = ((xe + np.float32(eb)).astype(np.float64) * inv).astype(np.int64)
    q = np.rint(de.astype(np.float64) * inv).astype(np.int64)
    boundary = int(np.count_nonzero(binv != q + 1))
    order = np.lexsort((xe, binv, de))
    same = (de[order][1:] == de[order][:-1]) & (binv[order][1:] == binv[order][:-1])
    dy = ye[order][1:] - ye[order][:-1]
    dx = xe[order][1:] - xe[order][:-1]
    bad = int(np.count_nonzero(same & ((dy < 0) | ((dx == 0) & (dy != 0)))))
    print(f"  extrema order: pairs={int(np.count_nonzero(same))} violations={bad} "
          f"reordered windows={info.windows_reordered} boundary extrema={boundary}")
    assert info.windows_reordered <= max(boundary, slack)
    assert bad <= 8 * info.windows_reordered


NOISY = [
    ("blobs2d_cfg0", lambda: fields.blobs2d(450, 900, nblobs=64, noise=0.01, seed=1), 1e-3),
    ("blobs2d_b", lambda: fields.blobs2d(300, 700, nblobs=64, noise=0.01, seed=7), 1e-3),
    ("ties_coarse", lambda: fields.ties2d(300, 300, seed=5, step=0.001), 5e-3),
    ("fbm3d", lambda: fields.fbm3d(48, 64, 80, seed=3), 3e-3),
    ("blobs3d", lambda: fields.blobs3d(40, 60, 70, seed=4, noise=2e-3), 2e-4),
]


@pytest.mark.parametrize("name,make,eb", NOISY, ids=[c[0] for c in NOISY])
def test_restore_zero_mismatches(name, make, eb):
    data = make()
    dec, y, types, info = _pipeline(data, eb)
    want = oracle.classify(data).ravel()
    np.testing.assert_array_equal(types, want)
    got = oracle.classify(y).ravel()
    raw = oracle.classify(dec.reshape(data.shape)).ravel()
    print(f"{name}: cps={np.count_nonzero(want)} raw mismatches={np.count_nonzero(raw != want)} "
          f"after={np.count_nonzero(got != want)} moves={info.moves} windows={info.windows} "
          f"rounds={info.rounds} maxdev={info.max_dev_over_eb:.3f}")
    assert info.mismatches() == np.count_nonzero(got != want)
    assert np.count_nonzero(got != want) == 0
    _check_cells(data, dec, y, eb)
    _check_extrema_order(data, dec, y, types, eb, info)


SMOOTH = [
    ("smooth2d", lambda: fields.blobs2d(400, 600, nblobs=20, noise=0.0005, seed=2), 1e-3),
    ("smooth2d_coarse", lambda: fields.blobs2d(400, 600, nblobs=20, noise=0.0, seed=3), 1e-2),
]


@pytest.mark.parametrize("name,make,eb", SMOOTH, ids=[c[0] for c in SMOOTH])
def test_restore_smooth_fields(name, make, eb):
    """Smooth fields with coarse bounds have large flat cells (plateaus of many points);
    the repair still removes every mismatch and never leaves the cells."""
    data = make()
    dec, y, types, info = _pipeline(data, eb)
    want = oracle.classify(data).ravel()
    got = oracle.classify(y).ravel()
    raw = oracle.classify(dec.reshape(data.shape)).ravel()
    after = np.count_nonzero(got != want)
    print(f"{name}: cps={np.count_nonzero(want)} raw={np.count_nonzero(raw != want)} after={after}")
    assert after == info.mismatches()
    assert after == 0
    _check_cells(data, dec, y, eb)
    _check_extrema_order(data, dec, y, types, eb, info, slack=8)


def test_verify_counts_match_oracle():
    import torch
    data = fields.blobs2d(200, 300, seed=9, noise=0.02)
    types = torch.from_numpy(oracle.classify(data).ravel()).cuda()
    other = fields.blobs2d(200, 300, seed=10, noise=0.02)
    counts = tz.topo_verify(torch.from_numpy(other).cuda(), types, data.shape)
    a, b = oracle.classify(data).ravel(), oracle.classify(other).ravel()
    assert counts == (int(np.count_nonzero((a != 0) & (b == 0))), int(np.count_nonzero((a == 0) & (b != 0))),
                      int(np.count_nonzero((a != 0) & (b != 0) & (a != b))))


@pytest.mark.parametrize("name,make,eb", [NOISY[0], NOISY[3]], ids=["blobs2d", "fbm3d"])
def test_restore_out_of_place_matches_in_place(name, make, eb):
    """Writing the result to a separate buffer gives the same values, bit for bit, as
    restoring in place; two runs give the same result."""
    import torch
    data = make()
    f = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    info, buf = tz.topo_compress(f, eb, 64)
    out, types = tz.topo_decompress(buf, data.size, eb, 64)
    sp = tz.sortpos_decompress(buf, info.ext_count, 64) if info.ext_count else None
    dec = out.clone()
    sep = torch.empty_like(dec)
    r1 = tz.topo_restore(dec, types, data.shape, eb, sp=sp, n_ext=info.ext_count, out=sep)
    np.testing.assert_array_equal(dec.cpu().numpy(), out.cpu().numpy())  # input untouched
    r2 = tz.topo_restore(out, types, data.shape, eb, sp=sp, n_ext=info.ext_count)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sep.cpu().numpy(), out.cpu().numpy())
    assert r1.mismatches() == r2.mismatches() == 0
    assert (r1.moves, r1.windows) == (r2.moves, r2.windows)


@pytest.mark.parametrize("make,eb", [(lambda: fields.fbm3d(40, 44, 48, seed=71), 3e-3),
                                     (lambda: fields.blobs2d(300, 310, seed=72, noise=0.05), 1e-3),
                                     (lambda: fields.fbm3d(64, 64, 64, seed=73, noise=0.2), 1e-2)],
                         ids=["fbm3d", "blobs2d", "fbm3d_noisy"])
def test_solve_one_launch_equals_phases(make, eb, monkeypatch):
    """The window solve in one dataflow launch (a tile waits for its neighbouring tiles of
    lower phases) gives the field of one launch per parity phase, bit for bit."""
    data = make()
    print(f"solve flow vs phases {data.shape} eb={eb}", flush=True)
    _, y_flow, _, info_flow = _pipeline(data, eb)
    monkeypatch.setenv("TSZP_SOLVE_PHASES", "1")
    _, y_ph, _, info_ph = _pipeline(data, eb)
    assert np.array_equal(y_flow.view(np.uint32), y_ph.view(np.uint32))
    assert info_flow.windows == info_ph.windows and info_flow.missing == info_ph.missing == 0


@pytest.mark.parametrize("make,eb", [(lambda: fields.fbm3d(40, 44, 48, seed=74), 3e-3),
                                     (lambda: fields.blobs2d(300, 310, seed=75, noise=0.05), 1e-3),
                                     (lambda: fields.fbm3d(64, 64, 64, seed=76, noise=0.2), 1e-2)],
                         ids=["fbm3d", "blobs2d", "fbm3d_noisy"])
def test_move_segments_equal_serial(make, eb, monkeypatch):
    """The local moves with sixteen lanes per point give the field of one lane per point,
    bit for bit (same candidates, the serial loop's choice among equal scores)."""
    data = make()
    print(f"move segments vs serial {data.shape} eb={eb}", flush=True)
    _, y_seg, _, info_seg = _pipeline(data, eb)
    monkeypatch.setenv("TSZP_MOVE_SERIAL", "1")
    _, y_ser, _, info_ser = _pipeline(data, eb)
    assert np.array_equal(y_seg.view(np.uint32), y_ser.view(np.uint32))
    assert info_seg.moves == info_ser.moves and info_seg.missing == info_ser.missing == 0
