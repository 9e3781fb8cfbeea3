"""The look-back kernels take their tile ids from workgroup ids; TSZP_TILE_COUNTER=1
hands tiles out from an atomic counter instead (the documented fallback for a
dispatcher that would not start the workgroups of an XCD in order).  Both ways must
give the same bytes and values: topology compress + decompress + restore, the radix
sort positions, the element-partitioned codecs, the classification and the plain
batch codec."""
import numpy as np
import pytest

import fields
import oracle
import toposzp_hip as tz

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _topo_run(data, eb, bs):
    torch = _torch()
    f = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    info, buf = tz.topo_compress(f, eb, bs)
    topo = buf.topo[:info.topo_bytes].cpu().numpy().copy()
    sp = buf.sp[:info.sortpos_bytes].cpu().numpy().copy()
    out, types = tz.topo_decompress(buf, data.size, eb, bs)
    spd = tz.sortpos_decompress(buf, info.ext_count, bs)
    y = torch.empty_like(out)
    r = tz.topo_restore(out, types, data.shape, eb, sp=spd, n_ext=info.ext_count, out=y)
    torch.cuda.synchronize()
    return topo, sp, out.cpu().numpy().copy(), types.cpu().numpy().copy(), y.cpu().numpy().copy(), r


CASES = [
    ("fbm3d", lambda: fields.fbm3d(48, 64, 80, seed=91), 3e-3, 64),
    ("blobs2d", lambda: fields.blobs2d(300, 410, seed=92, noise=0.02), 1e-3, 32),
    ("blobs2d_generic", lambda: fields.blobs2d(211, 333, seed=93, noise=0.02), 1e-3, 48),
]


@pytest.mark.parametrize("name,make,eb,bs", CASES, ids=[c[0] for c in CASES])
def test_topology_pipeline_with_tile_counter(name, make, eb, bs, monkeypatch):
    data = make()
    print(f"tile counter: {name} {data.shape} eb={eb} bs={bs}", flush=True)
    ref = _topo_run(data, eb, bs)
    monkeypatch.setenv("TSZP_TILE_COUNTER", "1")
    got = _topo_run(data, eb, bs)
    for a, b in zip(ref[:5], got[:5]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert got[5].mismatches() == 0
    # and the oracle's streams
    cps = oracle.find_cps_2d(data, eb) if data.ndim == 2 else oracle.find_cps_3d(data, eb)
    oracle.sort_positions(cps, data)
    types = oracle.mark_types(cps, data.shape)
    assert np.array_equal(got[0], oracle.topo_compress(data, eb, bs, types))
    assert np.array_equal(got[1], oracle.sortpos_compress(oracle.extrema_sort_positions(cps), bs))


def test_radix_sort_positions_with_tile_counter(monkeypatch):
    data = fields.fbm3d(40, 48, 56, seed=94)
    eb, bs = 3e-3, 64
    ref = _topo_run(data, eb, bs)
    monkeypatch.setenv("TSZP_SORTPOS", "radix")
    monkeypatch.setenv("TSZP_TILE_COUNTER", "1")
    got = _topo_run(data, eb, bs)
    assert np.array_equal(ref[1], got[1])


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_element_codecs_with_tile_counter(dtype, monkeypatch):
    torch = _torch()
    rng = np.random.default_rng(95)
    x = np.cumsum(rng.normal(0, 0.01, 300_001)).astype(np.float32 if dtype == "f32" else np.float64)
    f = torch.from_numpy(x).cuda()
    eb, bs, T = 1e-3, 64, 7

    def run():
        nb, s, blk = tz.threadblock_compress(f, eb, bs, T)
        dec = tz.threadblock_decompress(s, blk, x.size, eb, bs, T, dtype=f.dtype)
        out = [s[:nb].cpu().numpy().copy(), dec.cpu().numpy().copy()]
        if dtype == "f64":
            nb2, s2, blk2 = tz.plain_compress_f64(f, eb, bs, T)
            out += [s2[:nb2].cpu().numpy().copy(), tz.plain_decompress_f64(s2, blk2, x.size, eb, bs, T).cpu().numpy()]
        return out

    ref = run()
    monkeypatch.setenv("TSZP_TILE_COUNTER", "1")
    got = run()
    for a, b in zip(ref, got):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_classify_and_plain_batch_with_tile_counter(monkeypatch):
    torch = _torch()
    data = fields.blobs2d(257, 300, seed=96, noise=0.02)
    f = torch.from_numpy(data).cuda()
    fl = [torch.from_numpy(fields.blobs2d(64 + 37 * i, 90 + 11 * i, seed=97 + i).ravel()).cuda() for i in range(9)]
    ebs = [1e-3 * (1 + i % 3) for i in range(9)]
    bss = [16, 32, 64, 128, 48, 7, 256, 64, 32]

    def run():
        t, cnt = tz.classify(f)
        sizes, streams, blks = tz.plain_compress_batch(fl, ebs, bss)
        return t.cpu().numpy().copy(), cnt, [s[:n].cpu().numpy().copy() for s, n in zip(streams, sizes)]

    ref = run()
    monkeypatch.setenv("TSZP_TILE_COUNTER", "1")
    got = run()
    assert np.array_equal(ref[0], got[0]) and ref[1] == got[1]
    assert all(np.array_equal(a, b) for a, b in zip(ref[2], got[2]))
    assert np.array_equal(got[0], oracle.classify(data).ravel())
