"""Batched topology pipeline (tszp_topo_batch_*): many 2D / 3D fields of mixed sizes,
error bounds and block sizes through one launch per stage.  Every field's topology and
sort-position streams and side-cars equal the one-field call (tszp_topo_compress with one
chunk) byte for byte, and the oracle's; the decoded values, types and sort positions
equal the one-field decoders'; every field restores with zero missing, spurious or
changed critical points inside its cells; compress + decompress replayed from a
captured graph give the same bytes every time."""
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


def _mixed_fields():
    """Shapes and block sizes that reach every kernel class: the 16-byte classifying
    path in 2D and 3D, fastest axes that are not multiples of 4, generic block sizes,
    one-element last blocks, fields smaller than one tile and a field without extrema."""
    cases = [
        (fields.blobs2d(300, 410, seed=201, noise=0.02), 1e-3, 64),
        (fields.blobs2d(211, 333, seed=202, noise=0.02), 2e-3, 64),      # fastest axis odd
        (fields.fbm3d(40, 48, 56, seed=203), 3e-3, 64),
        (fields.fbm3d(33, 37, 41, seed=204), 3e-3, 32),                  # fastest axis odd
        (fields.blobs3d(30, 40, 52, seed=205, noise=2e-3), 2e-4, 128),
        (fields.blobs2d(97, 131, seed=206, noise=0.01), 1e-3, 7),        # generic geometry
        (fields.blobs2d(64, 65, seed=207, noise=0.01), 5e-3, 48),
        (fields.blobs2d(5, 7, seed=208, noise=0.05), 1e-3, 1),            # block size 1
        (fields.fbm3d(3, 3, 3, seed=209), 1e-2, 16),                      # one interior point
        (np.full((40, 50), 0.25, np.float32), 1e-3, 64),                  # no critical points
        (fields.blobs2d(257, 129, seed=210, noise=0.02), 1e-3, 257),      # largest block size
        (fields.ties2d(90, 97, seed=211), 1e-3, 16),                      # ties and -0.0
    ]
    return cases


def _one_field(torch, data, eb, bs):
    f = torch.from_numpy(np.ascontiguousarray(data)).cuda()
    info, buf = tz.topo_compress(f, eb, bs, T=1)
    out, types = tz.topo_decompress(buf, data.size, eb, bs)
    sp = tz.sortpos_decompress(buf, info.ext_count, bs) if info.ext_count else None
    torch.cuda.synchronize()
    return (info, buf.topo[:info.topo_bytes].cpu().numpy().copy(), buf.topo_blk.cpu().numpy().copy(),
            buf.sp[:info.sortpos_bytes].cpu().numpy().copy(), buf.sp_blk.cpu().numpy().copy(),
            out.cpu().numpy().copy(), types.cpu().numpy().copy(),
            sp.cpu().numpy().copy() if sp is not None else np.zeros(0, np.int32))


def _check_batch_against_one_field(torch, batch, cases, oracle_fields):
    res = batch.results()
    for i, (data, eb, bs) in enumerate(cases):
        info, rinfo = res[i]
        one = _one_field(torch, data, eb, bs)
        tp, tb, sp, sb = batch.streams[i]
        assert (info.topo_bytes, info.sortpos_bytes, info.cp_count, info.ext_count) == \
            (one[0].topo_bytes, one[0].sortpos_bytes, one[0].cp_count, one[0].ext_count), f"field {i}: sizes"
        assert np.array_equal(tp[:info.topo_bytes].cpu().numpy(), one[1]), f"field {i}: topology stream"
        assert np.array_equal(tb.cpu().numpy(), one[2]), f"field {i}: topology side-car"
        assert np.array_equal(sp[:info.sortpos_bytes].cpu().numpy(), one[3]), f"field {i}: sort-position stream"
        if info.ext_count:
            nb = tz.lib().tszp_blocks(info.ext_count, bs) + 1
            assert np.array_equal(sb.cpu().numpy()[:nb], one[4][:nb]), f"field {i}: sort-position side-car"
        assert np.array_equal(batch.copy(i, "decoded").cpu().numpy().view(np.uint32), one[5].view(np.uint32))
        assert np.array_equal(batch.copy(i, "types").cpu().numpy(), one[6])
        assert np.array_equal(batch.copy(i, "sortpos", info.ext_count).cpu().numpy(), one[7])
        if i in oracle_fields:
            cps = oracle.find_cps_2d(data, eb) if data.ndim == 2 else oracle.find_cps_3d(data, eb)
            oracle.sort_positions(cps, data)
            types = oracle.mark_types(cps, data.shape)
            assert np.array_equal(one[1], oracle.topo_compress(data, eb, bs, types)), f"field {i}: oracle topology"
            spv = oracle.extrema_sort_positions(cps)
            if spv.size:
                assert np.array_equal(one[3], oracle.sortpos_compress(spv, bs)), f"field {i}: oracle sort positions"
    return res


def _check_restored(batch, cases, res):
    for i, (data, eb, bs) in enumerate(cases):
        y = batch.copy(i, "restored").cpu().numpy().reshape(data.shape)
        dec = batch.copy(i, "decoded").cpu().numpy()
        want = oracle.classify(data).ravel()
        got = oracle.classify(y).ravel()
        bad = int(np.count_nonzero(got != want))
        info, rinfo = res[i]
        assert rinfo.mismatches() == bad, f"field {i}: counts {rinfo.mismatches()} vs oracle {bad}"
        assert bad == 0, f"field {i} {data.shape}: {bad} mismatches"
        # cells: [d, d + 0.9 eb] above zero, [d - 0.9 eb, d] below, |y| <= eb 2^-16 at zero
        yv, d = y.ravel().astype(np.float64), dec.astype(np.float64)
        ebf = float(np.float32(eb))
        slack = np.abs(d) * 2.0 ** -22
        pos, neg, zero = d > 0, d < 0, d == 0
        assert np.all(yv[pos] >= d[pos]) and np.all(yv[pos] <= d[pos] + 0.9 * ebf * (1 + 1e-5) + slack[pos])
        assert np.all(yv[neg] <= d[neg]) and np.all(yv[neg] >= d[neg] - 0.9 * ebf * (1 + 1e-5) - slack[neg])
        assert np.all(np.abs(yv[zero]) <= ebf * 2.0 ** -16)
        assert rinfo.max_dev_over_eb <= 0.9 * (1 + 1e-5)


def test_topo_batch_mixed_fields_equal_one_field_calls():
    torch = _torch()
    cases = _mixed_fields()
    ts = [torch.from_numpy(np.ascontiguousarray(d)).cuda() for d, _, _ in cases]
    batch = tz.TopoBatch(ts, [c[1] for c in cases], [c[2] for c in cases])
    print(f"topo batch: {len(cases)} fields", flush=True)
    batch.compress()
    batch.decompress()
    res = _check_batch_against_one_field(torch, batch, cases, oracle_fields=set(range(len(cases))))
    batch.restore()
    res = batch.results()
    _check_restored(batch, cases, res)
    # a second run over the same batch gives the same bytes
    first = [batch.streams[i][0][:res[i][0].topo_bytes].clone() for i in range(len(cases))]
    batch.compress()
    batch.decompress()
    res2 = batch.results()
    for i in range(len(cases)):
        assert res2[i][0].topo_bytes == res[i][0].topo_bytes
        assert torch.equal(batch.streams[i][0][:res[i][0].topo_bytes], first[i])
    batch.close()


def test_topo_batch_graph_replays():
    """compress + decompress captured in a graph and replayed, with eager work between
    the replays, equal the eager run every time."""
    torch = _torch()
    cases = _mixed_fields()[:6]
    ts = [torch.from_numpy(np.ascontiguousarray(d)).cuda() for d, _, _ in cases]
    batch = tz.TopoBatch(ts, [c[1] for c in cases], [c[2] for c in cases])
    batch.compress()
    batch.decompress()
    res = batch.results()
    ref = [(batch.streams[i][0][:res[i][0].topo_bytes].clone(), batch.streams[i][2][:res[i][0].sortpos_bytes].clone(),
            batch.copy(i, "decoded")) for i in range(len(cases))]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            batch.compress()
            batch.decompress()
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        for st in batch.streams:  # clobber the outputs between replays
            st[0].fill_(0xA5)
            st[2].fill_(0x5A)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        res_r = batch.results()
        for i in range(len(cases)):
            assert res_r[i][0].topo_bytes == res[i][0].topo_bytes and res_r[i][0].sortpos_bytes == res[i][0].sortpos_bytes
            assert torch.equal(batch.streams[i][0][:res[i][0].topo_bytes], ref[i][0]), f"replay {rep} field {i}"
            assert torch.equal(batch.streams[i][2][:res[i][0].sortpos_bytes], ref[i][1]), f"replay {rep} field {i}"
            assert torch.equal(batch.copy(i, "decoded"), ref[i][2])
    del g
    batch.close()


def test_topo_batch_tuned_block_sizes_and_tile_counter(monkeypatch):
    torch = _torch()
    cases = _mixed_fields()[:5]
    ts = [torch.from_numpy(np.ascontiguousarray(d)).cuda() for d, _, _ in cases]
    ebs = [c[1] for c in cases]
    batch = tz.TopoBatch(ts, ebs, None)
    assert all(b in (16, 32, 64, 128) for b in batch.bss)
    tuned = [(d, eb, bs) for (d, eb, _), bs in zip(cases, batch.bss)]
    batch.compress()
    batch.decompress()
    _check_batch_against_one_field(torch, batch, tuned, oracle_fields={0, 3})
    batch.close()
    monkeypatch.setenv("TSZP_TILE_COUNTER", "1")
    batch = tz.TopoBatch(ts, ebs, [c[2] for c in cases])
    batch.compress()
    batch.decompress()
    _check_batch_against_one_field(torch, batch, cases, oracle_fields=set())
    batch.close()


def test_topo_batch_rejects_bad_fields():
    torch = _torch()
    f = torch.zeros((10, 10), dtype=torch.float32, device="cuda")
    with pytest.raises(tz.TszpError):
        tz.TopoBatch([f], [0.0], [64])
    with pytest.raises(tz.TszpError):
        tz.TopoBatch([f], [1e-3], [300])


def test_topo_batch_mixed_workload_256_fields():
    """The mixed-size workload at its real shapes: 128 2D fields (rows, cols log-uniform
    in [64, 2048]) and 128 3D fields (dims log-uniform in [32, 256]), blobs + noise, abs
    eb log-uniform in [1e-4, 1e-2], block sizes from the tuner.  Every field's streams and
    side-cars equal the one-field call's, the 16 smallest also the oracle's; after the
    batched restoration every field's classification (by the one-field classifier, and
    by the oracle on the 16 smallest) equals the original's and every value stays in
    its cell."""
    import workloads
    torch = _torch()
    wl = workloads.mixed_fields(20251203)
    ts = [f for f, _ in wl]
    ebs = [eb for _, eb in wl]
    print(f"mixed workload: {len(ts)} fields, {sum(t.numel() for t in ts)} elements", flush=True)
    batch = tz.TopoBatch(ts, ebs, None)
    batch.compress()
    batch.decompress()
    batch.restore()
    res = batch.results()
    small = set(sorted(range(len(ts)), key=lambda i: ts[i].numel())[:16])
    for i, f in enumerate(ts):
        info, rinfo = res[i]
        bs, eb = batch.bss[i], ebs[i]
        ref, buf = tz.topo_compress(f, eb, bs, T=1)
        assert (info.topo_bytes, info.sortpos_bytes, info.cp_count, info.ext_count) == \
            (ref.topo_bytes, ref.sortpos_bytes, ref.cp_count, ref.ext_count), f"field {i} {tuple(f.shape)}"
        tp, tb, sp, sb = batch.streams[i]
        assert torch.equal(tp[:info.topo_bytes], buf.topo[:ref.topo_bytes]), f"field {i}: topology stream"
        assert torch.equal(tb, buf.topo_blk), f"field {i}: topology side-car"
        assert torch.equal(sp[:info.sortpos_bytes], buf.sp[:ref.sortpos_bytes]), f"field {i}: sort-position stream"
        out, types = tz.topo_decompress(buf, f.numel(), eb, bs)
        assert torch.equal(batch.copy(i, "decoded"), out) and torch.equal(batch.copy(i, "types"), types)
        # restoration: zero mismatches by the one-field classifier, values in their cells
        y = batch.copy(i, "restored")
        ty, _ = tz.classify(y.view(f.shape))
        assert torch.equal(ty, types), f"field {i} {tuple(f.shape)}: critical points differ after restoration"
        ebf = float(np.float32(eb))
        # |y - d| <= 0.9 eb up to the float spacing of the largest |d| (d + 0.9 eb rounds)
        amax = float(f.abs().max())
        assert rinfo.mismatches() == 0 and rinfo.max_dev_over_eb <= 0.9 * (1 + 1e-5) + amax * 2.0 ** -22 / ebf
        err = float((y.double() - f.view(-1).double()).abs().max())
        assert err <= ebf * (1 + 2.0 ** -16) + float(f.abs().max()) * 2.0 ** -23, f"field {i}: error {err / ebf} eb"
        if i in small:
            data = f.cpu().numpy()
            cps = oracle.find_cps_2d(data, eb) if data.ndim == 2 else oracle.find_cps_3d(data, eb)
            oracle.sort_positions(cps, data)
            otypes = oracle.mark_types(cps, data.shape)
            assert np.array_equal(tp[:info.topo_bytes].cpu().numpy(), oracle.topo_compress(data, eb, bs, otypes,
                                                                                             adaptive=data.ndim == 2))
            spv = oracle.extrema_sort_positions(cps)
            if spv.size:
                assert np.array_equal(sp[:info.sortpos_bytes].cpu().numpy(), oracle.sortpos_compress(spv, bs))
            assert np.array_equal(oracle.classify(y.cpu().numpy().reshape(data.shape)), oracle.classify(data))
        del buf
    batch.close()
