"""GPU: batches of plain float fields (tszp_plain_compress_batch / _decompress_batch /
tszp_batch_tune) against the one-field native calls and the CPU oracle.

Each field of a batch must give the one-chunk plain stream of the oracle byte for byte
(SZp/src/szp_float.cc:2061-2292) and decode to the oracle's reconstruction
(SZp/src/szpd_float.cc:431-524), whatever the other fields of the batch are.
"""
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


def _eq(a, b, msg=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{msg}: shape {a.shape} vs {b.shape}"
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a.ravel() != b.ravel())
        raise AssertionError(f"{msg}: {bad.size} mismatches, first at {bad[:8]}")


def _mixed(seed, count):
    """Fields of uneven 2D / 3D sizes (flattened), error bounds and block sizes that cover
    every block-size class of the codec (1 element per lane up to 32 per lane)."""
    rng = np.random.default_rng(seed)
    sizes = [2, 7, 16, 17, 32, 48, 64, 65, 100, 128, 256, 1025]
    out = []
    for i in range(count):
        if i % 2:
            r, c = (int(v) for v in np.exp(rng.uniform(np.log(8), np.log(300), 2)))
            data = fields.blobs2d(r, c, seed=seed * 1000 + i).ravel()
        else:
            d = [int(v) for v in np.exp(rng.uniform(np.log(4), np.log(40), 3))]
            data = fields.blobs3d(*d, seed=seed * 1000 + i).ravel()
        eb = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        out.append((data, eb, sizes[i % len(sizes)]))
    return out


@pytest.mark.parametrize("seed,count", [(1, 1), (2, 12), (3, 40)])
def test_batch_matches_oracle_and_single_calls(seed, count):
    torch = _torch()
    cases = _mixed(seed, count)
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    ebs = [eb for _, eb, _ in cases]
    bss = [bs for _, _, bs in cases]
    sizes, streams, blks = tz.plain_compress_batch(dev, ebs, bss)
    outs = tz.plain_decompress_batch(streams, blks, [d.numel() for d in dev], ebs, bss)
    torch.cuda.synchronize()
    for i, (data, eb, bs) in enumerate(cases):
        ref = oracle.plain_compress(data, eb, bs, T=1)
        _eq(streams[i][:sizes[i]].cpu().numpy(), ref, f"field {i} (n {data.size}, bs {bs}) stream")
        nb, one, one_blk = tz.plain_compress(dev[i], eb, bs, T=1)
        assert nb == sizes[i]
        _eq(blks[i].cpu().numpy(), one_blk.cpu().numpy(), f"field {i} side-car")
        _eq(outs[i].cpu().numpy(), oracle.plain_decompress(ref[4:], data.size, eb, bs, T=1), f"field {i} decode")


def test_batch_large_and_small_fields_together():
    torch = _torch()
    big = tz.gen_fbm((128, 256, 256), seed=3, device="cuda").reshape(-1)
    small = [torch.from_numpy(fields.blobs2d(31, 45, seed=k).ravel()).cuda() for k in range(9)]
    dev = [small[0], big] + small[1:]
    ebs = [1e-3] * len(dev)
    bss = [64, 64, 32, 128, 64, 16, 64, 300, 64, 64]
    sizes, streams, blks = tz.plain_compress_batch(dev, ebs, bss)
    outs = tz.plain_decompress_batch(streams, blks, [d.numel() for d in dev], ebs, bss)
    for i, f in enumerate(dev):
        nb, one, one_blk = tz.plain_compress(f, ebs[i], bss[i], T=1)
        assert nb == sizes[i], f"field {i}"
        assert torch.equal(streams[i][:nb], one[:nb]), f"field {i} stream"
        assert torch.equal(blks[i], one_blk), f"field {i} side-car"
        dec = tz.plain_decompress(one, one_blk, f.numel(), ebs[i], bss[i], T=1)
        assert torch.equal(outs[i].view(torch.int32), dec.view(torch.int32)), f"field {i} decode"
        assert float((outs[i] - f).abs().max()) <= ebs[i] * 1.0001


def test_batch_aligned_and_unaligned_inputs():
    """Block sizes 32 and 64 on 16-byte aligned inputs take the 16-byte path, the same
    sizes on inputs 4 bytes off take the generic one; both give the oracle's bytes."""
    torch = _torch()
    base = [fields.blobs2d(37, 53 + k, seed=20 + k).ravel() for k in range(6)]
    dev, cases = [], []
    for k, data in enumerate(base):
        t = torch.from_numpy(data).cuda()
        if k % 2:
            t = t[1:]  # 4 bytes past an aligned allocation
            data = data[1:]
        dev.append(t)
        cases.append((data, 2e-3, 64 if k < 3 else 32))
    sizes, streams, blks = tz.plain_compress_batch(dev, [c[1] for c in cases], [c[2] for c in cases])
    for i, (data, eb, bs) in enumerate(cases):
        assert (dev[i].data_ptr() % 16 == 0) == (i % 2 == 0)
        _eq(streams[i][:sizes[i]].cpu().numpy(), oracle.plain_compress(data, eb, bs, T=1), f"field {i} stream")


def test_plan_matches_batch_calls_and_replays_in_a_graph():
    """A plan gives the batch calls' streams and reconstructions; its compress and
    decompress launches captured in a graph give them again on every replay, with new
    field contents written eagerly between replays."""
    torch = _torch()
    cases = _mixed(7, 20)
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    ebs = [eb for _, eb, _ in cases]
    bss = [bs for _, _, bs in cases]
    ns = [d.numel() for d in dev]

    def reference():
        sizes, streams, blks = tz.plain_compress_batch(dev, ebs, bss)
        return sizes, streams, blks, tz.plain_decompress_batch(streams, blks, ns, ebs, bss)

    def check(got_sizes, ref):
        sizes, streams, blks, outs = ref
        assert got_sizes == sizes
        for i in range(len(dev)):
            assert torch.equal(pstreams[i][:sizes[i]], streams[i][:sizes[i]]), f"field {i} stream"
            assert torch.equal(pblks[i], blks[i]), f"field {i} side-car"
            assert torch.equal(pouts[i].view(torch.int32), outs[i].view(torch.int32)), f"field {i} decode"

    ref = reference()
    pstreams = [torch.empty_like(x) for x in ref[1]]
    pblks = [torch.empty_like(x) for x in ref[2]]
    pouts = [torch.empty_like(d) for d in dev]
    plan = tz.BatchPlan(dev, ebs, bss, pstreams, pblks, pouts)
    plan.compress()
    plan.decompress()
    check(plan.results(), ref)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.compress()
        plan.decompress()
    for rep in range(3):
        for d in dev:
            d.mul_(1.0 + 0.25 * (rep + 1))
        torch.cuda.synchronize()
        g.replay()
        got = plan.results()
        check(got, reference())
    plan.close()


def _payload(data, eb, bs):
    return len(oracle.plain_compress(data, eb, bs, T=1)) - 12


@pytest.mark.parametrize("seed", [4, 5])
def test_tune_picks_the_smallest_stream(seed):
    """Fields of 128 k elements with k <= 512 are sampled whole, so the estimate is the
    payload itself: the pick must be the candidate whose oracle stream is smallest
    (64 on a tie, then the larger block size)."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(8):
        k = int(rng.integers(1, 513))
        smooth = fields.blobs2d(1, 128 * k, nblobs=4, noise=float(rng.choice([0.0, 1e-4, 1e-2])),
                                seed=seed * 50 + i).ravel()
        cases.append((smooth, float(np.exp(rng.uniform(np.log(1e-5), np.log(1e-2))))))
    dev = [torch.from_numpy(d).cuda() for d, _ in cases]
    picks = tz.batch_tune(dev, [eb for _, eb in cases])
    for (data, eb), got in zip(cases, picks):
        pay = {bs: _payload(data, eb, bs) for bs in (16, 32, 64, 128)}
        best = 64
        for bs in (128, 64, 32, 16):
            if pay[bs] < pay[best]:
                best = bs
        assert got == best, f"n {data.size} eb {eb}: picked {got}, payloads {pay}"


def test_tune_keeps_given_sizes_and_handles_short_fields():
    torch = _torch()
    a = torch.from_numpy(fields.blobs2d(5, 20, seed=1).ravel()).cuda()     # 100 < 128: not sampled
    b = torch.from_numpy(fields.blobs2d(64, 64, seed=2).ravel()).cuda()
    tab = tz._batch_table([a, b, b], [1e-3] * 3, [0, 0, 100])
    tz._check(tz.lib().tszp_batch_tune(tab, 3, tz._stream()))
    assert tab[0].block_size == 64
    assert tab[1].block_size in (16, 32, 64, 128)
    assert tab[2].block_size == 100


def test_tuned_batch_round_trip():
    torch = _torch()
    cases = _mixed(6, 16)
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    ebs = [eb for _, eb, _ in cases]
    bss = tz.batch_tune(dev, ebs)
    sizes, streams, blks = tz.plain_compress_batch(dev, ebs, bss)
    outs = tz.plain_decompress_batch(streams, blks, [d.numel() for d in dev], ebs, bss)
    for i, (data, eb, _) in enumerate(cases):
        ref = oracle.plain_compress(data, eb, bss[i], T=1)
        _eq(streams[i][:sizes[i]].cpu().numpy(), ref, f"field {i} stream")
        _eq(outs[i].cpu().numpy(), oracle.plain_decompress(ref[4:], data.size, eb, bss[i], T=1), f"field {i} decode")


def test_batch_rejects_bad_fields():
    torch = _torch()
    f = torch.from_numpy(fields.blobs2d(20, 30, seed=3).ravel()).cuda()
    s = torch.empty(tz.lib().tszp_plain_bound(f.numel(), 64, 1) + 8, dtype=torch.uint8, device="cuda")
    blk = torch.empty(tz.lib().tszp_blocks(f.numel(), 64) + 1, dtype=torch.int32, device="cuda")
    L = tz.lib()
    with pytest.raises(tz.TszpError, match="blockSize 0"):
        tz.plain_compress_batch([f], [1e-3], [0], [s], [blk])
    with pytest.raises(tz.TszpError, match="error bound"):
        tz.plain_compress_batch([f], [0.0], [64], [s], [blk])
    with pytest.raises(tz.TszpError, match="4-byte aligned"):
        tz.plain_compress_batch([f], [1e-3], [64], [s[1:]], [blk])
    with pytest.raises(tz.TszpError, match="field 1: device error bits"):  # capacity
        tz.plain_compress_batch([f, f], [1e-3, 1e-3], [64, 64], [s, s[:64]], [blk, blk.clone()])
    with pytest.raises(tz.TszpError, match="no fields"):
        tz._check(L.tszp_plain_compress_batch(None, 0, tz._stream()))


def test_batch_decode_rejects_corrupt_side_car():
    torch = _torch()
    # 10000 elements at bs 64: 157 blocks, tiles of 64 blocks start at blocks 0, 64, 128
    a = torch.from_numpy(fields.blobs2d(100, 100, seed=8).ravel()).cuda()
    sizes, streams, blks = tz.plain_compress_batch([a, a], [1e-3, 1e-3], [64, 64])
    bad = blks[1].clone()
    bad[128] = bad[64] - 4  # the second tile ends before it starts
    with pytest.raises(tz.TszpError, match="field 1: device error bits"):
        tz.plain_decompress_batch(streams, [blks[0], bad], [a.numel()] * 2, [1e-3] * 2, [64, 64])


def test_plan_rejects_bad_fields():
    torch = _torch()
    f = torch.from_numpy(fields.blobs2d(20, 30, seed=5).ravel()).cuda()
    s = torch.empty(tz.lib().tszp_plain_bound(f.numel(), 64, 1), dtype=torch.uint8, device="cuda")
    blk = torch.empty(tz.lib().tszp_blocks(f.numel(), 64) + 1, dtype=torch.int32, device="cuda")
    out = torch.empty_like(f)
    with pytest.raises(tz.TszpError, match="blockSize 0"):
        tz.BatchPlan([f], [1e-3], [0], [s], [blk], [out])
    with pytest.raises(tz.TszpError, match="stream capacity"):
        tz.BatchPlan([f], [1e-3], [64], [s[:8]], [blk], [out])
    with pytest.raises(ValueError, match="one output per field"):
        tz.BatchPlan([f], [1e-3], [64], [s], [blk], [out[:10]])
    plan = tz.BatchPlan([f], [1e-3], [64], [s], [blk], [out])
    plan.compress()
    plan.decompress()
    nb, s1, _ = tz.plain_compress(f, 1e-3, 64, T=1)
    assert plan.results() == [nb] and torch.equal(s[:nb], s1[:nb])
    plan.close()


def test_batch_of_tiny_fields():
    """Fields of 1 to 17 elements (one-element last blocks, fields shorter than a block)
    next to a larger one: the oracle's streams and reconstructions."""
    torch = _torch()
    rng = np.random.default_rng(12)
    sizes = [1, 2, 3, 15, 16, 17, 5, 1, 9]
    bss = [2, 64, 16, 16, 16, 16, 3, 128, 8]
    cases = [(rng.standard_normal(n).astype(np.float32), 1e-2, bs) for n, bs in zip(sizes, bss)]
    cases.append((fields.blobs2d(40, 41, seed=13).ravel(), 1e-3, 64))
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    ebs = [eb for _, eb, _ in cases]
    bsl = [bs for _, _, bs in cases]
    got, streams, blks = tz.plain_compress_batch(dev, ebs, bsl)
    outs = tz.plain_decompress_batch(streams, blks, [d.numel() for d in dev], ebs, bsl)
    for i, (data, eb, bs) in enumerate(cases):
        ref = oracle.plain_compress(data, eb, bs, T=1)
        _eq(streams[i][:got[i]].cpu().numpy(), ref, f"field {i} (n {data.size}, bs {bs})")
        _eq(outs[i].cpu().numpy(), oracle.plain_decompress(ref[4:], data.size, eb, bs, T=1), f"field {i} decode")
