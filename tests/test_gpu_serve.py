"""GPU: the request loop (tszp_server_*) over the batch codec.

Every request's result must be the one-field call's result whether it ran alone or
in a mix with others: the same stream bytes (the oracle's one-chunk plain stream) and
the same reconstruction.  Admission keeps the streams inside the HBM budget.
"""
import threading

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


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 3 == 0:
            d = [int(v) for v in np.exp(rng.uniform(np.log(6), np.log(48), 3))]
            data = fields.blobs3d(*d, seed=seed * 100 + i).ravel()
        else:
            r, c = (int(v) for v in np.exp(rng.uniform(np.log(16), np.log(400), 2)))
            data = fields.blobs2d(r, c, seed=seed * 100 + i).ravel()
        eb = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        out.append((data, eb, int(rng.choice([16, 32, 64, 128, 37]))))
    return out


def test_results_same_alone_and_in_the_mix():
    torch = _torch()
    cases = _cases(1, 36)
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    alone = []
    for f, (data, eb, bs) in zip(dev, cases):
        nb, s, blk = tz.plain_compress(f, eb, bs, T=1)
        alone.append((s[:nb].clone(), tz.plain_decompress(s, blk, f.numel(), eb, bs, T=1)))
    torch.cuda.synchronize()
    srv = tz.Server(1 << 30, max_batch=64)
    outs = [torch.empty_like(f) for f in dev]
    tickets = [None] * len(cases)
    errors = []

    def client(part):  # submit a compression and at once its decompression
        try:
            rng = np.random.default_rng(part)
            for i in range(part, len(cases), 3):
                c = srv.compress(dev[i], cases[i][1], cases[i][2])
                d = srv.decompress(c, outs[i])
                tickets[i] = (c, d)
                if rng.random() < 0.5:
                    threading.Event().wait(float(rng.uniform(0, 2e-3)))
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=client, args=(k,)) for k in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i, (data, eb, bs) in enumerate(cases):
        c, d = tickets[i]
        ic = srv.wait(c)
        srv.wait(d)
        got = srv.stream_copy(c)
        assert ic.stream_bytes == got.numel() == alone[i][0].numel(), f"request {i}"
        assert torch.equal(got, alone[i][0]), f"request {i}: stream differs from the one-field call"
        if i % 6 == 0:
            np.testing.assert_array_equal(got.cpu().numpy(), oracle.plain_compress(data, eb, bs, T=1))
        assert torch.equal(outs[i].view(torch.int32), alone[i][1].view(torch.int32)), f"request {i}: decode"
        srv.release(d)
        srv.release(c)
    st = srv.stats()
    assert st["requests"] == 2 * len(cases) and st["queued"] == 0 and st["hbm_in_use"] == 0
    assert st["batches"] < 2 * len(cases)  # requests shared batches
    srv.close()


def test_admission_keeps_streams_inside_the_budget():
    torch = _torch()
    cases = _cases(2, 16)
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    need = max(tz.lib().tszp_plain_bound(f.numel(), bs, 1) + 4 * (tz.lib().tszp_blocks(f.numel(), bs) + 1)
               for f, (_, _, bs) in zip(dev, cases))
    budget = 3 * (need + 1024)  # room for about three streams at a time
    srv = tz.Server(budget, max_batch=8)
    tickets = [srv.compress(f, eb, bs) for f, (_, eb, bs) in zip(dev, cases)]
    for i, t in enumerate(tickets):  # consume in order, releasing as we go
        info = srv.wait(t)
        nb, s, _ = tz.plain_compress(dev[i], cases[i][1], cases[i][2], T=1)
        assert info.stream_bytes == nb
        assert torch.equal(srv.stream_copy(t), s[:nb]), f"request {i}"
        srv.release(t)
    st = srv.stats()
    assert st["hbm_peak"] <= st["hbm_budget"] and st["hbm_in_use"] == 0
    assert st["batches"] >= 2
    srv.close()


def test_request_larger_than_the_budget_fails_alone():
    torch = _torch()
    srv = tz.Server(1 << 16, max_batch=4)
    big = torch.zeros(100000, device="cuda")
    small = torch.from_numpy(fields.blobs2d(20, 20, seed=1).ravel()).cuda()
    tb = srv.compress(big, 1e-3, 64)
    ts = srv.compress(small, 1e-3, 64)
    with pytest.raises(tz.TszpError, match="larger than the server's HBM budget"):
        srv.wait(tb)
    assert srv.wait(ts).stream_bytes > 0
    srv.release(tb)
    srv.release(ts)
    srv.close()


def test_tuned_block_size_matches_the_tuner():
    """block_size 0: the worker tunes the batch; the stream is the one-field call's at the
    block size tszp_batch_tune picks for the field."""
    torch = _torch()
    cases = _cases(3, 6)
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    ebs = [eb for _, eb, _ in cases]
    picks = tz.batch_tune(dev, ebs)
    srv = tz.Server(1 << 28, max_batch=16)
    tickets = [srv.compress(f, eb, 0) for f, eb in zip(dev, ebs)]
    for i, t in enumerate(tickets):
        srv.wait(t)
        _, _, nb, bs = srv.stream(t)
        assert bs == picks[i]
        n1, s, _ = tz.plain_compress(dev[i], ebs[i], bs, T=1)
        assert nb == n1 and torch.equal(srv.stream_copy(t), s[:n1]), f"request {i}"
        srv.release(t)
    srv.close()


def test_bad_requests_are_rejected_at_submission():
    torch = _torch()
    srv = tz.Server(1 << 20, max_batch=4)
    f = torch.ones(1000, device="cuda")
    with pytest.raises(tz.TszpError, match="blockSize"):
        srv.compress(f, 1e-3, 1)
    with pytest.raises(tz.TszpError, match="error bound"):
        srv.compress(f, 0.0, 64)
    with pytest.raises(tz.TszpError, match="not a live compress ticket"):
        srv.decompress(12345, f)
    t = srv.compress(f, 1e-3, 64)
    srv.wait(t)
    srv.release(t)
    with pytest.raises(tz.TszpError, match="not a live compress ticket"):
        srv.decompress(t, f)
    srv.close()


def test_release_while_its_decompression_is_queued():
    """Releasing a compress ticket whose stream a queued decompression still reads defers
    the free until that decompression has run; the result is the one-field call's."""
    torch = _torch()
    data = fields.blobs2d(300, 400, seed=9).ravel()
    f = torch.from_numpy(data).cuda()
    srv = tz.Server(1 << 26, max_batch=4)
    c = srv.compress(f, 1e-3, 64)
    srv.wait(c)
    out = torch.empty_like(f)
    d = srv.decompress(c, out)
    srv.release(c)  # the stream stays until d has read it
    with pytest.raises(tz.TszpError, match="not a live compress ticket"):
        srv.decompress(c, out)
    srv.wait(d)
    nb, s1, blk = tz.plain_compress(f, 1e-3, 64, T=1)
    ref = tz.plain_decompress(s1, blk, f.numel(), 1e-3, 64, T=1)
    assert torch.equal(out.view(torch.int32), ref.view(torch.int32))
    srv.release(d)
    assert srv.stats()["hbm_in_use"] == 0
    srv.close()


def test_many_clients_with_releases():
    """Eight threads each submit compressions and decompressions, wait and release as they
    go; every result equals the one-field call's and the arena ends empty."""
    torch = _torch()
    cases = _cases(4, 24)
    dev = [torch.from_numpy(d).cuda() for d, _, _ in cases]
    ref = []
    for f, (_, eb, bs) in zip(dev, cases):
        nb, s, blk = tz.plain_compress(f, eb, bs, T=1)
        ref.append((s[:nb].clone(), tz.plain_decompress(s, blk, f.numel(), eb, bs, T=1)))
    torch.cuda.synchronize()
    srv = tz.Server(1 << 28, max_batch=32)
    errors = []

    def client(k):
        try:
            rng = np.random.default_rng(100 + k)
            for _ in range(12):
                i = int(rng.integers(0, len(cases)))
                out = torch.empty_like(dev[i])
                c = srv.compress(dev[i], cases[i][1], cases[i][2])
                d = srv.decompress(c, out)
                srv.wait(c)
                assert torch.equal(srv.stream_copy(c), ref[i][0]), f"client {k}: stream of field {i}"
                srv.release(c)  # deferred while d may still read it
                srv.wait(d)
                assert torch.equal(out.view(torch.int32), ref[i][1].view(torch.int32)), f"client {k}: field {i}"
                srv.release(d)
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=client, args=(k,)) for k in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    st = srv.stats()
    assert st["requests"] == 8 * 12 * 2 and st["hbm_in_use"] == 0 and st["queued"] == 0
    srv.close()
