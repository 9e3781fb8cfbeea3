"""Request loop of the topology pipeline (tszp_topo_server_*).  Compress requests give
the one-field call's streams byte for byte (tszp_topo_compress with one chunk);
decompress requests give the one-field decoders' values and types, and with the
restoration a field whose classification equals the original's with every value in its
cell.  A Poisson-like stream of mixed requests from several threads, with a 512^3 field
among them, gives every request the same bytes as the same request served alone; lanes
run batches side by side; failures stay with their requests."""
import threading

import numpy as np
import pytest

import fields
import toposzp_hip as tz

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _cases(torch):
    specs = [
        (fields.blobs2d(300, 410, seed=301, noise=0.02), 1e-3, 64),
        (fields.fbm3d(40, 48, 56, seed=302), 3e-3, 0),          # tuned block size
        (fields.blobs3d(30, 40, 52, seed=303, noise=2e-3), 2e-4, 128),
        (fields.blobs2d(97, 131, seed=304, noise=0.01), 1e-3, 7),
        (fields.blobs2d(211, 333, seed=305, noise=0.02), 2e-3, 32),
        (np.full((40, 50), 0.25, np.float32), 1e-3, 64),          # no critical points
    ]
    return [(torch.from_numpy(np.ascontiguousarray(d)).cuda(), eb, bs) for d, eb, bs in specs]


def _alone(torch, f, eb, bs):
    info, buf = tz.topo_compress(f, eb, bs, T=1)
    out, types = tz.topo_decompress(buf, f.numel(), eb, bs)
    torch.cuda.synchronize()
    return info, buf.topo[:info.topo_bytes].clone(), buf.sp[:info.sortpos_bytes].clone(), out, types


def _check_restored(torch, f, eb, y, types):
    ty, _ = tz.classify(y.view(f.shape))
    assert torch.equal(ty, types), f"{tuple(f.shape)}: critical points differ after restoration"
    ebf = float(np.float32(eb))
    err = float((y.double() - f.view(-1).double()).abs().max())
    assert err <= ebf * (1 + 2.0 ** -16) + float(f.abs().max()) * 2.0 ** -23, f"{tuple(f.shape)}: error {err / ebf} eb"


def test_topo_server_requests_equal_one_field_calls():
    torch = _torch()
    cases = _cases(torch)
    torch.cuda.synchronize()
    srv = tz.TopoServer(8 << 30, max_batch=8, max_batch_elements=1 << 22, lanes=2)
    try:
        tickets = [srv.compress(f, eb, bs) for f, eb, bs in cases]
        for t in tickets:
            srv.wait(t)
        outs, dec, tys, dts = [], [], [], []
        for (f, eb, bs), t in zip(cases, tickets):
            info, _, chosen = srv.result(t)
            one = _alone(torch, f, eb, chosen)
            assert (info.topo_bytes, info.sortpos_bytes, info.cp_count, info.ext_count) == \
                (one[0].topo_bytes, one[0].sortpos_bytes, one[0].cp_count, one[0].ext_count)
            tp, sp = srv.streams(t)
            assert torch.equal(tp, one[1]) and torch.equal(sp, one[2]), f"{tuple(f.shape)}: streams"
            if bs:
                assert chosen == bs
            y = torch.empty(f.numel(), dtype=torch.float32, device="cuda")
            ty = torch.empty(f.numel(), dtype=torch.uint8, device="cuda")
            x = torch.empty(f.numel(), dtype=torch.float32, device="cuda")
            outs.append((y, ty, one))
            dec.append(x)
            dts.append(srv.decompress(t, y, ty, restore=True))
            dts.append(srv.decompress(t, x, None, restore=False))
        for t in dts:
            srv.wait(t)
        for k, ((f, eb, bs), (y, ty, one)) in enumerate(zip(cases, outs)):
            assert torch.equal(ty, one[4]), f"field {k}: types"
            assert torch.equal(dec[k], one[3]), f"field {k}: decoded values"
            _, rinfo, _ = srv.result(dts[2 * k])
            assert rinfo.mismatches() == 0
            _check_restored(torch, f, eb, y, one[4])
        for t in dts + tickets:
            srv.release(t)
        st = srv.stats()
        assert st["requests"] == 3 * len(cases) and st["queued"] == 0
        assert st["hbm_peak"] <= st["hbm_budget"]
    finally:
        srv.close()


def test_topo_server_stream_alone_and_in_the_mix():
    """Requests from three threads with random gaps (about half compressions, half
    decompressions of fields compressed earlier, one 512^3 field among them) give the
    bytes and decoded values of the same requests served alone, and restored fields that
    keep every critical point; several batches ran."""
    import time
    torch = _torch()
    rng = np.random.default_rng(17)
    pool = _cases(torch)
    big = tz.gen_fbm((512, 512, 512), seed=20251203, octaves=6, noise=0.05, device="cuda")
    beb, _ = tz.rel_to_abs(big, 1e-3)
    pool.append((big, beb, 64))
    torch.cuda.synchronize()
    alone = {}
    for i, (f, eb, bs) in enumerate(pool):
        if bs == 0:
            continue
        alone[i] = _alone(torch, f, eb, bs)
    srv = tz.TopoServer(64 << 30, max_batch=16, max_batch_elements=(1 << 27) + (1 << 22), lanes=2)
    lock = threading.Lock()
    done = []
    errors = []

    def client(seed):
        r = np.random.default_rng(seed)
        mine = []
        try:
            for _ in range(10):
                i = int(r.integers(0, len(pool) - 1)) if r.random() > 0.05 else len(pool) - 1
                f, eb, bs = pool[i]
                if bs == 0:
                    continue
                t = srv.compress(f, eb, bs)
                mine.append(("c", i, t, None))
                if r.random() < 0.6:
                    y = torch.empty(f.numel(), dtype=torch.float32, device="cuda")
                    ty = torch.empty(f.numel(), dtype=torch.uint8, device="cuda")
                    mine.append(("d", i, srv.decompress(t, y, ty, restore=True), (y, ty)))
                time.sleep(float(r.exponential(2e-4)))
        except Exception as e:  # noqa: BLE001 -- reported by the main thread
            errors.append(e)
        with lock:
            done.extend(mine)

    ths = [threading.Thread(target=client, args=(s,)) for s in (1, 2, 3)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors, errors
    # the 512^3 field at least once
    if not any(i == len(pool) - 1 for _, i, _, _ in done):
        f, eb, bs = pool[-1]
        t = srv.compress(f, eb, bs)
        y = torch.empty(f.numel(), dtype=torch.float32, device="cuda")
        ty = torch.empty(f.numel(), dtype=torch.uint8, device="cuda")
        done += [("c", len(pool) - 1, t, None), ("d", len(pool) - 1, srv.decompress(t, y, ty), (y, ty))]
    for kind, i, t, _ in done:
        srv.wait(t)
    for kind, i, t, outp in done:
        f, eb, bs = pool[i]
        one = alone[i]
        if kind == "c":
            info, _, _ = srv.result(t)
            assert (info.topo_bytes, info.sortpos_bytes) == (one[0].topo_bytes, one[0].sortpos_bytes)
            tp, sp = srv.streams(t)
            assert torch.equal(tp, one[1]) and torch.equal(sp, one[2]), f"pool {i}: streams in the mix"
        else:
            y, ty = outp
            assert torch.equal(ty, one[4]), f"pool {i}: types in the mix"
            _, rinfo, _ = srv.result(t)
            assert rinfo.mismatches() == 0
            _check_restored(torch, f, eb, y, one[4])
    st = srv.stats()
    assert st["batches"] >= 2 and st["requests"] == len(done)
    for kind, i, t, _ in done:
        if kind == "d":
            srv.release(t)
    for kind, i, t, _ in done:
        if kind == "c":
            srv.release(t)
    srv.close()


def test_topo_server_failures_stay_with_their_requests():
    torch = _torch()
    f = torch.from_numpy(fields.blobs2d(200, 300, seed=311, noise=0.02)).cuda()
    big = torch.from_numpy(fields.blobs2d(600, 800, seed=312, noise=0.02)).cuda()
    torch.cuda.synchronize()
    srv = tz.TopoServer(4 << 30, max_batch=4, max_batch_elements=200 * 300, lanes=2)
    try:
        t_big = srv.compress(big, 1e-3, 64)   # larger than a batch: fails alone
        t_ok = srv.compress(f, 1e-3, 64)
        with pytest.raises(tz.TszpError, match="larger than the server's batch"):
            srv.wait(t_big)
        srv.wait(t_ok)
        with pytest.raises(tz.TszpError):
            srv.decompress(12345, torch.empty(10, device="cuda"))
        with pytest.raises(tz.TszpError):
            srv.compress(f.view(-1), 1e-3, 64)   # one-dimensional
        y = torch.empty(f.numel(), dtype=torch.float32, device="cuda")
        t_d = srv.decompress(t_ok, y)
        srv.release(t_ok)                          # deferred until the decompression is done
        srv.wait(t_d)
        _, types = tz.classify(f)
        _check_restored(torch, f, 1e-3, y, tz.classify(f)[0])
        srv.release(t_d)
        srv.release(t_big)
    finally:
        srv.close()
    with pytest.raises(tz.TszpError):
        tz.TopoServer(1 << 20, max_batch=4, max_batch_elements=1 << 22, lanes=2)  # budget below the lanes'
