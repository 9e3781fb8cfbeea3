"""GPU: block offsets of streams without a side-car, parsed on the device
(the host walk of SZp/src/szp_cuda_decompress.cu:391-477 and the chunk-count detection
of SZp/include/szp_detect_threads.h:229-309).

Streams come from the oracle (the reference's CPU writer restated) with T = 1 .. 32
chunks; the parsed offsets must equal the oracle's block walk, the chunk count the
oracle's detection, and the native decoders must reproduce the oracle's values from
the parsed side-car."""
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


def _dev(a, pad=0):
    torch = _torch()
    b = np.concatenate([a, np.full(pad, 0x5a, np.uint8)]) if pad else a
    return torch.from_numpy(np.ascontiguousarray(b)).cuda()


TOPO = [
    ("fbm3d_T1", lambda: fields.fbm3d(30, 40, 50, seed=71), 3e-3, 64, 1),
    ("fbm3d_T16", lambda: fields.fbm3d(33, 41, 52, seed=72), 3e-3, 64, 16),
    ("blobs2d_bs32_T5", lambda: fields.blobs2d(210, 333, seed=73, noise=0.05), 1e-3, 32, 5),
    ("blobs2d_bs7_T3", lambda: fields.blobs2d(97, 101, seed=74), 1e-3, 7, 3),
    ("ragged_bs128", lambda: fields.blobs2d(37, 41, seed=75, noise=0.05), 1e-3, 128, 2),
]


@pytest.mark.parametrize("name,make,eb,bs,T", TOPO, ids=[c[0] for c in TOPO])
def test_parse_topology_stream(name, make, eb, bs, T):
    torch = _torch()
    data = np.ascontiguousarray(make(), dtype=np.float32)
    n = data.size
    types = oracle.classify(data)
    cmp = oracle.topo_compress(data, eb, bs, types, T=T, adaptive=data.ndim == 2)
    blk, t = tz.parse_blocks(_dev(cmp, pad=300), n, bs, "topo")
    torch.cuda.synchronize()
    assert t == T == oracle.detect_T_topo(cmp, n, bs)
    ref = oracle.block_offsets(cmp[8 * T:], n, bs, True)
    assert np.array_equal(blk.cpu().numpy().astype(np.uint64), ref)
    # decode with the parsed side-car
    buf = tz.TopoBuffers(n, bs, T, torch.device("cuda"))
    buf.topo[:cmp.size] = torch.from_numpy(cmp).cuda()
    buf.topo_blk.copy_(blk)
    out, ty = tz.topo_decompress(buf, n, eb, bs, T=T)
    dec, fn = oracle.topo_decompress(cmp, n, eb, bs)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), dec.view(np.uint32))
    assert np.array_equal(ty.cpu().numpy().astype(np.int32), fn)


@pytest.mark.parametrize("bs,T", [(64, 1), (64, 8), (32, 3), (16, 32)])
def test_parse_sortpos_and_plain_streams(bs, T):
    torch = _torch()
    rng = np.random.default_rng(bs * 100 + T)
    sp = rng.integers(0, 5000, size=123457).astype(np.int32)
    cmp = oracle.sortpos_compress(sp, bs, T=T, adaptive=False)
    blk, t = tz.parse_blocks(_dev(cmp), sp.size, bs, "sortpos")
    assert t == T
    ref = oracle.block_offsets(cmp[8 + 8 * T:], sp.size, bs, False, bc_always=False)
    assert np.array_equal(blk.cpu().numpy().astype(np.uint64), ref)

    data = fields.blobs2d(301, 299, seed=bs + T).ravel()
    pc = oracle.plain_compress(data, 1e-3, bs, T=T, adaptive=False)
    blk, t = tz.parse_blocks(_dev(pc, pad=17), data.size, bs, "plain")
    assert t == T
    ref = oracle.block_offsets(pc[4 + 8 * T:], data.size, bs, False, bc_always=False)
    assert np.array_equal(blk.cpu().numpy().astype(np.uint64), ref)
    dec = tz.plain_decompress(_dev(pc), blk, data.size, 1e-3, bs, T=T)
    # the oracle's reader takes a single-chunk stream after the 4-byte header; the
    # decoded values do not depend on the chunk count
    ref1 = oracle.plain_compress(data, 1e-3, bs, T=1)
    ref_dec = oracle.plain_decompress(ref1[4:], data.size, 1e-3, bs)
    assert np.array_equal(dec.cpu().numpy().view(np.uint32), ref_dec.view(np.uint32))


@pytest.mark.parametrize("T", [1, 8])
def test_parse_matches_codec_side_car_large(T):
    """8.4 M points: the side-car the compressor writes equals the parsed one."""
    torch = _torch()
    f = tz.gen_fbm((128, 256, 256), seed=5)
    eb, _ = tz.rel_to_abs(f, 1e-3)
    info, buf = tz.topo_compress(f, eb, 64, T=T)
    blk, t = tz.parse_blocks(buf.topo[:info.topo_bytes], f.numel(), 64, "topo")
    torch.cuda.synchronize()
    assert t == T
    nb = tz.lib().tszp_blocks(f.numel(), 64)
    assert torch.equal(blk, buf.topo_blk[:nb + 1])
    blk2, t2 = tz.parse_blocks(buf.sp[:info.sortpos_bytes], info.ext_count, 64, "sortpos")
    assert t2 == T
    assert torch.equal(blk2, buf.sp_blk[:tz.lib().tszp_blocks(info.ext_count, 64) + 1])


def test_parse_rejects_corrupt_stream():
    torch = _torch()
    data = fields.blobs2d(64, 80, seed=76)
    types = oracle.classify(data)
    cmp = oracle.topo_compress(data, 1e-3, 64, types, T=1).copy()
    cmp[8 + 4] = 200  # an impossible bit count in the first block
    with pytest.raises(tz.TszpError):
        tz.parse_blocks(_dev(cmp), data.size, 64, "topo")
    # a stream cut short
    good = oracle.topo_compress(data, 1e-3, 64, types, T=1)
    with pytest.raises(tz.TszpError):
        tz.parse_blocks(_dev(good[: good.size // 2]), data.size, 64, "topo")
