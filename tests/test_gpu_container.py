"""GPU: the .SZp container of the reference's CPU drivers
(example/testfloat_compress_fastmode1_topology.cc:105-139,
example/testfloat_decompress_fastmode1_topology.cc:722-816).

Streams written by the native path are packed into a container whose bytes equal the
oracle's container for the same field; a container written from the oracle's streams
is unpacked and decoded by the library's readers (chunk-count detection included)."""
import struct

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


def _oracle_container(data, eb, bs, T):
    cps = oracle.find_cps_3d(data, eb) if data.ndim == 3 else oracle.find_cps_2d(data, eb)
    oracle.sort_positions(cps, data)
    types = oracle.mark_types(cps, data.shape)
    sp = oracle.extrema_sort_positions(cps)
    adaptive = data.ndim == 2
    topo = oracle.topo_compress(data, eb, bs, types, T=T, adaptive=adaptive)
    sps = oracle.sortpos_compress(sp, bs, T=T, adaptive=adaptive) if sp.size else np.zeros(0, np.uint8)
    footer = struct.pack("<QQ", topo.size, sps.size)
    return np.concatenate([topo, sps, np.frombuffer(footer, np.uint8)]), types, sp, topo, sps


CASES = [
    ("fbm3d_T1", lambda: fields.fbm3d(24, 36, 40, seed=61), 3e-3, 64, 1),
    ("fbm3d_T3", lambda: fields.fbm3d(20, 30, 33, seed=62), 3e-3, 64, 3),
    ("blobs2d_T4", lambda: fields.blobs2d(150, 161, seed=63, noise=0.03), 1e-3, 32, 4),
]


@pytest.mark.parametrize("name,make,eb,bs,T", CASES, ids=[c[0] for c in CASES])
def test_pack_matches_oracle_container(name, make, eb, bs, T):
    torch = _torch()
    data = np.ascontiguousarray(make(), dtype=np.float32)
    ref, _, _, _, _ = _oracle_container(data, eb, bs, T)
    info, buf = tz.topo_compress(torch.from_numpy(data).cuda(), eb, bs, T=T)
    out, nbytes = tz.szp_pack(buf.topo, info.topo_bytes, buf.sp, info.sortpos_bytes)
    torch.cuda.synchronize()
    assert nbytes == ref.size
    got = out[:nbytes].cpu().numpy()
    assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} bytes differ"
    # the footer read back from device memory
    assert tz.szp_unpack(out, nbytes) == (info.topo_bytes, info.topo_bytes, info.sortpos_bytes, info.ext_count)


@pytest.mark.parametrize("name,make,eb,bs,T", CASES, ids=[c[0] for c in CASES])
def test_unpack_and_decode_oracle_container(name, make, eb, bs, T):
    _torch()
    data = np.ascontiguousarray(make(), dtype=np.float32)
    cont, types, sp, topo, sps = _oracle_container(data, eb, bs, T)
    tb, so, sb, e = tz.szp_unpack(cont)  # host memory
    assert (tb, so, sb, e) == (topo.size, topo.size, sps.size, sp.size)
    # the decompress driver's readers: topology stream with its chunk header, sort
    # positions after their u64 count
    dec, fn = tz.decompress_topology(cont[:tb], data.size, eb, bs)
    odec, ofn = oracle.topo_decompress(topo, data.size, eb, bs)
    assert np.array_equal(dec.view(np.uint32), odec.view(np.uint32))
    assert np.array_equal(fn, ofn)
    assert np.array_equal(fn.astype(np.uint8), types.ravel())
    got_sp = tz.decompress_sort_positions(cont[so + 8:so + sb], e, bs)
    assert np.array_equal(got_sp, sp)


def test_unpack_without_valid_footer_is_one_topology_stream():
    _torch()
    # an old single-stream file: the last 16 bytes are not a consistent footer
    buf = np.arange(200, dtype=np.uint8)
    assert tz.szp_unpack(buf) == (200, 200, 0, 0)
    # sizes larger than the buffer fall back the same way
    bad = np.concatenate([np.zeros(40, np.uint8), np.frombuffer(struct.pack("<QQ", 1000, 8), np.uint8)])
    assert tz.szp_unpack(bad) == (bad.size, bad.size, 0, 0)
    # a sort-position size whose sum with the topology size wraps around 2^64 must not
    # pass the check: sp_offset + sp_bytes stays within the buffer
    wrap = np.concatenate([np.zeros(40, np.uint8), np.frombuffer(struct.pack("<QQ", 20, 2**64 - 30), np.uint8)])
    assert tz.szp_unpack(wrap) == (wrap.size, wrap.size, 0, 0)
    ok = np.concatenate([np.zeros(40, np.uint8), np.frombuffer(struct.pack("<QQ", 20, 20), np.uint8)])
    assert tz.szp_unpack(ok)[:3] == (20, 20, 20)


def test_pack_rejects_small_output():
    torch = _torch()
    topo = torch.zeros(100, dtype=torch.uint8, device="cuda")
    sp = torch.zeros(30, dtype=torch.uint8, device="cuda")
    small = torch.empty(100, dtype=torch.uint8, device="cuda")
    with pytest.raises(tz.TszpError):
        tz.szp_pack(topo, 100, sp, 30, out=small)
