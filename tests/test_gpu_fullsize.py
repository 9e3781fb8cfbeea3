"""GPU: the field workloads at their full sizes.

1800 x 3600 (2D blobs + 0.01 noise, abs 1e-3, bs 64) and 100 x 500 x 500 (blobs + sine
waves + 1e-3 noise, rel 1e-4, bs 64) are compared with the oracle byte for byte.  512^3
(fBm, rel 1e-3, bs 64 and 128) is beyond what the oracle finishes in seconds, so it is
checked by properties that do not depend on the size:
the type map that comes out of the topology stream is the field's own classification
(SZp/src/szp_topology_3d.cc:79-98), every decoded value is within the error bound,
the sort-position stream decodes to the positions the compressor computed, the
restoration leaves no critical point missing, spurious or changed, the device parse of
the stream without its side-car finds the side-car the compressor wrote, and a second
compression gives the same bytes.
"""
import numpy as np
import pytest

import oracle
import toposzp_hip as tz

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def test_blob_plane_1800x3600_matches_oracle():
    torch = _torch()
    shape = (1800, 3600)
    f = tz.gen_blobs(shape, 100, 64, 20.0, 200.0, 0.01)
    eb, bs = 1e-3, 64
    info, buf = tz.topo_compress(f, eb, bs)
    data = f.cpu().numpy()
    cps = oracle.find_cps_2d(data, eb)
    oracle.sort_positions(cps, data)
    types = oracle.mark_types(cps, data.shape)
    sp = oracle.extrema_sort_positions(cps)
    topo = oracle.topo_compress(data, eb, bs, types, T=1, adaptive=True)
    sps = oracle.sortpos_compress(sp, bs, T=1, adaptive=True)
    assert info.cp_count == cps.size and info.ext_count == sp.size
    assert np.array_equal(buf.topo[:info.topo_bytes].cpu().numpy(), topo), "topology stream"
    assert np.array_equal(buf.sp[:info.sortpos_bytes].cpu().numpy(), sps), "sort-position stream"
    out, ty = tz.topo_decompress(buf, data.size, eb, bs)
    dec, _ = oracle.topo_decompress(topo, data.size, eb, bs)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), dec.view(np.uint32))
    assert np.array_equal(ty.cpu().numpy(), types.ravel())
    pos = tz.sortpos_decompress(buf, info.ext_count, bs)
    r = tz.topo_restore(out, ty, shape, eb, sp=pos, n_ext=info.ext_count, out=torch.empty_like(out))
    assert r.mismatches() == 0 and r.max_dev_over_eb < 1.0


def test_blob_volume_100x500x500_matches_oracle():
    torch = _torch()
    shape = (100, 500, 500)
    f = tz.gen_blobs(shape, 101, 128, 5.0, 60.0, 1e-3, waves=True)
    eb, _ = tz.rel_to_abs(f, 1e-4)
    bs = 64
    info, buf = tz.topo_compress(f, eb, bs)
    data = f.cpu().numpy()
    cps = oracle.find_cps_3d(data, eb)
    oracle.sort_positions(cps, data)
    types = oracle.mark_types(cps, data.shape)
    sp = oracle.extrema_sort_positions(cps)
    topo = oracle.topo_compress(data, eb, bs, types, T=1, adaptive=False)
    sps = oracle.sortpos_compress(sp, bs, T=1, adaptive=False)
    assert info.cp_count == cps.size and info.ext_count == sp.size
    assert np.array_equal(buf.topo[:info.topo_bytes].cpu().numpy(), topo), "topology stream"
    assert np.array_equal(buf.sp[:info.sortpos_bytes].cpu().numpy(), sps), "sort-position stream"
    out, ty = tz.topo_decompress(buf, data.size, eb, bs)
    dec, _ = oracle.topo_decompress(topo, data.size, eb, bs)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), dec.view(np.uint32))
    assert np.array_equal(ty.cpu().numpy(), types.ravel())
    pos = tz.sortpos_decompress(buf, info.ext_count, bs)
    r = tz.topo_restore(out, ty, shape, eb, sp=pos, n_ext=info.ext_count, out=torch.empty_like(out))
    assert r.mismatches() == 0 and r.max_dev_over_eb < 1.0


@pytest.mark.parametrize("bs", [64, 128])
def test_fbm_512_cube_properties(bs):
    torch = _torch()
    shape = (512, 512, 512)
    n = 512 ** 3
    f = tz.gen_fbm(shape, seed=20251203)
    eb, _ = tz.rel_to_abs(f, 1e-3)
    info, buf = tz.topo_compress(f, eb, bs)
    first = buf.topo[:info.topo_bytes].clone()
    sp_vals = buf.sp_values[:info.ext_count].clone()
    types, cnt = tz.classify(f)
    assert cnt == info.cp_count
    out, ty = tz.topo_decompress(buf, n, eb, bs)
    assert torch.equal(ty, types), "type map of the stream differs from the field's classification"
    err = float((out - f.view(-1)).abs().max())
    assert err < eb * (1 + 1e-6), f"error {err / eb} eb"
    pos = tz.sortpos_decompress(buf, info.ext_count, bs)
    assert torch.equal(pos, sp_vals), "sort positions"
    blk, T = tz.parse_blocks(first, n, bs, "topo")
    assert T == 1 and torch.equal(blk, buf.topo_blk), "parsed side-car"
    info2, buf2 = tz.topo_compress(f, eb, bs)
    assert info2.topo_bytes == info.topo_bytes and torch.equal(buf2.topo[:info.topo_bytes], first)
    del buf2
    y = torch.empty_like(out)
    r = tz.topo_restore(out, ty, shape, eb, sp=pos, n_ext=info.ext_count, out=y)
    assert r.mismatches() == 0 and r.max_dev_over_eb < 1.0
    # pointwise bound of the restored field: eb, and eb (1 + 2^-16) in the zero cell
    yerr = float((y.double() - f.view(-1).double()).abs().max())
    assert yerr <= eb * (1 + 2.0 ** -16), f"restored error {yerr / eb} eb"


@pytest.mark.parametrize("bs", [64, 128])
def test_plain_512_cube_properties(bs):
    """Plain SZp on the 512^3 fBm field (rel 1e-3): every decoded value within the error
    bound (the reference's plain decoder rounds (float)q * eb the same way), the device
    parse of the stream without its side-car finds the compressor's side-car, and a
    second compression gives the same bytes."""
    torch = _torch()
    shape = (512, 512, 512)
    n = 512 ** 3
    f = tz.gen_fbm(shape, seed=20251203)
    eb, _ = tz.rel_to_abs(f, 1e-3)
    print(f"plain 512^3 bs={bs} eb={eb}", flush=True)
    nbytes, out, blk = tz.plain_compress(f, eb, bs)
    first = out[:nbytes].clone()
    dec = tz.plain_decompress(out, blk, n, eb, bs)
    err = float((dec - f.view(-1)).abs().max())
    assert err < eb * (1 + 1e-6), f"error {err / eb} eb"
    pblk, T = tz.parse_blocks(first, n, bs, "plain")
    assert T == 1 and torch.equal(pblk, blk), "parsed side-car"
    nbytes2, out2, _ = tz.plain_compress(f, eb, bs)
    assert nbytes2 == nbytes and torch.equal(out2[:nbytes], first)
    # the header: eb big-endian, one chunk at offset 0
    h = first[:12].cpu().numpy()
    assert h[:4].tobytes() == np.float32(eb).byteswap().tobytes() and not h[4:].any()
