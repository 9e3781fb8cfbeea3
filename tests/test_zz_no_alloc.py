"""tszp_plain_compress_noalloc (scratch in a static device array, no allocation) gives
the bytes, side-car and size of tszp_plain_compress, on two calls in a row with the
second smaller than the first (the shared scratch keeps the last call's contents, so
the second call must clear what it needs), and rejects fields beyond its scratch."""
import ctypes as C

import numpy as np
import pytest

import fields
import toposzp_hip as tz

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _noalloc():
    L = tz.lib()
    fn = L.tszp_plain_compress_noalloc
    fn.argtypes = [C.c_void_p, C.c_size_t, C.c_float, C.c_int, C.c_uint, C.c_void_p, C.c_size_t, C.c_void_p,
                   C.POINTER(C.c_uint64), C.c_void_p]
    fn.restype = C.c_int
    return fn


def _run_noalloc(field, eb, bs, T):
    torch = _torch()
    L = tz.lib()
    n = field.numel()
    out = torch.full((L.tszp_plain_bound(n, bs, T),), 0xAB, dtype=torch.uint8, device=field.device)
    blk = torch.full((L.tszp_blocks(n, bs) + 1,), -1, dtype=torch.int32, device=field.device)
    nbytes = C.c_uint64()
    rc = _noalloc()(field.data_ptr(), n, eb, bs, T, out.data_ptr(), out.numel(), blk.data_ptr(), C.byref(nbytes),
                    C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, L.tszp_last_error().decode()
    return nbytes.value, out, blk


@pytest.mark.parametrize("T", [1, 3])
def test_noalloc_equals_plain_compress_two_calls(T):
    torch = _torch()
    big = torch.from_numpy(fields.blobs3d(64, 160, 300, seed=31, noise=2e-3)).cuda().view(-1)
    small = torch.from_numpy(fields.blobs2d(301, 517, seed=32, noise=0.01)).cuda().view(-1)
    for f, eb, bs in ((big, 1e-4, 64), (small, 1e-3, 32)):
        print(f"noalloc n={f.numel()} eb={eb} bs={bs} T={T}", flush=True)
        nb, out, blk = _run_noalloc(f, eb, bs, T)
        rb, rout, rblk = tz.plain_compress(f, eb, bs, T)
        assert nb == rb
        assert torch.equal(out[:nb], rout[:rb])
        assert torch.equal(blk, rblk)
        # the plain codec quantizes with the float reciprocal of eb, so its error can pass
        # eb slightly; the reference's own round-trip test allows 1.01 eb
        # (tests/test_compress_decompress.cu:88)
        dec = tz.plain_decompress(out, blk, f.numel(), eb, bs, T)
        assert float((dec - f).abs().max()) < 1.01 * eb


def test_noalloc_generic_block_size_and_tile_counter(monkeypatch):
    torch = _torch()
    f = torch.from_numpy(fields.blobs2d(257, 389, seed=33, noise=0.02)).cuda().view(-1)
    ref = tz.plain_compress(f, 2e-3, 48, 1)
    monkeypatch.setenv("TSZP_TILE_COUNTER", "1")
    nb, out, blk = _run_noalloc(f, 2e-3, 48, 1)
    assert nb == ref[0] and torch.equal(out[:nb], ref[1][:nb]) and torch.equal(blk, ref[2])


def test_noalloc_rejects_fields_beyond_its_scratch():
    torch = _torch()
    L = tz.lib()
    dummy = torch.empty(1024, dtype=torch.uint8, device="cuda")
    nbytes = C.c_uint64()
    # 2^31 - 2 elements need 2^19 tiles at block size 64: rejected before any launch
    rc = _noalloc()(dummy.data_ptr(), 2 ** 31 - 2, 1e-3, 64, 1, dummy.data_ptr(), dummy.numel(), dummy.data_ptr(),
                    C.byref(nbytes), None)
    assert rc == -1 and b"static scratch" in L.tszp_last_error()
