"""GPU quality metrics (example/test_cuda_topology.cu:540-868) against the oracle's
restatement, on 2D and 3D fields: the restored output of the pipeline and noisy copies
that lose, change and create critical points."""
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


def _cmp(got, ref):
    for k, v in ref.items():
        g = got[k]
        if isinstance(v, list):
            assert list(g) == v, f"{k}: {list(g)} vs {v}"
        elif isinstance(v, float):
            if v == 1e300:
                assert g >= 1e299, k
            else:
                assert abs(g - v) <= 1e-9 * max(1.0, abs(v)), f"{k}: {g} vs {v}"
        else:
            assert g == v, f"{k}: {g} vs {v}"


CASES = [
    ("blobs2d", lambda: fields.blobs2d(180, 211, seed=81, noise=0.02), 1e-3),
    ("fbm3d", lambda: fields.fbm3d(30, 34, 40, seed=82), 3e-3),
    ("ties2d", lambda: fields.ties2d(90, 97, seed=83), 1e-3),
]


@pytest.mark.parametrize("name,make,eb", CASES, ids=[c[0] for c in CASES])
def test_quality_of_noisy_copy(name, make, eb):
    torch = _torch()
    x = np.ascontiguousarray(make(), dtype=np.float32)
    rng = np.random.default_rng(7)
    y = (x + rng.uniform(-1.5 * eb, 1.5 * eb, size=x.shape)).astype(np.float32)
    got = tz.quality_metrics(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), x.shape, eb)
    _cmp(got, oracle.quality_metrics(x, y, eb))


@pytest.mark.parametrize("name,make,eb", CASES, ids=[c[0] for c in CASES])
def test_quality_of_restored_field(name, make, eb):
    torch = _torch()
    x = np.ascontiguousarray(make(), dtype=np.float32)
    f = torch.from_numpy(x).cuda()
    info, buf = tz.topo_compress(f, eb, 64)
    out, types = tz.topo_decompress(buf, x.size, eb, 64)
    sp = tz.sortpos_decompress(buf, info.ext_count, 64)
    yt = torch.empty_like(out)
    tz.topo_restore(out, types, x.shape, eb, sp=sp, n_ext=info.ext_count, out=yt)
    y = yt.view(x.shape)
    got = tz.quality_metrics(f, y, x.shape, eb)
    ref = oracle.quality_metrics(x, y.cpu().numpy(), eb)
    _cmp(got, ref)
    # the restoration keeps every critical point: none lost, changed or new
    assert got["lost"][0] == got["changed"][0] == got["new_cps"] == 0
    assert got["preserved"] == got["orig_cp"][0]
