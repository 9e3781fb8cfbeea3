"""Synthetic workloads of the benchmark configurations, generated on the device.

mixed_fields(): 256 independent float32 fields, half 2D with rows and cols log-uniform
in [64, 2048] (Gaussian blobs of sigma 20-200 px + 0.01 N(0,1) noise), half 3D with each
dimension log-uniform in [32, 256] (blobs of sigma 5-60 voxels + sine waves + 1e-3 N(0,1)
noise), absolute error bounds log-uniform in [1e-4, 1e-2]; the same draw for a seed.
"""
import numpy as np

import toposzp_hip as tz


def mixed_shapes(seed, count=256):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 2 == 0:
            shape = tuple(int(round(v)) for v in np.exp(rng.uniform(np.log(64), np.log(2048), 2)))
        else:
            shape = tuple(int(round(v)) for v in np.exp(rng.uniform(np.log(32), np.log(256), 3)))
        eb = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        out.append((shape, eb))
    return out


def mixed_fields(seed, count=256, device="cuda"):
    """[(tensor of the field's shape, eb)] of the mixed-size workload."""
    fields = []
    for i, (shape, eb) in enumerate(mixed_shapes(seed, count)):
        if len(shape) == 2:
            f = tz.gen_blobs(list(shape), seed + i, 64, 20.0, 200.0, 0.01, device=device)
        else:
            f = tz.gen_blobs(list(shape), seed + i, 128, 5.0, 60.0, 1e-3, waves=True, device=device)
        fields.append((f, eb))
    return fields
