"""Small deterministic synthetic fields for tests (numpy PCG64 streams).

Shapes follow the workloads: Gaussian blobs + white noise (2D/3D), fractal value
noise (3D), and adversarial cases (ties, negative zero, huge values, constant).
"""
import numpy as np


def blobs2d(rows, cols, nblobs=16, noise=0.01, seed=0, amp=1.0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:rows, 0:cols].astype(np.float64)
    f = np.zeros((rows, cols), dtype=np.float64)
    for _ in range(nblobs):
        a = rng.uniform(-amp, amp)
        cy, cx = rng.uniform(0, rows), rng.uniform(0, cols)
        sy = rng.uniform(0.05, 0.4) * rows
        sx = rng.uniform(0.05, 0.4) * cols
        th = rng.uniform(0, np.pi)
        dy, dx = y - cy, x - cx
        u = np.cos(th) * dx + np.sin(th) * dy
        v = -np.sin(th) * dx + np.cos(th) * dy
        f += a * np.exp(-0.5 * ((u / sx) ** 2 + (v / sy) ** 2))
    f += noise * rng.standard_normal((rows, cols))
    return f.astype(np.float32)


def blobs3d(d1, d2, d3, nblobs=16, noise=1e-3, seed=0):
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:d1, 0:d2, 0:d3].astype(np.float64)
    f = np.zeros((d1, d2, d3), dtype=np.float64)
    for _ in range(nblobs):
        a = rng.uniform(-1, 1)
        c = rng.uniform(0, 1, 3) * np.array([d1, d2, d3])
        s = rng.uniform(0.1, 0.5) * min(d1, d2, d3)
        f += a * np.exp(-0.5 * (((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2) / s ** 2))
    k = rng.uniform(0.02, 0.2, 3) * 2 * np.pi
    f += 0.3 * np.sin(k[0] * z) * np.sin(k[1] * y) * np.sin(k[2] * x)
    f += noise * rng.standard_normal((d1, d2, d3))
    return f.astype(np.float32)


def fbm3d(d1, d2, d3, octaves=6, noise=0.05, seed=0):
    """Value-noise fBm (lacunarity 2, gain 0.5) plus white noise."""
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:d1, 0:d2, 0:d3].astype(np.float64)
    f = np.zeros((d1, d2, d3))
    amp, freq = 1.0, 1.0 / 16
    for _ in range(octaves):
        g = rng.uniform(-1, 1, (int(d1 * freq) + 2, int(d2 * freq) + 2, int(d3 * freq) + 2))
        zz, yy, xx = z * freq, y * freq, x * freq
        z0, y0, x0 = zz.astype(int), yy.astype(int), xx.astype(int)
        tz, ty, tx = zz - z0, yy - y0, xx - x0
        acc = 0
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    w = (tz if a else 1 - tz) * (ty if b else 1 - ty) * (tx if c else 1 - tx)
                    acc = acc + w * g[z0 + a, y0 + b, x0 + c]
        f += amp * acc
        amp *= 0.5
        freq *= 2
    f += noise * rng.standard_normal((d1, d2, d3))
    return f.astype(np.float32)


def ties2d(rows, cols, seed=0, step=0.01):
    """Values on a coarse lattice: many exact ties among critical points, with
    negative values and negative zeros."""
    f = blobs2d(rows, cols, nblobs=6, noise=0.05, seed=seed)
    f = (np.round(f / step) * step).astype(np.float32)
    f[f == 0] = np.float32(-0.0)
    return f


def huge2d(rows, cols, seed=0):
    """Large magnitudes so that block deltas need 20-33 bits."""
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1.0e6, 1.0e6, (rows, cols))).astype(np.float32)
