"""Shared test helpers (not a test module)."""
import math

import numpy as np

TWOPI = 2.0 * math.acos(-1.0)


def seeded_field(shape, seed, lengths=(TWOPI, TWOPI, TWOPI), nmodes=8, kmax=4):
    """U[-1,1] noise plus `nmodes` random low-wavenumber sine modes
    (|k|_inf <= kmax, amplitude U[0.5,1], phase U[0,2pi)) on a (nz, ny, nx)
    grid spanning `lengths` = (xlx, yly, zlz) without the end point."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    u = rng.uniform(-1.0, 1.0, size=shape)
    x = np.arange(nx) * lengths[0] / nx
    y = np.arange(ny) * lengths[1] / ny
    z = np.arange(nz) * lengths[2] / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    for _ in range(nmodes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        a = rng.uniform(0.5, 1.0)
        ph = rng.uniform(0.0, TWOPI)
        u += a * np.sin(k[0] * X * TWOPI / lengths[0] + k[1] * Y * TWOPI / lengths[1]
                        + k[2] * Z * TWOPI / lengths[2] + ph)
    return u


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
