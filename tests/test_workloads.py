"""Synthetic workloads: closed forms and reproducibility (CPU only)."""
import math
import os

import numpy as np

from tmc_amd import workloads


def test_genus2_closed_form():
    f, g, iso = workloads.genus2((5, 6, 7))
    nz, ny, nx = f.shape
    x = -1.5 + (np.arange(nx) + 0.5) * 3.0 / nx
    y = -1.5 + (np.arange(ny) + 0.5) * 3.0 / ny
    z = -1.5 + (np.arange(nz) + 0.5) * 3.0 / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ref = 2 * Y * (Y * Y - 3 * X * X) * (1 - Z * Z) + (X * X + Y * Y) * (X * X + Y * Y) - (9 * Z * Z - 1) * (1 - Z * Z)
    assert np.array_equal(f, ref.astype(np.float32))
    assert (g.nx, g.ny, g.nz) == (7, 6, 5) and iso == 0.0
    assert math.isclose(g.x0, x[0]) and math.isclose(g.dx, 3.0 / 7)


def test_schwarz_without_noise_closed_form():
    f, g, iso = workloads.schwarz(9, sigma=0.0)
    t = -8 * math.pi + np.arange(9) * 16 * math.pi / 8
    c = np.cos(t)
    ref = (c[None, None, :] + c[None, :, None]) + c[:, None, None]
    assert np.array_equal(f, ref.astype(np.float32))
    assert iso == 0.3 and g.x0 == -8 * math.pi


def test_noise_is_standard_normal():
    z = np.array([workloads.normal(7, i) for i in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    u = np.array([workloads.uniform(7, i) for i in range(20000)])
    assert 0 < u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


def test_deterministic_and_thread_independent():
    a, _, _ = workloads.blobs(24, seed=5)
    os.environ["TMC_WORKLOAD_THREADS"] = "3"
    try:
        b, _, _ = workloads.blobs(24, seed=5)
        c, _, _ = workloads.schwarz(20, seed=4)
    finally:
        del os.environ["TMC_WORKLOAD_THREADS"]
    d, _, _ = workloads.schwarz(20, seed=4)
    assert np.array_equal(a, b) and np.array_equal(c, d)
    e, _, _ = workloads.blobs(24, seed=6)
    assert not np.array_equal(a, e)
