#!/usr/bin/env python3
"""Write the topologically correct MC fixtures (tests/golden/tmc_*.npz) from
the reference's t_mc compiled from mc/MarchingCubes.cpp
(oracle/_ref/libtmcref.so; build it with `make -C oracle ref`).

Each fixture holds the field (float64, (nz, ny, nx)), the grid
(nx, ny, nz, x0, y0, z0, dx, dy, dz), the iso-value and the reference's
output: v, n (float64, 3 per vertex) and tris (int32, 3 per triangle).
The fields are small and chosen for p_slice's branches: random fields (many
ambiguous cells, tunnels, 12-vertex contours), small integers (face
deciders exactly at the iso-value), and a smooth surface on an offset,
anisotropic grid."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import Grid, TmcReference  # noqa: E402


def cases():
    rng = np.random.default_rng(2024)
    yield "tmc_uniform_14", rng.uniform(0, 1, (14, 14, 14)), 0.5, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    yield "tmc_normal_9x11x13", rng.standard_normal((9, 11, 13)), 0.05, (1.0, -2.0, 0.25), (0.5, 1.25, 0.75)
    yield "tmc_ints_12", rng.integers(0, 5, (12, 12, 12)).astype(np.float64), 2.0, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    z, y, x = np.meshgrid(np.linspace(-2, 2, 17), np.linspace(-2, 2, 19), np.linspace(-2, 2, 21), indexing="ij")
    g2 = 2 * y * (y * y - 3 * x * x) * (1 - z * z) + (x * x + y * y) ** 2 - (9 * z * z - 1) * (1 - z * z)
    yield "tmc_genus2_21x19x17", g2, 0.0, (-2.0, -2.0, -2.0), (4.0 / 20, 4.0 / 18, 4.0 / 16)


def main():
    ref = TmcReference()
    for name, f, iso, origin, spacing in cases():
        g = Grid.make(f.shape, origin, spacing)
        out = ref.tmc(g, f, iso)
        grid = np.array([g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.dx, g.dy, g.dz], np.float64)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), field=f.astype(np.float64), grid=grid,
                            iso=np.float64(iso), v=out["v"], n=out["n"], tris=out["tris"])
        print(name, f.shape, len(out["v"]), len(out["tris"]))


if __name__ == "__main__":
    main()
