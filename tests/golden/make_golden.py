#!/usr/bin/env python3
"""Write the golden fixtures in tests/golden/ from the compiled reference.

    make -C oracle ref && python tests/golden/make_golden.py

The vectors come from oracle/_ref/libdmcref.so, i.e. the reference cpp_dmc
built from its own sources.  Each fixture stores its inputs (field, grid,
iso) and the reference's outputs, so the tests need neither the reference
nor a regenerated field.  Files:

  dualmc_<case>_s<0|1>.npz  whole dualMC (sFlag off/on): v, n, quads, tris and
                            the final halfedges (he, he_v, he_f, bnd)
  stages_<case>.npz         cells (v, n, quads, colors, failed projections),
                            halfedges, colorFaces, vertexValence, mark3X3Y,
                            both mergeVertices3X3Y variants, mark3333
  cells_catalogue.npz       single cells over all 256 cases: polygons and
                            representatives (local coordinates)
  tables.npz                the reference lookup tables r_pattern, t_ambig
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import Grid, Reference  # noqa: E402
from tmc_amd import workloads  # noqa: E402


def grid_arr(g):
    return np.array([g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.dx, g.dy, g.dz], np.float64)


def cases():
    out = {}
    f, g, iso = workloads.genus2(24)
    out["genus2_24"] = (f, g, iso, None)
    rng = np.random.default_rng(11)
    f = rng.standard_normal((9, 11, 13)).astype(np.float32)
    out["noise_13x11x9"] = (f, Grid.make(f.shape, (0.5, -1.0, 2.0), (0.1, 0.2, 0.3)), 0.1, None)
    rng = np.random.default_rng(12)
    f = rng.integers(-2, 3, size=(10, 9, 8)).astype(np.float32)
    out["ties_8x9x10"] = (f, Grid.make(f.shape, (0, 0, 0), (1, 1, 1)), 0.0, None)
    f, g, iso = workloads.blobs(32, seed=2)
    out["blobs_32"] = (f, g, iso, None)
    f, g, iso = workloads.schwarz(25, seed=3, lo=-2 * np.pi, hi=2 * np.pi)
    out["schwarz_25"] = (f, g, iso, None)
    rng = np.random.default_rng(13)
    f = rng.standard_normal((8, 10, 12))            # float64 field, not float-representable
    out["f64_12x10x8"] = (f, Grid.make(f.shape, (-1, -1, -1), (0.25, 0.25, 0.25)), 0.3, None)
    f, g, iso = workloads.genus2(20)
    grad = np.random.default_rng(14).standard_normal(f.size * 3)
    out["genus2_20_grad"] = (f, g, iso, grad)
    return out


def main():
    R = Reference()
    rp, ta = R.tables()
    np.savez_compressed(os.path.join(HERE, "tables.npz"), r_pattern=rp, t_ambig=ta)
    for name, (f, g, iso, grad) in cases().items():
        s = f.astype(np.float64).ravel()
        gr = grad if grad is not None else R.gradient(g, s)
        for sflag in (0, 1):
            m = R.dualmc(g, s, gr, iso, sflag)
            h = R.halfedges(m["v"].shape[0], m["quads"])
            np.savez_compressed(os.path.join(HERE, f"dualmc_{name}_s{sflag}.npz"),
                                field=f, grid=grid_arr(g), iso=iso, sflag=sflag,
                                grad=grad if grad is not None else np.zeros(0),
                                v=m["v"], n=m["n"], quads=m["quads"], tris=m["tris"],
                                he=h["he"], he_v=h["he_v"], he_f=h["he_f"], bnd=h["bnd"])
        if name in ("noise_13x11x9", "genus2_24", "ties_8x9x10"):
            c = R.cells(g, s, gr, iso)
            nv = c["v"].shape[0]
            h = R.halfedges(nv, c["quads"])
            col = R.color_faces(h["he"], h["he_f"], c["colors"])
            val = R.valence(nv, h["he"])
            p = R.mark3x3y(c["quads"], val)
            mc = R.merge3x3y(c["v"], c["n"], p, val, col, h["he"], h["he_f"])
            mi = R.merge3x3y(c["v"], c["n"], p, val, None, h["he"], h["he_f"])
            m33 = R.mark3333(nv, c["quads"], h["he"], h["he_f"], val)
            np.savez_compressed(os.path.join(HERE, f"stages_{name}.npz"),
                                field=f, grid=grid_arr(g), iso=iso,
                                v=c["v"], n=c["n"], quads=c["quads"], colors=c["colors"],
                                nr_proj=np.array([c["nr_proj1"], c["nr_proj2"], c["nv_before_unused"]]),
                                he=h["he"], he_v=h["he_v"], bnd=h["bnd"], colors_final=col, valence=val,
                                p3x3y=p, mc_v=mc[0], mc_n=mc[1], mc_vf=mc[2], mc_vs=mc[3], mc_em=mc[4],
                                mi_v=mi[0], mi_n=mi[1], mi_vf=mi[2], mi_vs=mi[3], mi_em=mi[4],
                                m33_vf=m33[0], m33_vt=m33[1], m33_p=m33[2], m33_r=m33[3])
    # single-cell catalogue: every case, random corner values around iso,
    # plus near-ties on ambiguous faces
    rng = np.random.default_rng(21)
    F, ISO = [], []
    for icase in range(256):
        for rep in range(4):
            iso = float(rng.normal() * 0.2)
            mag = rng.uniform(0.01, 1.0, 8)
            if rep == 3:
                mag = rng.uniform(1e-9, 1e-3, 8)
            sign = np.array([1.0 if (icase >> c) & 1 else -1.0 for c in range(8)])
            vals = np.where(sign > 0, iso + mag, iso - mag)
            F.append(vals)
            ISO.append(iso)
    F = np.array(F)
    ISO = np.array(ISO)
    cnt = np.zeros(len(F), np.int32)
    sizes = np.zeros((len(F), 4), np.int32)
    edges = np.zeros((len(F), 12), np.int32)
    loc = np.full((len(F), 4, 3), np.nan)
    fails = np.zeros((len(F), 2), np.int64)
    for a in range(len(F)):
        cnt[a], sizes[a], edges[a] = R.cell_polygons(F[a], ISO[a])
        l, f1, f2 = R.cell_vertices(F[a], ISO[a])
        loc[a, :l.shape[0]] = l
        fails[a] = (f1, f2)
    np.savez_compressed(os.path.join(HERE, "cells_catalogue.npz"), f=F, iso=ISO, cnt=cnt, sizes=sizes,
                        edges=edges, local=loc, fails=fails)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
