"""Known answers for the full-size workload of 64 Gaussian blobs + noise at
512^3 (tmc_amd.workloads.blobs(512, seed=2), iso 0.5), written by the
compiled references (oracle/_ref: cpp_dmc's dualMC with sFlag and gradients
from estimateGradient + flip_gradient; mc/'s t_mc without the orientation
pass): counts and SHA-256 digests of the output arrays in little-endian
bytes (DMC: positions, normals, quads; TMC: positions, normals, triangles).
DMC triangles are left out: cpp_dmc picks each split with float acos, whose
last bit may differ between math libraries at ties.

Run in a tree where oracle/_ref is built:
    python tests/golden/make_fullsize_digests.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def digest(*arrays):
    h = hashlib.sha256()
    for a, dt in arrays:
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
    return h.hexdigest()


def main():
    from oracle import Grid, Reference, TmcReference
    from tmc_amd import workloads
    f, g, iso = workloads.CONFIGS["blobs_512"]()
    grid = Grid(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.dx, g.dy, g.dz)
    s = f.astype(np.float64).ravel()
    out = {"workload": "blobs_512", "iso": iso}
    t0 = time.time()
    ref = Reference()
    m = ref.dualmc(grid, s, ref.gradient(grid, s), iso, True)
    out["dmc"] = {"nv": int(m["v"].shape[0]), "nq": int(m["quads"].shape[0]),
                  "sha256_v_n_quads": digest((m["v"], "<f8"), (m["n"], "<f8"), (m["quads"], "<i4"))}
    print(f"dmc {time.time() - t0:.1f} s {out['dmc']}", flush=True)
    del m
    t0 = time.time()
    r = TmcReference().tmc(grid, s, iso)
    out["tmc"] = {"nv": int(r["v"].shape[0]), "nt": int(r["tris"].shape[0]),
                  "sha256_v_n_tris": digest((r["v"], "<f8"), (r["n"], "<f8"), (r["tris"], "<i4"))}
    print(f"tmc {time.time() - t0:.1f} s {out['tmc']}", flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "fullsize_digests.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
