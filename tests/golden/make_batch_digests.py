"""Known answers for the batched and served workloads: per-volume counts and
SHA-256 digests written by the oracle's C restatement (liboracle.so,
dmco_dualmc with sFlag and triangles; every volume is below the size at
which the restatement and the compiled cpp_dmc are pinned bit for bit).

* batch_4096 -- the 4,096 mixed volumes of tmc_amd.workloads
  .batch_specs(4096, seed=4): each axis round(exp(U[ln 16, ln 96])),
  generator drawn from {genus-2, Schwarz-P, Gaussian blobs}, its own iso.
* serve_pool -- the serving pool of bench.serve_pool(5): 14 volumes of
  U{32..128}^3, 5 of 256^3, 1 of 512^3.

Each volume's field is the float32 array the workload generator writes,
widened to double for the oracle (the engine reads the same floats).
Digest per volume: positions and normals (<f8), quads (<i4), halfedge rows
{origin, face, next, twin, flag} and he_v (<i4), boundary vertices (<i4),
in that order; triangles are left out (float acos ties).

    python tests/golden/make_batch_digests.py [batch_4096] [serve_pool]
"""
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(ROOT, "tests", "golden", "batch_digests.json")


def mesh_digest(m):
    """Digest of one mesh dict (v, n, quads, he, he_v, bnd)."""
    h = hashlib.sha256()
    for k, dt in (("v", "<f8"), ("n", "<f8"), ("quads", "<i4"), ("he", "<i4"), ("he_v", "<i4"), ("bnd", "<i4")):
        h.update(np.ascontiguousarray(m[k], dtype=dt).tobytes())
    return h.hexdigest()[:32]


def _one(spec):
    from oracle import Grid, Oracle
    from tmc_amd import workloads
    f, g, iso = workloads.make(spec, threads=1)
    grid = Grid(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.dx, g.dy, g.dz)
    orc = Oracle()
    s = f.astype(np.float64).ravel()
    m = orc.dualmc(grid, s, orc.gradient(grid, s), iso, True)
    return {"nv": int(m["v"].shape[0]), "nq": int(m["quads"].shape[0]), "nbnd": int(m["bnd"].shape[0]),
            "digest": mesh_digest(m)}


def _run(specs, workers):
    with ProcessPoolExecutor(workers) as ex:
        return list(ex.map(_one, specs, chunksize=8))


def batch_4096(workers):
    from tmc_amd import workloads
    specs = workloads.batch_specs(4096, seed=4)
    return {"spec": "workloads.batch_specs(4096, seed=4)", "volumes": _run(specs, workers)}


def serve_pool(workers):
    sys.path.insert(0, ROOT)
    from bench import serve_pool as pool
    specs = pool(5)
    return {"spec": "bench.serve_pool(5)", "shapes": [list(s["shape"]) for s in specs],
            "volumes": _run(specs, workers)}


MAKERS = {"batch_4096": batch_4096, "serve_pool": serve_pool}


def main(names, workers=4):
    known = {}
    if os.path.exists(OUT):
        with open(OUT) as fh:
            known = json.load(fh)
    for name in names:
        t0 = time.time()
        known[name] = MAKERS[name](workers)
        print(f"{name}: {len(known[name]['volumes'])} volumes in {time.time() - t0:.1f} s", flush=True)
        with open(OUT, "w") as fh:
            json.dump(known, fh, separators=(",", ":"))
            fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:] or list(MAKERS))
