"""Known answers for the two full-size single-volume workloads: counts and
SHA-256 digests of the output arrays in little-endian bytes (DMC:
positions, normals, quads, and the halfedge rows where written; TMC:
positions, normals, triangles).  DMC triangles are left out: cpp_dmc picks
each split with float acos, whose last bit may differ between math
libraries at ties.

* blobs_512 -- 64 Gaussian blobs + noise at 512^3 (tmc_amd.workloads
  .blobs(512, seed=2), iso 0.5), written by the compiled references
  (oracle/_ref: cpp_dmc's dualMC with sFlag and gradients from
  estimateGradient + flip_gradient; mc/'s t_mc without the orientation
  pass).
* schwarz_1024 -- Schwarz-P + noise at 1024^3 (workloads.schwarz(1024,
  seed=3), iso 0.3), written by the oracle's C restatement (liboracle.so,
  dmco_dualmc with sFlag and tmco_tmc).  The compiled references cannot
  answer here: 1024^3 = 1,073,741,824 points is above the 715,827,882 at
  which cpp_dmc's int edge key 3*gl+dir overflows
  (cpp_dmc/DualMarchingCubes.h:215-223) and above the 357,913,941 at which
  t_mc's 6*point+axis key does (mc/MarchingCubes.cpp).  The restatement
  keeps those keys 64-bit, which is what they mean without the overflow;
  it is pinned to the compiled references bit for bit below that size
  (tests/test_oracle.py, tests/test_tmc_oracle.py, and blobs_512 here).
  It needs about 45 GB of host memory and several minutes.

    python tests/golden/make_fullsize_digests.py [blobs_512] [schwarz_1024]

Each named workload's entry in fullsize_digests.json is rewritten; the
others are kept.
"""
import gc
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(ROOT, "tests", "golden", "fullsize_digests.json")


def digest(*arrays):
    h = hashlib.sha256()
    for a, dt in arrays:
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
    return h.hexdigest()


def _grid(g):
    from oracle import Grid
    return Grid(g.nx, g.ny, g.nz, g.x0, g.y0, g.z0, g.dx, g.dy, g.dz)


def blobs_512():
    from oracle import Reference, TmcReference
    from tmc_amd import workloads
    f, g, iso = workloads.CONFIGS["blobs_512"]()
    grid = _grid(g)
    s = f.astype(np.float64).ravel()
    del f
    out = {"iso": iso, "source": "compiled references (oracle/_ref: cpp_dmc dualMC, mc/ t_mc)"}
    t0 = time.time()
    ref = Reference()
    m = ref.dualmc(grid, s, ref.gradient(grid, s), iso, True)
    out["dmc"] = {"nv": int(m["v"].shape[0]), "nq": int(m["quads"].shape[0]),
                  "sha256_v_n_quads": digest((m["v"], "<f8"), (m["n"], "<f8"), (m["quads"], "<i4"))}
    print(f"blobs_512 dmc {time.time() - t0:.1f} s {out['dmc']}", flush=True)
    del m
    t0 = time.time()
    r = TmcReference().tmc(grid, s, iso)
    out["tmc"] = {"nv": int(r["v"].shape[0]), "nt": int(r["tris"].shape[0]),
                  "sha256_v_n_tris": digest((r["v"], "<f8"), (r["n"], "<f8"), (r["tris"], "<i4"))}
    print(f"blobs_512 tmc {time.time() - t0:.1f} s {out['tmc']}", flush=True)
    return out


def schwarz_1024():
    from oracle import Oracle
    from tmc_amd import workloads
    f, g, iso = workloads.CONFIGS["schwarz_1024"]()
    grid = _grid(g)
    s = f.astype(np.float64).ravel()
    del f
    gc.collect()
    out = {"iso": iso, "source": "oracle restatement (liboracle.so, 64-bit edge keys)"}
    orc = Oracle()
    t0 = time.time()
    grad = orc.gradient(grid, s)
    print(f"schwarz_1024 gradient {time.time() - t0:.1f} s", flush=True)
    m = orc.dualmc(grid, s, grad, iso, True)
    del grad
    gc.collect()
    he = m["he"]
    out["dmc"] = {"nv": int(m["v"].shape[0]), "nq": int(m["quads"].shape[0]),
                  "nonmanifold": int(m["nonmanifold"]), "nbnd": int(m["bnd"].shape[0]),
                  "sha256_v_n_quads": digest((m["v"], "<f8"), (m["n"], "<f8"), (m["quads"], "<i4")),
                  "sha256_he_hev": digest((he, "<i4"), (m["he_v"], "<i4"))}
    print(f"schwarz_1024 dmc {time.time() - t0:.1f} s {out['dmc']}", flush=True)
    del m, he
    gc.collect()
    t0 = time.time()
    r = orc.tmc(grid, s, iso)
    out["tmc"] = {"nv": int(r["v"].shape[0]), "nt": int(r["tris"].shape[0]),
                  "sha256_v_n_tris": digest((r["v"], "<f8"), (r["n"], "<f8"), (r["tris"], "<i4"))}
    print(f"schwarz_1024 tmc {time.time() - t0:.1f} s {out['tmc']}", flush=True)
    return out


MAKERS = {"blobs_512": blobs_512, "schwarz_1024": schwarz_1024}


def main(names):
    known = {}
    if os.path.exists(OUT):
        with open(OUT) as fh:
            known = json.load(fh)
    for name in names:
        known[name] = MAKERS[name]()
        with open(OUT, "w") as fh:
            json.dump(known, fh, indent=1)
            fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:] or list(MAKERS))
