"""Full-size workloads on the GPU, checked against known answers and through
properties that hold at any size.  Known answers (tests/golden/
fullsize_digests.json, tests/golden/make_fullsize_digests.py): counts and
SHA-256 digests of positions, normals, quads (and for 1024^3 the halfedge
rows and he_v) or triangles, written for the 512^3 blobs by the compiled
references and for the 1024^3 Schwarz-P field by the oracle's C
restatement (the compiled references' int keys overflow at that size).
Properties:

* 1024^3 Schwarz-P + noise and 512^3 Gaussian blobs + noise, Dual Marching
  Cubes with every simplification: halfedge structure (next cycles, twin
  symmetry, twins over the same vertex pair, manifold twins opposite),
  mesh boundary only in the grid's boundary layer (no holes inside), unit
  normals, edge count from the halfedges equal to the unique-edge
  extraction, and the same bytes under another launch shape.
* 512^3 blobs, topologically correct MC: every undirected edge in one or
  two triangles (more only at inner vertices of ambiguous cells, as in the
  reference), the one-triangle edges on the grid's boundary faces, and
  every vertex on a grid edge but the inner vertices of ambiguous cells.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from meshcheck import assert_same, euler, halfedge_invariants

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402
from tmc_amd import SIMPLIFY_ALL, TRIANGLES, workloads  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = tmc_amd.Engine(0)
    yield e
    e.close()


# counts and digests per workload (tests/golden/make_fullsize_digests.py)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_digests.json")) as _fh:
    KNOWN = json.load(_fh)


def _digest(*arrays):
    h = hashlib.sha256()
    for a, dt in arrays:
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
    return h.hexdigest()


def _box(g):
    lo = np.array([g.x0, g.y0, g.z0])
    hi = lo + np.array([(g.nx - 1) * g.dx, (g.ny - 1) * g.dy, (g.nz - 1) * g.dz])
    step = np.array([g.dx, g.dy, g.dz])
    return lo, hi, step


@pytest.mark.parametrize("which", ["schwarz_1024", "blobs_512"])
def test_dmc_full_size_properties(eng, which):
    _say(f"[gpu] full size DMC {which}: generating")
    f, g, iso = workloads.CONFIGS[which]()
    m = eng.extract(f, iso, g, flags=SIMPLIFY_ALL | TRIANGLES)
    s = m.summary()
    _say(f"[gpu] {which}: V {s['nv']} Q {s['nq']} non-manifold {s['nonmanifold_edges']} errors {s['edge_errors']}")
    assert s["nv"] > 0 and s["nq"] > 0 and s["invalid_refs"] == 0 and s["edge_errors"] == 0
    d = eng.download(halfedges=True)
    quads, he, v, n = d["quads"].reshape(-1, 4), d["he"].reshape(-1, 5), d["v"].reshape(-1, 3), d["n"].reshape(-1, 3)
    nv = s["nv"]
    k = KNOWN[which]["dmc"]
    assert (nv, s["nq"]) == (k["nv"], k["nq"])
    assert _digest((v, "<f8"), (n, "<f8"), (quads, "<i4")) == k["sha256_v_n_quads"], \
        f"positions, normals or quads differ from the {KNOWN[which]['source']}"
    if "sha256_he_hev" in k:
        assert (s["nonmanifold_edges"], s["nbnd"]) == (k["nonmanifold"], k["nbnd"])
        assert _digest((he, "<i4"), (d["he_v"], "<i4")) == k["sha256_he_hev"], \
            f"halfedges differ from the {KNOWN[which]['source']}"
    _say(f"[gpu] {which}: bit-exact against the {KNOWN[which]['source']}")
    assert quads.min() >= 0 and quads.max() < nv
    assert np.all(np.bincount(quads.reshape(-1), minlength=nv) > 0), "unreferenced vertex"
    halfedge_invariants(nv, quads, he)
    # the mesh boundary lies in the grid's boundary layer: a missing quad
    # inside the volume would leave a boundary edge there
    bnd = (he[:, 3] == -1) & (he[:, 4] == 0)
    lo, hi, step = _box(g)
    vb = v[he[bnd, 0]]
    near = np.any((vb < lo + 2 * step) | (vb > hi - 2 * step), axis=1)
    assert np.all(near), f"{int((~near).sum())} boundary halfedges inside the volume"
    # vertices stay inside the grid's box; normals are unit vectors
    assert np.all(v >= lo - 1e-9) and np.all(v <= hi + 1e-9)
    ln = np.linalg.norm(n, axis=1)
    ok = np.isfinite(ln)
    assert ok.mean() > 0.999999
    assert np.all(np.abs(ln[ok] - 1) < 1e-12)
    # unique undirected edges: one per vertex pair; with only two-face and
    # boundary edges that is the halfedge count (4Q + boundary) / 2
    e = eng.edges()
    if s["nonmanifold_edges"] == 0 and int((he[:, 4] != 0).sum()) == 0:
        E, chi = euler(nv, quads, he)
        assert len(e) == E, f"edges {len(e)} vs {E} from the halfedges"
        _say(f"[gpu] {which}: E {E}, chi {chi}")
    # the same bytes under another launch shape
    keep = {k: d[k] for k in ("v", "n", "quads", "tris", "he")}
    del d, he, quads, v, n, vb
    before = eng.launch
    eng.launch = {"grid_cap": 4096, "sweep_waves": 8192}
    try:
        m2 = eng.extract(f, iso, g, flags=SIMPLIFY_ALL | TRIANGLES)
        assert m2.summary() == s
        d2 = eng.download(halfedges=True)
        for k, a in keep.items():
            assert_same(d2[k], a, f"{which} under another launch shape: {k}")
    finally:
        eng.launch = before


@pytest.mark.parametrize("which", ["blobs_512", "schwarz_1024"])
def test_tmc_full_size_properties(eng, which):
    if "tmc" not in KNOWN.get(which, {}):
        pytest.skip(f"no known TMC answer for {which}")
    _say(f"[gpu] full size TMC {which}: generating")
    f, g, iso = workloads.CONFIGS[which]()
    r = eng.tmc(f, iso, g)
    nv, nt = r["nv"], r["nt"]
    _say(f"[gpu] TMC {which}: V {nv} T {nt} ambiguous {r['ambiguous_cells']} failures {r['face_failures']}")
    assert nv > 0 and nt > 0
    t = r["tris"].reshape(-1, 3).astype(np.int64)
    v = r["v"].reshape(-1, 3)
    k = KNOWN[which]["tmc"]
    assert (nv, nt) == (k["nv"], k["nt"])
    assert _digest((v, "<f8"), (r["n"], "<f8"), (r["tris"], "<i4")) == k["sha256_v_n_tris"], \
        f"positions, normals or triangles differ from the {KNOWN[which]['source']}"
    assert t.min() >= 0 and t.max() < nv
    assert np.all(np.bincount(t.reshape(-1), minlength=nv) > 0), "unreferenced vertex"
    # every vertex on a grid edge (two coordinates on grid lines), except
    # the inner vertices the asymptotic decider's slicing adds inside
    # ambiguous cells
    lo, hi, step = _box(g)
    x = (v - lo) / step
    on = np.abs(x - np.round(x)) < 1e-6
    innerv = on.sum(axis=1) < 2
    inner = int(innerv.sum())
    assert inner <= 8 * r["ambiguous_cells"], f"{inner} vertices off the grid edges"
    # edge use: each undirected edge in two triangles, or in one on the
    # grid's boundary faces (both ends on the same face plane); a cell the
    # decider cannot triangulate leaves a hole, counted in face_failures
    a = t.reshape(-1)
    b = t[:, [1, 2, 0]].reshape(-1)
    key = np.minimum(a, b) * nv + np.maximum(a, b)
    del a, b
    key.sort()
    starts = np.concatenate(([0], np.nonzero(np.diff(key))[0] + 1))
    counts = np.diff(np.concatenate((starts, [len(key)])))
    # more than two triangles only at an inner vertex, where the reference's
    # slicing of an ambiguous cell may join two contours (the compiled mc/
    # t_mc gives the same 4 such edges on this field)
    many = key[starts[counts > 2]]
    assert len(many) <= 1e-5 * len(starts), f"{len(many)} edges in more than two triangles"
    assert np.all(innerv[many // nv] | innerv[many % nv]), "an edge in more than two triangles off inner vertices"
    one = key[starts[counts == 1]]
    p, q = v[one // nv], v[one % nv]
    tol = 1e-9
    face = np.zeros(len(one), bool)
    for ax in range(3):
        for plane in (lo[ax], hi[ax]):
            face |= (np.abs(p[:, ax] - plane) < tol) & (np.abs(q[:, ax] - plane) < tol)
    if r["face_failures"] == 0:
        assert np.all(face), f"{int((~face).sum())} open edges inside the volume"
    _say(f"[gpu] TMC {which}: {len(starts)} edges, {len(one)} on the boundary, {len(many)} at inner vertices")
