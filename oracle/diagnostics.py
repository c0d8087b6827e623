"""CPU restatements of the reference's mesh diagnostics and OBJ output --
TEST INFRASTRUCTURE ONLY (numpy; the engine never imports this).

Each function states the reference behaviour it follows and how it is
pinned (tests/test_diagnostics_oracle.py):

* ``mesh_edges``        dmc/HalfedgeMesh.cu:430-450 with the edge key of
                        dmc/EdgeHashTable.h:101-127 ((min, max) vertex pair).
                        Pinned by Euler's formula on closed and open meshes
                        and by brute force on small quad sets.
* ``valence_histogram`` dmc/VertexValence.cu:1-125 (one count per key and
                        end vertex, 100 bins, trimmed after the last
                        non-empty bin above 0).  Pinned by sum = 2 E and a
                        regular grid's known distribution.
* ``quality_quads`` / ``quality_tris``
                        dmc/ElementQuality.h:65-91 in float32 with the
                        reference's operand order.  Pinned by closed forms
                        (square, rectangle, equilateral and right triangles).
* ``obj_text``          cpp_dmc/main.cpp:35-87; the C++ stream's default
                        double format is printf's "%g" (pinned against a
                        g++-built std::ostream on edge values).
"""
import numpy as np


def _keys(a, b):
    lo = np.minimum(a, b).astype(np.int64)
    hi = np.maximum(a, b).astype(np.int64)
    return (lo << 32) | hi


def halfedge_ends(faces):
    """(origin, destination) of halfedge c*f + k of each polygon face."""
    faces = np.asarray(faces, np.int64)
    k = faces.shape[1]
    return faces.reshape(-1), np.roll(faces, -1, axis=1).reshape(-1)


def mesh_edges(quads):
    """Unique undirected edges of a quad mesh, each as the (origin,
    destination) of the first halfedge (in halfedge order) with its vertex
    pair, in that order."""
    o, d = halfedge_ends(quads)
    if o.size == 0:
        return np.zeros((0, 2), np.int32)
    _, first = np.unique(_keys(o, d), return_index=True)
    first.sort()
    return np.stack([o[first], d[first]], axis=1).astype(np.int32)


def vertex_valence(nv, quads):
    """Distinct edges at each vertex: every unique pair adds one to both of
    its end vertices (a loop edge twice to its vertex)."""
    e = mesh_edges(quads).astype(np.int64)
    val = np.zeros(nv, np.int64)
    np.add.at(val, e[:, 0], 1)
    np.add.at(val, e[:, 1], 1)
    return val


def valence_histogram(nv, quads, bins=100):
    """(hist trimmed like the reference's, number of vertices beyond the
    last bin)."""
    val = vertex_valence(nv, quads)
    h = np.bincount(np.minimum(val, bins), minlength=bins + 1)
    hist, beyond = h[:bins], int(h[bins])
    nz = np.nonzero(hist[1:])[0]
    n = int(nz[-1]) + 2 if nz.size else 1
    return hist[:n].astype(np.uint32), beyond


def _cross(a, b):
    # the reference's component formulas (x has the opposite sign)
    return np.stack([a[..., 2] * b[..., 1] - a[..., 1] * b[..., 2],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _norm2(v):
    return v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2]


def _norm(v):
    return np.sqrt(_norm2(v))


def quality_quads(positions, quads):
    """Least corner ratio 2 |e1 x e2| / (|e1|^2 + |e2|^2) of each quad, in
    float32 on the positions rounded to float."""
    v = np.asarray(positions, np.float64).astype(np.float32)
    q = np.asarray(quads, np.int64)
    two = np.float32(2)
    with np.errstate(all="ignore"):
        m = None
        for c in range(4):
            p = v[q[:, c]]
            e1 = v[q[:, (c + 1) % 4]] - p
            e2 = v[q[:, (c + 3) % 4]] - p
            r = (two * _norm(_cross(e1, e2))) / (_norm2(e1) + _norm2(e2))
            m = r if m is None else np.where(m < r, m, r)
    return m.astype(np.float32) if m is not None else np.zeros(0, np.float32)


def quality_tris(positions, tris):
    """2 sqrt(3) |e1 x e2| / (l0 + l1 + l2) of each triangle (squared edge
    lengths), 0 when an edge has length 0; float32."""
    v = np.asarray(positions, np.float64).astype(np.float32)
    t = np.asarray(tris, np.int64)
    factor = np.float32(2) * np.sqrt(np.float32(3))
    v0, v1, v2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    with np.errstate(all="ignore"):
        a = _norm(_cross(v1 - v0, v2 - v0))
        l0, l1, l2 = _norm2(v1 - v0), _norm2(v2 - v1), _norm2(v0 - v2)
        q = (factor * a) / (l0 + l1 + l2)
    ok = (l0 != 0) & (l1 != 0) & (l2 != 0)
    return np.where(ok, q, np.float32(0)).astype(np.float32)


def fmt_g(x):
    """printf "%g" as the C library writes it: like Python's, except that a
    NaN with its sign bit set is "-nan"."""
    x = float(x)
    if x != x:
        return "-nan" if np.signbit(x) else "nan"
    return "%g" % x


def obj_text(positions, normals, faces, triangles):
    """OBJ text as cpp_dmc's main program writes it."""
    out = ["#Dual Marching Cubes, triangle mesh\n" if triangles else "#Dual Marching Cubes, quad only mesh\n"]
    out += ["v %s %s %s\n" % (fmt_g(p[0]), fmt_g(p[1]), fmt_g(p[2])) for p in np.asarray(positions, np.float64)]
    out += ["vn %s %s %s\n" % (fmt_g(n[0]), fmt_g(n[1]), fmt_g(n[2])) for n in np.asarray(normals, np.float64)]
    for f in np.asarray(faces, np.int64):
        out.append("f" + "".join(" %d//%d" % (x + 1, x + 1) for x in f) + "\n")
    return "".join(out)
