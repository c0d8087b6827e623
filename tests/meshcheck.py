"""Mesh comparisons and invariants shared by the CPU and GPU tests."""
import numpy as np


def same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        return bool(np.array_equal(a, b, equal_nan=True))
    return bool(np.array_equal(a, b))


def first_diff(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else (a == b)))
    if len(bad) == 0:
        return "equal"
    i = tuple(bad[0])
    return f"{len(bad)} differ, first at {i}: {a[i]!r} vs {b[i]!r}"


def assert_same(a, b, what):
    assert same(a, b), f"{what}: {first_diff(a, b)}"


def euler(nv, quads, he):
    """(E, chi) with E counted from halfedges: interior edges once, boundary once."""
    nq = len(quads)
    nb = int((he[:, 3] == -1).sum())
    E = (4 * nq + nb) // 2
    return E, nv - E + nq


def halfedge_invariants(nv, quads, he):
    """Structural invariants of a {origin, face, next, twin, flag} table:
    next cycles the quad, twin(twin(h)) = h, twins run opposite ways (the
    manifold ones) and share the vertex pair."""
    nq = len(quads)
    h = np.arange(4 * nq)
    assert np.array_equal(he[:, 0], quads.reshape(-1)), "origin"
    assert np.array_equal(he[:, 1], h // 4), "face"
    assert np.array_equal(he[:, 2], (h & ~3) | ((h + 1) & 3)), "next"
    tw = he[:, 3]
    has = tw >= 0
    assert np.all(tw[tw[has]] == h[has]), "twin(twin(h)) != h"
    dest = quads.reshape(-1)[he[:, 2]]
    orig = he[:, 0]
    pair_a = np.minimum(orig, dest)
    pair_b = np.maximum(orig, dest)
    assert np.all(pair_a[tw[has]] == pair_a[has]) and np.all(pair_b[tw[has]] == pair_b[has]), "twin vertex pair"
    manifold = has & (he[:, 4] == 0)
    assert np.all(orig[tw[manifold]] == dest[manifold]), "manifold twins must run opposite ways"
    return True


def valence_sum(nv, quads, he):
    """sum of reference valences = halfedges + boundary halfedges."""
    v = np.bincount(he[:, 0], minlength=nv)
    b = he[:, 3] == -1
    v += np.bincount(quads.reshape(-1)[he[b, 2]], minlength=nv)
    return int(v.sum())


def _min_angle(a, b, c):
    la = np.linalg.norm(b - a, axis=1)
    lb = np.linalg.norm(c - b, axis=1)
    lc = np.linalg.norm(a - c, axis=1)
    with np.errstate(all="ignore"):
        A = np.arccos(np.clip((lb * lb + lc * lc - la * la) / (2 * lb * lc), -1, 1))
        B = np.arccos(np.clip((la * la + lc * lc - lb * lb) / (2 * la * lc), -1, 1))
        C = np.arccos(np.clip((lb * lb + la * la - lc * lc) / (2 * lb * la), -1, 1))
    return np.minimum(np.minimum(A, B), C)


def triangle_mismatches(t_exp, t_act, quads, v, tol=1e-5):
    """collectTriangles splits each quad along the diagonal with the larger
    minimum angle, computed with float acos.  Different acosf implementations
    may disagree only where the two splits are within float rounding, so a
    split may differ from the expected one only where the two candidates'
    minimum angles agree to `tol`.  Returns the quads that break this."""
    t_exp = np.asarray(t_exp).reshape(-1, 6)
    t_act = np.asarray(t_act).reshape(-1, 6)
    assert t_exp.shape == t_act.shape
    diff = np.nonzero(np.any(t_exp != t_act, axis=1))[0]
    if len(diff) == 0:
        return diff
    q = quads[diff]
    p = [v[q[:, k]] for k in range(4)]
    b1 = np.minimum(_min_angle(p[0], p[1], p[2]), _min_angle(p[0], p[2], p[3]))
    c1 = np.minimum(_min_angle(p[1], p[3], p[0]), _min_angle(p[1], p[2], p[3]))
    near = np.abs(b1 - c1) <= tol
    # the only allowed outputs are the two splits of the same quad
    s1 = np.stack([q[:, 1], q[:, 3], q[:, 0], q[:, 1], q[:, 2], q[:, 3]], 1)
    s2 = np.stack([q[:, 0], q[:, 1], q[:, 2], q[:, 0], q[:, 2], q[:, 3]], 1)
    valid = np.all(t_act[diff] == s1, axis=1) | np.all(t_act[diff] == s2, axis=1)
    return diff[~(near & valid)]
