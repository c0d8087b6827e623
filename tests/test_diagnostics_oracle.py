"""CPU checks that pin oracle/diagnostics.py (the restatements the GPU
diagnostics are compared with) by brute force, closed forms and invariants,
and the number format of the OBJ writer against a g++-built std::ostream."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import diagnostics as D


def _brute_edges(faces):
    """First-occurrence edge list by plain Python loops."""
    seen, out = set(), []
    for f in faces:
        k = len(f)
        for c in range(k):
            a, b = int(f[c]), int(f[(c + 1) % k])
            key = (min(a, b), max(a, b))
            if key not in seen:
                seen.add(key)
                out.append((a, b))
    return np.array(out, np.int32).reshape(-1, 2)


def _brute_valence(nv, faces):
    keys = {tuple(sorted(map(int, e))) for e in _brute_edges(faces)}
    val = [0] * nv
    for a, b in keys:
        val[a] += 1
        val[b] += 1
    return val


def _cube():
    # closed quad surface of a cube: V 8, E 12, F 6 (Euler characteristic 2)
    return np.array([[0, 2, 3, 1], [4, 5, 7, 6], [0, 1, 5, 4], [2, 6, 7, 3], [0, 4, 6, 2], [1, 3, 7, 5]], np.int32)


def _grid(n):
    # n x n open quad grid on (n+1)^2 vertices
    idx = lambda i, j: i + (n + 1) * j
    return np.array([[idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)] for j in range(n) for i in range(n)],
                    np.int32)


def test_edges_brute_force_and_euler():
    rng = np.random.default_rng(11)
    for it in range(30):
        nv = int(rng.integers(3, 12))
        faces = rng.integers(0, nv, size=(int(rng.integers(1, 20)), 4)).astype(np.int32)   # degenerate too
        assert np.array_equal(D.mesh_edges(faces), _brute_edges(faces))
        assert list(D.vertex_valence(nv, faces)) == _brute_valence(nv, faces)
        tris = faces[:, :3]
        assert np.array_equal(D.mesh_edges(tris), _brute_edges(tris))
        assert list(D.vertex_valence(nv, tris)) == _brute_valence(nv, tris)
    cube = _cube()
    e = D.mesh_edges(cube)
    assert 8 - len(e) + 6 == 2
    assert list(D.vertex_valence(8, cube)) == [3] * 8
    n = 5
    g = _grid(n)
    assert len(D.mesh_edges(g)) == 2 * n * (n + 1)
    assert (n + 1) ** 2 - len(D.mesh_edges(g)) + n * n == 1


def test_valence_histogram_known_distributions():
    h, beyond = D.valence_histogram(8, _cube())
    assert list(h) == [0, 0, 0, 8] and beyond == 0
    n = 6
    h, beyond = D.valence_histogram((n + 1) ** 2, _grid(n))
    # 4 corners of valence 2, 4(n-1) border vertices of 3, (n-1)^2 of 4
    assert list(h) == [0, 0, 4, 4 * (n - 1), (n - 1) ** 2] and beyond == 0
    # sum of valences is twice the edge count; trimmed after the last non-empty bin
    val = D.vertex_valence((n + 1) ** 2, _grid(n))
    assert val.sum() == 2 * len(D.mesh_edges(_grid(n)))
    h, _ = D.valence_histogram(5, np.zeros((0, 4), np.int32))
    assert list(h) == [5]
    # a loop edge counts twice at its vertex; valences past the last bin
    assert list(D.vertex_valence(2, np.array([[0, 0, 1, 1]], np.int32))) == [3, 3]
    star = np.array([[0, k, k + 1, k + 2] for k in range(1, 300, 2)], np.int32)
    h, beyond = D.valence_histogram(302, star, bins=100)
    assert beyond == 1 and h.sum() == 301


def test_quality_closed_forms():
    sq = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float64)
    assert D.quality_quads(sq, np.array([[0, 1, 2, 3]]))[0] == np.float32(1)
    rect = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], np.float64)
    assert D.quality_quads(rect, np.array([[0, 1, 2, 3]]))[0] == np.float32(0.8)   # 2*2/(4+1) at every corner
    # a right corner and a flat one: the least ratio wins
    kite = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0]], np.float64)
    assert D.quality_quads(kite, np.array([[0, 1, 2, 3]]))[0] == np.float32(0)
    eq = np.array([[0, 0, 0], [1, 0, 0], [0.5, math.sqrt(3) / 2, 0]], np.float64)
    assert abs(float(D.quality_tris(eq, np.array([[0, 1, 2]]))[0]) - 1.0) < 2e-7
    rt = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    assert abs(float(D.quality_tris(rt, np.array([[0, 1, 2]]))[0]) - math.sqrt(3) / 2) < 1e-7
    assert D.quality_tris(rt, np.array([[0, 0, 2]]))[0] == 0   # a zero-length edge


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_obj_number_format_is_the_streams(tmp_path):
    exe = str(tmp_path / "fmt")
    r = subprocess.run(["g++", "-O1", os.path.join(ROOT, "tests", "cpp", "ostream_format.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(5)
    vals = [0.0, -0.0, 1.0, -1.0, 0.5, 1e-5, 1.5e-5, 123456.0, 1234567.0, 999999.5, 999999.4, 0.1 + 0.2,
            3.14159265358979, -2.5e-7, 1e-300, 7.0e22, 0.000123456789, 9.9999995e-5]
    vals += list(rng.normal(0, 10, 200)) + list(rng.uniform(-1, 1, 200) * 10.0 ** rng.integers(-8, 8, 200))
    vals += [float("inf"), float("-inf"), float("nan"), -float("nan")]
    text = "\n".join(("-nan" if np.signbit(v) else "nan") if v != v else float(v).hex() for v in vals) + "\n"
    r = subprocess.run([exe], input=text, capture_output=True, text=True)
    assert r.returncode == 0
    got = r.stdout.splitlines()
    assert got == [D.fmt_g(v) for v in vals]


def test_obj_text_layout():
    v = np.array([[0.5, -0.0, 1e-5], [1.0, 2.0, 3.0]])
    n = np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])
    t = D.obj_text(v, n, np.array([[0, 1, 1]]), True)
    assert t == ("#Dual Marching Cubes, triangle mesh\nv 0.5 -0 1e-05\nv 1 2 3\nvn 0 0 1\nvn 0 1 0\n"
                 "f 1//1 2//2 2//2\n")
    q = D.obj_text(v, n, np.array([[0, 1, 1, 0]]), False)
    assert q.startswith("#Dual Marching Cubes, quad only mesh\n") and q.endswith("f 1//1 2//2 2//2 1//1\n")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tetrahedron_surface_closed_forms(tmp_path):
    """Volumes' Tetrahedron (cpp_dmc/Volumes.h:248-335) in include/tmc_dmc.hpp:
    the signed distance to the nearest face of the tetrahedron
    (-.8,-.8,-.8), (.8,-.8,-.8), (.8,.8,-.8), (0,0,.8), positive outside."""
    exe = str(tmp_path / "probe")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "volumes_probe.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    tet = 9   # Volumes::Surface::Tetrahedron
    pts = [(0.3, -0.3, -0.9), (0.3, -0.3, -0.7), (0.0, 0.0, 0.9), (-0.8, -0.8, -0.9), (0.3, -0.3, -0.8)]
    r = subprocess.run([exe], input="".join("%d %r %r %r\n" % (tet, *p) for p in pts), capture_output=True,
                       text=True)
    got = [float(x) for x in r.stdout.split()]
    # below the base face (inside its triangle): +0.1; above it, inside the
    # solid: -0.1 (the base is nearest); above the apex: distance to the
    # apex; below a base vertex: distance to that vertex; on the base: 0
    assert abs(got[0] - 0.1) < 1e-12 and abs(got[1] + 0.1) < 1e-12
    assert abs(got[2] - 0.1) < 1e-12 and abs(got[3] - 0.1) < 1e-12 and abs(got[4]) < 1e-12
