"""CPU checks that pin the oracle's topologically correct MC
(oracle/tmc_oracle.c) to the reference mc/MarchingCubes.cpp compiled from its
own sources (oracle/_ref/libtmcref.so): vertices, normals and triangles bit
for bit, on fields chosen to reach every branch of p_slice (random fields
full of ambiguous faces, tunnels and 12-vertex contours; small integer
fields whose face deciders hit the iso-value exactly)."""
import os

import numpy as np
import pytest

from oracle import TMC_REF_SO, Grid, Oracle, TmcReference


def _fields():
    rng = np.random.default_rng(17)
    out = []
    out.append(("uniform16", rng.uniform(0, 1, (16, 16, 16)), 0.5, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
    out.append(("normal_12x14x18", rng.standard_normal((12, 14, 18)), 0.1, (-1.0, 2.0, 0.5), (0.5, 0.25, 1.5)))
    out.append(("ints", rng.integers(0, 4, (14, 13, 12)).astype(np.float64), 1.5, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
    out.append(("ints_on_iso", rng.integers(0, 5, (14, 14, 14)).astype(np.float64), 2.0, (0.0, 0.0, 0.0),
                (1.0, 1.0, 1.0)))
    z, y, x = np.meshgrid(*[np.linspace(-1.6, 1.6, 40)] * 3, indexing="ij")
    g2 = 2 * y * (y * y - 3 * x * x) * (1 - z * z) + (x * x + y * y) ** 2 - (9 * z * z - 1) * (1 - z * z)
    out.append(("genus2_40", g2, 0.0, (-1.6, -1.6, -1.6), (3.2 / 39,) * 3))
    z, y, x = np.meshgrid(*[np.linspace(-3 * np.pi, 3 * np.pi, 33)] * 3, indexing="ij")
    out.append(("schwarz_33", np.cos(x) + np.cos(y) + np.cos(z), 0.2, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
    return out


@pytest.mark.skipif(not os.path.exists(TMC_REF_SO), reason="compiled reference (oracle/_ref) not built")
@pytest.mark.parametrize("case", _fields(), ids=lambda c: c[0])
def test_tmc_oracle_matches_reference(case):
    name, f, iso, origin, spacing = case
    g = Grid.make(f.shape, origin, spacing)
    exp = TmcReference().tmc(g, f, iso)
    got = Oracle().tmc(g, f, iso)
    assert len(got["v"]) == len(exp["v"]) and len(got["tris"]) == len(exp["tris"]), name
    assert np.array_equal(got["tris"], exp["tris"]), name
    assert np.array_equal(got["v"].view(np.uint64), exp["v"].view(np.uint64)), name
    assert np.array_equal(got["n"].view(np.uint64), exp["n"].view(np.uint64)), name


def test_tmc_oracle_reaches_the_special_branches():
    """The random fields above exercise the decider paths: count the cells
    with inner vertices (positions strictly inside a cell) and the
    ambiguous cells."""
    rng = np.random.default_rng(17)
    f = rng.uniform(0, 1, (16, 16, 16))
    got = Oracle().tmc(Grid.make(f.shape), f, 0.5)
    assert got["ambiguous"] > 100
    v = got["v"]
    inner = np.all(np.abs(v - np.round(v)) > 1e-12, axis=1)
    assert inner.sum() > 10


def test_tmc_oracle_closed_surface_invariants():
    """On a closed surface away from the grid border every edge is shared by
    exactly two triangles (t_mc does not orient them consistently; the
    reference leaves that to its optional connectivity pass), and V - E + F
    is the Euler characteristic (genus 2: -2)."""
    z, y, x = np.meshgrid(*[np.linspace(-2.0, 2.0, 48)] * 3, indexing="ij")
    f = 2 * y * (y * y - 3 * x * x) * (1 - z * z) + (x * x + y * y) ** 2 - (9 * z * z - 1) * (1 - z * z)
    assert min(f[0].min(), f[-1].min(), f[:, 0].min(), f[:, -1].min(), f[:, :, 0].min(), f[:, :, -1].min()) > 0
    got = Oracle().tmc(Grid.make(f.shape, (-2.0,) * 3, (4.0 / 47,) * 3), f, 0.0)
    t = got["tris"].astype(np.int64)
    he = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    key = np.minimum(he[:, 0], he[:, 1]) * (1 << 32) + np.maximum(he[:, 0], he[:, 1])
    _, count = np.unique(key, return_counts=True)
    assert np.all(count == 2)
    nv, ne, nf = len(got["v"]), len(count), len(t)
    assert nv - ne + nf == -2


def test_tmc_oracle_matches_fixtures():
    """The committed fixtures of the compiled reference
    (tests/golden/make_tmc_golden.py), checked without the reference."""
    from conftest import golden_files, grid_from, load_golden
    names = golden_files("tmc_")
    assert len(names) >= 4
    for name in names:
        z = load_golden(name)
        got = Oracle().tmc(grid_from(z["grid"]), z["field"], float(z["iso"]))
        assert np.array_equal(got["tris"], z["tris"]), name
        assert np.array_equal(got["v"].view(np.uint64), z["v"].view(np.uint64)), name
        assert np.array_equal(got["n"].view(np.uint64), z["n"].view(np.uint64)), name


@pytest.mark.skipif(not os.path.exists(TMC_REF_SO), reason="compiled reference (oracle/_ref) not built")
def test_dropin_orientation_matches_reference_connectivity(tmp_path):
    """tmc::MarchingCubes::connectivity() of include/tmc_dmc.hpp flips the
    same triangles as the reference's connectivity() (operator()'s c_flag),
    on meshes with many inconsistently oriented ambiguous cells."""
    import subprocess
    from conftest import ROOT
    exe = str(tmp_path / "orient")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "tmc_orient_probe.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ref = TmcReference()
    total = 0
    # integer fields with values on the iso-value (degenerate triangles,
    # which t_mc leaves inconsistently oriented: seeds found by search), and
    # the fields above (already consistent)
    flipping = []
    for seed in (43, 62, 67, 78, 83, 107):
        rng = np.random.default_rng(seed)
        sh = tuple(int(x) for x in rng.integers(4, 12, 3))
        flipping.append((f"ints_on_iso_seed{seed}", rng.integers(0, 3, sh).astype(np.float64), 1.0, (0.0,) * 3,
                         (1.0,) * 3))
    for name, f, iso, origin, spacing in flipping + list(_fields()):
        g = Grid.make(f.shape, origin, spacing)
        plain = ref.tmc(g, f, iso)
        exp = ref.tmc(g, f, iso, orient=True)
        t = plain["tris"]
        text = f"{len(plain['v'])} {len(t)}\n" + "".join(f"{a} {b} {c}\n" for a, b, c in t)
        r = subprocess.run([exe], input=text, capture_output=True, text=True)
        assert r.returncode == 0, name
        lines = r.stdout.split("\n")
        flips = int(lines[0])
        got = np.array([[int(x) for x in l.split()] for l in lines[1:] if l], np.int32).reshape(-1, 3)
        assert np.array_equal(got, exp["tris"]), name
        total += flips
    assert total > 100
