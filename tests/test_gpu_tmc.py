"""Parity of the HIP topologically correct MC (tmc_extract, through the C
ABI) with the oracle's restatement of the reference's t_mc
(oracle/tmc_oracle.c, itself pinned bit for bit against
mc/MarchingCubes.cpp): vertex positions and normals (FP64), triangles and
the reference's vertex numbering must be bit-exact."""
import sys

import numpy as np
import pytest

from conftest import golden_files, grid_from, load_golden
from oracle import Grid as OGrid

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(engine, oracle, f, iso, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), what=""):
    og = OGrid.make(f.shape, origin, spacing)
    exp = oracle.tmc(og, f.astype(np.float64), iso)
    g = tmc_amd.Grid.make(f.shape, origin, spacing)
    got = engine.tmc(f, iso, g)
    assert got["nv"] == len(exp["v"]) and got["nt"] == len(exp["tris"]), (what, got["nv"], len(exp["v"]),
                                                                           got["nt"], len(exp["tris"]))
    assert got["ambiguous_cells"] == exp["ambiguous"] and got["face_failures"] == exp["face_failures"], what
    assert np.array_equal(got["tris"], exp["tris"]), what
    assert np.array_equal(got["v"].view(np.uint64), exp["v"].view(np.uint64)), what
    assert np.array_equal(got["n"].view(np.uint64), exp["n"].view(np.uint64)), what
    return got


def _smooth(kind, n, lo, hi):
    z, y, x = np.meshgrid(*[np.linspace(lo, hi, n)] * 3, indexing="ij")
    if kind == "genus2":
        return 2 * y * (y * y - 3 * x * x) * (1 - z * z) + (x * x + y * y) ** 2 - (9 * z * z - 1) * (1 - z * z)
    return np.cos(x) + np.cos(y) + np.cos(z)


@pytest.mark.parametrize("name", golden_files("tmc_"))
def test_tmc_matches_reference_fixture(engine, name):
    """Fixtures of the reference's t_mc compiled from mc/MarchingCubes.cpp."""
    z = load_golden(name)
    g = grid_from(z["grid"])
    _say(f"[gpu] tmc fixture {name}")
    got = engine.tmc(z["field"], float(z["iso"]), g)
    assert np.array_equal(got["tris"], z["tris"]), name
    assert np.array_equal(got["v"].view(np.uint64), z["v"].view(np.uint64)), name
    assert np.array_equal(got["n"].view(np.uint64), z["n"].view(np.uint64)), name


CASES = [
    ("uniform16", lambda r: r.uniform(0, 1, (16, 16, 16)), 0.5),
    ("normal_12x14x18", lambda r: r.standard_normal((12, 14, 18)), 0.1),
    ("ints_1.5", lambda r: r.integers(0, 4, (14, 13, 12)).astype(np.float64), 1.5),
    ("ints_on_iso", lambda r: r.integers(0, 5, (14, 14, 14)).astype(np.float64), 2.0),
    ("tiny_2x2x2", lambda r: r.uniform(0, 1, (2, 2, 2)), 0.5),
    ("slab_3x40x2", lambda r: r.uniform(0, 1, (3, 40, 2)), 0.5),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_tmc_matches_oracle_f64(engine, oracle, case):
    name, make, iso = case
    f = make(np.random.default_rng(17))
    _say(f"[gpu] tmc {name} {f.shape} iso={iso}")
    _check(engine, oracle, f, iso, what=name)


@pytest.mark.parametrize("seed", range(6))
def test_tmc_matches_oracle_random_f32(engine, oracle, seed):
    """Float fields (the engine reads them as float, the oracle gets the
    same values as doubles), odd shapes, origins and spacings."""
    rng = np.random.default_rng(100 + seed)
    sh = tuple(int(x) for x in rng.integers(2, 23, size=3))
    f = rng.standard_normal(sh).astype(np.float32)
    origin = tuple(float(x) for x in rng.uniform(-3, 3, 3))
    spacing = tuple(float(x) for x in rng.uniform(0.1, 2, 3))
    _say(f"[gpu] tmc random f32 seed={seed} shape={sh}")
    _check(engine, oracle, f, 0.05, origin, spacing, what=f"seed {seed}")


@pytest.mark.parametrize("kind,n,iso", [("genus2", 96, 0.0), ("schwarz", 130, 0.3)])
def test_tmc_matches_oracle_surfaces(engine, oracle, kind, n, iso):
    lo, hi = (-2.0, 2.0) if kind == "genus2" else (-4 * np.pi, 4 * np.pi)
    f = _smooth(kind, n, lo, hi).astype(np.float32)
    _say(f"[gpu] tmc {kind} {n}^3")
    got = _check(engine, oracle, f, iso, (lo,) * 3, ((hi - lo) / (n - 1),) * 3, what=kind)
    if kind == "genus2":
        # closed surface inside the grid: V - E + F = -2, every edge in two triangles
        t = got["tris"].astype(np.int64)
        he = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
        key = np.minimum(he[:, 0], he[:, 1]) * (1 << 32) + np.maximum(he[:, 0], he[:, 1])
        _, count = np.unique(key, return_counts=True)
        assert np.all(count == 2)
        assert got["nv"] - len(count) + got["nt"] == -2


def test_tmc_device_field_and_repeat(engine, oracle):
    """The field already in HBM (DMC_FIELD_ON_DEVICE), twice in a row with a
    smaller grid second: the same results as from host memory."""
    rng = np.random.default_rng(9)
    for sh in ((20, 21, 22), (9, 8, 7)):
        f = rng.uniform(-1, 1, sh).astype(np.float32)
        exp = oracle.tmc(OGrid.make(sh), f.astype(np.float64), 0.0)
        buf = engine.alloc(f.nbytes).upload(f)
        got = engine.tmc(buf, 0.0, shape=sh, dtype=np.float32)
        buf.free()
        assert np.array_equal(got["tris"], exp["tris"])
        assert np.array_equal(got["v"].view(np.uint64), exp["v"].view(np.uint64))


def test_tmc_then_dmc_on_one_context(engine, oracle):
    """A DMC extraction after a TMC one on the same context is unaffected."""
    rng = np.random.default_rng(3)
    f = rng.uniform(0, 1, (12, 12, 12)).astype(np.float32)
    engine.tmc(f, 0.5)
    m = engine.extract(f, 0.5, flags=tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES)
    got = engine.download()
    engine.extract(f, 0.5, flags=tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES)
    again = engine.download()
    assert m.nv == len(got["v"])
    for k in ("v", "quads", "tris"):
        assert np.array_equal(got[k], again[k])


def _ushort_volume(path, f, spacing):
    with open(path, "wb") as fh:
        np.array([f.shape[2], f.shape[1], f.shape[0]], np.uint16).tofile(fh)
        np.array(spacing, np.float32).tofile(fh)
        np.ascontiguousarray(f, np.uint16).tofile(fh)


def _f7(x):
    """std::fixed with precision 7 as the C library writes it (a NaN with its
    sign bit set is "-nan")."""
    x = float(x)
    if x != x:
        return "-nan" if np.signbit(x) else "nan"
    return "%.7f" % x


def _obj_text(v, n, t):
    out = ["# Topologically correct and manifold isosurface\n"]
    out += ["v %s %s %s\n" % (_f7(p[0]), _f7(p[1]), _f7(p[2])) for p in v]
    out += ["vn %s %s %s\n" % (_f7(q[0]), _f7(q[1]), _f7(q[2])) for q in n]
    out += ["f %d//%d %d//%d %d//%d\n" % (a + 1, a + 1, b + 1, b + 1, c + 1, b + 1) for a, b, c in t]
    return "".join(out)


def _off_text(v, t):
    from oracle.diagnostics import fmt_g
    out = ["OFF\n", "%d %d 0\n" % (len(v), len(t))]
    out += ["%s %s %s\n" % (fmt_g(p[0]), fmt_g(p[1]), fmt_g(p[2])) for p in v]
    out += ["3 %d %d %d\n" % (a, b, c) for a, b, c in t]
    return "".join(out)


def test_tmc_cpp_dropin_driver(tmp_path):
    """mc/tmc.cpp's driver against include/tmc_dmc.hpp: tmc::MarchingCubes
    reads a ushort volume file, extracts, orients (c_flag) and writes OBJ and
    OFF text identical to what the reference's t_mc + connectivity() mesh
    gives in the reference's formats (checked against the compiled reference
    where it is built, else against the oracle without orientation)."""
    import os
    import subprocess
    from conftest import ROOT
    from oracle import TMC_REF_SO, TmcReference
    exe = str(tmp_path / "tmc_dropin")
    libdir = os.path.join(ROOT, "tmc_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "tmc_dropin_main.cpp"), "-L" + libdir, "-ltmc_dmc",
                        "-Wl,-rpath," + libdir, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    z, y, x = np.meshgrid(np.linspace(-2, 2, 22), np.linspace(-2, 2, 25), np.linspace(-2, 2, 28), indexing="ij")
    g2 = 2 * y * (y * y - 3 * x * x) * (1 - z * z) + (x * x + y * y) ** 2 - (9 * z * z - 1) * (1 - z * z)
    smooth = np.clip((g2 + 4) * 3000, 0, 65535).astype(np.uint16)
    rng = np.random.default_rng(43)
    sh = tuple(int(v) for v in rng.integers(4, 12, 3))
    ints = rng.integers(0, 3, sh).astype(np.uint16)
    have_ref = os.path.exists(TMC_REF_SO)
    for name, f, iso, sp, cflag in (("smooth", smooth, 12000.5, (0.1, 0.2, 0.3), 1), ("ints", ints, 1.0, (1, 1, 1), 1),
                                    ("ints_plain", ints, 1.0, (0.5, 0.5, 0.5), 0)):
        vol = str(tmp_path / (name + ".bin"))
        _ushort_volume(vol, f, sp)
        obj, off = str(tmp_path / (name + ".obj")), str(tmp_path / (name + ".off"))
        _say(f"[gpu] tmc drop-in {name} c_flag={cflag}")
        r = subprocess.run([exe, vol, repr(iso), str(cflag), obj, off], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        d = [float(s_) * (n - 1.0) / (n - 1) for s_, n in zip(np.float32(sp).astype(np.float64),
                                                               (f.shape[2], f.shape[1], f.shape[0]))]
        g = OGrid.make(f.shape, (0.0, 0.0, 0.0), d)
        if have_ref:
            exp = TmcReference().tmc(g, f.astype(np.float64), iso, orient=bool(cflag))
        else:
            if cflag:
                continue
            from oracle import Oracle
            exp = Oracle().tmc(g, f.astype(np.float64), iso)
        assert open(obj).read() == _obj_text(exp["v"], exp["n"], exp["tris"]), name
        assert open(off).read() == _off_text(exp["v"], exp["tris"]), name
