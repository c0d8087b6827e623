"""Parity of the HIP engine (through the C ABI) with the oracle and with the
golden fixtures of the compiled reference.

Integer and topological outputs (quads, halfedges {origin, face, next,
twin, flag}, he_v, boundary vertices, counts) must be bit-exact.  Positions
and normals are FP64 and must be bit-exact as well (the engine reproduces
the reference's rounding points; the stated tolerance of 1e-6 grid units is
therefore met with margin 0).  Triangles may differ only where the two
diagonal splits tie to 1e-5 rad in minimum angle (float acos; see
meshcheck.triangle_mismatches).
"""
import sys

import numpy as np
import pytest

from conftest import golden_files, grid_from, load_golden
from meshcheck import assert_same, euler, halfedge_invariants, triangle_mismatches

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402
from tmc_amd import SIMPLIFY_3333, SIMPLIFY_3X3Y_COLOR, SIMPLIFY_3X3Y_ISOLATED, SIMPLIFY_ALL, TRIANGLES  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def engine_mesh(engine, field, grid, iso, flags, grad=None):
    engine.extract(field, iso, grid, flags=flags, gradients=grad)
    return engine.download()


def check_mesh(got, exp, what, tris=True):
    for k in ("v", "n", "quads", "he", "he_v", "bnd"):
        assert_same(got[k], exp[k], f"{what}:{k}")
    if tris:
        bad = triangle_mismatches(exp["tris"], got["tris"], exp["quads"], exp["v"])
        assert len(bad) == 0, f"{what}: triangles differ off ties at quads {bad[:10]}"


@pytest.mark.parametrize("name", golden_files("dualmc_"))
def test_engine_matches_reference_fixture(engine, name):
    _say(f"[gpu] fixture {name}")
    z = load_golden(name)
    g = grid_from(z["grid"])
    flags = TRIANGLES | (SIMPLIFY_ALL if int(z["sflag"]) else 0)
    grad = z["grad"] if z["grad"].size else None
    got = engine_mesh(engine, z["field"], g, float(z["iso"]), flags, grad)
    check_mesh(got, z, name)


def _fields(seed, count, max_n=14):
    rng = np.random.default_rng(seed)
    for it in range(count):
        kind = it % 5
        sh = tuple(int(x) for x in rng.integers(3, max_n, size=3))
        if kind == 0:
            f, iso = rng.integers(-2, 3, size=sh).astype(np.float32), float(rng.integers(-1, 2))
        elif kind == 1:
            f, iso = rng.standard_normal(sh).astype(np.float32), float(rng.normal() * 0.3)
        elif kind == 2:
            f, iso = (rng.random(sh) < 0.5).astype(np.float32), 0.5
        elif kind == 3:
            f, iso = rng.standard_normal(sh), 0.3
        else:
            f, iso = rng.standard_normal(sh).astype(np.float32), 0.3
        g = tmc_amd.Grid.make(sh, tuple(rng.normal(size=3)), tuple(rng.uniform(0.1, 2, size=3)))
        yield it, f, g, iso


FLAG_SETS = [0, SIMPLIFY_3X3Y_COLOR, SIMPLIFY_3X3Y_ISOLATED, SIMPLIFY_3333,
             SIMPLIFY_3X3Y_COLOR | SIMPLIFY_3X3Y_ISOLATED, SIMPLIFY_3X3Y_COLOR | SIMPLIFY_3333,
             SIMPLIFY_3X3Y_ISOLATED | SIMPLIFY_3333, SIMPLIFY_ALL]


def test_engine_matches_oracle_random_fields(engine, oracle):
    """Random, tie-heavy, binary, FP64 and non-float-iso fields, every
    combination of simplification passes."""
    nfail = nnm = nbnd = 0
    for it, f, g, iso in _fields(3, 60):
        s = np.asarray(f, np.float64).ravel()
        grad = oracle.gradient(g, s)
        fl = FLAG_SETS[it % len(FLAG_SETS)]
        _say(f"[gpu] random field {it} shape {f.shape} dtype {f.dtype} iso {iso} flags {fl}")
        exp = oracle.dualmc(g, s, grad, iso, flags=fl | 8)
        got = engine_mesh(engine, f, g, iso, fl | TRIANGLES)
        check_mesh(got, exp, f"field {it}")
        assert got["nonmanifold_edges"] == exp["nonmanifold"]
        c = oracle.cells(g, s, grad, iso)
        assert (got["failed_projections_l1"], got["failed_projections_l2"]) == (c["nr_proj1"], c["nr_proj2"])
        assert got["nv_emitted"] == c["nv_before_unused"]
        assert (got["nv_cells"], got["nq_cells"]) == (c["v"].shape[0], c["quads"].shape[0])
        nfail += c["nr_proj1"]
        nnm += exp["nonmanifold"]
        nbnd += len(exp["bnd"])
    assert nfail > 0 and nnm > 0 and nbnd > 0, "sample must reach failed projections and non-manifold edges"


def test_engine_supplied_gradients(engine, oracle):
    rng = np.random.default_rng(4)
    f = rng.standard_normal((7, 9, 11)).astype(np.float32)
    g = tmc_amd.Grid.make(f.shape, (0, 0, 0), (1, 1, 1))
    grad = rng.standard_normal(f.size * 3)
    exp = oracle.dualmc(g, f.astype(np.float64).ravel(), grad, 0.0, True)
    got = engine_mesh(engine, f, g, 0.0, SIMPLIFY_ALL | TRIANGLES, grad=grad)
    check_mesh(got, exp, "grad")


def test_config1_genus2_128(engine, oracle):
    """Config 1: genus-2, 128^3 cell-centred on [-1.5,1.5]^3, iso 0."""
    from tmc_amd import workloads
    _say("[gpu] genus2 128")
    f, g, iso = workloads.genus2(128)
    s = f.astype(np.float64).ravel()
    exp = oracle.dualmc(g, s, oracle.gradient(g, s), iso, True)
    got = engine_mesh(engine, f, g, iso, SIMPLIFY_ALL | TRIANGLES)
    check_mesh(got, exp, "genus2_128")
    assert (got["nv_cells"], got["nq_cells"]) == (65965, 65752)
    assert (len(got["v"]), len(got["quads"])) == (59177, 58964)
    E, chi = euler(len(got["v"]), got["quads"], got["he"])
    assert (E, chi) == (118146, -5)


@pytest.mark.parametrize("which", ["blobs_160", "schwarz_161"])
def test_benchmark_fields_sampled(engine, oracle, which):
    """The blob and Schwarz-P generators at a size the oracle finishes in
    seconds (same parameters as the 512^3 and 1024^3 workloads)."""
    from tmc_amd import workloads
    _say(f"[gpu] {which}")
    if which == "blobs_160":
        f, g, iso = workloads.blobs(160, seed=2)
    else:
        f, g, iso = workloads.schwarz(161, seed=3)
    s = f.astype(np.float64).ravel()
    exp = oracle.dualmc(g, s, oracle.gradient(g, s), iso, True)
    got = engine_mesh(engine, f, g, iso, SIMPLIFY_ALL | TRIANGLES)
    check_mesh(got, exp, which)


def test_repeat_and_reuse_is_deterministic(engine):
    """Same input twice, with a different extraction in between, gives the
    same bytes (no state leaks through the reused buffers)."""
    from tmc_amd import workloads
    f, g, iso = workloads.genus2(48)
    a = engine_mesh(engine, f, g, iso, SIMPLIFY_ALL | TRIANGLES)
    f2, g2, iso2 = workloads.blobs(40, seed=9)
    engine_mesh(engine, f2, g2, iso2, SIMPLIFY_ALL)
    b = engine_mesh(engine, f, g, iso, SIMPLIFY_ALL | TRIANGLES)
    for k in ("v", "n", "quads", "tris", "he", "he_v"):
        assert_same(a[k], b[k], k)


def test_empty_and_degenerate_inputs(engine):
    f = np.zeros((4, 5, 6), np.float32)
    got = engine_mesh(engine, f, tmc_amd.Grid.make(f.shape), 1.0, SIMPLIFY_ALL | TRIANGLES)
    assert len(got["v"]) == 0 and len(got["quads"]) == 0
    f = np.zeros((2, 2, 2), np.float32)
    f[0, 0, 0] = 1
    got = engine_mesh(engine, f, tmc_amd.Grid.make(f.shape), 0.5, SIMPLIFY_ALL | TRIANGLES)
    assert len(got["quads"]) == 0 and got["nv_emitted"] == 1 and len(got["v"]) == 0
    with pytest.raises(tmc_amd.DMCError):
        engine.extract(np.zeros((1, 4, 4), np.float32), 0.0)


def test_device_resident_field(engine, oracle):
    from tmc_amd import workloads
    f, g, iso = workloads.genus2(40)
    buf = engine.alloc(f.nbytes).upload(f)
    try:
        engine.extract(buf, iso, g, flags=SIMPLIFY_ALL | TRIANGLES, shape=f.shape)
        got = engine.download()
    finally:
        buf.free()
    s = f.astype(np.float64).ravel()
    exp = oracle.dualmc(g, s, oracle.gradient(g, s), iso, True)
    check_mesh(got, exp, "device field")


def test_fp32_export(engine):
    from tmc_amd import workloads
    f, g, iso = workloads.genus2(32)
    engine.extract(f, iso, g)
    m = engine.download(fp32=True)
    assert np.array_equal(m["v32"], m["v"].astype(np.float32))
    assert np.array_equal(m["n32"], m["n"].astype(np.float32))


def test_halfedge_invariants_on_gpu_mesh(engine):
    from tmc_amd import workloads
    f, g, iso = workloads.blobs(96, seed=2)
    m = engine_mesh(engine, f, g, iso, SIMPLIFY_ALL)
    halfedge_invariants(len(m["v"]), m["quads"], m["he"])


def test_cpp_dropin_programs(oracle, tmp_path):
    """The reference's own driver programs (cpp_dmc/main.cpp, dmc/kernel.cu)
    written against include/tmc_dmc.hpp: cpp_dmc/main.cpp's 64^3 GenusTwo
    gives V 8,825 / Q 8,827 in the reference; the C++ surface must return
    the oracle's mesh bit for bit, also through the binary-file loader."""
    import json
    import os
    import subprocess
    from conftest import ROOT
    exe = str(tmp_path / "dropin")
    src = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
    libdir = os.path.join(ROOT, "tmc_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), src, "-L" + libdir,
                        "-ltmc_dmc", "-Wl,-rpath," + libdir, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    _say("[gpu] running C++ drop-in program")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = {d["case"]: d for d in (json.loads(l) for l in r.stdout.splitlines() if l.startswith("{"))}
    assert (res["cpp_genus2_64"]["nv"], res["cpp_genus2_64"]["nq"]) == (8825, 8827)
    assert res["cpp_file_64"]["on_device_gradients"] == 1
    assert res["cuda_genus2_64"]["twins_ok"] == 1
    rd = lambda n, t: np.fromfile(str(tmp_path / n), dtype=t)
    geo = rd("geo.f64", np.float64)
    from oracle import Grid
    g = Grid(64, 64, 64, *geo)
    s = rd("scalars.f64", np.float64)
    exp = oracle.dualmc(g, s, rd("grads.f64", np.float64), 0.0, True)
    assert_same(rd("v.f64", np.float64).reshape(-1, 3), exp["v"], "cpp v")
    assert_same(rd("n.f64", np.float64).reshape(-1, 3), exp["n"], "cpp n")
    assert_same(rd("quads.i32", np.int32).reshape(-1, 4), exp["quads"], "cpp quads")
    # file path: float field, gradients estimated + flipped; the reference's
    # loader puts the grid origin at 0 (cpp_dmc/UniformGrid.cpp:21-29)
    s32 = s.astype(np.float32).astype(np.float64)
    g0 = Grid(64, 64, 64, 0.0, 0.0, 0.0, *geo[3:])
    exp2 = oracle.dualmc(g0, s32, oracle.gradient(g0, s32), 0.0, True)
    assert_same(rd("file_v.f64", np.float64).reshape(-1, 3), exp2["v"], "file v")
    assert_same(rd("file_quads.i32", np.int32).reshape(-1, 4), exp2["quads"], "file quads")
    # the CUDA path's diagnostics switches: every stage's distribution counts
    # every vertex of its mesh; the final ones, the quality and the edges
    # match the restatements on the returned mesh
    from oracle import diagnostics as D
    cd = res["cuda_diag_64"]
    assert cd["dist_3333"] == cd["nv"] and cd["dist_tris"] == cd["nv"]
    assert cd["dist_dmc"] >= cd["dist_color"] >= cd["dist_old"] >= cd["dist_3333"] > 0
    assert cd["nquality_q"] == cd["nq"] and cd["nquality_t"] == 2 * cd["nq"]
    cv = rd("cu_v.f32", np.float32).reshape(-1, 3).astype(np.float64)
    cq = rd("cu_quads.i32", np.int32).reshape(-1, 4)
    ct = rd("cu_tris.i32", np.int32).reshape(-1, 3)
    assert np.array_equal(rd("cu_dist_3333.u32", np.uint32), D.valence_histogram(len(cv), cq)[0])
    assert np.array_equal(rd("cu_dist_tris.u32", np.uint32), D.valence_histogram(len(cv), ct)[0])
    assert np.array_equal(rd("cu_quality_q.f32", np.float32).view(np.uint32), D.quality_quads(cv, cq).view(np.uint32))
    assert np.array_equal(rd("cu_quality_t.f32", np.float32).view(np.uint32), D.quality_tris(cv, ct).view(np.uint32))
    assert np.array_equal(rd("cu_edges.i32", np.int32).reshape(-1, 2), D.mesh_edges(cq))



# ---------------------------------------------------------------------------
# diagnostics of the finished mesh (the CUDA reference's edges, valence
# distributions and element quality) and the OBJ writer, against the
# restatements of oracle/diagnostics.py on the downloaded mesh
# ---------------------------------------------------------------------------
def _diag_meshes():
    from tmc_amd import workloads
    f, g, iso = workloads.genus2(40)
    yield "genus2_40", f, g, iso
    f, g, iso = workloads.blobs(36, seed=3)
    yield "blobs_36", f, g, iso
    for it, f, g, iso in _fields(17, 10):
        yield f"random_{it}", f, g, iso


def _same_f32(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_mesh_diagnostics_match_oracle(engine, tmp_path):
    from oracle import diagnostics as D
    nnm = 0
    for name, f, g, iso in _diag_meshes():
        _say(f"[gpu] diagnostics {name} shape {f.shape}")
        m = engine.extract(f, iso, g, flags=SIMPLIFY_ALL | TRIANGLES)
        got = engine.download(halfedges=False)
        nv, quads, tris = len(got["v"]), got["quads"], got["tris"]
        nnm += got["nonmanifold_edges"]
        assert np.array_equal(engine.edges(), D.mesh_edges(quads)), name
        for tri, faces in ((False, quads), (True, tris)):
            h, beyond = engine.valence_histogram(triangles=tri)
            eh, eb = D.valence_histogram(nv, faces)
            assert np.array_equal(h, eh) and beyond == eb, (name, tri, h, eh)
        assert _same_f32(engine.element_quality(), D.quality_quads(got["v"], quads)), name
        assert _same_f32(engine.element_quality(triangles=True), D.quality_tris(got["v"], tris)), name
        for tri, faces in ((False, quads), (True, tris)):
            path = tmp_path / f"{name}_{int(tri)}.obj"
            engine.write_obj(str(path), triangles=tri)
            assert path.read_text() == D.obj_text(got["v"], got["n"], faces, tri), (name, tri)
    assert nnm > 0, "sample must reach non-manifold edges"


def test_mesh_diagnostics_need_a_single_extraction(engine):
    """After a batch the mesh ids are per volume: diagnostics refuse."""
    from tmc_amd import workloads
    vols = [workloads.genus2(16), workloads.blobs(16, seed=5)]
    engine.extract_batch([(f, iso, g) for f, g, iso in vols], flags=SIMPLIFY_ALL | TRIANGLES)
    for call in (engine.edges, engine.valence_histogram, engine.element_quality):
        with pytest.raises(tmc_amd.DMCError):
            call()


# ---------------------------------------------------------------------------
# batched extraction: every volume of a batch must come back exactly as the
# same volume extracted alone (and the alone path is pinned to the oracle
# above), whatever else is in the batch
# ---------------------------------------------------------------------------
SUMMARY_KEYS = ("nv", "nq", "nt", "nbnd", "nonmanifold_edges", "edge_errors", "failed_projections_l1",
                "failed_projections_l2", "active_cells", "nv_emitted", "nv_cells", "nq_cells")


def _batch_volumes(seed, count):
    from tmc_amd import workloads
    vols = []
    for it, f, g, iso in _fields(seed, count, max_n=12):
        if f.dtype != np.float32:
            f = f.astype(np.float32)
        vols.append((f, iso, g))
    # mixed generators of the batched workload, small sizes
    for s in workloads.batch_specs(8, seed=seed)[:8]:
        shape = tuple(max(2, x // 3) for x in s["shape"])
        f, g, iso = workloads.make(dict(s, shape=shape))
        vols.append((f, iso, g))
    # degenerate members: no surface, minimum size, all equal to iso
    vols.append((np.zeros((2, 2, 2), np.float32), 0.5, tmc_amd.Grid.make((2, 2, 2))))
    vols.append((np.full((3, 4, 5), 0.25, np.float32), 0.25, tmc_amd.Grid.make((3, 4, 5))))
    return vols


def _alone(engine, f, iso, g, flags):
    m = engine.extract(f, iso, g, flags=flags)
    return m.summary(), engine.download()


@pytest.mark.parametrize("flags", [SIMPLIFY_ALL | TRIANGLES, TRIANGLES, SIMPLIFY_3X3Y_ISOLATED | SIMPLIFY_3333])
def test_batch_matches_alone(engine, flags):
    vols = _batch_volumes(41, 25)
    _say(f"[gpu] batch of {len(vols)} volumes, flags {flags}")
    alone = [_alone(engine, f, iso, g, flags) for f, iso, g in vols]
    meshes = engine.extract_batch(vols, flags=flags)
    assert len(meshes) == len(vols)
    for i, (exp_sum, exp) in enumerate(alone):
        got_sum = meshes[i].summary()
        for k in SUMMARY_KEYS:
            assert got_sum[k] == exp_sum[k], f"volume {i}: {k} {got_sum[k]} vs {exp_sum[k]}"
        got = engine.download_batch(i)
        for k in ("v", "n", "quads", "tris", "he", "he_v", "he_f", "bnd"):
            assert_same(got[k], exp[k], f"volume {i}:{k}")


def test_batch_order_and_reuse(engine):
    """The same volumes in another order, and a batch after a bigger one,
    give the same meshes (no state leaks between batches or members)."""
    vols = _batch_volumes(43, 10)
    flags = SIMPLIFY_ALL | TRIANGLES
    engine.extract_batch(vols, flags=flags)
    first = [engine.download_batch(i) for i in range(len(vols))]
    order = list(range(len(vols)))[::-1]
    engine.extract_batch([vols[i] for i in order], flags=flags)
    for j, i in enumerate(order):
        got = engine.download_batch(j)
        for k in ("v", "n", "quads", "tris", "he", "he_v", "bnd"):
            assert_same(got[k], first[i][k], f"reordered volume {i}:{k}")
    engine.extract_batch(vols[:3], flags=flags)
    for i in range(3):
        got = engine.download_batch(i)
        for k in ("v", "quads", "he"):
            assert_same(got[k], first[i][k], f"smaller batch volume {i}:{k}")


def test_batch_device_fields(engine):
    """Device-resident fields of a batch (one allocation, volumes at offsets)."""
    vols = _batch_volumes(47, 6)
    flags = SIMPLIFY_ALL | TRIANGLES
    offs, total = [], 0
    for f, _, _ in vols:
        offs.append(total)
        total += (f.nbytes + 255) // 256 * 256
    buf = engine.alloc(total)
    host = np.zeros(total, np.uint8)
    for (f, _, _), o in zip(vols, offs):
        host[o:o + f.nbytes] = np.frombuffer(f.tobytes(), np.uint8)
    buf.upload(host)
    dvols = [(buf.ptr.value + o, iso, g) for (f, iso, g), o in zip(vols, offs)]
    exp = [_alone(engine, f, iso, g, flags)[1] for f, iso, g in vols]
    engine.extract_batch(dvols, flags=flags | tmc_amd.FIELD_ON_DEVICE)
    for i in range(len(vols)):
        got = engine.download_batch(i)
        for k in ("v", "n", "quads", "he", "bnd"):
            assert_same(got[k], exp[i][k], f"device volume {i}:{k}")
    buf.free()


# ---------------------------------------------------------------------------
# serving loop: results identical alone and in the mix, whatever the batch
# the server formed; admission under a budget
# ---------------------------------------------------------------------------
def _server_results(srv, vols, threads=1):
    import threading
    tickets = [None] * len(vols)

    def sub(idx):
        for i in idx:
            f, iso, g = vols[i]
            tickets[i] = srv.submit(f, iso, g)

    parts = [list(range(t, len(vols), threads)) for t in range(threads)]
    ts = [threading.Thread(target=sub, args=(p,)) for p in parts]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    out = []
    for t in tickets:
        info = srv.wait(t)
        got = srv.download(t)
        got["info"] = info
        srv.release(t)
        out.append(got)
    return out


@pytest.mark.parametrize("max_batch,threads,workers",
                         [(1, 1, 1), (1 << 20, 1, 1), (1 << 20, 4, 1), (3, 2, 1), (3, 4, 3), (1, 2, 2)])
def test_server_matches_alone(engine, max_batch, threads, workers):
    """Every request's mesh equals its own dualMC, whatever the batching,
    the submitting threads and the concurrent passes."""
    flags = SIMPLIFY_ALL | TRIANGLES
    vols = _batch_volumes(53, 12)
    _say(f"[gpu] server max_batch={max_batch} threads={threads} workers={workers}, {len(vols)} volumes")
    exp = [_alone(engine, f, iso, g, flags) for f, iso, g in vols]
    srv = tmc_amd.Server(0, max_batch=max_batch, flags=flags, workers=workers)
    try:
        got = _server_results(srv, vols, threads)
        for i, (es, e) in enumerate(exp):
            info = got[i]["info"]
            assert 1 <= info["batch_size"] <= max_batch
            assert info["t_submit"] <= info["t_start"] <= info["t_done"]
            for k in SUMMARY_KEYS:
                assert info[k] == es[k], f"request {i}: {k}"
            for k in ("v", "n", "quads", "tris", "he", "he_v", "he_f", "bnd"):
                assert_same(got[i][k], e[k], f"request {i}:{k}")
        st = srv.stats()
        assert st["served"] == len(vols) and st["held_bytes"] == 0
    finally:
        srv.close()


@pytest.mark.parametrize("workers", [1, 2])
def test_server_budget_admission(engine, workers):
    """A small budget splits the waiting requests into several batches and a
    request larger than the budget fails with DMC_ERR_BUDGET; the others
    are unaffected.  With two workers the passes in flight share the
    budget."""
    flags = SIMPLIFY_ALL | TRIANGLES
    rng = np.random.default_rng(5)
    vols = [(rng.standard_normal((10, 10, 10)).astype(np.float32), 0.2, tmc_amd.Grid.make((10, 10, 10)))
            for _ in range(6)]
    big = (rng.standard_normal((40, 40, 40)).astype(np.float32), 0.2, tmc_amd.Grid.make((40, 40, 40)))
    exp = [_alone(engine, f, iso, g, flags)[1] for f, iso, g in vols]
    # one 10^3 volume: 81 rows x 1 segment x 1032 B + 40 B x 1000 points + 4,000 B of field = 127,592 B
    srv = tmc_amd.Server(0, budget_bytes=300_000, flags=flags, workers=workers)
    try:
        tk = [srv.submit(f, iso, g) for f, iso, g in vols[:3]]
        tb = srv.submit(*big)
        tk += [srv.submit(f, iso, g) for f, iso, g in vols[3:]]
        with pytest.raises(tmc_amd.DMCError, match="budget"):
            srv.wait(tb)
        srv.release(tb)
        for i, t in enumerate(tk):
            info = srv.wait(t)
            assert info["batch_size"] <= 2, info["batch_size"]
            got = srv.download(t)
            for k in ("v", "quads", "he", "tris"):
                assert_same(got[k], exp[i][k], f"request {i}:{k}")
            srv.release(t)
    finally:
        srv.close()
