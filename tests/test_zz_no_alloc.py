"""dmc_extract_noalloc against dmc_extract, through the C ABI.

The no-alloc entry point runs the same kernels with its working buffers in
the library's fixed scratch array and writes the mesh into buffers the
caller owns.  Its output must equal dmc_extract's bit for bit (positions,
normals, quads, triangles, twins, halfedge flags, he_v, boundary vertices
and every count), on two calls in a row, the second with a smaller volume
(the scratch keeps the first call's contents).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402
from tmc_amd import SIMPLIFY_ALL, TRIANGLES, workloads  # noqa: E402

SUMMARY = ("nv", "nq", "nt", "nbnd", "nonmanifold_edges", "edge_errors", "invalid_refs", "valence_zero",
           "failed_projections_l1", "failed_projections_l2", "active_cells", "nv_emitted", "nv_cells", "nq_cells")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = tmc_amd.Engine(0)
    yield e
    e.close()


def _outs(eng, cap_v, cap_q):
    sizes = [24 * cap_v, 24 * cap_v, 16 * cap_q, 24 * cap_q, 16 * cap_q, 4 * cap_q, 4 * cap_v]
    bufs = [eng.alloc(max(256, s)) for s in sizes]
    o = tmc_amd.OutBuffers()
    o.positions, o.normals, o.quads, o.tris, o.twins, o.he_flags, o.he_v = (b.ptr.value for b in bufs)
    o.cap_v, o.cap_q = cap_v, cap_q
    return o, bufs


def _host(eng, m, bnd):
    nv, nq = m.nv, m.nq
    return {"v": eng.copy_to_host(np.empty((nv, 3), np.float64), m.d_positions),
            "n": eng.copy_to_host(np.empty((nv, 3), np.float64), m.d_normals),
            "quads": eng.copy_to_host(np.empty((nq, 4), np.int32), m.d_quads),
            "tris": eng.copy_to_host(np.empty((m.nt, 3), np.int32), m.d_tris) if m.nt else np.empty((0, 3), np.int32),
            "twins": eng.copy_to_host(np.empty(4 * nq, np.int32), m.d_twins),
            "flags": eng.copy_to_host(np.empty(4 * nq, np.uint8), m.d_he_flags),
            "he_v": eng.copy_to_host(np.empty(nv, np.int32), m.d_he_v),
            "bnd": bnd}


def _expected(eng, dev, iso, g, flags):
    m = eng.extract(dev, iso, g, flags=flags)
    d = eng.download(halfedges=True)
    return m.summary(), {"v": d["v"], "n": d["n"], "quads": d["quads"], "tris": d["tris"],
                         "twins": d["he"][:, 3].copy(), "flags": d["he"][:, 4].astype(np.uint8),
                         "he_v": d["he_v"], "bnd": d["bnd"]}


def test_noalloc_matches_extract(eng):
    flags = SIMPLIFY_ALL | TRIANGLES
    cases = [("genus2 64^3", workloads.genus2(64)), ("blobs 40x36x32", workloads.blobs((32, 36, 40), seed=7))]
    outs, bufs = _outs(eng, 1 << 17, 1 << 17)
    try:
        for name, (f, g, iso) in cases:
            _say(f"[gpu] no-alloc {name}")
            dev = eng.alloc(f.nbytes).upload(f)
            m, bnd = eng.extract_noalloc(dev, iso, g, outs, flags=flags)
            got_sum = m.summary()
            got = _host(eng, m, bnd)
            exp_sum, exp = _expected(eng, dev, iso, g, flags | tmc_amd.FIELD_ON_DEVICE)
            for k in SUMMARY:
                assert got_sum[k] == exp_sum[k], f"{name}: {k} {got_sum[k]} vs {exp_sum[k]}"
            for k, a in exp.items():
                b = got[k]
                assert a.shape == b.shape and np.array_equal(a, b), f"{name}: {k} differs from dmc_extract"
            assert got_sum["nv"] > 1000 and got_sum["nq"] > 1000
            dev.free()
    finally:
        for b in bufs:
            b.free()


def test_noalloc_capacity_errors(eng):
    # output buffers too small: the call fails and reports the sizes needed
    f, g, iso = workloads.genus2(32)
    dev = eng.alloc(f.nbytes).upload(f)
    outs, bufs = _outs(eng, 64, 64)
    try:
        with pytest.raises(tmc_amd.DMCError, match="capacity"):
            eng.extract_noalloc(dev, iso, g, outs, flags=SIMPLIFY_ALL)
        # a grid above the entry point's 2^21 points is refused before any work
        big = tmc_amd.Grid.make((130, 130, 130))
        with pytest.raises(tmc_amd.DMCError, match="capacity"):
            eng.extract_noalloc(dev, iso, big, outs, flags=SIMPLIFY_ALL)
        # the context still extracts normally afterwards
        m = eng.extract(f, iso, g, flags=SIMPLIFY_ALL)
        assert m.nv > 0
    finally:
        dev.free()
        for b in bufs:
            b.free()
