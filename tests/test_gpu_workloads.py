"""The batched and served workloads at their stated sizes, against the
per-volume digests the oracle wrote (tests/golden/batch_digests.json,
tests/golden/make_batch_digests.py; digest = SHA-256 over positions,
normals, quads, halfedge rows, he_v and boundary vertices, bit-exact):

* all 4,096 volumes of workloads.batch_specs(4096, seed=4) (16-96 points
  per axis, mixed generators and iso-values) in one dmc_extract_batch, the
  fields resident in HBM;
* a Poisson stream of requests drawn from the serving pool (14 volumes of
  32-128^3, 5 of 256^3, 1 of 512^3) through a server with two workers and
  a 96 GiB budget, as the benchmark runs it;
* a served mix with requests the engine cannot take: each fails alone and
  every other request comes back exact.
"""
import hashlib
import json
import os
import sys
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402
from tmc_amd import SIMPLIFY_ALL, TRIANGLES, workloads  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "batch_digests.json")) as _fh:
    KNOWN = json.load(_fh)

FLAGS = SIMPLIFY_ALL | TRIANGLES | tmc_amd.FIELD_ON_DEVICE


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mesh_digest(m):
    h = hashlib.sha256()
    for k, dt in (("v", "<f8"), ("n", "<f8"), ("quads", "<i4"), ("he", "<i4"), ("he_v", "<i4"), ("bnd", "<i4")):
        h.update(np.ascontiguousarray(m[k], dtype=dt).tobytes())
    return h.hexdigest()[:32]


def _pack(eng, fields):
    offs, total = [], 0
    for f in fields:
        offs.append(total)
        total += (f.nbytes + 255) // 256 * 256
    host = np.zeros(total, np.uint8)
    for f, o in zip(fields, offs):
        host[o:o + f.nbytes] = np.frombuffer(f.tobytes(), np.uint8)
    buf = eng.alloc(total).upload(host)
    return buf, [buf.ptr.value + o for o in offs]


@pytest.fixture(scope="module")
def eng():
    e = tmc_amd.Engine(0)
    yield e
    e.close()


def test_batch_4096_matches_oracle(eng):
    specs = workloads.batch_specs(4096, seed=4)
    known = KNOWN["batch_4096"]["volumes"]
    assert len(known) == len(specs)
    t0 = time.time()
    vols = workloads.make_many(specs)
    _say(f"[gpu] batch_4096: {len(vols)} volumes generated in {time.time() - t0:.1f} s")
    buf, ptrs = _pack(eng, [f for f, _, _ in vols])
    try:
        dv = [(p, iso, g) for p, (f, g, iso) in zip(ptrs, vols)]
        meshes = eng.extract_batch(dv, flags=FLAGS)
        _say(f"[gpu] batch_4096: extracted, V {sum(m.nv for m in meshes)} Q {sum(m.nq for m in meshes)}")
        bad = []
        for i, (m, k) in enumerate(zip(meshes, known)):
            if (m.nv, m.nq, m.nbnd) != (k["nv"], k["nq"], k["nbnd"]):
                bad.append((i, "counts", (m.nv, m.nq, m.nbnd), (k["nv"], k["nq"], k["nbnd"])))
                continue
            if mesh_digest(eng.download_batch(i)) != k["digest"]:
                bad.append((i, "digest"))
        assert not bad, f"{len(bad)} of 4096 volumes differ from the oracle: {bad[:5]}"
    finally:
        buf.free()


def _serve_pool():
    sys.path.insert(0, os.path.dirname(HERE))
    from bench import serve_pool
    return serve_pool(5)


def test_serving_stream_matches_oracle(eng):
    specs = _serve_pool()
    known = KNOWN["serve_pool"]["volumes"]
    assert [list(s["shape"]) for s in specs] == KNOWN["serve_pool"]["shapes"]
    vols = [workloads.make(s) for s in specs]
    buf, ptrs = _pack(eng, [f for f, _, _ in vols])
    pool = [(p, iso, g) for p, (f, g, iso) in zip(ptrs, vols)]
    rng = np.random.default_rng(11)
    n = 120
    pick = rng.integers(0, len(pool), size=n)
    pick[:len(pool)] = np.arange(len(pool))       # every pool volume at least once
    rng.shuffle(pick)
    gaps = rng.exponential(1.0 / 400.0, size=n)    # Poisson arrivals, 400 requests/s
    srv = tmc_amd.Server(0, flags=FLAGS, budget_bytes=96 << 30, max_batch_points=64 << 20, workers=2)
    _say(f"[gpu] serving stream: {n} requests from a pool of {len(pool)}")
    tickets = [None] * n

    def submit():
        t0 = time.perf_counter()
        at = np.cumsum(gaps)
        for i in range(n):
            w = at[i] - (time.perf_counter() - t0)
            if w > 0:
                time.sleep(w)
            p, iso, g = pool[pick[i]]
            tickets[i] = srv.submit(p, iso, g)

    try:
        th = threading.Thread(target=submit)
        th.start()
        th.join()
        bad, batched = [], 0
        for i, t in enumerate(tickets):
            info = srv.wait(t)
            k = known[pick[i]]
            batched += info["batch_size"] > 1
            if (info["nv"], info["nq"], info["nbnd"]) != (k["nv"], k["nq"], k["nbnd"]):
                bad.append((i, int(pick[i]), "counts"))
            elif mesh_digest(srv.download(t)) != k["digest"]:
                bad.append((i, int(pick[i]), "digest"))
            srv.release(t)
        st = srv.stats()
        _say(f"[gpu] serving stream: {st['batches']} passes, {batched} requests served in a batch of 2+")
        assert not bad, f"{len(bad)} of {n} served results differ from the oracle: {bad[:5]}"
        assert st["served"] == n and st["held_bytes"] == 0
    finally:
        srv.close()
        buf.free()


def test_server_bad_requests_fail_alone(eng):
    """Requests the engine cannot take are refused at submit, alone; the
    requests around them come back exact."""
    specs = _serve_pool()[:6]
    known = KNOWN["serve_pool"]["volumes"][:6]
    vols = [workloads.make(s) for s in specs]
    buf, ptrs = _pack(eng, [f for f, _, _ in vols])
    srv = tmc_amd.Server(0, flags=FLAGS, workers=1)
    try:
        tk = []
        for i, (p, (f, g, iso)) in enumerate(zip(ptrs, vols)):
            tk.append(srv.submit(p, iso, g))
            if i == 2:
                # an axis above 2^24 points, then more than 2^32 cells: refused
                # before they reach a pass (the pointer is never read)
                with pytest.raises(tmc_amd.DMCError, match="invalid argument"):
                    srv.submit(ptrs[0], 0.0, tmc_amd.Grid.make((2, 2, (1 << 24) + 1)))
                with pytest.raises(tmc_amd.DMCError, match="capacity"):
                    srv.submit(ptrs[0], 0.0, tmc_amd.Grid.make((2049, 2049, 2049)))
        with pytest.raises(tmc_amd.DMCError, match="field has"):
            tmc_amd.Server(0, flags=SIMPLIFY_ALL, workers=1).submit(np.zeros(8, np.float32), 0.0,
                                                                   tmc_amd.Grid.make((4, 4, 4)))
        for i, t in enumerate(tk):
            info = srv.wait(t)
            assert (info["nv"], info["nq"]) == (known[i]["nv"], known[i]["nq"])
            assert mesh_digest(srv.download(t)) == known[i]["digest"]
            srv.release(t)
    finally:
        srv.close()
        buf.free()


def test_capacities_per_volume_and_trim():
    """A context sizes a pass's buffers per volume: after the pool's densest
    small volume (36^3, 0.65 vertices per point) the 512^3 volume takes no
    more memory than in a fresh context.  dmc_context_trim frees every pool,
    and the next pass allocates anew and is still exact."""
    specs = _serve_pool()
    known = KNOWN["serve_pool"]["volumes"]
    dense, large, small = 13, 19, 0
    vols = {i: workloads.make(specs[i]) for i in (dense, large, small)}
    flags = SIMPLIFY_ALL | TRIANGLES

    def run(e, i):
        f, g, iso = vols[i]
        m = e.extract(f, iso, g, flags=flags)
        assert (m.nv, m.nq) == (known[i]["nv"], known[i]["nq"]), i
        assert mesh_digest(e.download()) == known[i]["digest"], i

    fresh = tmc_amd.Engine(0)
    try:
        run(fresh, large)
        alone = fresh.context_bytes()
    finally:
        fresh.close()
    e = tmc_amd.Engine(0)
    try:
        run(e, dense)
        run(e, large)
        after = e.context_bytes()
        _say(f"[gpu] 512^3 context: {alone / 2**30:.2f} GiB fresh, {after / 2**30:.2f} GiB after the dense volume")
        assert after <= alone * 5 // 4
        e.trim()
        assert e.context_bytes() == 0
        with pytest.raises(tmc_amd.DMCError):
            e.download()
        run(e, small)
        assert 0 < e.context_bytes() < after // 4
        run(e, large)
    finally:
        e.close()


def test_server_trims_worker_pools_over_budget(eng):
    """With a budget just above a 512^3 request's working-set estimate, the
    pools a 512^3 pass leaves keep the server over budget when a 256^3
    request comes: the worker frees them first, so the server's pool
    shrinks, and every result stays exact."""
    specs = _serve_pool()
    known = KNOWN["serve_pool"]["volumes"]
    idx = (19, 14, 19)

    def estimate(shape):
        # the server's admission estimate: (row, segment) lists and counts,
        # 40 bytes per grid point
        nz, ny, nx = shape
        return (ny - 1) * (nz - 1) * ((nx - 1 + 255) // 256) * (256 * 4 + 8) + 40 * nx * ny * nz

    vols = [workloads.make(specs[i]) for i in idx]
    buf, ptrs = _pack(eng, [f for f, _, _ in vols])
    srv = tmc_amd.Server(0, flags=FLAGS, budget_bytes=estimate(specs[19]["shape"]) + (16 << 20), workers=1)
    try:
        pools = []
        for k, (p, (f, g, iso)) in enumerate(zip(ptrs, vols)):
            t = srv.submit(p, iso, g)
            info = srv.wait(t)
            i = idx[k]
            assert (info["nv"], info["nq"]) == (known[i]["nv"], known[i]["nq"]), i
            assert mesh_digest(srv.download(t)) == known[i]["digest"], i
            srv.release(t)
            pools.append(srv.stats()["pool_bytes"])
        _say(f"[gpu] server pool after 512^3, 256^3, 512^3: {[round(x / 2**30, 2) for x in pools]} GiB")
        assert pools[1] < pools[0] // 2
    finally:
        srv.close()
        buf.free()
