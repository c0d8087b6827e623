"""Launch shape and autotuner on the GPU.

The launch shape (block cap of the grid-stride kernels, waves of the field
sweep) may change only the time of an extraction: every array must come
out bit-identical under every shape, alone and in a batch.  The autotuner
times candidates and keeps the fastest; its result is checked the same way.
"""
import sys

import numpy as np
import pytest

from meshcheck import assert_same

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402
from tmc_amd import SIMPLIFY_ALL, TRIANGLES, workloads  # noqa: E402

KEYS = ("v", "n", "quads", "tris", "he", "he_v", "he_f", "bnd")
FLAGS = SIMPLIFY_ALL | TRIANGLES


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _volumes():
    vols = []
    for s in workloads.batch_specs(12, seed=9):
        f, g, iso = workloads.make(s)
        vols.append((f, iso, g))
    # one volume large enough that the sweep's plane count follows the
    # wave target (255 planes, 16 row tiles)
    f, g, iso = workloads.blobs(256, seed=7)
    vols.append((f, iso, g))
    return vols


def _batch(engine, vols):
    meshes = engine.extract_batch(vols, flags=FLAGS)
    return [(m.summary(), engine.download_batch(i)) for i, m in enumerate(meshes)]


def _same(got, exp, what):
    assert len(got) == len(exp)
    for i, ((gs, g), (es, e)) in enumerate(zip(got, exp)):
        assert gs == es, f"{what}: volume {i} summary differs"
        for k in KEYS:
            assert_same(g[k], e[k], f"{what}: volume {i}:{k}")


@pytest.fixture(scope="module")
def own_engine():
    from tmc_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def test_launch_shape_does_not_change_results(own_engine):
    eng = own_engine
    default = eng.launch
    assert default == {"grid_cap": 32768, "sweep_waves": 32768}
    vols = _volumes()
    _say(f"[gpu] launch shapes: batch of {len(vols)} volumes, default {default}")
    ref = _batch(eng, vols)
    f, iso, g = vols[-1]
    eng.extract(f, iso, g, flags=FLAGS)
    alone_ref = eng.download()
    try:
        for cap, waves in ((64, 256), (2048, 8192), (65536, 1 << 20)):
            _say(f"[gpu] grid_cap {cap} sweep_waves {waves}")
            eng.launch = {"grid_cap": cap, "sweep_waves": waves}
            assert eng.launch == {"grid_cap": cap, "sweep_waves": waves}
            _same(_batch(eng, vols), ref, f"cap {cap} waves {waves}")
            eng.extract(f, iso, g, flags=FLAGS)
            alone = eng.download()
            for k in KEYS:
                assert_same(alone[k], alone_ref[k], f"alone, cap {cap} waves {waves}:{k}")
    finally:
        eng.launch = default


def test_launch_shape_range_is_checked(own_engine):
    eng = own_engine
    before = eng.launch
    for bad in ({"grid_cap": 63}, {"grid_cap": (1 << 24) + 1}, {"sweep_waves": 255}, {"sweep_waves": 1 << 25}):
        with pytest.raises(tmc_amd.DMCError):
            eng.launch = bad
        assert eng.launch == before


def test_autotune_keeps_fastest_and_same_results(own_engine):
    eng = own_engine
    vols = _volumes()
    ref = _batch(eng, vols)
    _say(f"[gpu] autotune over a batch of {len(vols)} volumes")
    res = eng.autotune(vols, flags=FLAGS, reps=2)
    trials = res["trials"]
    # four sweep targets, then the four block caps other than the start's
    assert len(trials) == 8
    assert {t["sweep_waves"] for t in trials[:4]} == {8192, 16384, 32768, 65536}
    assert {t["grid_cap"] for t in trials[4:]} == {4096, 8192, 16384, 65536}
    best = res["best"]
    assert best["ms"] == min(t["ms"] for t in trials)
    assert any(t["grid_cap"] == best["grid_cap"] and t["sweep_waves"] == best["sweep_waves"] for t in trials)
    assert eng.launch == {"grid_cap": best["grid_cap"], "sweep_waves": best["sweep_waves"]}
    _say(f"[gpu] autotune chose {best}")
    _same(_batch(eng, vols), ref, "after autotune")
    eng.launch = {"grid_cap": 16384, "sweep_waves": 32768}
