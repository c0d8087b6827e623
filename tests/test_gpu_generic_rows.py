"""The generic halfedge rows against the usual builds.

A pass pairs its first halfedge build through the rows the cell stage
leaves and carries the build over every compaction.  When the device finds
either unusable (an unresolved neighbour cell, an invalid vertex
reference), the pass runs again with every build paired through generic
rows counted from the quadrilaterals.  An engine created with
TMC_GENERIC_ROWS=1 takes that path on every pass; its meshes must equal the
usual ones bit for bit, and so must the diagnostics that read the rows.
"""
import os
import sys

import numpy as np
import pytest

from meshcheck import assert_same

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402
from tmc_amd import SIMPLIFY_ALL, TRIANGLES, workloads  # noqa: E402

KEYS = ("v", "n", "quads", "tris", "he", "he_v", "bnd")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _generic_engine():
    os.environ["TMC_GENERIC_ROWS"] = "1"
    try:
        return tmc_amd.Engine(0)
    finally:
        del os.environ["TMC_GENERIC_ROWS"]


CASES = [("genus2 64^3", lambda: workloads.genus2(64)),
         ("schwarz 96^3", lambda: workloads.schwarz(96, seed=5)),
         ("blobs 48x40x56", lambda: workloads.blobs((56, 40, 48), seed=11))]


@pytest.mark.parametrize("flags", [SIMPLIFY_ALL | TRIANGLES, TRIANGLES, tmc_amd.SIMPLIFY_3X3Y_COLOR])
def test_generic_rows_match(flags):
    usual, generic = tmc_amd.Engine(0), _generic_engine()
    try:
        for name, make in CASES:
            f, g, iso = make()
            _say(f"[gpu] generic rows: {name}, flags {flags}")
            mu = usual.extract(f, iso, g, flags=flags)
            du = usual.download()
            eu = usual.edges()
            mg = generic.extract(f, iso, g, flags=flags)
            dg = generic.download()
            eg = generic.edges()
            su, sg = mu.summary(), mg.summary()
            for k in su:
                assert su[k] == sg[k], f"{name}: {k} {su[k]} vs {sg[k]}"
            for k in KEYS:
                assert_same(dg[k], du[k], f"{name}: {k}")
            assert np.array_equal(eu, eg), name
    finally:
        usual.close()
        generic.close()
