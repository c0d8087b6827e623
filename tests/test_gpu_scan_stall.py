"""The single-pass scan's bounded waits.

Scans of up to 64 tiles run in one launch: each tile looks back over its
predecessors' published sums.  A tile that polls more times than the
context allows sums its predecessors' items straight from the input
instead, so its prefix is exact whichever way it is found.  An engine
created with TMC_SCAN_SPIN=1 gives every wait a single poll, so that path
runs often (how often depends on timing): every extraction must still
return the usual mesh bit for bit, and must leave the scan state clear for
the next one on the same context.
"""
import os
import sys

import pytest

from meshcheck import assert_same

pytestmark = pytest.mark.gpu

import tmc_amd  # noqa: E402
from tmc_amd import SIMPLIFY_ALL, TRIANGLES, workloads  # noqa: E402

KEYS = ("v", "n", "quads", "tris", "he", "he_v", "bnd")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [128, 96])
def test_scan_waits_of_one_poll_stay_exact(n):
    f, g, iso = workloads.genus2(n)
    flags = SIMPLIFY_ALL | TRIANGLES
    usual = tmc_amd.Engine(0)
    os.environ["TMC_SCAN_SPIN"] = "1"
    try:
        tight = tmc_amd.Engine(0)
    finally:
        del os.environ["TMC_SCAN_SPIN"]
    try:
        usual.extract(f, iso, g, flags=flags)
        ref = usual.download()
        for rep in range(6):
            _say(f"[gpu] genus2 {n}^3, scan waits of one poll: extraction {rep}")
            tight.extract(f, iso, g, flags=flags)
            d = tight.download()
            for k in KEYS:
                assert_same(d[k], ref[k], f"extraction {rep}: {k}")
    finally:
        usual.close()
        tight.close()
