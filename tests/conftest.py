import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: full-size workloads (minutes)")


def _ensure_built():
    from tmc_amd import build
    if not os.path.exists(build.WORKLOADS_SO):
        build.build_workloads()
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build.build_oracle(with_reference=False)


@pytest.fixture(scope="session")
def oracle():
    _ensure_built()
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import REF_SO, Reference
    if not os.path.exists(REF_SO):
        pytest.skip("compiled reference (oracle/_ref) not available on this machine")
    return Reference()


@pytest.fixture(scope="session")
def engine():
    from tmc_amd import Engine, build
    if not os.path.exists(build.ENGINE_SO):
        build.build_engine()
    _ensure_built()
    e = Engine(0)
    yield e
    e.close()


GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden_files(prefix):
    return sorted(f for f in os.listdir(GOLDEN) if f.startswith(prefix) and f.endswith(".npz"))


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def grid_from(arr):
    from oracle import Grid
    a = [float(x) for x in arr]
    return Grid(int(a[0]), int(a[1]), int(a[2]), *a[3:])
