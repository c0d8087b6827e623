"""Synthetic workloads of the benchmark configurations (libtmc_workloads.so).

Each generator returns (field float32 array (nz, ny, nx), Grid, iso):

* ``genus2(n)``        -- genus-2 implicit on [-1.5,1.5]^3, cell-centred, iso 0
* ``blobs(n, seed)``   -- 64 Gaussian blobs + N(0, 0.05^2) noise, cell-centred
                          on [0,1]^3, iso 0.5
* ``schwarz(n, seed)`` -- cos x + cos y + cos z on [-8pi,8pi]^3, node-aligned
                          (x_i = -8pi + i*16pi/(n-1)), + N(0, 0.01^2), iso 0.3
* ``batch_specs``      -- the mixed batch of small volumes (4,096 volumes,
                          16-96 points per axis, log-uniform)
* ``serve_specs``      -- the serving mix (70% 32-128^3, 25% 256^3, 5% 512^3)

Mixed volumes draw their generator uniformly from the three above, each on
its own domain sampled at the volume's size, with iso U[-0.2, 0.2]
(genus-2, Schwarz-P) or U[0.3, 0.7] (blobs).
"""
import ctypes as C
import math
import os

import numpy as np

from . import Grid
from .build import WORKLOADS_SO

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(WORKLOADS_SO):
            raise RuntimeError(f"{WORKLOADS_SO} not built: run `python -m tmc_amd.build`")
        L = C.CDLL(WORKLOADS_SO)
        fp = C.POINTER(C.c_float)
        L.wl_genus2.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, fp, C.c_int]
        L.wl_schwarz.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                 fp, C.c_int]
        L.wl_blobs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_uint64, fp, C.c_int]
        L.wl_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.wl_uniform.restype = C.c_double
        L.wl_normal.argtypes = [C.c_uint64, C.c_uint64]
        L.wl_normal.restype = C.c_double
        _lib = L
    return _lib


def _shape3(n):
    return (n, n, n) if np.isscalar(n) else tuple(int(x) for x in n)   # (nz, ny, nx)


def _threads(threads=None):
    """Generator threads (0: all hardware threads)."""
    if threads is not None:
        return int(threads)
    return int(os.environ.get("TMC_WORKLOAD_THREADS", "0"))


def genus2(n=128, lo=-1.5, hi=1.5, threads=None):
    nz, ny, nx = _shape3(n)
    f = np.empty((nz, ny, nx), np.float32)
    _load().wl_genus2(nx, ny, nz, lo, hi, f.ctypes.data_as(C.POINTER(C.c_float)), _threads(threads))
    h = [(hi - lo) / m for m in (nx, ny, nz)]
    g = Grid(nx, ny, nz, lo + 0.5 * h[0], lo + 0.5 * h[1], lo + 0.5 * h[2], h[0], h[1], h[2])
    return f, g, 0.0


def schwarz(n=1024, seed=3, sigma=0.01, lo=-8 * math.pi, hi=8 * math.pi, iso=0.3, threads=None):
    nz, ny, nx = _shape3(n)
    f = np.empty((nz, ny, nx), np.float32)
    _load().wl_schwarz(nx, ny, nz, lo, hi, sigma, seed, f.ctypes.data_as(C.POINTER(C.c_float)), _threads(threads))
    g = Grid(nx, ny, nz, lo, lo, lo, (hi - lo) / (nx - 1), (hi - lo) / (ny - 1), (hi - lo) / (nz - 1))
    return f, g, iso


def blobs(n=512, seed=2, nblobs=64, sigma=0.05, iso=0.5, threads=None):
    nz, ny, nx = _shape3(n)
    f = np.empty((nz, ny, nx), np.float32)
    _load().wl_blobs(nx, ny, nz, nblobs, 0.02, 0.08, 0.5, 1.5, sigma, seed,
                     f.ctypes.data_as(C.POINTER(C.c_float)), _threads(threads))
    g = Grid(nx, ny, nz, 0.5 / nx, 0.5 / ny, 0.5 / nz, 1.0 / nx, 1.0 / ny, 1.0 / nz)
    return f, g, iso


CONFIGS = {
    "genus2_128": lambda: genus2(128),
    "blobs_512": lambda: blobs(512),
    "schwarz_1024": lambda: schwarz(1024),
}


def uniform(seed, k):
    return _load().wl_uniform(seed, k)


def normal(seed, idx):
    return _load().wl_normal(seed, idx)


KINDS = ("genus2", "schwarz", "blobs")


def _mixed(kind, shape, seed, iso, threads=None):
    if kind == "genus2":
        f, g, _ = genus2(shape, threads=threads)
    elif kind == "schwarz":
        f, g, _ = schwarz(shape, seed=seed, threads=threads)
    else:
        f, g, _ = blobs(shape, seed=seed, threads=threads)
    return f, g, iso


def _spec(seed, i, shape):
    """Generator, iso and noise seed of mixed volume i: counter-based
    uniforms u_k = uniform(seed, 16 i + k)."""
    u = [uniform(seed, 16 * i + k) for k in range(6)]
    kind = KINDS[min(int(u[3] * 3), 2)]
    iso = 0.3 + 0.4 * u[4] if kind == "blobs" else -0.2 + 0.4 * u[4]
    return {"kind": kind, "shape": shape, "iso": iso, "seed": (seed * 1000003 + i) & 0xFFFFFFFFFFFF}


def batch_specs(count=4096, seed=4):
    """The batched workload: each axis round(exp(U[ln 16, ln 96]))."""
    lo, hi = math.log(16), math.log(96)
    out = []
    for i in range(count):
        shape = tuple(int(round(math.exp(lo + (hi - lo) * uniform(seed, 16 * i + k)))) for k in (2, 1, 0))
        out.append(_spec(seed, i, shape))
    return out


def serve_specs(count, seed=5):
    """The serving mix: 70% U{32..128}^3, 25% 256^3, 5% 512^3."""
    out = []
    for i in range(count):
        u = uniform(seed, 16 * i + 7)
        if u < 0.70:
            n = 32 + min(int(uniform(seed, 16 * i + 8) * 97), 96)
        elif u < 0.95:
            n = 256
        else:
            n = 512
        out.append(_spec(seed, i, (n, n, n)))
    return out


def make(spec, threads=None):
    """(field, grid, iso) of one mixed-volume spec."""
    return _mixed(spec["kind"], spec["shape"], spec["seed"], spec["iso"], threads)


def make_many(specs, workers=16):
    """Generate many (small) volumes in parallel, one generator thread each."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda s: make(s, threads=1), specs))
