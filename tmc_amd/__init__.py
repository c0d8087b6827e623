"""tmc_amd -- MI355X Dual Marching Cubes engine (Python binding over the C ABI).

The engine is libtmc_dmc.so (HIP, gfx950), declared in include/tmc_dmc.h.
This module loads it with ctypes and exposes:

* ``Engine`` -- one device context: ``extract`` runs the whole DMC path on
  the GPU, ``download`` copies the mesh back (positions/normals FP64, quads,
  triangles, halfedges {origin, face, next, twin, flag}, he_v, he_f and the
  boundary vertices of getBndVertices).
* ``DualMarchingCubes`` -- the reference call surface
  ``dualMC(i0, ugrid, sFlag)`` (cpp_dmc/DualMarchingCubes.h:163-165).

There is no CPU fallback: if the library is missing or no GPU is present,
the calls raise.
"""
import ctypes as C
import os

import numpy as np

from .build import ENGINE_SO

__all__ = ["Engine", "Server", "DualMarchingCubes", "UniformGrid", "DMCError", "SIMPLIFY_3X3Y_COLOR",
           "SIMPLIFY_3X3Y_ISOLATED", "SIMPLIFY_3333", "SIMPLIFY_ALL", "TRIANGLES", "GRADIENTS_SUPPLIED",
           "FIELD_ON_DEVICE", "STAGE_TIMING", "load_library"]

SIMPLIFY_3X3Y_COLOR = 1 << 0
SIMPLIFY_3X3Y_ISOLATED = 1 << 1
SIMPLIFY_3333 = 1 << 2
TRIANGLES = 1 << 3
GRADIENTS_SUPPLIED = 1 << 4
FIELD_ON_DEVICE = 1 << 5
STAGE_TIMING = 1 << 6
SIMPLIFY_ALL = SIMPLIFY_3X3Y_COLOR | SIMPLIFY_3X3Y_ISOLATED | SIMPLIFY_3333

FIELD_F32 = 0
FIELD_F64 = 1

# every symbol include/tmc_dmc.h declares
EXPORTS = ["dmc_create", "dmc_destroy", "dmc_extract", "dmc_extract_batch", "dmc_mesh_download",
           "dmc_batch_download", "dmc_server_create", "dmc_server_destroy", "dmc_server_submit", "dmc_server_poll",
           "dmc_server_wait", "dmc_server_download", "dmc_server_release", "dmc_server_stats",
           "dmc_server_last_error", "dmc_mesh_download_f32", "dmc_stage_times", "dmc_device_alloc",
           "dmc_device_free", "dmc_copy_to_device", "dmc_copy_to_host", "dmc_synchronize", "dmc_status_str",
           "dmc_last_error", "dmc_mesh_edges", "dmc_valence_histogram", "dmc_element_quality", "dmc_write_obj",
           "tmc_extract", "tmc_download", "dmc_get_launch", "dmc_set_launch", "dmc_autotune",
           "dmc_extract_noalloc", "dmc_context_bytes", "dmc_context_trim"]


class DMCError(RuntimeError):
    pass


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]

    @classmethod
    def make(cls, shape, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0)):
        """shape is (nz, ny, nx), the numpy order of an x-fastest field."""
        nz, ny, nx = (int(s) for s in shape)
        return cls(nx, ny, nz, *[float(o) for o in origin], *[float(s) for s in spacing])


class Volume(C.Structure):
    """dmc_volume: one volume of a batch."""
    _fields_ = [("grid", Grid), ("field", C.c_void_p), ("gradients", C.c_void_p), ("iso", C.c_double)]


class TmcMesh(C.Structure):
    """tmc_mesh: the topologically correct MC result (counts and device
    pointers)."""
    _fields_ = [("nv", C.c_int64), ("nt", C.c_int64), ("d_positions", C.c_void_p), ("d_normals", C.c_void_p),
                ("d_tris", C.c_void_p), ("active_cells", C.c_int64), ("ambiguous_cells", C.c_int64),
                ("face_failures", C.c_int64)]


class Mesh(C.Structure):
    _fields_ = [("nv", C.c_int64), ("nq", C.c_int64), ("nt", C.c_int64),
                ("d_positions", C.c_void_p), ("d_normals", C.c_void_p), ("d_quads", C.c_void_p),
                ("d_tris", C.c_void_p), ("d_twins", C.c_void_p), ("d_he_flags", C.c_void_p),
                ("d_he_v", C.c_void_p),
                ("nonmanifold_edges", C.c_int64), ("nbnd", C.c_int64), ("edge_errors", C.c_int64),
                ("invalid_refs", C.c_int64), ("valence_zero", C.c_int64),
                ("failed_projections_l1", C.c_uint32), ("failed_projections_l2", C.c_uint32),
                ("active_cells", C.c_int64), ("nv_emitted", C.c_int64), ("nv_cells", C.c_int64),
                ("nq_cells", C.c_int64), ("host_syncs", C.c_int64), ("pass_mode", C.c_int64)]

    # how the pass ran, not what it produced
    RUN_FIELDS = ("host_syncs", "pass_mode")

    def summary(self):
        """Counts and diagnostics of the mesh (device pointers and the
        run fields left out)."""
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("d_") and f not in self.RUN_FIELDS}


class OutBuffers(C.Structure):
    """dmc_out_buffers: caller-owned device outputs of dmc_extract_noalloc."""
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("quads", C.c_void_p), ("tris", C.c_void_p),
                ("twins", C.c_void_p), ("he_flags", C.c_void_p), ("he_v", C.c_void_p), ("bnd", C.c_void_p),
                ("cap_v", C.c_int64), ("cap_q", C.c_int64), ("cap_bnd", C.c_int64)]


class Launch(C.Structure):
    _fields_ = [("grid_cap", C.c_int32), ("sweep_waves", C.c_int32)]


class TuneTrial(C.Structure):
    _fields_ = [("launch", Launch), ("ms", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("field_type", C.c_int), ("flags", C.c_uint32), ("max_batch", C.c_int32),
                ("budget_bytes", C.c_size_t), ("max_batch_points", C.c_int64), ("workers", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("mesh", Mesh), ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("batch_size", C.c_int32), ("status", C.c_int)]


_lib = None


def load_library(path=ENGINE_SO):
    """Load libtmc_dmc.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DMCError(f"{path} not built: run `python -m tmc_amd.build` (or __graft_entry__.build())")
    L = C.CDLL(path)
    vp, i32p, f64p, f32p = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_float)
    L.dmc_create.argtypes = [C.POINTER(vp), C.c_int]
    L.dmc_destroy.argtypes = [vp]
    L.dmc_extract.argtypes = [vp, C.POINTER(Grid), vp, C.c_int, C.c_double, C.c_uint32, vp, C.POINTER(Mesh)]
    L.dmc_mesh_download.argtypes = [vp, f64p, f64p, i32p, i32p, i32p, i32p, i32p, i32p]
    L.dmc_extract_batch.argtypes = [vp, C.POINTER(Volume), C.c_int32, C.c_int, C.c_uint32, C.POINTER(Mesh)]
    L.dmc_extract_noalloc.argtypes = [vp, C.POINTER(Grid), vp, C.c_int, C.c_double, C.c_uint32, vp,
                                      C.POINTER(OutBuffers), C.POINTER(Mesh)]
    L.dmc_context_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.dmc_context_trim.argtypes = [vp]
    L.dmc_batch_download.argtypes = [vp, C.c_int32, f64p, f64p, i32p, i32p, i32p, i32p, i32p, i32p]
    u64 = C.c_uint64
    L.dmc_server_create.argtypes = [C.POINTER(vp), C.POINTER(ServerOpts)]
    L.dmc_server_destroy.argtypes = [vp]
    L.dmc_server_submit.argtypes = [vp, C.POINTER(Volume), C.POINTER(u64)]
    L.dmc_server_poll.argtypes = [vp, u64, C.POINTER(C.c_int)]
    L.dmc_server_wait.argtypes = [vp, u64, C.POINTER(Result)]
    L.dmc_server_download.argtypes = [vp, u64, f64p, f64p, i32p, i32p, i32p, i32p, i32p, i32p]
    L.dmc_server_release.argtypes = [vp, u64]
    L.dmc_server_stats.argtypes = [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_size_t)]
    L.dmc_server_last_error.argtypes = [vp]
    L.dmc_server_last_error.restype = C.c_char_p
    L.dmc_mesh_download_f32.argtypes = [vp, f32p, f32p, C.c_int]
    L.dmc_stage_times.argtypes = [vp, C.POINTER(C.c_char_p), f32p, C.POINTER(C.c_int)]
    L.dmc_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    L.dmc_device_free.argtypes = [vp, vp]
    L.dmc_copy_to_device.argtypes = [vp, vp, vp, C.c_size_t]
    L.dmc_copy_to_host.argtypes = [vp, vp, vp, C.c_size_t]
    L.dmc_synchronize.argtypes = [vp]
    L.dmc_mesh_edges.argtypes = [vp, i32p, C.POINTER(C.c_int64)]
    L.dmc_valence_histogram.argtypes = [vp, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int64)]
    L.dmc_element_quality.argtypes = [vp, C.c_int, f32p]
    L.dmc_write_obj.argtypes = [vp, C.c_char_p, C.c_int]
    L.tmc_extract.argtypes = [vp, C.POINTER(Grid), vp, C.c_int, C.c_double, C.c_uint32, C.POINTER(TmcMesh)]
    L.tmc_download.argtypes = [vp, f64p, f64p, i32p]
    L.dmc_get_launch.argtypes = [vp, C.POINTER(Launch)]
    L.dmc_set_launch.argtypes = [vp, C.POINTER(Launch)]
    L.dmc_autotune.argtypes = [vp, C.POINTER(Volume), C.c_int32, C.c_int, C.c_uint32, C.c_int32,
                               C.POINTER(TuneTrial), C.POINTER(C.c_int32), C.POINTER(TuneTrial)]
    L.dmc_status_str.argtypes = [C.c_int]
    L.dmc_status_str.restype = C.c_char_p
    L.dmc_last_error.argtypes = [vp]
    L.dmc_last_error.restype = C.c_char_p
    for name in ("dmc_create", "dmc_destroy", "dmc_extract", "dmc_extract_batch", "dmc_mesh_download",
                 "dmc_batch_download", "dmc_server_create", "dmc_server_destroy", "dmc_server_submit",
                 "dmc_server_poll", "dmc_server_wait", "dmc_server_download", "dmc_server_release",
                 "dmc_server_stats", "dmc_mesh_download_f32",
                 "dmc_stage_times", "dmc_device_alloc", "dmc_device_free", "dmc_copy_to_device",
                 "dmc_copy_to_host", "dmc_synchronize", "dmc_mesh_edges", "dmc_valence_histogram",
                 "dmc_element_quality", "dmc_write_obj", "dmc_extract_noalloc", "dmc_context_bytes",
                 "dmc_context_trim"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _check_size(arr, grid, what="field", per_point=1):
    """A host array must hold exactly the grid's points (times per_point):
    the engine copies nx*ny*nz elements from its address."""
    need = int(grid.nx) * int(grid.ny) * int(grid.nz) * per_point
    if arr.size != need:
        raise DMCError(f"{what} has {arr.size} elements, the grid {grid.nx}x{grid.ny}x{grid.nz} needs {need}")


def _download(m, halfedges, call):
    """Host arrays for mesh `m`, filled by call(positions, normals, quads,
    tris, halfedges, he_v, he_f, bnd)."""
    v = np.empty((m.nv, 3), np.float64)
    n = np.empty((m.nv, 3), np.float64)
    q = np.empty((m.nq, 4), np.int32)
    t = np.empty((m.nt, 3), np.int32)
    he = np.empty((4 * m.nq, 5), np.int32) if halfedges else None
    hv = np.empty(m.nv, np.int32)
    hf = np.empty(m.nq, np.int32)
    bnd = np.empty(m.nbnd, np.int32)
    i32 = C.c_int32
    call(_p(v, C.c_double), _p(n, C.c_double), _p(q, i32), _p(t, i32) if m.nt else None,
         _p(he, i32) if he is not None else None, _p(hv, i32), _p(hf, i32), _p(bnd, i32) if m.nbnd else None)
    out = dict(v=v, n=n, quads=q, tris=t, he_v=hv, he_f=hf, bnd=bnd)
    if he is not None:
        out["he"] = he
    return out


class DeviceBuffer:
    """Device allocation owned by an Engine (freed on close/GC of the engine)."""

    def __init__(self, engine, nbytes):
        self.engine = engine
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        engine._check(engine.lib.dmc_device_alloc(engine.ctx, self.nbytes, C.byref(self.ptr)))

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.engine._check(self.engine.lib.dmc_copy_to_device(self.engine.ctx, self.ptr, arr.ctypes.data, arr.nbytes))
        return self

    def free(self):
        if self.ptr and self.ptr.value:
            self.engine.lib.dmc_device_free(self.engine.ctx, self.ptr)
            self.ptr = C.c_void_p()


class _RawDevice:
    def __init__(self, ptr):
        self.ptr = C.c_void_p(int(ptr))


class Engine:
    """One device context (one HIP stream + reusable device buffers)."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.ctx = C.c_void_p()
        s = self.lib.dmc_create(C.byref(self.ctx), int(device))
        if s != 0:
            raise DMCError(f"dmc_create: {self.lib.dmc_status_str(s).decode()}")
        self.mesh = None
        self.batch = None

    def _check(self, s):
        if s != 0:
            msg = self.lib.dmc_last_error(self.ctx).decode() if self.ctx else ""
            raise DMCError(f"{self.lib.dmc_status_str(s).decode()}: {msg}")

    def close(self):
        if self.ctx and self.ctx.value:
            self.lib.dmc_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def extract(self, field, iso, grid=None, flags=SIMPLIFY_ALL | TRIANGLES, gradients=None, shape=None,
                dtype=np.float32):
        """Run the DMC path.  `field` is a numpy array (nz, ny, nx) or a
        DeviceBuffer holding one (then pass `shape` and `dtype`)."""
        if isinstance(field, int):   # raw device pointer
            field = _RawDevice(field)
        if isinstance(field, (DeviceBuffer, _RawDevice)):
            fptr = field.ptr
            flags |= FIELD_ON_DEVICE
            if grid is None:
                grid = Grid.make(shape)
            ftype = FIELD_F32 if np.dtype(dtype) == np.float32 else FIELD_F64
            gptr = gradients.ptr if isinstance(gradients, DeviceBuffer) else None
            if gradients is not None and not isinstance(gradients, DeviceBuffer):
                raise DMCError("device field needs device gradients")
            keep = None
        else:
            arr = np.ascontiguousarray(field)
            if arr.dtype not in (np.float32, np.float64):
                arr = arr.astype(np.float32)
            ftype = FIELD_F32 if arr.dtype == np.float32 else FIELD_F64
            if grid is None:
                grid = Grid.make(arr.shape)
            _check_size(arr, grid)
            fptr = C.c_void_p(arr.ctypes.data)
            keep = [arr]
            gptr = None
            if gradients is not None:
                g = np.ascontiguousarray(gradients, dtype=np.float64)
                _check_size(g, grid, "gradients", 3)
                keep.append(g)
                gptr = C.c_void_p(g.ctypes.data)
        if gradients is not None:
            flags |= GRADIENTS_SUPPLIED
        if not isinstance(grid, Grid):
            grid = Grid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)
        m = Mesh()
        self._check(self.lib.dmc_extract(self.ctx, C.byref(grid), fptr, ftype, float(iso), int(flags), gptr,
                                         C.byref(m)))
        del keep
        self.mesh = m
        return m

    def tmc(self, field, iso, grid=None, shape=None, dtype=np.float32, download=True, timing=False):
        """Topologically correct Marching Cubes (tmc_extract).  `field` is a
        numpy array (nz, ny, nx) or a DeviceBuffer / device pointer holding
        one (then pass `shape` and `dtype`).  Returns a dict with the counts
        and, with download=True, v / n (float64, 3 per vertex) and tris
        (int32, 3 per triangle)."""
        flags = STAGE_TIMING if timing else 0
        if isinstance(field, int):
            field = _RawDevice(field)
        if isinstance(field, (DeviceBuffer, _RawDevice)):
            fptr = field.ptr
            flags |= FIELD_ON_DEVICE
            if grid is None:
                grid = Grid.make(shape)
            ftype = FIELD_F32 if np.dtype(dtype) == np.float32 else FIELD_F64
            keep = None
        else:
            arr = np.ascontiguousarray(field)
            if arr.dtype not in (np.float32, np.float64):
                arr = arr.astype(np.float32)
            ftype = FIELD_F32 if arr.dtype == np.float32 else FIELD_F64
            if grid is None:
                grid = Grid.make(arr.shape)
            _check_size(arr, grid)
            fptr = C.c_void_p(arr.ctypes.data)
            keep = [arr]
        if not isinstance(grid, Grid):
            grid = Grid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)
        m = TmcMesh()
        self._check(self.lib.tmc_extract(self.ctx, C.byref(grid), fptr, ftype, float(iso), int(flags), C.byref(m)))
        del keep
        out = {f: getattr(m, f) for f, _ in m._fields_ if not f.startswith("d_")}
        if download:
            v = np.zeros((m.nv, 3), np.float64)
            n = np.zeros((m.nv, 3), np.float64)
            t = np.zeros((m.nt, 3), np.int32)
            self._check(self.lib.tmc_download(self.ctx, _p(v, C.c_double), _p(n, C.c_double), _p(t, C.c_int32)))
            out.update(v=v, n=n, tris=t)
        return out

    @property
    def launch(self):
        """Launch shape of this context: {"grid_cap", "sweep_waves"}."""
        l = Launch()
        self._check(self.lib.dmc_get_launch(self.ctx, C.byref(l)))
        return {"grid_cap": l.grid_cap, "sweep_waves": l.sweep_waves}

    @launch.setter
    def launch(self, value):
        cur = self.launch
        cur.update(value)
        l = Launch(int(cur["grid_cap"]), int(cur["sweep_waves"]))
        self._check(self.lib.dmc_set_launch(self.ctx, C.byref(l)))

    def autotune(self, volumes, flags=SIMPLIFY_ALL | TRIANGLES, dtype=np.float32, reps=3):
        """Time extract_batch(volumes) under candidate launch shapes and keep
        the fastest on this context.  Returns {"best": {...}, "trials": [...]}
        with each shape and its median ms."""
        vols, ftype, keep = self._volumes(volumes, flags, dtype)
        cap = 16
        trials = (TuneTrial * cap)()
        nt = C.c_int32(cap)
        best = TuneTrial()
        self._check(self.lib.dmc_autotune(self.ctx, vols, len(volumes), ftype, int(flags), int(reps), trials,
                                          C.byref(nt), C.byref(best)))
        del keep
        self.mesh = None
        self.batch = None

        def rec(t):
            return {"grid_cap": t.launch.grid_cap, "sweep_waves": t.launch.sweep_waves, "ms": t.ms}

        return {"best": rec(best), "trials": [rec(trials[i]) for i in range(nt.value)]}

    def _volumes(self, volumes, flags, dtype):
        n = len(volumes)
        vols = (Volume * n)()
        keep = []
        ftype = FIELD_F32 if np.dtype(dtype) == np.float32 else FIELD_F64
        for i, (field, iso, grid) in enumerate(volumes):
            if flags & FIELD_ON_DEVICE:
                ptr = field.ptr.value if isinstance(field, DeviceBuffer) else int(field)
                if grid is None:
                    raise DMCError("a device field needs its grid")
            else:
                arr = np.ascontiguousarray(field, dtype=np.dtype(dtype))
                keep.append(arr)
                ptr = arr.ctypes.data
                if grid is None:
                    grid = Grid.make(arr.shape)
                _check_size(arr, grid, f"field of volume {i}")
            if not isinstance(grid, Grid):
                grid = Grid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)
            vols[i].grid = grid
            vols[i].field = ptr
            vols[i].gradients = None
            vols[i].iso = float(iso)
        return vols, ftype, keep

    def extract_batch(self, volumes, flags=SIMPLIFY_ALL | TRIANGLES, dtype=np.float32):
        """Extract many volumes in one pass.  `volumes` is a list of
        (field, iso, grid) with numpy fields (nz, ny, nx) -- or
        (device_pointer, iso, grid) with FIELD_ON_DEVICE in `flags`.  All
        fields share one element type.  Returns one Mesh per volume."""
        n = len(volumes)
        vols, ftype, keep = self._volumes(volumes, flags, dtype)
        meshes = (Mesh * n)()
        self._check(self.lib.dmc_extract_batch(self.ctx, vols, n, ftype, int(flags), meshes))
        del keep
        self.mesh = None
        self.batch = list(meshes)
        return self.batch

    def extract_noalloc(self, field, iso, grid, outs, flags=SIMPLIFY_ALL | TRIANGLES, dtype=np.float32,
                        gradients=None, bnd_cap=1 << 16):
        """dmc_extract_noalloc: `field` is a device pointer (int) or
        DeviceBuffer holding the volume, `outs` an OutBuffers of device
        buffers the caller owns.  Returns (Mesh, boundary vertices)."""
        fptr = field.ptr if isinstance(field, DeviceBuffer) else C.c_void_p(int(field))
        gptr = None
        if gradients is not None:
            gptr = gradients.ptr if isinstance(gradients, DeviceBuffer) else C.c_void_p(int(gradients))
            flags |= GRADIENTS_SUPPLIED
        ftype = FIELD_F32 if np.dtype(dtype) == np.float32 else FIELD_F64
        if not isinstance(grid, Grid):
            grid = Grid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)
        bnd = np.empty(bnd_cap, np.int32)
        outs.bnd = bnd.ctypes.data
        outs.cap_bnd = bnd_cap
        m = Mesh()
        self._check(self.lib.dmc_extract_noalloc(self.ctx, C.byref(grid), fptr, ftype, float(iso), int(flags), gptr,
                                                 C.byref(outs), C.byref(m)))
        self.mesh = None
        self.batch = None
        return m, bnd[:m.nbnd].copy()

    def context_bytes(self):
        """Device bytes held by this context's buffer pools."""
        n = C.c_size_t()
        self._check(self.lib.dmc_context_bytes(self.ctx, C.byref(n)))
        return n.value

    def trim(self):
        """Free this context's buffer pools and captured passes (the last
        mesh goes with them)."""
        self._check(self.lib.dmc_context_trim(self.ctx))
        self.mesh = None
        self.batch = None

    def copy_to_host(self, dst, src_ptr):
        """Copy dst.nbytes from device address src_ptr into numpy array dst."""
        if dst.nbytes:
            self._check(self.lib.dmc_copy_to_host(self.ctx, dst.ctypes.data, C.c_void_p(int(src_ptr)), dst.nbytes))
        return dst

    def download_batch(self, index, halfedges=True):
        """Copy volume `index` of the last extract_batch to host arrays."""
        return self.download(halfedges=halfedges, index=index)

    def download(self, halfedges=True, fp32=False, index=None):
        """Copy the last mesh (or volume `index` of the last batch) to host
        numpy arrays."""
        m = self.mesh if index is None else self.batch[index]
        if m is None:
            raise DMCError("no mesh: call extract first")
        if index is None:
            out = _download(m, halfedges, lambda *a: self._check(self.lib.dmc_mesh_download(self.ctx, *a)))
        else:
            out = _download(m, halfedges,
                            lambda *a: self._check(self.lib.dmc_batch_download(self.ctx, int(index), *a)))
        if fp32:
            v32 = np.empty((m.nv, 3), np.float32)
            n32 = np.empty((m.nv, 3), np.float32)
            self._check(self.lib.dmc_mesh_download_f32(self.ctx, _p(v32, C.c_float), _p(n32, C.c_float), 0))
            out.update(v32=v32, n32=n32)
        out.update(m.summary())
        return out

    def stage_times(self):
        cap = 256
        n = C.c_int(cap)
        names = (C.c_char_p * cap)()
        ms = (C.c_float * cap)()
        self._check(self.lib.dmc_stage_times(self.ctx, names, ms, C.byref(n)))
        return {names[i].decode(): float(ms[i]) for i in range(min(n.value, cap))}

    def synchronize(self):
        self._check(self.lib.dmc_synchronize(self.ctx))

    # -- diagnostics of the last mesh ------------------------------------
    def edges(self):
        """Unique undirected edges (n, 2) int32, one per vertex pair."""
        n = C.c_int64(0)
        self._check(self.lib.dmc_mesh_edges(self.ctx, None, C.byref(n)))
        e = np.empty((n.value, 2), np.int32)
        if n.value:
            self._check(self.lib.dmc_mesh_edges(self.ctx, e.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)))
        return e

    def valence_histogram(self, triangles=False):
        """(hist trimmed like the reference's, vertices of valence >= 100)
        of the quadrilateral mesh, or of the triangle mesh."""
        h = (C.c_uint32 * 100)()
        n = C.c_int32(0)
        beyond = C.c_int64(0)
        self._check(self.lib.dmc_valence_histogram(self.ctx, 1 if triangles else 0, h, C.byref(n),
                                                   C.byref(beyond)))
        return np.array(h[:n.value], np.uint32), beyond.value

    def element_quality(self, triangles=False):
        """Per-element quality (float32): quadrilaterals, or triangles."""
        meshes = [self.mesh] if self.mesh is not None else list(self.batch or [])
        n = sum(m.nt if triangles else m.nq for m in meshes)
        q = np.empty(n, np.float32)
        self._check(self.lib.dmc_element_quality(self.ctx, 1 if triangles else 0,
                                                 q.ctypes.data_as(C.POINTER(C.c_float))))
        return q

    def write_obj(self, path, triangles=False):
        """The mesh as OBJ text in cpp_dmc's main program layout."""
        self._check(self.lib.dmc_write_obj(self.ctx, os.fsencode(path), 1 if triangles else 0))


class Server:
    """Serving loop (dmc_server_*): submit volumes, wait for them, download,
    release.  Requests waiting together are extracted as one batch (up to
    `max_batch`, within `budget_bytes` of HBM); `workers` passes run
    concurrently, each on its own context and stream; max_batch=1 with one
    worker serves them one at a time."""

    def __init__(self, device=0, max_batch=1 << 20, budget_bytes=0, flags=SIMPLIFY_ALL | TRIANGLES,
                 dtype=np.float32, max_batch_points=0, workers=1):
        self.lib = load_library()
        self.srv = C.c_void_p()
        self.dtype = np.dtype(dtype)
        self.flags = int(flags)
        opts = ServerOpts(int(device), FIELD_F32 if self.dtype == np.float32 else FIELD_F64, self.flags,
                          int(max_batch), int(budget_bytes), int(max_batch_points), int(workers))
        s = self.lib.dmc_server_create(C.byref(self.srv), C.byref(opts))
        if s != 0:
            raise DMCError(f"dmc_server_create: {self.lib.dmc_status_str(s).decode()}")
        self._keep = {}

    def _check(self, s):
        if s != 0:
            raise DMCError(f"{self.lib.dmc_status_str(s).decode()}: "
                           f"{self.lib.dmc_server_last_error(self.srv).decode()}")

    def submit(self, field, iso, grid=None):
        """`field`: numpy array (kept alive until release) or, with
        FIELD_ON_DEVICE in the server's flags, a device pointer (int)."""
        vol = Volume()
        keep = None
        if self.flags & FIELD_ON_DEVICE:
            vol.field = field.ptr.value if isinstance(field, DeviceBuffer) else int(field)
        else:
            keep = np.ascontiguousarray(field, dtype=self.dtype)
            vol.field = keep.ctypes.data
            if grid is None:
                grid = Grid.make(keep.shape)
            _check_size(keep, grid)
        if not isinstance(grid, Grid):
            grid = Grid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)
        vol.grid = grid
        vol.gradients = None
        vol.iso = float(iso)
        t = C.c_uint64()
        self._check(self.lib.dmc_server_submit(self.srv, C.byref(vol), C.byref(t)))
        self._keep[t.value] = keep
        return t.value

    def poll(self, ticket):
        d = C.c_int()
        self._check(self.lib.dmc_server_poll(self.srv, int(ticket), C.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        r = Result()
        self._check(self.lib.dmc_server_wait(self.srv, int(ticket), C.byref(r)))
        out = r.mesh.summary()
        out.update(t_submit=r.t_submit, t_start=r.t_start, t_done=r.t_done, batch_size=r.batch_size,
                   latency=r.t_done - r.t_submit, mesh=r.mesh)
        return out

    def download(self, ticket, halfedges=True):
        r = Result()
        self._check(self.lib.dmc_server_wait(self.srv, int(ticket), C.byref(r)))
        return _download(r.mesh, halfedges,
                         lambda *a: self._check(self.lib.dmc_server_download(self.srv, int(ticket), *a)))

    def release(self, ticket):
        self._check(self.lib.dmc_server_release(self.srv, int(ticket)))
        self._keep.pop(int(ticket), None)

    def stats(self):
        b, n = C.c_uint64(), C.c_uint64()
        h, p = C.c_size_t(), C.c_size_t()
        self._check(self.lib.dmc_server_stats(self.srv, C.byref(b), C.byref(n), C.byref(h), C.byref(p)))
        return {"batches": b.value, "served": n.value, "held_bytes": h.value, "pool_bytes": p.value}

    def close(self):
        if self.srv and self.srv.value:
            self.lib.dmc_server_destroy(self.srv)
            self.srv = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UniformGrid:
    """The reference's grid description (cpp_dmc/UniformGrid.h:18-354):
    nx*ny*nz scalars, x fastest, bounding box origin and spacing."""

    def __init__(self, scalars, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), gradients=None):
        self.scalars = np.ascontiguousarray(scalars)
        self.origin = tuple(float(o) for o in origin)
        self.spacing = tuple(float(s) for s in spacing)
        self.gradients = gradients

    def grid(self):
        return Grid.make(self.scalars.shape, self.origin, self.spacing)


class DualMarchingCubes:
    """Reference call surface: dualMC(i0, ugrid, sFlag) -> (v, n, tris, quads)
    as cpp_dmc/DualMarchingCubes.h:163-165 returns them (flat lists become
    (n,3)/(n,4) arrays)."""

    def __init__(self, device=0):
        self.engine = Engine(device)
        self.bnd = np.zeros(0, np.int32)

    def dualMC(self, i0, ugrid, sFlag=True):
        flags = TRIANGLES | (SIMPLIFY_ALL if sFlag else 0)
        self.engine.extract(ugrid.scalars, i0, ugrid.grid(), flags=flags, gradients=ugrid.gradients)
        m = self.engine.download(halfedges=False)
        self.bnd = m["bnd"]
        return m["v"], m["n"], m["tris"], m["quads"]

    def getBndVertices(self):
        """cpp_dmc/DualMarchingCubes.h:171-182"""
        return self.bnd.copy()
