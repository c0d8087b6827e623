"""Parity checkers for the DMC engine -- TEST INFRASTRUCTURE ONLY.

``Oracle`` wraps liboracle.so, the C restatement of the reference cpp_dmc
path (dmc_oracle.c) and of the reference's topologically correct MC
(tmc_oracle.c).  ``Reference`` wraps oracle/_ref/libdmcref.so, the
reference cpp_dmc compiled from its own sources (ref_harness.cpp);
``TmcReference`` wraps oracle/_ref/libtmcref.so, mc/MarchingCubes.cpp
compiled from its own sources (tmc_ref_harness.cpp).  Only
tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py use
this package; the engine (tmc_amd) never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdmcref.so")
TMC_REF_SO = os.path.join(HERE, "_ref", "libtmcref.so")

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]

    @classmethod
    def make(cls, shape, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0)):
        """shape is (nz, ny, nx) like the numpy field."""
        nz, ny, nx = shape
        return cls(nx, ny, nz, *[float(o) for o in origin], *[float(s) for s in spacing])


class _Mesh(C.Structure):
    _fields_ = [("nv", C.c_int64), ("v", f64p), ("n", f64p), ("nq", C.c_int64),
                ("quads", i32p), ("colors", i32p), ("nt", C.c_int64), ("tris", i32p),
                ("nr_proj1", C.c_uint32), ("nr_proj2", C.c_uint32),
                ("nv_before_unused", C.c_int64)]


class _HE(C.Structure):
    _fields_ = [("nhe", C.c_int64), ("he", i32p), ("he_v", i32p), ("he_f", i32p),
                ("nbnd", C.c_int64), ("bnd", i32p), ("nonmanifold", C.c_int64),
                ("errors", C.c_int64)]


def _g(grid):
    """Accept any object with the grid fields (e.g. tmc_amd.Grid)."""
    if isinstance(grid, Grid):
        return grid
    return Grid(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).astype(dtype, copy=True)


def _p(a, t):
    return a.ctypes.data_as(t)


class Mesh(dict):
    """Plain dict of numpy arrays: v (nv,3) f64, n (nv,3), quads (nq,4) i32, ..."""


def _mesh_from(m):
    out = Mesh()
    out["v"] = _arr(m.v, 3 * m.nv, np.float64).reshape(-1, 3)
    out["n"] = _arr(m.n, 3 * m.nv, np.float64).reshape(-1, 3)
    out["quads"] = _arr(m.quads, 4 * m.nq, np.int32).reshape(-1, 4)
    if m.colors:
        out["colors"] = _arr(m.colors, m.nq, np.int32)
    if m.tris and m.nt:
        out["tris"] = _arr(m.tris, 3 * m.nt, np.int32).reshape(-1, 3)
    out["nr_proj1"] = int(m.nr_proj1)
    out["nr_proj2"] = int(m.nr_proj2)
    out["nv_before_unused"] = int(m.nv_before_unused)
    return out


class Oracle:
    """ctypes view of liboracle.so (the C restatement)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` or __graft_entry__.build()")
        L = self.lib = C.CDLL(path)
        L.dmco_estimate_gradient.argtypes = [C.POINTER(Grid), f64p, f64p, C.c_int]
        L.dmco_cells.argtypes = [C.POINTER(Grid), f64p, f64p, C.c_double, C.POINTER(_Mesh)]
        L.dmco_mesh_free.argtypes = [C.POINTER(_Mesh)]
        L.dmco_halfedges_build.argtypes = [C.c_int64, C.c_int64, i32p, C.POINTER(_HE)]
        L.dmco_halfedges_free.argtypes = [C.POINTER(_HE)]
        L.dmco_color_faces.argtypes = [C.c_int64, i32p, i32p, i32p]
        L.dmco_vertex_valence.argtypes = [C.c_int64, C.c_int64, i32p, i32p]
        L.dmco_mark3x3y.argtypes = [C.c_int64, i32p, i32p, u8p]
        L.dmco_merge3x3y.argtypes = [C.c_int64, C.c_int64, f64p, f64p, u8p, i32p, i32p, i32p,
                                     i32p, u8p, i32p, u8p]
        L.dmco_mark3333.argtypes = [C.c_int64, C.c_int64, i32p, i32p, i32p, u8p, i32p, u8p, u8p]
        L.dmco_simplify3x3y.argtypes = [C.POINTER(_Mesh)]
        L.dmco_simplify3333.argtypes = [C.POINTER(_Mesh)]
        L.dmco_triangles.argtypes = [C.POINTER(_Mesh)]
        L.dmco_dualmc.argtypes = [C.POINTER(Grid), f64p, f64p, C.c_double, C.c_int,
                                  C.POINTER(_Mesh), C.POINTER(_HE)]
        L.dmco_dualmc_flags.argtypes = [C.POINTER(Grid), f64p, f64p, C.c_double, C.c_int,
                                        C.POINTER(_Mesh), C.POINTER(_HE)]
        L.dmco_cell_polygons.argtypes = [f64p, C.c_double, i32p, i32p]
        L.dmco_cell_vertices.argtypes = [f64p, C.c_double, f64p, u32p, u32p]
        L.tmco_tmc.argtypes = [C.POINTER(Grid), f64p, C.c_double, i64p, C.POINTER(f64p), C.POINTER(f64p), i64p,
                               C.POINTER(i32p), i64p]
        L.tmco_tmc.restype = C.c_int

    def tmc(self, grid, scalars, iso):
        """Topologically correct MC (mc/MarchingCubes.cpp t_mc): dict with
        v, n (float64, 3 per vertex), tris (int32, 3 per triangle),
        ambiguous (cells sliced with the decider), face_failures."""
        s = np.ascontiguousarray(scalars, dtype=np.float64).reshape(-1)
        nv, nt = C.c_int64(), C.c_int64()
        v, n, t = f64p(), f64p(), i32p()
        st = (C.c_int64 * 2)()
        grid = _g(grid)
        if self.lib.tmco_tmc(C.byref(grid), _p(s, f64p), float(iso), C.byref(nv), C.byref(v), C.byref(n),
                             C.byref(nt), C.byref(t), st) != 0:
            raise MemoryError("tmco_tmc")
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        out = {"v": _arr(v, 3 * nv.value, np.float64).reshape(-1, 3).copy(),
               "n": _arr(n, 3 * nv.value, np.float64).reshape(-1, 3).copy(),
               "tris": _arr(t, 3 * nt.value, np.int32).reshape(-1, 3).copy(),
               "ambiguous": int(st[0]), "face_failures": int(st[1])}
        for ptr in (v, n, t):
            libc.free(C.cast(ptr, C.c_void_p))
        return out

    # -- grid ---------------------------------------------------------------
    def gradient(self, grid, scalars, flip=True):
        s = np.ascontiguousarray(scalars, dtype=np.float64)
        g = np.zeros(s.size * 3, np.float64)
        self.lib.dmco_estimate_gradient(C.byref(_g(grid)), _p(s, f64p), _p(g, f64p), int(flip))
        return g

    # -- stages -------------------------------------------------------------
    def cells(self, grid, scalars, grad, iso):
        s = np.ascontiguousarray(scalars, dtype=np.float64)
        m = _Mesh()
        self.lib.dmco_cells(C.byref(_g(grid)), _p(s, f64p), _p(grad, f64p), float(iso), C.byref(m))
        out = _mesh_from(m)
        self.lib.dmco_mesh_free(C.byref(m))
        return out

    def halfedges(self, nv, quads):
        q = np.ascontiguousarray(quads, dtype=np.int32)
        h = _HE()
        self.lib.dmco_halfedges_build(int(nv), q.shape[0], _p(q, i32p), C.byref(h))
        out = {"he": _arr(h.he, 5 * h.nhe, np.int32).reshape(-1, 5),
               "he_v": _arr(h.he_v, nv, np.int32),
               "he_f": _arr(h.he_f, q.shape[0], np.int32),
               "bnd": _arr(h.bnd, h.nbnd, np.int32),
               "nonmanifold": int(h.nonmanifold), "errors": int(h.errors)}
        self.lib.dmco_halfedges_free(C.byref(h))
        return out

    def color_faces(self, he, he_f, colors):
        c = np.array(colors, dtype=np.int32)
        he = np.ascontiguousarray(he, dtype=np.int32)
        hf = np.ascontiguousarray(he_f, dtype=np.int32)
        self.lib.dmco_color_faces(c.size, _p(he, i32p), _p(hf, i32p), _p(c, i32p))
        return c

    def valence(self, nv, he):
        he = np.ascontiguousarray(he, dtype=np.int32)
        out = np.zeros(nv, np.int32)
        self.lib.dmco_vertex_valence(int(nv), he.shape[0], _p(he, i32p), _p(out, i32p))
        return out

    def mark3x3y(self, quads, valence):
        q = np.ascontiguousarray(quads, dtype=np.int32)
        vv = np.ascontiguousarray(valence, dtype=np.int32)
        p = np.zeros(q.shape[0], np.uint8)
        self.lib.dmco_mark3x3y(q.shape[0], _p(q, i32p), _p(vv, i32p), _p(p, u8p))
        return p

    def merge3x3y(self, v, n, p, valence, colors, he, he_f):
        v = np.array(v, dtype=np.float64, order="C")
        n = np.array(n, dtype=np.float64, order="C")
        nv, nq = v.shape[0], p.shape[0]
        vf = np.zeros(nv, np.uint8)
        vs = np.zeros(nv, np.int32)
        em = np.zeros(nq, np.uint8)
        col = None if colors is None else np.ascontiguousarray(colors, dtype=np.int32)
        self.lib.dmco_merge3x3y(nv, nq, _p(v, f64p), _p(n, f64p), _p(np.ascontiguousarray(p, np.uint8), u8p),
                                _p(np.ascontiguousarray(valence, np.int32), i32p),
                                None if col is None else _p(col, i32p),
                                _p(np.ascontiguousarray(he, np.int32), i32p),
                                _p(np.ascontiguousarray(he_f, np.int32), i32p),
                                _p(vf, u8p), _p(vs, i32p), _p(em, u8p))
        return v, n, vf, vs, em

    def mark3333(self, nv, he, he_f, valence):
        he = np.ascontiguousarray(he, np.int32)
        hf = np.ascontiguousarray(he_f, np.int32)
        nq = hf.size
        vf = np.zeros(nv, np.uint8)
        vt = np.zeros(nv, np.int32)
        p = np.zeros(nq, np.uint8)
        r = np.zeros(nq, np.uint8)
        self.lib.dmco_mark3333(int(nv), nq, _p(he, i32p), _p(hf, i32p),
                               _p(np.ascontiguousarray(valence, np.int32), i32p),
                               _p(vf, u8p), _p(vt, i32p), _p(p, u8p), _p(r, u8p))
        return vf, vt, p, r

    def _mesh_in(self, mesh):
        keep = []

        def hold(a):
            keep.append(a)
            return a
        m = _Mesh()
        # the oracle frees/reallocs these, so hand it malloc'd copies via the C helpers
        libc = C.CDLL(None)
        libc.malloc.restype = C.c_void_p
        libc.malloc.argtypes = [C.c_size_t]

        def cpy(a, t):
            a = np.ascontiguousarray(a)
            ptr = libc.malloc(max(a.nbytes, 8))
            C.memmove(ptr, a.ctypes.data, a.nbytes)
            return C.cast(ptr, t)
        m.nv = mesh["v"].shape[0]
        m.v = cpy(mesh["v"].astype(np.float64), f64p)
        m.n = cpy(mesh["n"].astype(np.float64), f64p)
        m.nq = mesh["quads"].shape[0]
        m.quads = cpy(mesh["quads"].astype(np.int32), i32p)
        m.colors = cpy(mesh.get("colors", np.zeros(m.nq, np.int32)).astype(np.int32), i32p)
        m.nt = 0
        m.tris = None
        return m

    def simplify3x3y(self, mesh):
        m = self._mesh_in(mesh)
        self.lib.dmco_simplify3x3y(C.byref(m))
        out = _mesh_from(m)
        self.lib.dmco_mesh_free(C.byref(m))
        return out

    def simplify3333(self, mesh):
        m = self._mesh_in(mesh)
        self.lib.dmco_simplify3333(C.byref(m))
        out = _mesh_from(m)
        self.lib.dmco_mesh_free(C.byref(m))
        return out

    def triangles(self, mesh):
        m = self._mesh_in(mesh)
        self.lib.dmco_triangles(C.byref(m))
        out = _mesh_from(m)
        self.lib.dmco_mesh_free(C.byref(m))
        return out["tris"] if "tris" in out else np.zeros((0, 3), np.int32)

    def dualmc(self, grid, scalars, grad, iso, sflag=True, flags=None):
        """Whole path.  flags (engine bits: 1 colored 3X3Y, 2 isolated 3X3Y,
        4 3333, 8 triangles) overrides sflag."""
        s = np.ascontiguousarray(scalars, dtype=np.float64)
        m = _Mesh()
        h = _HE()
        if flags is None:
            flags = (7 if sflag else 0) | 8
        self.lib.dmco_dualmc_flags(C.byref(_g(grid)), _p(s, f64p), _p(grad, f64p), float(iso), int(flags),
                                   C.byref(m), C.byref(h))
        out = _mesh_from(m)
        out["he"] = _arr(h.he, 5 * h.nhe, np.int32).reshape(-1, 5)
        out["he_v"] = _arr(h.he_v, out["v"].shape[0], np.int32)
        out["bnd"] = _arr(h.bnd, h.nbnd, np.int32)
        out["nonmanifold"] = int(h.nonmanifold)
        out["he_errors"] = int(h.errors)
        self.lib.dmco_halfedges_free(C.byref(h))
        self.lib.dmco_mesh_free(C.byref(m))
        return out

    def cell_polygons(self, f, iso):
        f = np.ascontiguousarray(f, dtype=np.float64)
        sizes = np.zeros(4, np.int32)
        edges = np.zeros(12, np.int32)
        cnt = self.lib.dmco_cell_polygons(_p(f, f64p), float(iso), _p(sizes, i32p), _p(edges, i32p))
        return cnt, sizes, edges

    def cell_vertices(self, f, iso):
        f = np.ascontiguousarray(f, dtype=np.float64)
        loc = np.zeros(12, np.float64)
        a, b = C.c_uint32(0), C.c_uint32(0)
        cnt = self.lib.dmco_cell_vertices(_p(f, f64p), float(iso), _p(loc, f64p), C.byref(a), C.byref(b))
        return loc[:3 * cnt].reshape(-1, 3), a.value, b.value


class Reference:
    """ctypes view of oracle/_ref/libdmcref.so (the compiled reference)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        L.dmcref_free.argtypes = [C.c_void_p]
        L.dmcref_r_pattern.restype = C.POINTER(C.c_int8)
        L.dmcref_t_ambig.restype = C.POINTER(C.c_uint8)
        L.dmcref_estimate_gradient.argtypes = [C.POINTER(Grid), f64p, f64p, C.c_int]
        L.dmcref_cells.argtypes = [C.POINTER(Grid), f64p, f64p, C.c_double, i64p, C.POINTER(f64p),
                                   C.POINTER(f64p), i64p, C.POINTER(i32p), C.POINTER(i32p), u32p, u32p, i64p]
        L.dmcref_halfedges.argtypes = [C.c_int64, C.c_int64, i32p, C.POINTER(i32p), C.POINTER(i32p),
                                       C.POINTER(i32p), i64p, C.POINTER(i32p)]
        L.dmcref_color_faces.argtypes = [C.c_int64, i32p, i32p, i32p]
        L.dmcref_vertex_valence.argtypes = [C.c_int64, C.c_int64, i32p, i32p]
        L.dmcref_mark3x3y.argtypes = [C.c_int64, i32p, i32p, C.c_int64, u8p]
        L.dmcref_merge3x3y.argtypes = [C.c_int64, C.c_int64, f64p, f64p, u8p, i32p, i32p, i32p,
                                       i32p, u8p, i32p, u8p]
        L.dmcref_mark3333.argtypes = [C.c_int64, C.c_int64, i32p, i32p, i32p, i32p, u8p, i32p, u8p, u8p]
        L.dmcref_simplify.argtypes = [C.c_int, i64p, C.POINTER(f64p), C.POINTER(f64p), i64p,
                                      C.POINTER(i32p), i32p]
        L.dmcref_triangles.argtypes = [C.c_int64, f64p, C.c_int64, i32p, i32p]
        L.dmcref_dualmc.argtypes = [C.POINTER(Grid), f64p, f64p, C.c_double, C.c_int, i64p,
                                    C.POINTER(f64p), C.POINTER(f64p), i64p, C.POINTER(i32p), i64p,
                                    C.POINTER(i32p)]
        L.dmcref_run_float.argtypes = [C.POINTER(Grid), C.POINTER(C.c_float), C.c_double, C.c_int, i64p,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.dmcref_run_float.restype = C.c_int64
        L.dmcref_cell_polygons.argtypes = [f64p, C.c_double, i32p, i32p]
        L.dmcref_cell_vertices.argtypes = [f64p, C.c_double, f64p, u32p, u32p]

    def _take(self, ptr, n, dtype):
        a = _arr(ptr, n, dtype)
        self.lib.dmcref_free(C.cast(ptr, C.c_void_p))
        return a

    def tables(self):
        rp = np.ctypeslib.as_array(self.lib.dmcref_r_pattern(), shape=(4352,)).copy()
        ta = np.ctypeslib.as_array(self.lib.dmcref_t_ambig(), shape=(256,)).copy()
        return rp, ta

    def gradient(self, grid, scalars, flip=True):
        s = np.ascontiguousarray(scalars, dtype=np.float64)
        g = np.zeros(s.size * 3, np.float64)
        self.lib.dmcref_estimate_gradient(C.byref(_g(grid)), _p(s, f64p), _p(g, f64p), int(flip))
        return g

    def cells(self, grid, scalars, grad, iso):
        s = np.ascontiguousarray(scalars, dtype=np.float64)
        nv, nq, nvb = C.c_int64(), C.c_int64(), C.c_int64()
        v, n = f64p(), f64p()
        q, c = i32p(), i32p()
        p1, p2 = C.c_uint32(), C.c_uint32()
        self.lib.dmcref_cells(C.byref(_g(grid)), _p(s, f64p), _p(grad, f64p), float(iso), C.byref(nv),
                              C.byref(v), C.byref(n), C.byref(nq), C.byref(q), C.byref(c),
                              C.byref(p1), C.byref(p2), C.byref(nvb))
        out = Mesh()
        out["v"] = self._take(v, 3 * nv.value, np.float64).reshape(-1, 3)
        out["n"] = self._take(n, 3 * nv.value, np.float64).reshape(-1, 3)
        out["quads"] = self._take(q, 4 * nq.value, np.int32).reshape(-1, 4)
        out["colors"] = self._take(c, nq.value, np.int32)
        out["nr_proj1"], out["nr_proj2"] = p1.value, p2.value
        out["nv_before_unused"] = nvb.value
        return out

    def halfedges(self, nv, quads):
        q = np.ascontiguousarray(quads, dtype=np.int32)
        he, hv, hf, bnd = i32p(), i32p(), i32p(), i32p()
        nb = C.c_int64()
        self.lib.dmcref_halfedges(int(nv), q.shape[0], _p(q, i32p), C.byref(he), C.byref(hv),
                                  C.byref(hf), C.byref(nb), C.byref(bnd))
        return {"he": self._take(he, 20 * q.shape[0], np.int32).reshape(-1, 5),
                "he_v": self._take(hv, nv, np.int32),
                "he_f": self._take(hf, q.shape[0], np.int32),
                "bnd": self._take(bnd, nb.value, np.int32)}

    def color_faces(self, he, he_f, colors):
        c = np.array(colors, dtype=np.int32)
        self.lib.dmcref_color_faces(c.size, _p(np.ascontiguousarray(he, np.int32), i32p),
                                    _p(np.ascontiguousarray(he_f, np.int32), i32p), _p(c, i32p))
        return c

    def valence(self, nv, he):
        he = np.ascontiguousarray(he, dtype=np.int32)
        out = np.zeros(nv, np.int32)
        self.lib.dmcref_vertex_valence(int(nv), he.shape[0], _p(he, i32p), _p(out, i32p))
        return out

    def mark3x3y(self, quads, valence):
        q = np.ascontiguousarray(quads, dtype=np.int32)
        vv = np.ascontiguousarray(valence, dtype=np.int32)
        p = np.zeros(q.shape[0], np.uint8)
        self.lib.dmcref_mark3x3y(q.shape[0], _p(q, i32p), _p(vv, i32p), vv.size, _p(p, u8p))
        return p

    def merge3x3y(self, v, n, p, valence, colors, he, he_f):
        v = np.array(v, dtype=np.float64, order="C")
        n = np.array(n, dtype=np.float64, order="C")
        nv, nq = v.shape[0], p.shape[0]
        vf = np.zeros(nv, np.uint8)
        vs = np.zeros(nv, np.int32)
        em = np.zeros(nq, np.uint8)
        col = None if colors is None else np.ascontiguousarray(colors, dtype=np.int32)
        self.lib.dmcref_merge3x3y(nv, nq, _p(v, f64p), _p(n, f64p), _p(np.ascontiguousarray(p, np.uint8), u8p),
                                  _p(np.ascontiguousarray(valence, np.int32), i32p),
                                  None if col is None else _p(col, i32p),
                                  _p(np.ascontiguousarray(he, np.int32), i32p),
                                  _p(np.ascontiguousarray(he_f, np.int32), i32p),
                                  _p(vf, u8p), _p(vs, i32p), _p(em, u8p))
        return v, n, vf, vs, em

    def mark3333(self, nv, quads, he, he_f, valence):
        q = np.ascontiguousarray(quads, np.int32)
        he = np.ascontiguousarray(he, np.int32)
        hf = np.ascontiguousarray(he_f, np.int32)
        nq = hf.size
        vf = np.zeros(nv, np.uint8)
        vt = np.zeros(nv, np.int32)
        p = np.zeros(nq, np.uint8)
        r = np.zeros(nq, np.uint8)
        self.lib.dmcref_mark3333(int(nv), nq, _p(q, i32p), _p(he, i32p), _p(hf, i32p),
                                 _p(np.ascontiguousarray(valence, np.int32), i32p),
                                 _p(vf, u8p), _p(vt, i32p), _p(p, u8p), _p(r, u8p))
        return vf, vt, p, r

    def simplify(self, which, mesh):
        nv = C.c_int64(mesh["v"].shape[0])
        nq = C.c_int64(mesh["quads"].shape[0])
        libc = C.CDLL(None)
        libc.malloc.restype = C.c_void_p
        libc.malloc.argtypes = [C.c_size_t]

        def cpy(a, t):
            a = np.ascontiguousarray(a)
            ptr = libc.malloc(max(a.nbytes, 8))
            C.memmove(ptr, a.ctypes.data, a.nbytes)
            return C.cast(ptr, t)
        v = cpy(mesh["v"].astype(np.float64), f64p)
        n = cpy(mesh["n"].astype(np.float64), f64p)
        q = cpy(mesh["quads"].astype(np.int32), i32p)
        col = np.ascontiguousarray(mesh["colors"], np.int32) if which == 0 else None
        self.lib.dmcref_simplify(which, C.byref(nv), C.byref(v), C.byref(n), C.byref(nq), C.byref(q),
                                 None if col is None else _p(col, i32p))
        out = Mesh()
        out["v"] = self._take(v, 3 * nv.value, np.float64).reshape(-1, 3)
        out["n"] = self._take(n, 3 * nv.value, np.float64).reshape(-1, 3)
        out["quads"] = self._take(q, 4 * nq.value, np.int32).reshape(-1, 4)
        return out

    def triangles(self, mesh):
        v = np.ascontiguousarray(mesh["v"], np.float64)
        q = np.ascontiguousarray(mesh["quads"], np.int32)
        t = np.zeros(6 * q.shape[0], np.int32)
        self.lib.dmcref_triangles(v.shape[0], _p(v, f64p), q.shape[0], _p(q, i32p), _p(t, i32p))
        return t.reshape(-1, 3)

    def dualmc(self, grid, scalars, grad, iso, sflag=True):
        s = np.ascontiguousarray(scalars, dtype=np.float64)
        nv, nq, nt = C.c_int64(), C.c_int64(), C.c_int64()
        v, n = f64p(), f64p()
        q, t = i32p(), i32p()
        self.lib.dmcref_dualmc(C.byref(_g(grid)), _p(s, f64p), _p(grad, f64p), float(iso), int(sflag),
                               C.byref(nv), C.byref(v), C.byref(n), C.byref(nq), C.byref(q),
                               C.byref(nt), C.byref(t))
        out = Mesh()
        out["v"] = self._take(v, 3 * nv.value, np.float64).reshape(-1, 3)
        out["n"] = self._take(n, 3 * nv.value, np.float64).reshape(-1, 3)
        out["quads"] = self._take(q, 4 * nq.value, np.int32).reshape(-1, 4)
        out["tris"] = self._take(t, 3 * nt.value, np.int32).reshape(-1, 3)
        return out

    def run_float(self, grid, field32, iso, sflag=True, legs=False):
        """UniformGrid from the float field, estimateGradient +
        flip_gradient, dualMC: (V, Q), or with legs=True (V, Q, gradient
        seconds, dualMC seconds) timed inside the library."""
        f = np.ascontiguousarray(field32, dtype=np.float32)
        nv = C.c_int64()
        tg, td = C.c_double(), C.c_double()
        nq = self.lib.dmcref_run_float(C.byref(_g(grid)), f.ctypes.data_as(C.POINTER(C.c_float)),
                                       float(iso), int(sflag), C.byref(nv), C.byref(tg), C.byref(td))
        return (nv.value, nq, tg.value, td.value) if legs else (nv.value, nq)

    def cell_polygons(self, f, iso):
        f = np.ascontiguousarray(f, dtype=np.float64)
        sizes = np.zeros(4, np.int32)
        edges = np.zeros(12, np.int32)
        cnt = self.lib.dmcref_cell_polygons(_p(f, f64p), float(iso), _p(sizes, i32p), _p(edges, i32p))
        return cnt, sizes, edges

    def cell_vertices(self, f, iso):
        f = np.ascontiguousarray(f, dtype=np.float64)
        loc = np.zeros(12, np.float64)
        a, b = C.c_uint32(0), C.c_uint32(0)
        cnt = self.lib.dmcref_cell_vertices(_p(f, f64p), float(iso), _p(loc, f64p), C.byref(a), C.byref(b))
        return loc[:3 * cnt].reshape(-1, 3), a.value, b.value


class TmcReference:
    """ctypes view of oracle/_ref/libtmcref.so (the reference's t_mc
    compiled from mc/MarchingCubes.cpp)."""

    def __init__(self, path=TMC_REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        L.tmcref_run.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [f64p, C.c_double, C.c_int, i64p,
                                                                                    i64p]
        L.tmcref_fetch.argtypes = [f64p, f64p, i32p]

    def tmc(self, grid, scalars, iso, orient=False):
        """t_mc; with orient=True followed by connectivity() (operator()'s
        c_flag)."""
        s = np.ascontiguousarray(scalars, dtype=np.float64).reshape(-1)
        nv, nt = C.c_int64(), C.c_int64()
        self.lib.tmcref_run(grid.nx, grid.ny, grid.nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz,
                            _p(s, f64p), float(iso), int(orient), C.byref(nv), C.byref(nt))
        v = np.zeros((nv.value, 3), np.float64)
        n = np.zeros((nv.value, 3), np.float64)
        t = np.zeros((nt.value, 3), np.int32)
        self.lib.tmcref_fetch(_p(v, f64p), _p(n, f64p), _p(t, i32p))
        return {"v": v, "n": n, "tris": t}
