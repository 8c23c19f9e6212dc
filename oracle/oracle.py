"""ctypes binding of the FP64 oracle (oracle/femrt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"


class OrcMesh(C.Structure):
    _fields_ = [("n_elem", C.c_int), ("ng", C.c_int), ("nf", C.c_int), ("pad_", C.c_int),
                ("geom", C.c_void_p), ("field", C.c_void_p)]


class OrcQuads(C.Structure):
    _fields_ = [("n_quad", C.c_int), ("ng", C.c_int), ("nf", C.c_int), ("pad_", C.c_int),
                ("geom", C.c_void_p), ("field", C.c_void_p)]


class OrcPrisms(C.Structure):
    _fields_ = [("n", C.c_int), ("pad_", C.c_int), ("gm", C.c_int * 3), ("fm", C.c_int * 3),
                ("vertices", C.c_void_p), ("geom", C.c_void_p), ("field", C.c_void_p)]


class OrcTris(C.Structure):
    _fields_ = [("n", C.c_int), ("gm", C.c_int * 2), ("fm", C.c_int * 2), ("pad_", C.c_int),
                ("vertices", C.c_void_p), ("geom", C.c_void_p), ("field", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH}")
        L = C.CDLL(str(LIB_PATH))
        P, I, D = C.c_void_p, C.c_int, C.c_double
        sig = {
            "orc_modA": (D, [I, D]),
            "orc_modA_d": (D, [I, D]),
            "orc_eval_modal": (None, [P, I, P, P]),
            "orc_map": (None, [P, I, P, P, P]),
            "orc_invert": (I, [P, I, P, P]),
            "orc_field": (None, [P, I, P, P, P, P]),
            "orc_face_hit": (I, [P, I, I, P, P, D, D, P, P, P, P]),
            "orc_elem_box": (None, [P, I, I, D, P, P]),
            "orc_slab": (I, [P, P, P, P, D, D, P, P]),
            "orc_iso_segment": (I, [P, I, P, P, D, D, I, D, P, P, P, P]),
            "orc_trace": (None, [P, P, I, P, I, I, D, P, P]),
            "orc_camera_rays": (None, [P, P, P, D, I, I, P]),
            "orc_quad_map": (None, [P, I, D, D, P, P, P]),
            "orc_quad_field": (D, [P, I, D, D]),
            "orc_quad_trace": (None, [P, P, I, D, P]),
            "orc_jacobi": (D, [I, D, D, D]),
            "orc_modB": (D, [I, I, D]),
            "orc_prism_count": (I, [P]),
            "orc_prism_map": (None, [P, I, P, P]),
            "orc_prism_field": (D, [P, I, P]),
            "orc_tri_count": (I, [P]),
            "orc_tri_map": (None, [P, I, D, D, P]),
            "orc_tri_field": (D, [P, I, D, D]),
            "orc_geom_stride": (I, []),
            "orc_iso_stride": (I, []),
            "orc_face_acc": (D, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _d3(v):
    return np.ascontiguousarray(v, np.float64).reshape(3)


class Oracle:
    """FP64 oracle over a uniform-order hex mesh (float coefficients)."""

    def __init__(self, geom, field, ng, nf):
        self.geom = np.ascontiguousarray(geom, np.float32)
        self.field = np.ascontiguousarray(field, np.float32)
        self.ng, self.nf = int(ng), int(nf)
        self.n_elem = self.geom.size // (3 * self.ng ** 3)
        self.m = OrcMesh(self.n_elem, self.ng, self.nf, 0, self.geom.ctypes.data,
                         self.field.ctypes.data)

    @classmethod
    def from_mesh(cls, hm):
        return cls(hm.geom, hm.field, hm.ng, hm.nf)

    @property
    def mp(self):
        return C.byref(self.m)

    def map(self, e, rst):
        x = np.zeros(3)
        J = np.zeros(9)
        lib().orc_map(self.mp, int(e), _p(_d3(rst)), _p(x), _p(J))
        return x, J.reshape(3, 3)

    def invert(self, e, x):
        rst = np.zeros(3)
        it = lib().orc_invert(self.mp, int(e), _p(_d3(x)), _p(rst))
        return rst, it

    def field_at(self, e, rst):
        u = C.c_double(0)
        gr = np.zeros(3)
        gw = np.zeros(3)
        lib().orc_field(self.mp, int(e), _p(_d3(rst)), C.byref(u), _p(gr), _p(gw))
        return u.value, gr, gw

    def face_hit(self, e, f, o, d, tmin=1e-3, tmax=1e30):
        t = C.c_double(0)
        rst = np.zeros(3)
        mg = C.c_double(0)
        nm = C.c_double(0)
        ok = lib().orc_face_hit(self.mp, int(e), int(f), _p(_d3(o)), _p(_d3(d)), tmin, tmax,
                                C.byref(t), _p(rst), C.byref(mg), C.byref(nm))
        return bool(ok), t.value, rst, mg.value

    def elem_box(self, e, samples=11, pad=0.10):
        lo = np.zeros(3)
        hi = np.zeros(3)
        lib().orc_elem_box(self.mp, int(e), samples, pad, _p(lo), _p(hi))
        return lo, hi

    def iso_segment(self, e, o, d, ta, tb, iso, M=16):
        t = C.c_double(0)
        rst = np.zeros(3)
        u = C.c_double(0)
        mn = C.c_double(0)
        ok = lib().orc_iso_segment(self.mp, int(e), _p(_d3(o)), _p(_d3(d)), ta, tb, M, iso,
                                   C.byref(t), _p(rst), C.byref(u), C.byref(mn))
        return bool(ok), t.value, rst, u.value

    def trace(self, rays, isos=(), M=16, tmin=1e-3):
        rays = np.ascontiguousarray(rays, np.float32).reshape(6, -1)
        n = rays.shape[1]
        isos = np.ascontiguousarray(isos, np.float64)
        gs, is_ = lib().orc_geom_stride(), lib().orc_iso_stride()
        G = np.zeros((n, gs))
        I = np.zeros((n, max(1, len(isos)), is_))
        lib().orc_trace(self.mp, _p(rays), n, _p(isos) if len(isos) else None, len(isos), M,
                        tmin, _p(G), _p(I))
        geom = dict(elem=G[:, 0].astype(np.int64), face=G[:, 1].astype(np.int64), t=G[:, 2],
                    rst=G[:, 3:6].T.copy(), u=G[:, 6], t2=G[:, 7], margin=G[:, 8],
                    nearmiss=G[:, 9])
        iso = dict(elem=I[:, :, 0].T.astype(np.int64), t=I[:, :, 1].T.copy(),
                   rst=np.transpose(I[:, :, 2:5], (1, 2, 0)).copy(), u=I[:, :, 5].T.copy(),
                   minabs=I[:, :, 6].T.copy())
        return geom, iso


class QuadOracle:
    """FP64 oracle over curved quads (layout of femrt_mesh_upload_quads):
    nearest root over every quad from a dense seed grid."""

    def __init__(self, geom, field, ng, nf):
        self.geom = np.ascontiguousarray(geom, np.float32)
        self.field = np.ascontiguousarray(field, np.float32)
        self.ng, self.nf = int(ng), int(nf)
        self.n_quad = self.geom.size // (3 * self.ng ** 2)
        self.m = OrcQuads(self.n_quad, self.ng, self.nf, 0, self.geom.ctypes.data, self.field.ctypes.data)

    @classmethod
    def from_mesh(cls, qm):
        return cls(qm.geom, qm.field, qm.ng, qm.nf)

    def map(self, e, r, s):
        x, xr, xs = np.zeros(3), np.zeros(3), np.zeros(3)
        lib().orc_quad_map(C.byref(self.m), int(e), float(r), float(s), _p(x), _p(xr), _p(xs))
        return x, xr, xs

    def field_at(self, e, r, s):
        return lib().orc_quad_field(C.byref(self.m), int(e), float(r), float(s))

    def trace(self, rays, tmin=1e-3):
        rays = np.ascontiguousarray(rays, np.float32).reshape(6, -1)
        n = rays.shape[1]
        G = np.zeros((n, lib().orc_geom_stride()))
        lib().orc_quad_trace(C.byref(self.m), _p(rays), n, tmin, _p(G))
        return dict(elem=G[:, 0].astype(np.int64), face=G[:, 1].astype(np.int64), t=G[:, 2],
                    rst=G[:, 3:6].T.copy(), u=G[:, 6], t2=G[:, 7], margin=G[:, 8], nearmiss=G[:, 9])


class PrismOracle:
    """FP64 prism map and field in the reference's own basis (no hex
    embedding).  geom / field None: the linear map / a zero field."""

    def __init__(self, vertices, geom=None, field=None, gm=(0, 0, 0), fm=(0, 0, 0)):
        self.vertices = np.ascontiguousarray(vertices, np.float32)
        self.geom = None if geom is None else np.ascontiguousarray(geom, np.float32)
        self.field = None if field is None else np.ascontiguousarray(field, np.float32)
        n = self.vertices.size // 18
        self.m = OrcPrisms(n, 0, (C.c_int * 3)(*gm), (C.c_int * 3)(*fm), self.vertices.ctypes.data,
                           None if self.geom is None else self.geom.ctypes.data,
                           None if self.field is None else self.field.ctypes.data)

    def map(self, e, rst):
        x = np.zeros(3)
        lib().orc_prism_map(C.byref(self.m), int(e), _p(_d3(rst)), _p(x))
        return x

    def field_at(self, e, rst):
        return lib().orc_prism_field(C.byref(self.m), int(e), _p(_d3(rst)))


class TriOracle:
    """FP64 triangle map and field in the reference's own basis; geom None:
    the flat triangle in collapsed coordinates."""

    def __init__(self, vertices, geom=None, field=None, gm=(0, 0), fm=(0, 0)):
        self.vertices = np.ascontiguousarray(vertices, np.float32)
        self.geom = None if geom is None else np.ascontiguousarray(geom, np.float32)
        self.field = None if field is None else np.ascontiguousarray(field, np.float32)
        n = self.vertices.size // 9
        self.m = OrcTris(n, (C.c_int * 2)(*gm), (C.c_int * 2)(*fm), 0, self.vertices.ctypes.data,
                         None if self.geom is None else self.geom.ctypes.data,
                         None if self.field is None else self.field.ctypes.data)

    def map(self, e, r, s):
        x = np.zeros(3)
        lib().orc_tri_map(C.byref(self.m), int(e), float(r), float(s), _p(x))
        return x

    def field_at(self, e, r, s):
        return lib().orc_tri_field(C.byref(self.m), int(e), float(r), float(s))


def tri_count(m):
    return lib().orc_tri_count(_p(np.ascontiguousarray(m, np.int32)))


def jacobi(n, a, b, x):
    return lib().orc_jacobi(int(n), float(a), float(b), float(x))


def modB(i, k, x):
    return lib().orc_modB(int(i), int(k), float(x))


def prism_count(m):
    return lib().orc_prism_count(_p(np.ascontiguousarray(m, np.int32)))


def modA(i, x):
    return lib().orc_modA(int(i), float(x))


def modA_d(i, x):
    return lib().orc_modA_d(int(i), float(x))


def eval_modal(c, n, rst):
    c = np.ascontiguousarray(c, np.float32)
    out = np.zeros(4)
    lib().orc_eval_modal(_p(c), int(n), _p(_d3(rst)), _p(out))
    return out


def camera_rays(cam10, W, H):
    cam = np.asarray(cam10, np.float64)
    out = np.zeros((6, W * H))
    lib().orc_camera_rays(_p(_d3(cam[0:3])), _p(_d3(cam[3:6])), _p(_d3(cam[6:9])), float(cam[9]),
                          W, H, _p(out))
    return out
