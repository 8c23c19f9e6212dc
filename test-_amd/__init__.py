"""MI355X curved-hex FEM ray tracer: Python binding of the C ABI (include/femrt.h).

The product is ``_build/libfemrt.so`` (gfx950 HIP kernels + host C++); this
module only loads it with ctypes and moves numpy arrays across the boundary.
Nothing here computes the method itself, and there is no CPU fallback: when
the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "_build" / "libfemrt.so"

OK = 0
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EUNSUP", -5: "EIO", -6: "ESTATE"}
GUESS_PLAIN, GUESS_WARM, GUESS_LEARNED = 0, 1, 2
TRACE_RGBA, TRACE_BINNED_WALK, TRACE_LEARNED_EVERY_SAMPLE, TRACE_NO_SAMPLE_SKIP = 1, 2, 4, 8   # trace flag bits
TRACE_COOP_SCAN = 16
TRACE_ISO_BISECT, TRACE_ISO_LEARNED_BRACKET, TRACE_SEG_INFO, TRACE_NESTED_FIRST_HIT = 32, 64, 128, 256
TRACE_NO_HERMITE_START = 512
TRACE_NO_SKIP_JUMP = 1024
GRAD_MFMA, GRAD_VALU = 0, 1   # femrt_frame_iso_gradients engines
BVH_HOST_SAH, BVH_GPU_LBVH = 0, 1   # femrt_mesh_build_bvh builders
MAX_ISO = 8
NCOUNTERS = 34
NOALLOC_MAX_POINTS = 65536
LAUNCH_CFG_N = 5   # launch shape values (femrt_mesh_get_launch_cfg)
COUNTER_NAMES = ["rays", "hit_rays", "face_newton_it", "invert_it", "field_evals",
                 "geo64_it", "field64_evals", "slab_tests", "mlp_infer", "elems_marched",
                 "exit_pred", "exit_search", "exit_march", "reentry", "redo", "brackets",
                 "walk_t_start", "walk_t_drained", "walk_t_end",
                 "len_1", "len_2_3", "len_4_7", "len_8_15", "len_16_31", "len_32_63", "len_64_127",
                 "len_128_up", "scan_skipped", "samples_skipped", "walks_stalled",
                 "first_hit_face_it", "first_hit_slab_tests", "walk_skips", "walk_jumps"]


class FemrtError(RuntimeError):
    pass


class TraceParams(C.Structure):
    _fields_ = [("n_iso", C.c_int), ("iso_enable", C.c_int), ("iso", C.c_double * MAX_ISO),
                ("samples", C.c_int), ("guess", C.c_int), ("t_min", C.c_float),
                ("t_max", C.c_float), ("max_march", C.c_int), ("flags", C.c_int)]


class MlpTrainParams(C.Structure):
    _fields_ = [("n_samples", C.c_int), ("batch", C.c_int), ("epochs", C.c_int),
                ("lr", C.c_float), ("lr_final", C.c_float), ("seed", C.c_uint64)]


MLP_GUESS = 0
MLP_BRACKET = 1
MLP_SHAPE = {MLP_GUESS: (27, 3), MLP_BRACKET: (32, 2)}   # (inputs, outputs)
MLP_NPARAM = 6147
MLP_IN = 27


class Material(C.Structure):
    _fields_ = [("color", C.c_float * 3), ("emission", C.c_float * 3), ("roughness", C.c_float),
                ("ior", C.c_float), ("type", C.c_int)]


LAMBERT, METAL, DIELECTRIC = 0, 1, 2


def materials(specs):
    """[(type, color, emission, roughness, ior), ...] -> ctypes array."""
    arr = (Material * len(specs))()
    for i, (ty, col, em, rough, ior) in enumerate(specs):
        arr[i].type = ty
        arr[i].color[:] = list(col)
        arr[i].emission[:] = list(em)
        arr[i].roughness = rough
        arr[i].ior = ior
    return arr


# the reference's default scene materials (main.cu:39-41): grey and blue diffuse
DEFAULT_MATERIALS = [(LAMBERT, (0.5, 0.5, 0.5), (0, 0, 0), 0.0, 0.0),
                     (LAMBERT, (0.1, 0.2, 0.8), (0, 0, 0), 0.0, 0.0)]


def camera_basis(lookfrom, lookat, vup, vfov, aspect):
    """Camera of raytracer.h:31-48 (focus distance 1): origin,
    lower_left_corner, horizontal, vertical as 12 floats (float32 math)."""
    f = np.float32
    lf, la, up = (np.asarray(v, f) for v in (lookfrom, lookat, vup))
    h = f(np.tan(f(vfov) * f(3.14159265) / f(180.0) / f(2.0)))
    vh = f(2.0) * h
    vw = f(aspect) * vh
    w = lf - la
    w = w / f(np.sqrt(np.dot(w, w)))
    u = np.cross(up, w).astype(f)
    u = u / f(np.sqrt(np.dot(u, u)))
    v = np.cross(w, u).astype(f)
    hor = vw * u
    ver = vh * v
    ll = lf - hor / f(2) - ver / f(2) - w
    return np.concatenate([lf, ll, hor, ver]).astype(f)


class GenParams(C.Structure):
    _fields_ = [("n", C.c_int), ("q", C.c_int), ("p", C.c_int), ("pert", C.c_int),
                ("amp", C.c_double), ("seed", C.c_uint64)]


_lib = None


def lib():
    """Load the native library (raises FemrtError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FemrtError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(str(LIB_PATH))
    P, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
    sig = {
        "femrt_version": (I, []),
        "femrt_ctx_create": (I, [I, C.POINTER(P)]),
        "femrt_ctx_destroy": (I, [P]),
        "femrt_last_error": (C.c_char_p, [P]),
        "femrt_sync": (I, [P]),
        "femrt_timing_enable": (I, [P, I]),
        "femrt_timing_reset": (I, [P]),
        "femrt_timing_read": (I, [P, I, P, P, P]),
        "femrt_mesh_upload_hex": (I, [P, I, I, I, P, P, P, P, C.POINTER(P)]),
        "femrt_mesh_upload_quads": (I, [P, I, I, I, P, P, P, C.POINTER(P)]),
        "femrt_mesh_upload_prisms": (I, [P, I, P, P, P, P, P, C.POINTER(P)]),
        "femrt_host_prism_to_hex": (I, [I, P, P, P, P, P, C.POINTER(I), C.POINTER(I), P, P, P]),
        "femrt_mesh_upload_tris": (I, [P, I, P, P, P, P, P, C.POINTER(P)]),
        "femrt_host_tri_to_quad": (I, [I, P, P, P, P, P, C.POINTER(I), C.POINTER(I), P, P, P]),
        "femrt_mesh_destroy": (I, [P]),
        "femrt_mesh_boxes": (I, [P, P]),
        "femrt_mesh_build_bvh": (I, [P, I, C.POINTER(I)]),
        "femrt_mesh_get_bvh": (I, [P, P, P]),
        "femrt_eval_field": (I, [P, I, P, P, I, P]),
        "femrt_eval_field_noalloc": (I, [P, I, P, P, I, P]),
        "femrt_invert_points": (I, [P, I, P, P, P, P]),
        "femrt_trace_params_default": (None, [P]),
        "femrt_frame_create": (I, [P, I, C.POINTER(P)]),
        "femrt_frame_destroy": (I, [P]),
        "femrt_frame_set_rays": (I, [P, P]),
        "femrt_frame_camera_rays": (I, [P, P, I, I]),
        "femrt_frame_get_rays": (I, [P, P]),
        "femrt_trace": (I, [P, P, P]),
        "femrt_frame_get_hits": (I, [P, P, P, P, P, P]),
        "femrt_frame_get_iso": (I, [P, I, P, P, P, P]),
        "femrt_frame_get_rgba": (I, [P, P]),
        "femrt_frame_counters": (I, [P, I, P]),
        "femrt_frame_stalled_rays": (I, [P, I, P]),
        "femrt_gen_structured_hex": (I, [P, P, P, P, P]),
        "femrt_gen_field_waves": (I, [I, I, P, I, I, C.c_ulonglong, P]),
        "femrt_host_boxes": (I, [I, I, P, P]),
        "femrt_host_field_range": (I, [I, I, P, P]),
        "femrt_host_field_lipschitz": (I, [I, I, P, P]),
        "femrt_host_geom_jinv_bound": (I, [I, I, P, P]),
        "femrt_host_build_bvh": (I, [I, P, C.POINTER(I), P, P]),
        "femrt_load_xml_fld": (I, [C.c_char_p, C.c_char_p, I, C.POINTER(I), C.POINTER(I),
                                   C.POINTER(I), P, P, P, P]),
        "femrt_save_xml_fld": (I, [C.c_char_p, C.c_char_p, I, I, I, P, P, P]),
        "femrt_io_last_error": (C.c_char_p, []),
        "femrt_frame_camera_rays_jitter": (I, [P, P, I, I, I, C.c_ulonglong]),
        "femrt_frame_path_begin": (I, [P]),
        "femrt_frame_scatter": (I, [P, P, P, I, C.c_ulonglong, I, I]),
        "femrt_frame_resolve": (I, [P, I, I, I, P]),
        "femrt_render_path": (I, [P, P, P, I, I, I, I, P, I, C.c_ulonglong, P]),
        "femrt_mlp_num_params": (I, [I]),
        "femrt_mlp_create": (I, [P, I, C.c_ulonglong, C.POINTER(P)]),
        "femrt_mlp_destroy": (I, [P]),
        "femrt_mlp_get_params": (I, [P, P]),
        "femrt_mlp_set_params": (I, [P, P]),
        "femrt_mlp_forward": (I, [P, I, P, P]),
        "femrt_mlp_loss_grad": (I, [P, I, P, P, P, P]),
        "femrt_mlp_make_samples": (I, [P, I, C.c_ulonglong, P, P]),
        "femrt_mlp_train_params_default": (None, [P]),
        "femrt_mlp_train": (I, [P, P, P, P]),
        "femrt_mesh_set_guess_mlp": (I, [P, P]),
        "femrt_mesh_set_bracket_mlp": (I, [P, P]),
        "femrt_mlp_train_pairs": (I, [P, I, P, P, P, P]),
        "femrt_frame_bracket_pairs": (I, [P, P, I, P, I, P, P, P]),
        "femrt_frame_iso_gradients": (I, [P, P, I, I, P]),
        "femrt_frame_set_materials": (I, [P, P, I]),
        "femrt_batch_create": (I, [P, C.POINTER(P)]),
        "femrt_batch_destroy": (I, [P]),
        "femrt_batch_add_trace": (I, [P, P, P, P]),
        "femrt_batch_add_paths": (I, [P, P, P, P, I, I, I, I, P, I, C.c_ulonglong]),
        "femrt_batch_capture": (I, [P]),
        "femrt_batch_launch": (I, [P]),
        "femrt_mesh_get_launch_cfg": (I, [P, P, I]),
        "femrt_mesh_set_launch_cfg": (I, [P, P, I]),
        "femrt_autotune": (I, [P, P, P, I, P, I]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/femrt.h that the loaded library must export."""
    return ["femrt_version", "femrt_ctx_create", "femrt_ctx_destroy", "femrt_last_error",
            "femrt_sync", "femrt_timing_enable", "femrt_timing_reset", "femrt_timing_read",
            "femrt_mesh_upload_hex", "femrt_mesh_upload_quads", "femrt_mesh_upload_prisms",
            "femrt_host_prism_to_hex", "femrt_mesh_upload_tris", "femrt_host_tri_to_quad",
            "femrt_mesh_destroy", "femrt_mesh_boxes",
            "femrt_mesh_build_bvh", "femrt_mesh_get_bvh", "femrt_eval_field", "femrt_eval_field_noalloc",
            "femrt_invert_points", "femrt_trace_params_default", "femrt_frame_create",
            "femrt_frame_destroy", "femrt_frame_set_rays", "femrt_frame_camera_rays",
            "femrt_frame_get_rays", "femrt_trace", "femrt_frame_get_hits", "femrt_frame_get_iso",
            "femrt_frame_get_rgba", "femrt_frame_counters", "femrt_frame_stalled_rays", "femrt_gen_structured_hex",
            "femrt_gen_field_waves",
            "femrt_load_xml_fld", "femrt_host_boxes", "femrt_host_build_bvh", "femrt_host_field_range",
            "femrt_host_field_lipschitz", "femrt_host_geom_jinv_bound",
            "femrt_save_xml_fld", "femrt_io_last_error",
            "femrt_frame_camera_rays_jitter", "femrt_frame_path_begin", "femrt_frame_scatter",
            "femrt_frame_resolve", "femrt_render_path",
            "femrt_mlp_num_params", "femrt_mlp_create", "femrt_mlp_destroy",
            "femrt_mlp_get_params", "femrt_mlp_set_params", "femrt_mlp_forward",
            "femrt_mlp_loss_grad", "femrt_mlp_make_samples", "femrt_mlp_train_params_default",
            "femrt_mlp_train", "femrt_mesh_set_guess_mlp", "femrt_frame_set_materials",
            "femrt_mesh_set_bracket_mlp", "femrt_mlp_train_pairs", "femrt_frame_bracket_pairs",
            "femrt_frame_iso_gradients",
            "femrt_batch_create", "femrt_batch_destroy", "femrt_batch_add_trace",
            "femrt_batch_add_paths", "femrt_batch_capture", "femrt_batch_launch",
            "femrt_mesh_get_launch_cfg", "femrt_mesh_set_launch_cfg", "femrt_autotune"]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, ctx=None, what=""):
    if rc != OK:
        msg = ""
        if ctx is not None:
            m = lib().femrt_last_error(ctx)
            msg = m.decode() if m else ""
        raise FemrtError(f"{what}: {ERRORS.get(rc, rc)} {msg}")


# ----------------------------------------------------------------- host data

class HexMesh:
    """Uniform-order hex mesh in the SoA layout of femrt_mesh_upload_hex."""

    def __init__(self, vertices, geom, field, adjacency, ng, nf):
        self.vertices = np.ascontiguousarray(vertices, np.float32)
        self.geom = np.ascontiguousarray(geom, np.float32)
        self.field = np.ascontiguousarray(field, np.float32)
        self.adjacency = None if adjacency is None else np.ascontiguousarray(adjacency, np.int32)
        self.ng, self.nf = int(ng), int(nf)
        self.n_elem = self.vertices.shape[0]

    def geom_blocks(self):
        return self.geom.reshape(self.n_elem, 3, self.ng ** 3)


class QuadMesh:
    """Curved quadrilateral surface elements in the layout of
    femrt_mesh_upload_quads: vertices n x 4 x 3 (corner q = i + 2 j at
    (r, s) = (2 i - 1, 2 j - 1)), geometry n x 3 x ng^2 and field n x nf^2
    Modified_A coefficients, mode (i, j) at j * n + i (fem/QuadElement.cu)."""

    def __init__(self, vertices, geom, field, ng, nf):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4, 3)
        self.n_elem = self.vertices.shape[0]
        self.ng, self.nf = int(ng), int(nf)
        self.geom = np.ascontiguousarray(geom, np.float32).reshape(self.n_elem, 3, self.ng ** 2)
        self.field = np.ascontiguousarray(field, np.float32).reshape(self.n_elem, self.nf ** 2)


def gen_bumpy_quads(n, amp=0.2, seed=0, tilt=0.0) -> QuadMesh:
    """An n x n grid of curved quads over [0, 1]^2 at z = tilt * x, each
    raised by a seeded bump z += a (1 - r^2)(1 - s^2) (|a| <= amp per quad,
    the Modified_A mode (2, 2) with coefficient 16 a), at geometry order 2;
    the field has seeded order-2 coefficients uniform in [-1, 1]."""
    rng = np.random.default_rng(seed)
    ng, nf = 3, 3
    v = np.zeros((n * n, 4, 3), np.float32)
    g = np.zeros((n * n, 3, ng * ng), np.float32)
    f = rng.uniform(-1, 1, (n * n, nf * nf)).astype(np.float32)
    h = 1.0 / n
    for a in range(n):
        for b in range(n):
            e = a * n + b
            for q in range(4):
                i, j = q & 1, q >> 1
                x, y = (b + i) * h, (a + j) * h
                v[e, q] = (x, y, tilt * x)
                for c in range(3):
                    g[e, c, j * ng + i] = v[e, q, c]
            g[e, 2, 2 * ng + 2] = 16.0 * rng.uniform(-amp, amp) * h
    return QuadMesh(v, g, f, ng, nf)


def prism_count(m):
    """Coefficients per prism for modes (mx, my, mz): sum over i < mx of
    my (mz - i) (fem/PrismElement.cu:16-37)."""
    return sum(m[1] * max(m[2] - i, 0) for i in range(m[0]))


class PrismMesh:
    """Prisms in the layout of femrt_mesh_upload_prisms: vertices n x 6 x 3,
    geometry n x 3 x count(gm) (or None: linear), field n x count(fm) (or
    None: zero), the reference's Modified_A / Modified_B coefficient order."""

    def __init__(self, vertices, geom=None, field=None, gm=(0, 0, 0), fm=(0, 0, 0)):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 6, 3)
        self.n_elem = self.vertices.shape[0]
        self.gm, self.fm = tuple(int(v) for v in gm), tuple(int(v) for v in fm)
        self.geom = None if geom is None else np.ascontiguousarray(geom, np.float32).reshape(
            self.n_elem, 3, prism_count(self.gm))
        self.field = None if field is None else np.ascontiguousarray(field, np.float32).reshape(
            self.n_elem, prism_count(self.fm))

    def to_hex(self):
        """The hex embedding the library traces, in double:
        (vertices n x 8 x 3, geom n x 3 x ng^3, field n x nf^3, ng, nf)."""
        gm = np.array(self.gm, np.int32)
        fm = np.array(self.fm, np.int32)
        ng, nf = C.c_int(0), C.c_int(0)
        args = (self.n_elem, _ptr(gm), _ptr(fm), _ptr(self.vertices), _ptr(self.geom), _ptr(self.field))
        _check(lib().femrt_host_prism_to_hex(*args, C.byref(ng), C.byref(nf), None, None, None),
               what="host_prism_to_hex")
        v = np.zeros((self.n_elem, 8, 3), np.float32)
        g = np.zeros((self.n_elem, 3, ng.value ** 3))
        f = np.zeros((self.n_elem, nf.value ** 3))
        _check(lib().femrt_host_prism_to_hex(*args, C.byref(ng), C.byref(nf), _ptr(v), _ptr(g), _ptr(f)),
               what="host_prism_to_hex")
        return v, g, f, ng.value, nf.value


def tri_count(m):
    """Coefficients per triangle for modes (mx, my): sum over i < mx of
    my - i (fem/TriElement.cu:46-63)."""
    return sum(max(m[1] - i, 0) for i in range(m[0]))


class TriMesh:
    """Curved triangles in the layout of femrt_mesh_upload_tris: vertices
    n x 3 x 3, geometry n x 3 x count(gm) (or None: flat), field
    n x count(fm) (or None: zero), the reference's coefficient order."""

    def __init__(self, vertices, geom=None, field=None, gm=(0, 0), fm=(0, 0)):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3, 3)
        self.n_elem = self.vertices.shape[0]
        self.gm, self.fm = tuple(int(v) for v in gm), tuple(int(v) for v in fm)
        self.geom = None if geom is None else np.ascontiguousarray(geom, np.float32).reshape(
            self.n_elem, 3, tri_count(self.gm))
        self.field = None if field is None else np.ascontiguousarray(field, np.float32).reshape(
            self.n_elem, tri_count(self.fm))

    def to_quad(self):
        """The quad embedding the library traces, in double: (vertices
        n x 4 x 3, geom n x 3 x ng^2, field n x nf^2, ng, nf)."""
        gm = np.array(self.gm, np.int32)
        fm = np.array(self.fm, np.int32)
        ng, nf = C.c_int(0), C.c_int(0)
        args = (self.n_elem, _ptr(gm), _ptr(fm), _ptr(self.vertices), _ptr(self.geom), _ptr(self.field))
        _check(lib().femrt_host_tri_to_quad(*args, C.byref(ng), C.byref(nf), None, None, None),
               what="host_tri_to_quad")
        v = np.zeros((self.n_elem, 4, 3), np.float32)
        g = np.zeros((self.n_elem, 3, ng.value ** 2))
        f = np.zeros((self.n_elem, nf.value ** 2))
        _check(lib().femrt_host_tri_to_quad(*args, C.byref(ng), C.byref(nf), _ptr(v), _ptr(g), _ptr(f)),
               what="host_tri_to_quad")
        return v, g, f, ng.value, nf.value


def gen_tris(n, fm=(4, 4), amp=0.2, seed=0):
    """A piecewise-linear terrain over [0, 1]^2: an (n+1)^2 grid of seeded
    heights in [-amp, amp] / n, each cell split into two flat triangles, with
    seeded field coefficients uniform in [-1, 1] in the reference's
    triangle basis."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    z = rng.uniform(-amp, amp, (n + 1, n + 1)) * h
    P = lambda a, b: (b * h, a * h, z[a, b])  # noqa: E731
    v = []
    for a in range(n):
        for b in range(n):
            v.append([P(a, b), P(a, b + 1), P(a + 1, b)])
            v.append([P(a + 1, b + 1), P(a + 1, b), P(a, b + 1)])
    v = np.array(v, np.float32)
    f = rng.uniform(-1, 1, (len(v), tri_count(fm))).astype(np.float32)
    return TriMesh(v, None, f, (0, 0), fm)


def gen_prisms(n, fm=(4, 4, 4), seed=0):
    """An n^3 grid of unit-cube cells over [0, 1]^3, each split along its
    x-z diagonal into two linear prisms extruded along y (bottom quad, two
    triangle ends at s = -1 and s = +1, top edge v4 v5), with seeded field
    coefficients uniform in [-1, 1] in the reference's prism basis."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    v = []
    for a in range(n):
        for b in range(n):
            for c in range(n):
                x0, y0, z0 = c * h, b * h, a * h
                v.append([(x0, y0, z0), (x0 + h, y0, z0), (x0 + h, y0 + h, z0), (x0, y0 + h, z0),
                          (x0, y0, z0 + h), (x0, y0 + h, z0 + h)])
                v.append([(x0 + h, y0, z0 + h), (x0, y0, z0 + h), (x0, y0 + h, z0 + h),
                          (x0 + h, y0 + h, z0 + h), (x0 + h, y0, z0), (x0 + h, y0 + h, z0)])
    v = np.array(v, np.float32)
    f = rng.uniform(-1, 1, (len(v), prism_count(fm))).astype(np.float32)
    return PrismMesh(v, None, f, (0, 0, 0), fm)


def gen_structured_hex(n, q, p, pert=0, amp=0.0, seed=0) -> HexMesh:
    gp = GenParams(n, q, p, pert, amp, seed)
    ne, ng, nf = n ** 3, q + 1, p + 1
    v = np.zeros((ne, 8, 3), np.float32)
    g = np.zeros((ne, 3, ng ** 3), np.float32)
    f = np.zeros((ne, nf ** 3), np.float32)
    a = np.zeros((ne, 6, 2), np.int32)
    _check(lib().femrt_gen_structured_hex(C.byref(gp), _ptr(v), _ptr(g), _ptr(f), _ptr(a)),
           what="gen_structured_hex")
    return HexMesh(v, g, f, a, ng, nf)


def gen_field_waves(hm: HexMesh, p, n_waves=8, seed=0) -> HexMesh:
    """The mesh `hm` with its field replaced by a seeded multi-frequency
    field (a sum of `n_waves` plane waves of 1, 2, 4 and 8 periods over the
    unit cube, unit RMS) fitted at field order `p` on each element."""
    nf = p + 1
    f = np.zeros((hm.n_elem, nf ** 3), np.float32)
    _check(lib().femrt_gen_field_waves(hm.n_elem, hm.ng, _ptr(hm.geom), nf, n_waves, seed, _ptr(f)),
           what="gen_field_waves")
    return HexMesh(hm.vertices, hm.geom, f, hm.adjacency, hm.ng, nf)


NBOX = 61  # element box, 6 face boxes, 6 x 9 face sub-patch boxes


def host_boxes(hm: HexMesh):
    out = np.zeros((hm.n_elem, NBOX, 6), np.float32)
    _check(lib().femrt_host_boxes(hm.n_elem, hm.ng, _ptr(hm.geom), _ptr(out)), what="host_boxes")
    return out


def host_field_range(hm: HexMesh):
    """Per element (lo, hi) bounds of the field over the reference cube."""
    out = np.zeros((hm.n_elem, 2), np.float32)
    _check(lib().femrt_host_field_range(hm.n_elem, hm.nf, _ptr(hm.field), _ptr(out)), what="host_field_range")
    return out


def host_field_lipschitz(hm: HexMesh):
    """Per element, the bound of |u(a) - u(b)| / |a - b|_inf on the
    reference cube that the walk's early sign acceptance uses."""
    out = np.zeros(hm.n_elem, np.float32)
    _check(lib().femrt_host_field_lipschitz(hm.n_elem, hm.nf, _ptr(hm.field), _ptr(out)),
           what="host_field_lipschitz")
    return out


def host_geom_jinv_bound(hm: HexMesh):
    """Per element, a bound of |J^-1|_inf over the reference cube (0: none)."""
    out = np.zeros(hm.n_elem, np.float32)
    _check(lib().femrt_host_geom_jinv_bound(hm.n_elem, hm.ng, _ptr(hm.geom), _ptr(out)),
           what="host_geom_jinv_bound")
    return out


def host_build_bvh(boxes):
    boxes = np.ascontiguousarray(boxes, np.float32)
    ne = boxes.shape[0]
    nodes = np.zeros((2 * ne, 8), np.float32)
    prims = np.zeros(ne, np.int32)
    nn = C.c_int(0)
    _check(lib().femrt_host_build_bvh(ne, _ptr(boxes), C.byref(nn), _ptr(nodes), _ptr(prims)),
           what="host_build_bvh")
    return nodes[:nn.value].copy(), prims


LOAD_VERTEX0_DEFECT = 1   # replicate the reference's vertex-0 parse
LOAD_GEOM_MODES_FROM_FIELD = 2   # replicate geometry modes raised to field modes


def _io_check(rc, what):
    if rc != OK:
        m = lib().femrt_io_last_error()
        raise FemrtError(f"{what}: {ERRORS.get(rc, rc)} {m.decode() if m else ''}")


def save_xml_fld(hm: HexMesh, xml_path, fld_path):
    """Write a Nektar++ XML mesh (every face CURVED) and FLD field."""
    rc = lib().femrt_save_xml_fld(str(xml_path).encode(), str(fld_path).encode(), hm.n_elem,
                                  hm.ng, hm.nf, _ptr(hm.vertices), _ptr(hm.geom), _ptr(hm.field))
    _io_check(rc, "save_xml_fld")


def load_xml_fld(xml_path, fld_path, flags=0) -> HexMesh:
    """Drop-in of load_xml_fld (xml_fld_loader.h:14-16): hex meshes."""
    L = lib()
    ne, ng, nf = C.c_int(0), C.c_int(0), C.c_int(0)
    fp = None if fld_path is None else str(fld_path).encode()
    rc = L.femrt_load_xml_fld(str(xml_path).encode(), fp, flags,
                              C.byref(ne), C.byref(ng), C.byref(nf), None, None, None, None)
    _io_check(rc, "load_xml_fld(sizes)")
    v = np.zeros((ne.value, 8, 3), np.float32)
    g = np.zeros((ne.value, 3, ng.value ** 3), np.float32)
    f = np.zeros((ne.value, nf.value ** 3), np.float32)
    a = np.zeros((ne.value, 6, 2), np.int32)
    rc = L.femrt_load_xml_fld(str(xml_path).encode(), fp, flags,
                              C.byref(ne), C.byref(ng), C.byref(nf), _ptr(v), _ptr(g), _ptr(f),
                              _ptr(a))
    _io_check(rc, "load_xml_fld")
    return HexMesh(v, g, f, a, ng.value, nf.value)


def auto_camera(lo, hi):
    """Camera of main.cu:248-302: look at the box centre from
    centre + (2.5, -2.5, -2.0) * max_dim, vup (0,1,0), vfov 20 degrees."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    c = 0.5 * (lo + hi)
    md = float(np.max(hi - lo)) or 1.0
    lf = c + np.array([2.5, -2.5, -2.0]) * md
    return np.array([*lf, *c, 0.0, 1.0, 0.0, 20.0], np.float32)


# ------------------------------------------------------------- device objects

class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().femrt_ctx_create(device, C.byref(self.h)), what="ctx_create")

    def close(self):
        if self.h:
            lib().femrt_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().femrt_sync(self.h), self.h, "sync")

    def timing(self, on=True):
        _check(lib().femrt_timing_enable(self.h, int(on)), self.h, "timing_enable")

    def timing_reset(self):
        _check(lib().femrt_timing_reset(self.h), self.h, "timing_reset")

    def timing_read(self):
        cap = 64
        names = C.create_string_buffer(48 * cap)
        ms = np.zeros(cap, np.float64)
        cnt = np.zeros(cap, np.int64)
        n = lib().femrt_timing_read(self.h, cap, names, _ptr(ms), _ptr(cnt))
        if n < 0:
            _check(n, self.h, "timing_read")
        out = {}
        raw = names.raw
        for i in range(min(n, cap)):
            nm = raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(cnt[i]))
        return out


class Mesh:
    def __init__(self, ctx: Context, hm: HexMesh, build_bvh=True):
        self.ctx = ctx
        self.host = hm
        self.h = C.c_void_p()
        adj = hm.adjacency
        _check(lib().femrt_mesh_upload_hex(ctx.h, hm.n_elem, hm.ng, hm.nf, _ptr(hm.vertices),
                                           _ptr(hm.geom), _ptr(hm.field), _ptr(adj),
                                           C.byref(self.h)), ctx.h, "mesh_upload_hex")
        self.n_nodes = 0
        if build_bvh:
            self.build_bvh()

    @classmethod
    def from_quads(cls, ctx: Context, qm: QuadMesh, build_bvh=True):
        """A surface mesh of curved quads (femrt_mesh_upload_quads): first
        hits only, no isosurfaces."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.host = qm
        self.h = C.c_void_p()
        _check(lib().femrt_mesh_upload_quads(ctx.h, qm.n_elem, qm.ng, qm.nf, _ptr(qm.vertices),
                                             _ptr(qm.geom), _ptr(qm.field), C.byref(self.h)),
               ctx.h, "mesh_upload_quads")
        self.n_nodes = 0
        if build_bvh:
            self.build_bvh()
        return self

    @classmethod
    def from_prisms(cls, ctx: Context, pm: PrismMesh, build_bvh=True):
        """A prism mesh (femrt_mesh_upload_prisms): first hits only."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.host = pm
        self.h = C.c_void_p()
        gm = np.array(pm.gm, np.int32)
        fm = np.array(pm.fm, np.int32)
        _check(lib().femrt_mesh_upload_prisms(ctx.h, pm.n_elem, _ptr(gm), _ptr(fm), _ptr(pm.vertices),
                                              _ptr(pm.geom), _ptr(pm.field), C.byref(self.h)),
               ctx.h, "mesh_upload_prisms")
        self.n_nodes = 0
        if build_bvh:
            self.build_bvh()
        return self

    @classmethod
    def from_tris(cls, ctx: Context, tm: TriMesh, build_bvh=True):
        """A curved-triangle surface mesh (femrt_mesh_upload_tris): first
        hits only."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.host = tm
        self.h = C.c_void_p()
        gm = np.array(tm.gm, np.int32)
        fm = np.array(tm.fm, np.int32)
        _check(lib().femrt_mesh_upload_tris(ctx.h, tm.n_elem, _ptr(gm), _ptr(fm), _ptr(tm.vertices),
                                            _ptr(tm.geom), _ptr(tm.field), C.byref(self.h)),
               ctx.h, "mesh_upload_tris")
        self.n_nodes = 0
        if build_bvh:
            self.build_bvh()
        return self

    def build_bvh(self, builder=BVH_HOST_SAH):
        nn = C.c_int(0)
        _check(lib().femrt_mesh_build_bvh(self.h, builder, C.byref(nn)), self.ctx.h, "build_bvh")
        self.n_nodes = nn.value

    def boxes(self):
        out = np.zeros((self.host.n_elem, NBOX, 6), np.float32)
        _check(lib().femrt_mesh_boxes(self.h, _ptr(out)), self.ctx.h, "mesh_boxes")
        return out

    def bvh(self):
        nodes = np.zeros((self.n_nodes, 8), np.float32)
        prims = np.zeros(self.host.n_elem, np.int32)
        _check(lib().femrt_mesh_get_bvh(self.h, _ptr(nodes), _ptr(prims)), self.ctx.h, "get_bvh")
        return nodes, prims

    def eval_field(self, elem, rst, precision=1):
        elem = np.ascontiguousarray(elem, np.int32)
        rst = np.ascontiguousarray(rst, np.float32).reshape(3, -1)
        n = elem.size
        out = np.zeros((4, n), np.float64 if precision else np.float32)
        _check(lib().femrt_eval_field(self.h, n, _ptr(elem), _ptr(rst), precision, _ptr(out)),
               self.ctx.h, "eval_field")
        return out

    def eval_field_noalloc(self, elem, rst, precision=1):
        """eval_field staged in the library's fixed device array (no device
        allocation; at most NOALLOC_MAX_POINTS points)."""
        elem = np.ascontiguousarray(elem, np.int32)
        rst = np.ascontiguousarray(rst, np.float32).reshape(3, -1)
        n = elem.size
        out = np.zeros((4, n), np.float64 if precision else np.float32)
        _check(lib().femrt_eval_field_noalloc(self.h, n, _ptr(elem), _ptr(rst), precision, _ptr(out)),
               self.ctx.h, "eval_field_noalloc")
        return out

    def invert(self, elem, x):
        elem = np.ascontiguousarray(elem, np.int32)
        x = np.ascontiguousarray(x, np.float32).reshape(3, -1)
        n = elem.size
        rst = np.zeros((3, n), np.float64)
        it = np.zeros(n, np.int32)
        _check(lib().femrt_invert_points(self.h, n, _ptr(elem), _ptr(x), _ptr(rst), _ptr(it)),
               self.ctx.h, "invert_points")
        return rst, it

    def set_guess_mlp(self, mlp):
        """Network used by GUESS_LEARNED traces (None detaches)."""
        _check(lib().femrt_mesh_set_guess_mlp(self.h, None if mlp is None else mlp.h),
               self.ctx.h, "set_guess_mlp")
        self._mlp = mlp  # keep the network alive while the mesh uses it

    def set_bracket_mlp(self, mlp):
        self._bracket_mlp = mlp
        _check(lib().femrt_mesh_set_bracket_mlp(self.h, None if mlp is None else mlp.h),
               self.ctx.h, "set_bracket_mlp")

    def launch_cfg(self, cfg=None):
        """Launch shape: (first-hit block, walk workgroups per CU, polish
        workgroups per CU, tile edge in pixels, walk queue segments per XCD).
        Sets the first len(cfg) values when cfg is given; returns all five."""
        if cfg is not None:
            c = np.ascontiguousarray(cfg, np.int32)
            _check(lib().femrt_mesh_set_launch_cfg(self.h, _ptr(c), int(c.size)), self.ctx.h, "set_launch_cfg")
        c = np.zeros(LAUNCH_CFG_N, np.int32)
        _check(lib().femrt_mesh_get_launch_cfg(self.h, _ptr(c), LAUNCH_CFG_N), self.ctx.h, "get_launch_cfg")
        return tuple(int(x) for x in c)

    def autotune(self, frame, params, reps=2):
        """Times candidate launch shapes on this trace and keeps the fastest."""
        c = np.zeros(LAUNCH_CFG_N, np.int32)
        _check(lib().femrt_autotune(self.h, frame.h, C.byref(params), reps, _ptr(c), LAUNCH_CFG_N),
               self.ctx.h, "autotune")
        return tuple(int(x) for x in c)

    def close(self):
        if self.h:
            lib().femrt_mesh_destroy(self.h)
            self.h = C.c_void_p()


class Mlp:
    """The library's networks on BF16 MFMA: the Newton-guess network
    (MLP_GUESS, 27 -> 64 tanh -> 64 tanh -> 3) or the isosurface-bracket
    network (MLP_BRACKET, 32 -> 64 tanh -> 64 tanh -> 2 quantiles)."""

    def __init__(self, ctx: Context, seed=0, kind=MLP_GUESS):
        self.ctx = ctx
        self.h = C.c_void_p()
        self.kind = kind
        self.n_in, self.n_out = MLP_SHAPE[kind]
        _check(lib().femrt_mlp_create(ctx.h, kind, seed, C.byref(self.h)), ctx.h, "mlp_create")
        self.n_param = lib().femrt_mlp_num_params(kind)

    def params(self):
        out = np.zeros(self.n_param, np.float32)
        _check(lib().femrt_mlp_get_params(self.h, _ptr(out)), self.ctx.h, "mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_param
        _check(lib().femrt_mlp_set_params(self.h, _ptr(p)), self.ctx.h, "mlp_set_params")

    def forward(self, feats):
        feats = np.ascontiguousarray(feats, np.float32).reshape(-1, self.n_in)
        out = np.zeros((feats.shape[0], self.n_out), np.float32)
        _check(lib().femrt_mlp_forward(self.h, feats.shape[0], _ptr(feats), _ptr(out)),
               self.ctx.h, "mlp_forward")
        return out

    def loss_grad(self, feats, targets):
        feats = np.ascontiguousarray(feats, np.float32).reshape(-1, self.n_in)
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, self.n_out)
        loss = C.c_float(0)
        grad = np.zeros(self.n_param, np.float32)
        _check(lib().femrt_mlp_loss_grad(self.h, feats.shape[0], _ptr(feats), _ptr(targets),
                                         C.byref(loss), _ptr(grad)), self.ctx.h, "mlp_loss_grad")
        return float(loss.value), grad

    def train(self, mesh: Mesh, n_samples=None, batch=None, epochs=None, lr=None, lr_final=None,
              seed=None):
        p = MlpTrainParams()
        lib().femrt_mlp_train_params_default(C.byref(p))
        for k, v in dict(n_samples=n_samples, batch=batch, epochs=epochs, lr=lr,
                         lr_final=lr_final, seed=seed).items():
            if v is not None:
                setattr(p, k, v)
        losses = np.zeros(p.epochs, np.float32)
        _check(lib().femrt_mlp_train(self.h, mesh.h, C.byref(p), _ptr(losses)), self.ctx.h,
               "mlp_train")
        return losses

    def train_pairs(self, feats, targets, batch=4096, epochs=4, lr=3e-3, lr_final=1e-4, seed=1):
        """Adam over given pairs (feats n x n_in, targets n x n_out)."""
        feats = np.ascontiguousarray(feats, np.float32).reshape(-1, self.n_in)
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, self.n_out)
        p = MlpTrainParams()
        lib().femrt_mlp_train_params_default(C.byref(p))
        p.batch, p.epochs, p.lr, p.lr_final, p.seed = batch, epochs, lr, lr_final, seed
        losses = np.zeros(epochs, np.float32)
        _check(lib().femrt_mlp_train_pairs(self.h, feats.shape[0], _ptr(feats), _ptr(targets), C.byref(p),
                                           _ptr(losses)), self.ctx.h, "mlp_train_pairs")
        return losses

    def close(self):
        if self.h:
            lib().femrt_mlp_destroy(self.h)
            self.h = C.c_void_p()


def mlp_samples(mesh: Mesh, n, seed=0):
    """Training pairs from the tracer's point inversions: (features n x 27,
    Newton preimages n x 3)."""
    feats = np.zeros((n, MLP_IN), np.float32)
    tgt = np.zeros((n, 3), np.float32)
    _check(lib().femrt_mlp_make_samples(mesh.h, n, seed, _ptr(feats), _ptr(tgt)), mesh.ctx.h,
           "mlp_make_samples")
    return feats, tgt


def trace_params(isos=(), guess=GUESS_PLAIN, samples=16, iso_enable=True, max_march=4096,
                 flags=1):
    p = TraceParams()
    lib().femrt_trace_params_default(C.byref(p))
    p.n_iso = len(isos)
    p.iso_enable = int(iso_enable)
    for i, v in enumerate(isos):
        p.iso[i] = float(v)
    p.samples = samples
    p.guess = guess
    p.max_march = max_march
    p.flags = flags
    return p


class Frame:
    def __init__(self, ctx: Context, n_rays: int):
        self.ctx = ctx
        self.n = int(n_rays)
        self.h = C.c_void_p()
        _check(lib().femrt_frame_create(ctx.h, self.n, C.byref(self.h)), ctx.h, "frame_create")

    def set_rays(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(6, self.n)
        _check(lib().femrt_frame_set_rays(self.h, _ptr(rays)), self.ctx.h, "set_rays")

    def camera_rays(self, cam10, W, H):
        cam = np.ascontiguousarray(cam10, np.float32)
        _check(lib().femrt_frame_camera_rays(self.h, _ptr(cam), W, H), self.ctx.h, "camera_rays")

    def get_rays(self):
        out = np.zeros((6, self.n), np.float32)
        _check(lib().femrt_frame_get_rays(self.h, _ptr(out)), self.ctx.h, "get_rays")
        return out

    def trace(self, mesh: Mesh, params: TraceParams):
        _check(lib().femrt_trace(mesh.h, self.h, C.byref(params)), self.ctx.h, "trace")

    def hits(self):
        n = self.n
        e = np.zeros(n, np.int32)
        f = np.zeros(n, np.int32)
        t = np.zeros(n, np.float64)
        r = np.zeros((3, n), np.float64)
        u = np.zeros(n, np.float64)
        _check(lib().femrt_frame_get_hits(self.h, _ptr(e), _ptr(f), _ptr(t), _ptr(r), _ptr(u)),
               self.ctx.h, "get_hits")
        return dict(elem=e, face=f, t=t, rst=r, u=u)

    def iso(self, n_iso):
        n = self.n
        e = np.zeros((n_iso, n), np.int32)
        t = np.zeros((n_iso, n), np.float64)
        r = np.zeros((n_iso, 3, n), np.float64)
        u = np.zeros((n_iso, n), np.float64)
        _check(lib().femrt_frame_get_iso(self.h, n_iso, _ptr(e), _ptr(t), _ptr(r), _ptr(u)),
               self.ctx.h, "get_iso")
        return dict(elem=e, t=t, rst=r, u=u)

    def rgba(self):
        out = np.zeros(self.n, np.uint32)
        _check(lib().femrt_frame_get_rgba(self.h, _ptr(out)), self.ctx.h, "get_rgba")
        return out

    def camera_rays_jitter(self, cam12, W, H, spp, seed):
        cam = np.ascontiguousarray(cam12, np.float32)
        _check(lib().femrt_frame_camera_rays_jitter(self.h, _ptr(cam), W, H, spp, seed),
               self.ctx.h, "camera_rays_jitter")

    def path_begin(self):
        _check(lib().femrt_frame_path_begin(self.h), self.ctx.h, "path_begin")

    def set_materials(self, mats):
        _check(lib().femrt_frame_set_materials(self.h, mats, len(mats)), self.ctx.h, "set_materials")

    def scatter(self, mesh, mats, seed, depth, last):
        _check(lib().femrt_frame_scatter(mesh.h, self.h, mats, len(mats), seed, depth, int(last)),
               self.ctx.h, "scatter")

    def resolve(self, W, H, spp):
        out = np.zeros(W * H, np.uint32)
        _check(lib().femrt_frame_resolve(self.h, W, H, spp, _ptr(out)), self.ctx.h, "resolve")
        return out

    def render_path(self, mesh, cam12, W, H, spp, max_depth, mats, seed):
        cam = np.ascontiguousarray(cam12, np.float32)
        out = np.zeros(W * H, np.uint32)
        _check(lib().femrt_render_path(mesh.h, self.h, _ptr(cam), W, H, spp, max_depth, mats,
                                       len(mats), seed, _ptr(out)), self.ctx.h, "render_path")
        return out

    def counters(self, enable=None):
        out = np.zeros(NCOUNTERS, np.int64)
        en = 1 if enable is None else int(enable)
        _check(lib().femrt_frame_counters(self.h, en, _ptr(out)), self.ctx.h, "counters")
        return dict(zip(COUNTER_NAMES, out.tolist()))

    def stalled_rays(self):
        """(count, ids) of the walks of the last counted trace that stalled
        (skipped past a stall or ended stalled); at most 64 ids are kept."""
        ids = np.zeros(64, np.int32)
        n = lib().femrt_frame_stalled_rays(self.h, ids.size, _ptr(ids))
        if n < 0:
            _check(n, self.ctx.h, "stalled_rays")
        return n, np.sort(ids[:min(n, ids.size)])

    def iso_gradients(self, mesh, n_iso, engine=GRAD_MFMA):
        """World gradient of the field at each isosurface crossing of the last
        trace: (n_iso, 3, n) float32, NaN where there is none."""
        out = np.zeros((n_iso, 3, self.n), np.float32)
        _check(lib().femrt_frame_iso_gradients(mesh.h, self.h, n_iso, engine, _ptr(out)), self.ctx.h,
               "iso_gradients")
        return out

    def bracket_pairs(self, mesh, isos, max_pairs):
        """Training pairs of the bracket network from the last trace (which
        recorded segments): (features n x 32, targets n x 2)."""
        iso = np.ascontiguousarray(isos, np.float64)
        feats = np.zeros((max_pairs, 32), np.float32)
        tgt = np.zeros((max_pairs, 2), np.float32)
        got = C.c_int(0)
        _check(lib().femrt_frame_bracket_pairs(mesh.h, self.h, iso.size, _ptr(iso), max_pairs, _ptr(feats),
                                               _ptr(tgt), C.byref(got)), self.ctx.h, "bracket_pairs")
        return feats[:got.value], tgt[:got.value]

    def close(self):
        if self.h:
            lib().femrt_frame_destroy(self.h)
            self.h = C.c_void_p()


class Batch:
    """Traces and path-traced renders of many meshes recorded into one HIP
    graph; launch() replays them all with one graph launch."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = C.c_void_p()
        self._keep = []
        _check(lib().femrt_batch_create(ctx.h, C.byref(self.h)), ctx.h, "batch_create")

    def add_trace(self, mesh: Mesh, frame: Frame, params: TraceParams):
        self._keep.append((mesh, frame, params))
        _check(lib().femrt_batch_add_trace(self.h, mesh.h, frame.h, C.byref(params)), self.ctx.h,
               "batch_add_trace")

    def add_paths(self, mesh: Mesh, frame: Frame, cam12, W, H, spp, max_depth, mats, seed):
        cam = np.ascontiguousarray(cam12, np.float32)
        self._keep.append((mesh, frame, mats))
        _check(lib().femrt_batch_add_paths(self.h, mesh.h, frame.h, _ptr(cam), W, H, spp, max_depth,
                                           mats, len(mats), seed), self.ctx.h, "batch_add_paths")

    def capture(self):
        _check(lib().femrt_batch_capture(self.h), self.ctx.h, "batch_capture")

    def launch(self):
        _check(lib().femrt_batch_launch(self.h), self.ctx.h, "batch_launch")

    def close(self):
        if self.h:
            lib().femrt_batch_destroy(self.h)
            self.h = C.c_void_p()


def save_image_ppm(rgba, W, H, path):
    """P6 writer of raytracer.cu:480-494 (rows written in buffer order, y=0 is
    the bottom row of the image)."""
    px = np.asarray(rgba, np.uint32).reshape(H * W)
    rgb = np.stack([(px >> s) & 0xFF for s in (0, 8, 16)], axis=1).astype(np.uint8)
    with open(path, "wb") as fh:
        fh.write(f"P6\n{W} {H}\n255\n".encode())
        fh.write(rgb.tobytes())
