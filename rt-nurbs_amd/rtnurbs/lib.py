"""ctypes binding of librtnurbs.so (the C ABI declared in include/rtnurbs.h).

This is the Python host over the HIP library: device buffers, scenes, ray
generation and the trace entry points.  There is no CPU fallback: importing
the binding without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "librtnurbs.so")

RTN_OK, RTN_EINVAL, RTN_EHIP, RTN_ENOMEM, RTN_EUNSUPPORTED, RTN_EOVERFLOW = 0, -1, -2, -3, -4, -5
MAX_DEGREE = 5
MODE_EXACT, MODE_WIDE, MODE_PERSISTENT, MODE_PACKET, MODE_PERSISTENT_PACKET = 0, 1, 2, 3, 4
SEED_PLAIN, SEED_LEARNED = 0, 1
PREC_FP32, PREC_MIXED = 0, 1

NODE_DTYPE = np.dtype([("b", np.float32, 6), ("left_first", np.uint32), ("prim_count", np.uint32)])
INTERACT_DTYPE = np.dtype([("t_hit", np.float32), ("wo", np.float32, 3), ("uv", np.float32, 2),
                           ("dpdu", np.float32, 3), ("dpdv", np.float32, 3), ("dndu", np.float32, 3),
                           ("dndv", np.float32, 3)])
assert INTERACT_DTYPE.itemsize == 72 and NODE_DTYPE.itemsize == 32

EXPORTS = [
    "rtn_last_error", "rtn_version", "rtn_malloc", "rtn_malloc_managed", "rtn_free", "rtn_memcpy", "rtn_memset",
    "rtn_device_synchronize", "rtn_stream_create", "rtn_stream_destroy", "rtn_stream_synchronize",
    "rtn_camera_perspective", "rtn_rays_perspective", "rtn_scene_create", "rtn_scene_destroy",
    "rtn_scene_add_nurbs", "rtn_scene_add_patches", "rtn_scene_build", "rtn_scene_stats_get",
    "rtn_scene_download", "rtn_trace_params_default", "rtn_trace", "rtn_calculate_intersect", "rtn_get_patch",
    "rtn_build_bvh", "rtn_patch_bounds", "rtn_refine", "rtn_scene_refine", "rtn_trace_f64", "rtn_trace_f64_roots", "rtn_event_create", "rtn_event_destroy",
    "rtn_event_record", "rtn_event_elapsed_ms", "rtn_trace_pairs", "rtn_mlp_create", "rtn_mlp_destroy",
    "rtn_mlp_kind", "rtn_mlp_set_params", "rtn_mlp_get_params", "rtn_mlp_pair_features", "rtn_mlp_forward", "rtn_mlp_gradient",
    "rtn_mlp_train_params_default", "rtn_mlp_train", "rtn_rays_silhouette", "rtn_rays_secondary",
    "rtn_shade_lambert", "rtn_build_params_default", "rtn_scene_build_ex", "rtn_batch_create", "rtn_batch_destroy",
    "rtn_batch_patch_base", "rtn_batch_trace", "rtn_frame_params_default", "rtn_frame_create", "rtn_frame_destroy",
    "rtn_frame_run", "rtn_frame_capture", "rtn_frame_status", "rtn_frame_results", "rtn_frame_set_shape",
    "rtn_frame_autotune", "rtn_solve_pairs", "rtn_solve_pairs_roots", "rtn_frame_set_refill", "rtn_frame_set_packet", "rtn_scene_save", "rtn_scene_load",
    "rtn_scene_trace_overflow", "rtn_malloc_host", "rtn_free_host", "rtn_lookat", "rtn_camera_init",
    "rtn_camera_get_perspective", "rtn_camera_get_orthographic", "rtn_rays_generate",
    "rtn_calculate_intersect_noalloc", "rtn_frame_count",
]
BVH_EXACT_SAH, BVH_BINNED_SAH = 0, 1
MLP_NIN, MLP_HIDDEN, MLP_NOUT, MLP_NPARAM = 20, 64, 3, 9859
MLP_FEAT_PROJECTED, MLP_FEAT_BOX = 0, 1


class RtnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rtnurbs error {code}: {msg}")
        self.code = code


class RayQueue(C.Structure):
    _fields_ = [("origin_x", C.c_void_p), ("origin_y", C.c_void_p), ("origin_z", C.c_void_p),
                ("direction_x", C.c_void_p), ("direction_y", C.c_void_p), ("direction_z", C.c_void_p),
                ("idx_interact", C.c_void_p)]


class TraceParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("max_iter", C.c_int), ("thr_det", C.c_float), ("thr_res", C.c_float),
                ("flat_threshold", C.c_float), ("cam_k", C.c_float), ("seed_base", C.c_uint32),
                ("seed_mode", C.c_int), ("precision", C.c_int), ("cond_thr", C.c_float),
                ("hull_cull", C.c_int), ("count", C.c_int), ("learned_iters", C.c_int),
                ("any_hit", C.c_int), ("refill_min", C.c_int), ("fallback_iters", C.c_int),
                ("seed_share", C.c_int), ("stall_exit", C.c_int), ("seed_net", C.c_void_p),
                ("skip_patch", C.c_void_p), ("t_tol", C.c_float), ("hit_gate", C.c_float),
                ("hit_skip", C.c_float), ("learned_min_lanes", C.c_int)]


class MlpTrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float), ("lr_final", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("hit_weight", C.c_float),
                ("seed", C.c_uint64)]


class BuildParams(C.Structure):
    _fields_ = [("bvh", C.c_int), ("bins", C.c_int), ("max_leaf", C.c_int), ("reserved", C.c_int * 5)]


class FrameParams(C.Structure):
    _fields_ = [("trace", TraceParams), ("light", C.c_float * 3), ("eps", C.c_float), ("shadow", C.c_int),
                ("mirror", C.c_int), ("packet_primary", C.c_int), ("packet_secondary", C.c_int),
                ("skip_origin", C.c_int), ("reserved", C.c_int)]


class FrameStatus(C.Structure):
    _fields_ = [("n_primary", C.c_int64), ("n_secondary", C.c_int64), ("n_secondary_slots", C.c_int64),
                ("overflow", C.c_ulonglong),
                ("shape_lanes", C.c_int), ("shape_stack", C.c_int), ("shape_blocks_per_cu", C.c_int),
                ("captured", C.c_int), ("refill_min", C.c_int), ("packet", C.c_int)]


class TuneEntry(C.Structure):
    _fields_ = [("lanes", C.c_int), ("stack", C.c_int), ("blocks_per_cu", C.c_int), ("refill_min", C.c_int),
                ("ms", C.c_float), ("packet_primary", C.c_int), ("packet_secondary", C.c_int)]


class Hits(C.Structure):
    _fields_ = [("prim", C.c_void_p), ("t", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("dpdu", C.c_void_p), ("dpdv", C.c_void_p), ("counters", C.c_void_p)]


class TraceTotals(C.Structure):
    _fields_ = [(k, C.c_ulonglong) for k in ("newton_calls", "newton_evals", "skips", "interior", "leaves",
                                             "patch_tests", "overflow", "newton_flops", "fp64_steps",
                                             "hull_culled", "learned_fallbacks", "learned_skips")]


class CameraDesc(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("look", C.c_float * 3), ("up", C.c_float * 3), ("fov_deg", C.c_float),
                ("width", C.c_int), ("height", C.c_int)]


class SceneStats(C.Structure):
    _fields_ = [("n_surfaces", C.c_int64), ("n_patches", C.c_int64), ("n_cp", C.c_int64), ("n_nodes", C.c_int64),
                ("n_wide_nodes", C.c_int64), ("max_depth", C.c_int64), ("n_leaves", C.c_int64),
                ("deg_hist", (C.c_int64 * (MAX_DEGREE + 1)) * (MAX_DEGREE + 1)),
                ("build_ms_decompose", C.c_double), ("build_ms_bvh", C.c_double), ("build_ms_wide", C.c_double)]


_lib = None


def load():
    """Load librtnurbs.so; raises if it was not built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RtnError(RTN_EUNSUPPORTED, f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, i, f, sz, i64 = C.c_void_p, C.c_int, C.c_float, C.c_size_t, C.c_int64
    L.rtn_last_error.restype = C.c_char_p
    L.rtn_malloc.argtypes = [C.POINTER(vp), sz]
    L.rtn_malloc_managed.argtypes = [C.POINTER(vp), sz]
    L.rtn_free.argtypes = [vp]
    L.rtn_memcpy.argtypes = [vp, vp, sz, i]
    L.rtn_memset.argtypes = [vp, i, sz]
    L.rtn_stream_create.argtypes = [C.POINTER(vp)]
    L.rtn_stream_destroy.argtypes = [vp]
    L.rtn_stream_synchronize.argtypes = [vp]
    L.rtn_camera_perspective.argtypes = [vp, vp, vp, i, i, f, f, f, vp, vp]
    L.rtn_rays_perspective.argtypes = [vp, vp, i, i, i, i, C.c_uint32, C.POINTER(RayQueue), vp]
    L.rtn_scene_create.argtypes = [C.POINTER(vp)]
    L.rtn_scene_destroy.argtypes = [vp]
    L.rtn_scene_add_nurbs.argtypes = [vp, i, i, vp, i, i, vp, vp]
    L.rtn_scene_add_patches.argtypes = [vp, i, i, i, vp]
    L.rtn_scene_build.argtypes = [vp]
    L.rtn_scene_save.argtypes = [vp, C.POINTER(CameraDesc), C.c_char_p]
    L.rtn_scene_load.argtypes = [C.c_char_p, C.POINTER(vp), C.POINTER(CameraDesc), C.POINTER(i)]
    L.rtn_build_params_default.argtypes = [C.POINTER(BuildParams)]
    L.rtn_build_params_default.restype = None
    L.rtn_scene_build_ex.argtypes = [vp, C.POINTER(BuildParams)]
    L.rtn_batch_create.argtypes = [C.POINTER(vp), C.POINTER(vp), i]
    L.rtn_batch_destroy.argtypes = [vp]
    L.rtn_batch_destroy.restype = None
    L.rtn_batch_patch_base.argtypes = [vp, vp]
    L.rtn_batch_trace.argtypes = [vp, C.POINTER(vp * 6), vp, C.c_int64, C.POINTER(TraceParams), C.POINTER(Hits),
                                  C.POINTER(TraceTotals), vp]
    L.rtn_frame_params_default.argtypes = [C.POINTER(FrameParams)]
    L.rtn_frame_params_default.restype = None
    L.rtn_frame_create.argtypes = [C.POINTER(vp), vp, C.POINTER(vp * 6), vp, C.c_int64, C.POINTER(FrameParams)]
    L.rtn_frame_destroy.argtypes = [vp]
    L.rtn_frame_destroy.restype = None
    L.rtn_frame_run.argtypes = [vp, vp]
    L.rtn_frame_capture.argtypes = [vp, vp]
    L.rtn_frame_status.argtypes = [vp, C.POINTER(FrameStatus)]
    L.rtn_frame_results.argtypes = [vp, C.POINTER(Hits), C.POINTER(Hits), C.POINTER(vp), C.POINTER(vp),
                                    C.POINTER(vp * 6)]
    L.rtn_frame_set_shape.argtypes = [vp, i, i, i]
    L.rtn_frame_count.argtypes = [vp, C.POINTER(TraceTotals), C.POINTER(TraceTotals)]
    L.rtn_frame_set_refill.argtypes = [vp, i]
    L.rtn_frame_set_packet.argtypes = [vp, i, i]
    L.rtn_frame_autotune.argtypes = [vp, i, C.POINTER(TuneEntry), i, C.POINTER(i), vp]
    L.rtn_solve_pairs.argtypes = [vp, C.POINTER(vp * 6), i64, vp, vp, vp, i64, C.POINTER(TraceParams), vp, vp, vp,
                                  vp, vp]
    L.rtn_solve_pairs_roots.argtypes = [vp, C.POINTER(vp * 6), i64, vp, vp, vp, i64, C.POINTER(TraceParams), i, vp,
                                        vp, vp, vp, vp]
    L.rtn_scene_stats_get.argtypes = [vp, C.POINTER(SceneStats)]
    L.rtn_scene_download.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.rtn_trace_params_default.argtypes = [C.POINTER(TraceParams)]
    L.rtn_trace_params_default.restype = None
    L.rtn_trace.argtypes = [vp, C.POINTER(vp * 6), C.c_int64, C.POINTER(TraceParams), C.POINTER(Hits),
                            C.POINTER(TraceTotals), vp]
    L.rtn_scene_trace_overflow.argtypes = [vp, C.POINTER(C.c_ulonglong)]
    L.rtn_lookat.argtypes = [vp, vp, vp, vp]
    L.rtn_camera_init.argtypes = [vp, i, i, vp]
    L.rtn_camera_get_perspective.argtypes = [vp, f, f, f, vp]
    L.rtn_camera_get_orthographic.argtypes = [vp, f, f, vp]
    L.rtn_rays_generate.argtypes = [C.POINTER(RayQueue), i, vp, vp]
    L.rtn_malloc_host.argtypes = [C.POINTER(vp), sz]
    L.rtn_free_host.argtypes = [vp]
    L.rtn_calculate_intersect.argtypes = [C.POINTER(RayQueue), vp, vp, vp, i, i, f, f, vp, vp, vp]
    L.rtn_calculate_intersect_noalloc.argtypes = [C.POINTER(RayQueue), vp, vp, vp, i, i, f, f, vp, vp, vp]
    L.rtn_get_patch.argtypes = [i, i, vp, i, i, vp, vp, C.POINTER(i), C.POINTER(i), vp, vp]
    L.rtn_build_bvh.argtypes = [vp, i, vp, vp, C.POINTER(i)]
    L.rtn_patch_bounds.argtypes = [vp, i, i, i, vp]
    L.rtn_refine.argtypes = [i, i, i, vp, i, vp, f, C.POINTER(i), vp, vp]
    L.rtn_scene_refine.argtypes = [vp, f, f]
    L.rtn_trace_f64.argtypes = [vp, C.POINTER(vp * 6), C.c_int64, i, C.c_double, C.c_double, C.c_uint32,
                                vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.rtn_trace_f64_roots.argtypes = [vp, C.POINTER(vp * 6), C.c_int64, i, C.c_double, C.c_double, i, C.c_uint32,
                                      vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.rtn_event_create.argtypes = [C.POINTER(vp)]
    L.rtn_event_destroy.argtypes = [vp]
    L.rtn_event_record.argtypes = [vp, vp]
    L.rtn_event_elapsed_ms.argtypes = [vp, vp, C.POINTER(C.c_float)]
    L.rtn_trace_pairs.argtypes = [vp, C.POINTER(vp * 6), i64, C.POINTER(TraceParams), C.c_int, vp, vp, vp, i64,
                                  C.POINTER(i64), vp]
    L.rtn_mlp_create.argtypes = [C.POINTER(vp), C.c_uint64, C.c_int]
    L.rtn_mlp_kind.argtypes = [vp]
    L.rtn_mlp_destroy.argtypes = [vp]
    L.rtn_mlp_destroy.restype = None
    L.rtn_mlp_set_params.argtypes = [vp, vp]
    L.rtn_mlp_get_params.argtypes = [vp, vp]
    L.rtn_mlp_pair_features.argtypes = [vp, C.POINTER(vp * 6), i64, vp, vp, i64, C.c_int, vp, vp]
    L.rtn_mlp_forward.argtypes = [vp, vp, vp, i64, vp, vp, vp]
    L.rtn_mlp_gradient.argtypes = [vp, vp, vp, i64, C.c_float, vp, C.POINTER(C.c_float)]
    L.rtn_mlp_train_params_default.argtypes = [C.POINTER(MlpTrainParams)]
    L.rtn_mlp_train_params_default.restype = None
    L.rtn_mlp_train.argtypes = [vp, vp, vp, i64, C.POINTER(MlpTrainParams), vp, vp]
    L.rtn_rays_silhouette.argtypes = [vp, i64, C.c_uint32, f, f, f, C.POINTER(RayQueue), vp, vp, vp]
    L.rtn_rays_secondary.argtypes = [C.POINTER(vp * 6), C.POINTER(Hits), i64, vp, f, i, i, C.POINTER(RayQueue), vp,
                                     vp, i64, C.POINTER(i64), vp]
    L.rtn_shade_lambert.argtypes = [C.POINTER(vp * 6), C.POINTER(Hits), i, i, i, i, vp, f, vp, vp, vp]
    _lib = L
    return L


def check(rc):
    if rc != RTN_OK:
        msg = load().rtn_last_error()
        raise RtnError(rc, msg.decode() if msg else "")
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- buffers
class DeviceArray:
    """A hipMalloc'd buffer with numpy shape/dtype metadata."""

    def __init__(self, shape, dtype, managed=False):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        fn = load().rtn_malloc_managed if managed else load().rtn_malloc
        check(fn(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, managed=False):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, managed)
        if a.nbytes:
            check(load().rtn_memcpy(d.ptr, a.ctypes.data, a.nbytes, 0))
        return d

    def get(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(load().rtn_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def set(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        check(load().rtn_memcpy(self.ptr, a.ctypes.data, a.nbytes, 0))

    def fill_bytes(self, value):
        check(load().rtn_memset(self.ptr, value, self.nbytes))

    def free(self):
        if self.ptr:
            load().rtn_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(load().rtn_device_synchronize())


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(load().rtn_event_create(C.byref(e)))
        self.h = e.value

    def record(self, stream=None):
        check(load().rtn_event_record(self.h, stream))

    def elapsed_ms(self, later):
        ms = C.c_float()
        check(load().rtn_event_elapsed_ms(self.h, later.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            load().rtn_event_destroy(self.h)
        except Exception:
            pass


def stream_create():
    s = C.c_void_p()
    check(load().rtn_stream_create(C.byref(s)))
    return s.value


# ---------------------------------------------------------------- camera
def camera_matrices(cam):
    """(world_from_camera, camera_from_raster) for a scenes.Camera."""
    wfc = np.zeros(16, np.float32)
    cfr = np.zeros(16, np.float32)
    eye, look, up = (np.asarray(x, np.float32) for x in (cam.eye, cam.look, cam.up))
    check(load().rtn_camera_perspective(eye.ctypes.data, look.ctypes.data, up.ctypes.data, cam.W, cam.H,
                                        cam.fov_deg, cam.znear, cam.zfar, wfc.ctypes.data, cfr.ctypes.data))
    return wfc, cfr


class Rays:
    """Device SoA ray queue (soa.h:6-14) of n rays."""

    def __init__(self, n, with_idx=True):
        self.n = int(n)
        self.comp = [DeviceArray((self.n,), np.float32) for _ in range(6)]
        self.idx = DeviceArray((self.n,), np.int32) if with_idx else None

    def queue(self):
        q = RayQueue(*[c.ptr for c in self.comp], self.idx.ptr if self.idx else None)
        return q

    def ptr_array(self):
        return (C.c_void_p * 6)(*[c.ptr for c in self.comp])

    def get(self):
        return [c.get() for c in self.comp]

    @classmethod
    def from_host(cls, arrs):
        r = cls(len(arrs[0]))
        for d, a in zip(r.comp, arrs):
            d.set(np.asarray(a, np.float32))
        r.idx.fill_bytes(0xFF)
        return r


RAYS_JITTER, RAYS_TILED = 1, 2


def generate_rays(cam, spp=1, jitter=False, seed=0, stream=None, rays=None, offset=0, tiled=False):
    """Pinhole rays of a camera (W*H*spp, samples of a pixel consecutive, pixels
    row-major or in 8x8 Morton tiles); written at rays[offset:] when rays is given."""
    wfc, cfr = camera_matrices(cam)
    n = cam.W * cam.H * spp
    if rays is None:
        rays = Rays(n)
        offset = 0
    assert offset + n <= rays.n
    q = RayQueue(*[c.ptr + 4 * offset for c in rays.comp], rays.idx.ptr + 4 * offset if rays.idx else None)
    flags = (RAYS_JITTER if jitter else 0) | (RAYS_TILED if tiled else 0)
    check(load().rtn_rays_perspective(wfc.ctypes.data, cfr.ctypes.data, cam.W, cam.H, spp, flags, seed,
                                      C.byref(q), stream))
    return rays


# ---------------------------------------------------------------- scene
class Scene:
    def __init__(self, surfaces=(), patch_groups=()):
        L = load()
        h = C.c_void_p()
        check(L.rtn_scene_create(C.byref(h)))
        self.h = h.value
        for s in surfaces:
            self.add_nurbs(s)
        for (np_, p, q, cp) in patch_groups:
            self.add_patches(cp, p, q)

    def add_nurbs(self, s):
        U, V, Pw = (np.ascontiguousarray(a, np.float32) for a in (s.U, s.V, s.Pw))
        check(load().rtn_scene_add_nurbs(self.h, s.n, s.p, U.ctypes.data, s.m, s.q, V.ctypes.data, Pw.ctypes.data))

    def add_patches(self, cp, p, q):
        cp = np.ascontiguousarray(cp, np.float32)
        n = cp.size // (4 * (p + 1) * (q + 1))
        check(load().rtn_scene_add_patches(self.h, n, p, q, cp.ctypes.data))

    def save(self, path, camera=None):
        """Writes the surfaces and patch groups added so far (and the camera, a
        scenes.Camera, when given) to a scene file (rtn_scene_save)."""
        cd = None
        if camera is not None:
            cd = CameraDesc()
            cd.eye[:], cd.look[:], cd.up[:] = list(camera.eye), list(camera.look), list(camera.up)
            cd.fov_deg, cd.width, cd.height = camera.fov_deg, camera.W, camera.H
        check(load().rtn_scene_save(self.h, C.byref(cd) if cd is not None else None, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """(Scene, scenes.Camera or None) from a scene file (rtn_scene_load); the scene is unbuilt."""
        from .scenes import Camera
        L = load()
        h, cd, has = C.c_void_p(), CameraDesc(), C.c_int()
        check(L.rtn_scene_load(os.fsencode(path), C.byref(h), C.byref(cd), C.byref(has)))
        sc = cls.__new__(cls)
        sc.h = h.value
        cam = None
        if has.value:
            cam = Camera(tuple(cd.eye), tuple(cd.look), tuple(cd.up), float(cd.fov_deg), int(cd.width),
                         int(cd.height))
        return sc, cam

    def refine(self, C_u, C_v=None):
        """curvatureRefinementU then V (nurbs.h:290-486) of every surface added
        so far, on the GPU (rtn_scene_refine); C_v defaults to C_u."""
        check(load().rtn_scene_refine(self.h, C_u, C_u if C_v is None else C_v))
        return self

    def build(self, bvh=BVH_EXACT_SAH, bins=32, max_leaf=4):
        """Decomposition, bounds and BVH on the device (exhaustive SAH by default;
        BVH_BINNED_SAH for large scenes)."""
        if bvh == BVH_EXACT_SAH:
            check(load().rtn_scene_build(self.h))
        else:
            bp = BuildParams()
            load().rtn_build_params_default(C.byref(bp))
            bp.bvh, bp.bins, bp.max_leaf = bvh, bins, max_leaf
            check(load().rtn_scene_build_ex(self.h, C.byref(bp)))
        return self

    def trace_overflow(self):
        """Stack-overflow count of the last trace on this scene (synchronises)."""
        c = C.c_ulonglong(0)
        check(load().rtn_scene_trace_overflow(self.h, C.byref(c)))
        return int(c.value)

    def stats(self):
        st = SceneStats()
        check(load().rtn_scene_stats_get(self.h, C.byref(st)))
        d = {k: getattr(st, k) for k, _ in SceneStats._fields_ if k != "deg_hist"}
        d["deg_hist"] = np.array([[st.deg_hist[p][q] for q in range(MAX_DEGREE + 1)]
                                  for p in range(MAX_DEGREE + 1)])
        return d

    def download(self):
        st = self.stats()
        np_, ncp, nn = st["n_patches"], st["n_cp"], st["n_nodes"]
        out = dict(cp=np.zeros((ncp, 4), np.float32), off=np.zeros(np_, np.int32), pq=np.zeros((np_, 2), np.int32),
                   uv=np.zeros((np_, 4), np.float32), pid=np.zeros(np_, np.int32),
                   knot=np.zeros((np_, 2), np.int32), nodes=np.zeros(nn, NODE_DTYPE),
                   delta=np.zeros(np_, np.float32))
        check(load().rtn_scene_download(self.h, *[_ptr(out[k]) for k in
                                                  ("cp", "off", "pq", "uv", "pid", "knot", "nodes", "delta")]))
        return out

    def __del__(self):
        try:
            if self.h:
                load().rtn_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass


def params(**kw):
    """rtn_trace_params with defaults; seed_net takes an Mlp and skip_patch an int32
    DeviceArray of one patch per ray (both kept alive by the struct)."""
    p = TraceParams()
    load().rtn_trace_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        if k == "seed_net":
            p._net = v
            v = v.h if v is not None else None
        elif k == "skip_patch":
            p._skip = v
            v = v.ptr if v is not None else None
        setattr(p, k, v)
    return p


class HitBuffers:
    def __init__(self, n, counters=False):
        self.n = n
        self.prim = DeviceArray((n,), np.int32)
        self.t = DeviceArray((n,), np.float32)
        self.u = DeviceArray((n,), np.float32)
        self.v = DeviceArray((n,), np.float32)
        self.dpdu = DeviceArray((n, 3), np.float32)
        self.dpdv = DeviceArray((n, 3), np.float32)
        self.counters = DeviceArray((n, 8), np.uint32) if counters else None

    def struct(self):
        return Hits(self.prim.ptr, self.t.ptr, self.u.ptr, self.v.ptr, self.dpdu.ptr, self.dpdv.ptr,
                    self.counters.ptr if self.counters else None)

    def get(self):
        d = dict(hit=self.prim.get(), t=self.t.get(), u=self.u.get(), v=self.v.get(), dpdu=self.dpdu.get(),
                 dpdv=self.dpdv.get())
        if self.counters is not None:
            d["counters"] = self.counters.get()
        return d

    def window(self, wins):
        """hit, t, u, v of the index windows [(start, count), ...], concatenated."""
        return {k: grab_windows(a, wins) for k, a in (("hit", self.prim), ("t", self.t), ("u", self.u),
                                                      ("v", self.v))}


def grab_windows(arr, wins):
    """Host copy of index windows [(start, count), ...] of a 1-D DeviceArray."""
    it = np.dtype(arr.dtype).itemsize
    parts = []
    for a, n in wins:
        out = np.empty(n, arr.dtype)
        if n:
            check(load().rtn_memcpy(out.ctypes.data, arr.ptr + a * it, out.nbytes, 1))
        parts.append(out)
    return np.concatenate(parts) if parts else np.empty(0, arr.dtype)


def trace(scene, rays, prm, hits=None, totals=True, stream=None):
    """Runs rtn_trace; returns (HitBuffers, totals dict or None)."""
    if hits is None:
        hits = HitBuffers(rays.n, counters=bool(prm.count))
    hs = hits.struct()
    tot = TraceTotals() if totals else None
    arr = rays.ptr_array()
    check(load().rtn_trace(scene.h, C.byref(arr), rays.n, C.byref(prm), C.byref(hs),
                           C.byref(tot) if tot is not None else None, stream))
    t = {k: getattr(tot, k) for k, _ in TraceTotals._fields_} if tot is not None else None
    return hits, t


# ------------------------------------------------------- reference-shaped
def get_patch(s):
    """getPatch on the GPU: (cp [np,(p+1)(q+1),4], uv [np,4], nbU, nbV)."""
    L = load()
    a, b = C.c_int(), C.c_int()
    U, V, Pw = (np.ascontiguousarray(x, np.float32) for x in (s.U, s.V, s.Pw))
    check(L.rtn_get_patch(s.n, s.p, U.ctypes.data, s.m, s.q, V.ctypes.data, Pw.ctypes.data, C.byref(a), C.byref(b),
                          None, None))
    npch = a.value * b.value
    cp = np.zeros((npch, (s.p + 1) * (s.q + 1), 4), np.float32)
    uv = np.zeros((npch, 4), np.float32)
    check(L.rtn_get_patch(s.n, s.p, U.ctypes.data, s.m, s.q, V.ctypes.data, Pw.ctypes.data, C.byref(a), C.byref(b),
                          cp.ctypes.data, uv.ctypes.data))
    return cp, uv, a.value, b.value


def patch_bounds(cp, p, q):
    cp = np.ascontiguousarray(cp, np.float32)
    n = cp.size // (4 * (p + 1) * (q + 1))
    out = np.zeros((n, 6), np.float32)
    check(load().rtn_patch_bounds(cp.ctypes.data, n, p, q, out.ctypes.data))
    return out


def build_bvh(bounds):
    bounds = np.ascontiguousarray(bounds, np.float32)
    n = len(bounds)
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    perm = np.zeros(n, np.uint32)
    used = C.c_int()
    check(load().rtn_build_bvh(bounds.ctypes.data, n, nodes.ctypes.data, perm.ctypes.data, C.byref(used)))
    return nodes, perm, used.value


def refine(s, C_u=None, C_v=None):
    """curvatureRefinementU then V (nurbs.h:290-486) on a scenes.Nurbs."""
    from dataclasses import replace
    L = load()
    cur = s
    for d, Cc in ((0, C_u), (1, C_v)):
        if Cc is None or Cc <= 0:
            continue
        deg = cur.p if d == 0 else cur.q
        K = np.ascontiguousarray(cur.U if d == 0 else cur.V, np.float32)
        Pw = np.ascontiguousarray(cur.Pw, np.float32)
        ins = C.c_int()
        check(L.rtn_refine(d, cur.n, deg, K.ctypes.data, cur.m, Pw.ctypes.data, Cc, C.byref(ins), None, None))
        tot = ins.value
        nn = cur.n if d == 0 else cur.m
        nk = np.zeros(nn + tot + deg + 2, np.float32)
        if d == 0:
            P2 = np.zeros(((cur.n + tot + 1) * (cur.m + 1), 4), np.float32)
        else:
            P2 = np.zeros(((cur.n + 1) * (cur.m + tot + 1), 4), np.float32)
        check(L.rtn_refine(d, cur.n, deg, K.ctypes.data, cur.m, Pw.ctypes.data, Cc, C.byref(ins), nk.ctypes.data,
                           P2.ctypes.data))
        cur = replace(cur, n=cur.n + tot, U=nk, Pw=P2) if d == 0 else replace(cur, m=cur.m + tot, V=nk, Pw=P2)
    return cur


class HitBuffersF64:
    def __init__(self, n, counters=False):
        self.n = n
        self.prim = DeviceArray((n,), np.int32)
        self.t = DeviceArray((n,), np.float64)
        self.u = DeviceArray((n,), np.float64)
        self.v = DeviceArray((n,), np.float64)
        self.dpdu = DeviceArray((n, 3), np.float64)
        self.dpdv = DeviceArray((n, 3), np.float64)
        self.t2 = DeviceArray((n,), np.float64)
        self.counters = DeviceArray((n, 8), np.uint32) if counters else None

    def get(self):
        d = dict(hit=self.prim.get(), t=self.t.get(), u=self.u.get(), v=self.v.get(), dpdu=self.dpdu.get(),
                 dpdv=self.dpdv.get(), t2=self.t2.get())
        if self.counters is not None:
            d["counters"] = self.counters.get()
        return d


def trace_f64(scene, rays, max_iter=32, thr_det=1e-12, thr_res=1e-10, seed_base=0, counters=False, stream=None,
              grid=0):
    """FP64 reference trace on the device (plain method evaluated in FP64).
    grid > 0: the closest-root reference (rtn_trace_f64_roots): each tested
    patch also runs Newton from grid x grid starts and keeps its nearest
    converged root."""
    h = HitBuffersF64(rays.n, counters)
    arr = rays.ptr_array()
    check(load().rtn_trace_f64_roots(scene.h, C.byref(arr), rays.n, max_iter, thr_det, thr_res, grid, seed_base,
                                     h.prim.ptr, h.t.ptr, h.u.ptr, h.v.ptr, h.dpdu.ptr, h.dpdv.ptr, h.t2.ptr,
                                     h.counters.ptr if h.counters else None, stream))
    return h


# ---------------------------------------------------------- learned seed
class Mlp:
    """The seed network (rtn_mlp): BF16 MFMA inference and Adam training on the GPU."""

    def __init__(self, seed=0, params=None, kind=MLP_FEAT_PROJECTED):
        h = C.c_void_p()
        check(load().rtn_mlp_create(C.byref(h), seed, kind))
        self.h = h.value
        self.kind = kind
        if params is not None:
            self.set_params(params)

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == MLP_NPARAM
        check(load().rtn_mlp_set_params(self.h, w.ctypes.data))

    def get_params(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        check(load().rtn_mlp_get_params(self.h, w.ctypes.data))
        return w

    def forward(self, feats, valid=None, hit=False):
        """feats: DeviceArray (n, 20) float32 -> DeviceArray (n, 2) of (u,v)
        (and, with hit=True, a DeviceArray (n,) of hit probabilities); rows
        with valid == 0 (device int32 array, optional) are skipped (left zero)."""
        n = feats.shape[0]
        uv = DeviceArray((n, 2), np.float32)
        uv.fill_bytes(0)
        ph = None
        if hit:
            ph = DeviceArray((n,), np.float32)
            ph.fill_bytes(0)
        check(load().rtn_mlp_forward(self.h, feats.ptr, valid.ptr if valid is not None else None, n, uv.ptr,
                                     ph.ptr if hit else None, None))
        synchronize()
        return (uv, ph) if hit else uv

    def gradient(self, feats, targets, hit_weight=0.05):
        """Gradient and loss over device feats (n, 20) and targets (n, 3) = (u, v, hit)."""
        n = feats.shape[0]
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_float()
        check(load().rtn_mlp_gradient(self.h, feats.ptr, targets.ptr, n, hit_weight, g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def train(self, feats, targets, **kw):
        tp = MlpTrainParams()
        load().rtn_mlp_train_params_default(C.byref(tp))
        for k, v in kw.items():
            if not hasattr(tp, k):
                raise KeyError(k)
            setattr(tp, k, v)
        losses = np.zeros(tp.epochs, np.float32)
        check(load().rtn_mlp_train(self.h, feats.ptr, targets.ptr, feats.shape[0], C.byref(tp),
                                   losses.ctypes.data, None))
        return losses

    def __del__(self):
        try:
            if self.h:
                load().rtn_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass


def trace_pairs(scene, rays, prm, cap, all=False):
    """Plain-seed Newton solves of a trace: (ray idx, patch, uv) device arrays
    and the count found; converged solves with t > 0 only, or with all=True
    every solve (uv = (-1, -1) where it does not converge)."""
    pr = DeviceArray((cap,), np.int32)
    pp = DeviceArray((cap,), np.int32)
    puv = DeviceArray((cap, 2), np.float32)
    cnt = C.c_int64()
    arr = rays.ptr_array()
    check(load().rtn_trace_pairs(scene.h, C.byref(arr), rays.n, C.byref(prm), 1 if all else 0, pr.ptr, pp.ptr,
                                 puv.ptr, cap,
                                 C.byref(cnt), None))
    return pr, pp, puv, cnt.value


def pair_features(scene, rays, pair_ray, pair_patch, n, out=None, out_offset=0, kind=MLP_FEAT_PROJECTED):
    """Features (n, 20) of a kind (MLP_FEAT_*) of device pairs; written into
    out[out_offset:] when given."""
    if out is None:
        out = DeviceArray((n, MLP_NIN), np.float32)
        out_offset = 0
    assert out_offset + n <= out.shape[0]
    arr = rays.ptr_array()
    check(load().rtn_mlp_pair_features(scene.h, C.byref(arr), rays.n, pair_ray.ptr, pair_patch.ptr, n, kind,
                                       out.ptr + out_offset * MLP_NIN * 4, None))
    return out


# ---------------------------------------------------------- ray producers
def silhouette_rays(scene, n, seed=0, max_deg=5.0, t0=(1.0, 3.0), rays=None, offset=0, aim=False):
    """n grazing rays aimed at random surface points (written at rays[offset:]
    when rays is given).  With aim=True also returns (patch, [u, v, theta, t0])."""
    if rays is None:
        rays = Rays(n)
        offset = 0
    assert offset + n <= rays.n
    q = RayQueue(*[c.ptr + 4 * offset for c in rays.comp], None)
    ap = DeviceArray((n,), np.int32) if aim else None
    au = DeviceArray((n, 4), np.float32) if aim else None
    check(load().rtn_rays_silhouette(scene.h, n, seed, max_deg, t0[0], t0[1], C.byref(q),
                                     ap.ptr if aim else None, au.ptr if aim else None, None))
    return (rays, ap, au) if aim else rays


def secondary_rays(rays, hits, light, eps=1e-4, shadow=True, mirror=True):
    """Shadow and mirror rays of the primary hits in fixed kind-major slots
    (slot j*n + r; kind -1 = empty): (Rays, parent, kind, count of non-empty slots)."""
    per = int(shadow) + int(mirror)
    cap = max(1, per * rays.n)
    out = Rays(cap, with_idx=False)
    parent = DeviceArray((cap,), np.int32)
    kind = DeviceArray((cap,), np.int32)
    cnt = C.c_int64()
    lt = np.asarray(light, np.float32)
    hs = hits.struct()
    arr = rays.ptr_array()
    q = out.queue()
    check(load().rtn_rays_secondary(C.byref(arr), C.byref(hs), rays.n, lt.ctypes.data, eps, int(shadow), int(mirror),
                                    C.byref(q), parent.ptr, kind.ptr, cap, C.byref(cnt), None))
    return out, parent, kind, cnt.value


def shade_lambert(rays, hits, W, H, spp, light, albedo=0.8, shadow_hit=None, tiled=False):
    img = DeviceArray((H, W), np.float32)
    lt = np.asarray(light, np.float32)
    hs = hits.struct()
    arr = rays.ptr_array()
    check(load().rtn_shade_lambert(C.byref(arr), C.byref(hs), W, H, spp, RAYS_TILED if tiled else 0, lt.ctypes.data,
                                   albedo, shadow_hit.ptr if shadow_hit is not None else None, img.ptr, None))
    return img


def write_pgm(path, img):
    """8-bit PGM of a [0,1] float image (row 0 at the top)."""
    a = np.clip(np.asarray(img) * 255.0 + 0.5, 0, 255).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (a.shape[1], a.shape[0]))
        f.write(a.tobytes())


# ------------------------------------------------------ batches and frames
class Batch:
    """Built scenes merged for one launch (rtn_batch); keeps the scenes alive."""

    def __init__(self, scenes_):
        self.scenes = list(scenes_)
        arr = (C.c_void_p * len(self.scenes))(*[s.h for s in self.scenes])
        h = C.c_void_p()
        check(load().rtn_batch_create(C.byref(h), arr, len(self.scenes)))
        self.h = h.value
        base = np.zeros(len(self.scenes) + 1, np.int64)
        check(load().rtn_batch_patch_base(self.h, base.ctypes.data))
        self.patch_base = base

    def trace(self, rays, ray_scene, prm, hits=None, totals=True):
        if hits is None:
            hits = HitBuffers(rays.n, counters=bool(prm.count))
        hs = hits.struct()
        tot = TraceTotals() if totals else None
        arr = rays.ptr_array()
        check(load().rtn_batch_trace(self.h, C.byref(arr), ray_scene.ptr, rays.n, C.byref(prm), C.byref(hs),
                                     C.byref(tot) if tot is not None else None, None))
        t = {k: getattr(tot, k) for k, _ in TraceTotals._fields_} if tot is not None else None
        return hits, t

    def __del__(self):
        try:
            if self.h:
                load().rtn_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def frame_params(light=(0.3, -0.4, 1.0), eps=1e-4, shadow=1, mirror=1, packet_primary=0, packet_secondary=0,
                 skip_origin=1, **trace_kw):
    fp = FrameParams()
    load().rtn_frame_params_default(C.byref(fp))
    for k, v in trace_kw.items():
        if not hasattr(fp.trace, k):
            raise KeyError(k)
        if k == "seed_net":
            fp._net = v
            v = v.h if v is not None else None
        setattr(fp.trace, k, v)
    fp.light[:] = list(light)
    fp.eps, fp.shadow, fp.mirror = eps, shadow, mirror
    fp.packet_primary, fp.packet_secondary = packet_primary, packet_secondary
    fp.skip_origin = skip_origin
    return fp


class Frame:
    """Primary + secondary persistent traces over a batch (rtn_frame)."""

    def __init__(self, batch, rays, ray_scene, fp):
        self.batch, self.rays, self.ray_scene, self.fp = batch, rays, ray_scene, fp
        h = C.c_void_p()
        arr = rays.ptr_array()
        check(load().rtn_frame_create(C.byref(h), batch.h, C.byref(arr), ray_scene.ptr, rays.n, C.byref(fp)))
        self.h = h.value

    def run(self, stream=None):
        check(load().rtn_frame_run(self.h, stream))

    def capture(self, stream):
        check(load().rtn_frame_capture(self.h, stream))

    def status(self):
        st = FrameStatus()
        check(load().rtn_frame_status(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in FrameStatus._fields_}

    def count(self):
        """Work totals (dicts) of the primary and the secondary launch (rtn_frame_count)."""
        a, b = TraceTotals(), TraceTotals()
        check(load().rtn_frame_count(self.h, C.byref(a), C.byref(b)))
        return tuple({k: getattr(t, k) for k, _ in TraceTotals._fields_} for t in (a, b))

    def set_shape(self, lanes, stack, blocks_per_cu):
        check(load().rtn_frame_set_shape(self.h, lanes, stack, blocks_per_cu))

    def set_refill(self, refill_min):
        check(load().rtn_frame_set_refill(self.h, refill_min))

    def set_packet(self, primary, secondary):
        check(load().rtn_frame_set_packet(self.h, int(primary), int(secondary)))

    def autotune(self, reps=2, stream=None):
        table = (TuneEntry * 64)()
        n = C.c_int()
        check(load().rtn_frame_autotune(self.h, reps, table, 64, C.byref(n), stream))
        return [dict(lanes=e.lanes, stack=e.stack, blocks_per_cu=e.blocks_per_cu, refill_min=e.refill_min, ms=e.ms,
                     packet_primary=e.packet_primary, packet_secondary=e.packet_secondary) for e in table[:n.value]]

    def results(self):
        """Host copies: primary hits, secondary hits (first n_secondary), parent, kind."""
        h1, h2 = Hits(), Hits()
        par, knd = C.c_void_p(), C.c_void_p()
        sr = (C.c_void_p * 6)()
        check(load().rtn_frame_results(self.h, C.byref(h1), C.byref(h2), C.byref(par), C.byref(knd), C.byref(sr)))
        st = self.status()
        n1, n2 = st["n_primary"], st["n_secondary_slots"]

        def grab(ptr, shape, dt):
            out = np.empty(shape, dt)
            if out.nbytes:
                check(load().rtn_memcpy(out.ctypes.data, ptr, out.nbytes, 1))
            return out

        def hits(h, n):
            return dict(hit=grab(h.prim, (n,), np.int32), t=grab(h.t, (n,), np.float32),
                        u=grab(h.u, (n,), np.float32), v=grab(h.v, (n,), np.float32),
                        dpdu=grab(h.dpdu, (n, 3), np.float32), dpdv=grab(h.dpdv, (n, 3), np.float32))
        out = dict(primary=hits(h1, n1), secondary=hits(h2, n2), n_secondary=st["n_secondary"],
                   parent=grab(par.value, (n2,), np.int32), kind=grab(knd.value, (n2,), np.int32),
                   secondary_rays=[grab(sr[k], (n2,), np.float32) for k in range(6)])
        return out

    def sample(self, windows_primary, windows_secondary):
        """Host copies of index windows [(start, count), ...] of the primary and
        of the secondary outputs (hit, t, u, v; parent and kind for the
        secondaries), concatenated, without downloading whole frames."""
        h1, h2 = Hits(), Hits()
        par, knd = C.c_void_p(), C.c_void_p()
        sr = (C.c_void_p * 6)()
        check(load().rtn_frame_results(self.h, C.byref(h1), C.byref(h2), C.byref(par), C.byref(knd), C.byref(sr)))

        def grab(ptr, wins, dt):
            parts = []
            for a, n in wins:
                out = np.empty(n, dt)
                if n:
                    check(load().rtn_memcpy(out.ctypes.data, ptr + a * out.itemsize, out.nbytes, 1))
                parts.append(out)
            return np.concatenate(parts) if parts else np.empty(0, dt)

        def hits(h, wins):
            return dict(hit=grab(h.prim, wins, np.int32), t=grab(h.t, wins, np.float32),
                        u=grab(h.u, wins, np.float32), v=grab(h.v, wins, np.float32))
        sec = hits(h2, windows_secondary)
        sec.update(parent=grab(par.value, windows_secondary, np.int32), kind=grab(knd.value, windows_secondary, np.int32))
        return dict(primary=hits(h1, windows_primary), secondary=sec)

    def __del__(self):
        try:
            if self.h:
                load().rtn_frame_destroy(self.h)
                self.h = None
        except Exception:
            pass


def solve_pairs(scene, rays, pair_ray, pair_patch, n, prm, seed_uv=None, grid=0):
    """Plain-method Newton on device pairs from the plain seed or seed_uv:
    host dict of iters, converged, uv, t.  grid > 0: also from grid x grid
    starts, the nearest converged root with t > 0 reported
    (rtn_solve_pairs_roots)."""
    it = DeviceArray((n,), np.int32)
    cv = DeviceArray((n,), np.int32)
    uv = DeviceArray((n, 2), np.float32)
    t = DeviceArray((n,), np.float32)
    arr = rays.ptr_array()
    check(load().rtn_solve_pairs_roots(scene.h, C.byref(arr), rays.n, pair_ray.ptr, pair_patch.ptr,
                                       seed_uv.ptr if seed_uv is not None else None, n, C.byref(prm), grid, it.ptr,
                                       cv.ptr, uv.ptr, t.ptr, None))
    synchronize()
    return dict(iters=it.get(), converged=cv.get().astype(bool), uv=uv.get(), t=t.get())
