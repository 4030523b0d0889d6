"""interflect_amd — MI355X (gfx950) transfer -> solve -> shade path of Interflect.

Python host mirror of the C ABI in include/interflect.h, loaded with ctypes
from the in-tree ``libinterflect_hip.so``.  The stage methods carry the names
of the reference's stage functions (Cherie05/interflect v0.1.3, src/*.rs):

    scene.parse            -> Context.load_scene        (src/scene.rs:365)
    surfel.generate        -> Context.generate_surfels  (src/surfel.rs:147)
    formfactor.build_clusters + build -> Context.transfer (src/formfactor.rs:91,153)
    solve.light_surfels + solve        -> Context.solve    (src/solve.rs:27,47)
    render_frame + Film.encode         -> Context.shade    (src/main.rs:229, src/film.rs:51)

Beyond the reference: Context.backward (implicit adjoint of the frozen-geometry
pipeline: gradients of a film loss w.r.t. material albedo and light emission).

There is no CPU fallback: if the library is missing, every entry raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libinterflect_hip.so")
_LIB = None

MODES = {"beauty": 0, "direct": 1, "indirect": 2, "normals": 3, "depth": 4, "steps": 5, "reference": 6}

STATUS = {0: "ok", -1: "invalid argument", -2: "parse error", -3: "bad state", -4: "device error",
          -5: "out of memory", -6: "limit exceeded"}

# Every symbol include/interflect.h declares.
EXPORTS = [
    "ifx_create", "ifx_destroy", "ifx_last_error", "ifx_load_scene", "ifx_get_scene_info", "ifx_set_render",
    "ifx_set_use_bvh", "ifx_generate_surfels", "ifx_set_surfels", "ifx_num_surfels", "ifx_get_surfels",
    "ifx_transfer", "ifx_get_transfer", "ifx_get_clusters", "ifx_solve", "ifx_get_radiosity", "ifx_set_radiosity",
    "ifx_shade", "ifx_dist_mat", "ifx_trace_rays", "ifx_soft_shadows", "ifx_set_profiling", "ifx_kernel_time",
    "ifx_reset_kernel_times", "ifx_set_counting", "ifx_kernel_ops", "ifx_set_tuning",
    "ifx_set_material_albedo", "ifx_set_light_emit", "ifx_set_solve_tolerance", "ifx_set_adjoint", "ifx_backward",
    "ifx_backward_surfels", "ifx_get_surfel_materials", "ifx_get_direct_geometry",
    "ifx_batch_create", "ifx_batch_destroy", "ifx_batch_last_error", "ifx_batch_add_scene",
    "ifx_batch_generate_surfels", "ifx_batch_set_graph", "ifx_batch_set_tuning", "ifx_batch_autotune",
    "ifx_batch_run", "ifx_batch_get_stats", "ifx_batch_get_image", "ifx_batch_get_radiosity",
    "ifx_batch_get_transfer", "ifx_shade_views", "ifx_turntable_cameras", "ifx_solve_noalloc",
    "ifx_set_scene", "ifx_scene_digest", "ifx_set_surfel_materials",
]


class IfxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(status, "error"), status, msg))
        self.status = status


class SceneInfo(C.Structure):
    _fields_ = [("n_objects", C.c_uint32), ("n_planes", C.c_uint32), ("n_quad_lights", C.c_uint32),
                ("n_materials", C.c_uint32), ("n_bvh_nodes", C.c_uint32), ("n_sphere_lights", C.c_uint32),
                ("n_point_lights", C.c_uint32), ("max_stack", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("bounces", C.c_uint32), ("reserved", C.c_uint32),
                ("surfels", C.c_uint64), ("exposure", C.c_float), ("blend_k", C.c_float),
                ("lo", C.c_float * 3), ("hi", C.c_float * 3), ("cam_pos", C.c_float * 3),
                ("cam_look", C.c_float * 3), ("cam_up", C.c_float * 3), ("cam_fov", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("look", C.c_float * 3), ("up", C.c_float * 3), ("fov", C.c_float)]


class SceneDesc(C.Structure):
    """ifx_scene_desc: array pointers (arrays in the layout of ifx_object,
    ifx_prim, ifx_bvh_node, ifx_material and the light structs) and counts."""
    _fields_ = [("objects", C.c_void_p), ("n_objects", C.c_uint32), ("planes", C.c_void_p), ("n_planes", C.c_uint32),
                ("subtractions", C.c_void_p), ("n_subtractions", C.c_uint32), ("bvh_nodes", C.c_void_p),
                ("n_bvh_nodes", C.c_uint32), ("bvh_order", C.c_void_p), ("materials", C.c_void_p),
                ("n_materials", C.c_uint32), ("quad_lights", C.c_void_p), ("n_quad_lights", C.c_uint32),
                ("sphere_lights", C.c_void_p), ("n_sphere_lights", C.c_uint32), ("point_lights", C.c_void_p),
                ("n_point_lights", C.c_uint32), ("lo", C.c_float * 3), ("hi", C.c_float * 3), ("camera", Camera),
                ("width", C.c_uint32), ("height", C.c_uint32), ("bounces", C.c_uint32), ("reserved", C.c_uint32),
                ("surfels", C.c_uint64), ("exposure", C.c_float)]


class TransferStats(C.Structure):
    _fields_ = [("clusters", C.c_uint32), ("reserved", C.c_uint32), ("candidates", C.c_uint64),
                ("links", C.c_uint64), ("evals", C.c_uint64), ("mean_row_sum", C.c_float), ("ms", C.c_float)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("residual", C.c_float), ("evals", C.c_uint64), ("ms", C.c_float),
                ("reserved", C.c_float)]


class ShadeStats(C.Structure):
    _fields_ = [("evals", C.c_uint64), ("ms", C.c_float), ("reserved", C.c_float)]


class BatchSceneStats(C.Structure):
    _fields_ = [("surfels", C.c_uint64), ("candidates", C.c_uint64), ("links", C.c_uint64),
                ("evals_transfer", C.c_uint64), ("evals_solve", C.c_uint64), ("evals_shade", C.c_uint64),
                ("clusters", C.c_uint32), ("iterations", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("residual", C.c_float), ("mean_row_sum", C.c_float)]


class BackwardStats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("residual", C.c_float), ("ms", C.c_float), ("link_tiles", C.c_uint32)]


def lib():
    """Load libinterflect_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libinterflect_hip.so is not built: run `python -m interflect_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, f32, u32, u64, i32 = C.c_void_p, C.c_float, C.c_uint32, C.c_uint64, C.c_int32
    P = C.c_void_p  # raw buffer pointers (numpy .ctypes.data)
    sig = {
        "ifx_create": [C.POINTER(vp), i32],
        "ifx_destroy": [vp],
        "ifx_load_scene": [vp, C.c_char_p],
        "ifx_set_scene": [vp, C.POINTER(SceneDesc)],
        "ifx_scene_digest": [vp, C.POINTER(u64)],
        "ifx_set_surfel_materials": [vp, P, u64],
        "ifx_get_scene_info": [vp, C.POINTER(SceneInfo)],
        "ifx_set_render": [vp, u32, u32, u64, u32],
        "ifx_set_use_bvh": [vp, i32],
        "ifx_generate_surfels": [vp, u64, C.POINTER(u64)],
        "ifx_set_surfels": [vp, P, P, P, P, u64],
        "ifx_num_surfels": [vp, C.POINTER(u64)],
        "ifx_get_surfels": [vp, P, P, P, P],
        "ifx_transfer": [vp, i32, C.POINTER(TransferStats)],
        "ifx_get_transfer": [vp, P, P, P, P, P],
        "ifx_get_clusters": [vp, P, P, P, P, P, P],
        "ifx_solve": [vp, u32, C.POINTER(SolveStats)],
        "ifx_solve_noalloc": [vp, u32, P, P, P, C.POINTER(SolveStats)],
        "ifx_get_radiosity": [vp, P, P, P],
        "ifx_set_radiosity": [vp, P, P, P],
        "ifx_shade": [vp, C.POINTER(Camera), u32, u32, i32, P, P, C.POINTER(ShadeStats)],
        "ifx_dist_mat": [vp, P, u64, P, P, P],
        "ifx_trace_rays": [vp, P, P, u64, f32, f32, P, P, P, P, P],
        "ifx_soft_shadows": [vp, P, P, u64, f32, P, P, P, P],
        "ifx_set_profiling": [vp, i32],
        "ifx_kernel_time": [vp, C.c_char_p, C.POINTER(u64), C.POINTER(C.c_double)],
        "ifx_reset_kernel_times": [vp],
        "ifx_set_counting": [vp, i32],
        "ifx_kernel_ops": [vp, C.c_char_p, C.POINTER(u64)],
        "ifx_set_tuning": [vp, C.c_char_p, C.c_int64],
        "ifx_set_material_albedo": [vp, u32, P],
        "ifx_set_light_emit": [vp, u32, P],
        "ifx_set_solve_tolerance": [vp, f32],
        "ifx_set_adjoint": [vp, u32, f32],
        "ifx_backward": [vp, P, P, P, C.POINTER(BackwardStats)],
        "ifx_backward_surfels": [vp, P, P, P, P, C.POINTER(BackwardStats)],
        "ifx_get_surfel_materials": [vp, P],
        "ifx_get_direct_geometry": [vp, P],
        "ifx_batch_create": [C.POINTER(vp), i32],
        "ifx_batch_destroy": [vp],
        "ifx_batch_add_scene": [vp, C.c_char_p, i32, C.POINTER(u32)],
        "ifx_batch_generate_surfels": [vp],
        "ifx_batch_set_graph": [vp, i32],
        "ifx_batch_set_tuning": [vp, u32, u32, u32],
        "ifx_batch_autotune": [vp, C.POINTER(f32)],
        "ifx_batch_run": [vp, C.POINTER(f32)],
        "ifx_batch_get_stats": [vp, u32, C.POINTER(BatchSceneStats)],
        "ifx_batch_get_image": [vp, u32, P, P],
        "ifx_batch_get_radiosity": [vp, u32, P, P, P],
        "ifx_batch_get_transfer": [vp, u32, P, P, P],
        "ifx_shade_views": [vp, P, u32, u32, u32, i32, P, P, C.POINTER(ShadeStats)],
        "ifx_turntable_cameras": [vp, u32, P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int32
    L.ifx_last_error.argtypes = [vp]
    L.ifx_last_error.restype = C.c_char_p
    L.ifx_batch_last_error.argtypes = [vp]
    L.ifx_batch_last_error.restype = C.c_char_p
    _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f3(a):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 3)


class Context:
    """One scene on one device (device=-1: host-only, loader and surfels only)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.ifx_create(C.byref(h), device)
        self._h = h
        if rc != 0:
            msg = L.ifx_last_error(h).decode() if h.value else "create failed"
            if h.value:
                L.ifx_destroy(h)
            self._h = None
            raise IfxError(rc, msg)
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib().ifx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise IfxError(rc, lib().ifx_last_error(self._h).decode())

    # ---- scene
    def load_scene(self, rad_text):
        self._check(lib().ifx_load_scene(self._h, rad_text.encode()))
        return self.info()

    def set_scene(self, desc, with_bvh=True):
        """ifx_set_scene from arrays in the descriptor's layout (keys objects,
        planes, subtractions, bvh_nodes, bvh_order, materials, quad_lights,
        sphere_lights, point_lights: 2-D float32 arrays, integers as their bit
        patterns; bounds[6]; camera[10]; render {width, height, bounces,
        surfels, exposure}).  with_bvh=False lets the library build the BVH."""
        keep = {k: np.ascontiguousarray(v) for k, v in desc.items() if isinstance(v, np.ndarray)}

        def arr(k):
            a = keep[k]
            return (a.ctypes.data if a.size else None), (a.shape[0] if a.ndim > 1 else a.size)

        d = SceneDesc()
        d.objects, d.n_objects = arr("objects")
        d.planes, d.n_planes = arr("planes")
        d.subtractions, d.n_subtractions = arr("subtractions")
        if with_bvh:
            d.bvh_nodes, d.n_bvh_nodes = arr("bvh_nodes")
            d.bvh_order = arr("bvh_order")[0]
        d.materials, d.n_materials = arr("materials")
        d.quad_lights, d.n_quad_lights = arr("quad_lights")
        d.sphere_lights, d.n_sphere_lights = arr("sphere_lights")
        d.point_lights, d.n_point_lights = arr("point_lights")
        b = keep["bounds"]
        d.lo, d.hi = (C.c_float * 3)(*b[:3]), (C.c_float * 3)(*b[3:6])
        c = keep["camera"]
        d.camera = Camera((C.c_float * 3)(*c[0:3]), (C.c_float * 3)(*c[3:6]), (C.c_float * 3)(*c[6:9]), float(c[9]))
        r = desc["render"]
        d.width, d.height, d.bounces = int(r["width"]), int(r["height"]), int(r["bounces"])
        d.surfels, d.exposure = int(r["surfels"]), float(r["exposure"])
        self._check(lib().ifx_set_scene(self._h, C.byref(d)))
        return self.info()

    def scene_digest(self):
        v = C.c_uint64()
        self._check(lib().ifx_scene_digest(self._h, C.byref(v)))
        return int(v.value)

    def set_surfel_materials(self, mat):
        m = np.ascontiguousarray(mat, np.uint32)
        self._check(lib().ifx_set_surfel_materials(self._h, _ptr(m), m.size))

    def info(self):
        s = SceneInfo()
        self._check(lib().ifx_get_scene_info(self._h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in SceneInfo._fields_ if k != "reserved"}
        for k in ("lo", "hi", "cam_pos", "cam_look", "cam_up"):
            d[k] = np.array(d[k][:], np.float32)
        return d

    def set_render(self, width, height, surfels, bounces):
        self._check(lib().ifx_set_render(self._h, width, height, surfels, bounces))

    def set_use_bvh(self, on):
        self._check(lib().ifx_set_use_bvh(self._h, 1 if on else 0))

    # ---- surfels
    def generate_surfels(self, target=None):
        target = int(target or self.info()["surfels"])
        placed = C.c_uint64()
        self._check(lib().ifx_generate_surfels(self._h, target, C.byref(placed)))
        return int(placed.value)

    def set_surfels(self, p, n, albedo, area):
        p, n, albedo = _f3(p), _f3(n), _f3(albedo)
        N = p.shape[0]
        area = np.ascontiguousarray(np.broadcast_to(np.float32(area), (N,)) if np.ndim(area) == 0 else area,
                                    dtype=np.float32)
        self._check(lib().ifx_set_surfels(self._h, _ptr(p), _ptr(n), _ptr(albedo), _ptr(area), N))

    def num_surfels(self):
        n = C.c_uint64()
        self._check(lib().ifx_num_surfels(self._h, C.byref(n)))
        return int(n.value)

    def surfels(self):
        N = self.num_surfels()
        p, n, a = (np.zeros((N, 3), np.float32) for _ in range(3))
        ar = np.zeros(N, np.float32)
        self._check(lib().ifx_get_surfels(self._h, _ptr(p), _ptr(n), _ptr(a), _ptr(ar)))
        return p, n, a, ar

    # ---- stages
    def transfer(self, cluster_res=8):
        st = TransferStats()
        self._check(lib().ifx_transfer(self._h, cluster_res, C.byref(st)))
        self._links, self._nc = int(st.links), int(st.clusters)
        return {"clusters": int(st.clusters), "candidates": int(st.candidates), "links": int(st.links),
                "evals": int(st.evals), "mean_row_sum": float(st.mean_row_sum), "ms": float(st.ms)}

    def transfer_arrays(self):
        N = self.num_surfels()
        L = self._links
        starts = np.zeros(N + 1, np.uint32)
        idx = np.zeros(max(L, 1), np.uint32)
        w = np.zeros(max(L, 1), np.float32)
        sky = np.zeros(max(N, 1), np.float32)
        rs = np.zeros(max(N, 1), np.float32)
        self._check(lib().ifx_get_transfer(self._h, _ptr(starts), _ptr(idx), _ptr(w), _ptr(sky), _ptr(rs)))
        return starts, idx[:L], w[:L], sky[:N], rs[:N]

    def clusters(self):
        N = self.num_surfels()
        nc = self._nc
        p = np.zeros((max(nc, 1), 3), np.float32)
        n = np.zeros((max(nc, 1), 3), np.float32)
        a = np.zeros(max(nc, 1), np.float32)
        r = np.zeros(max(nc, 1), np.float32)
        ms = np.zeros(nc + 1, np.uint32)
        mem = np.zeros(max(N, 1), np.uint32)
        self._check(lib().ifx_get_clusters(self._h, _ptr(p), _ptr(n), _ptr(a), _ptr(r), _ptr(ms), _ptr(mem)))
        return p[:nc], n[:nc], a[:nc], r[:nc], ms, mem[: int(ms[-1])]

    def solve(self, max_bounces=None):
        st = SolveStats()
        b = int(max_bounces or self.info()["bounces"])
        self._check(lib().ifx_solve(self._h, b, C.byref(st)))
        return {"iterations": int(st.iterations), "residual": float(st.residual), "evals": int(st.evals),
                "ms": float(st.ms)}

    def solve_noalloc(self, d_direct, d_irr, d_rad, max_bounces=None):
        """ifx_solve_noalloc: outputs into caller-owned device arrays (integer
        device addresses of N x 4 float32 each); allocates nothing."""
        st = SolveStats()
        b = int(max_bounces or self.info()["bounces"])
        self._check(lib().ifx_solve_noalloc(self._h, b, C.c_void_p(d_direct), C.c_void_p(d_irr), C.c_void_p(d_rad),
                                            C.byref(st)))
        return {"iterations": int(st.iterations), "residual": float(st.residual), "evals": int(st.evals),
                "ms": float(st.ms)}

    def radiosity(self):
        N = self.num_surfels()
        d, i, r = (np.zeros((N, 3), np.float32) for _ in range(3))
        self._check(lib().ifx_get_radiosity(self._h, _ptr(d), _ptr(i), _ptr(r)))
        return d, i, r

    def set_radiosity(self, direct, irr, rad):
        d, i, r = _f3(direct), _f3(irr), _f3(rad)
        self._check(lib().ifx_set_radiosity(self._h, _ptr(d), _ptr(i), _ptr(r)))

    def shade(self, mode="beauty", width=None, height=None, camera=None, film=True, rgb8=True):
        info = self.info()
        w = int(width or info["width"])
        h = int(height or info["height"])
        cam = None
        if camera is not None:
            c = np.asarray(camera, np.float32).ravel()
            cam = Camera((C.c_float * 3)(*c[0:3]), (C.c_float * 3)(*c[3:6]), (C.c_float * 3)(*c[6:9]), float(c[9]))
        f = np.zeros((h, w, 3), np.float32) if film else None
        b = np.zeros((h, w, 3), np.uint8) if rgb8 else None
        st = ShadeStats()
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        self._check(lib().ifx_shade(self._h, C.byref(cam) if cam is not None else None, w, h, m, _ptr(f), _ptr(b),
                                    C.byref(st)))
        return {"film": f, "rgb8": b, "evals": int(st.evals), "ms": float(st.ms)}

    def shade_views(self, cameras, mode="beauty", width=None, height=None, film=True, rgb8=True):
        """One launch over many cameras (n x 10 floats: pos, look, up, fov)."""
        info = self.info()
        w = int(width or info["width"])
        h = int(height or info["height"])
        cams = np.ascontiguousarray(cameras, np.float32).reshape(-1, 10)
        n = cams.shape[0]
        f = np.zeros((n, h, w, 3), np.float32) if film else None
        b = np.zeros((n, h, w, 3), np.uint8) if rgb8 else None
        st = ShadeStats()
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        self._check(lib().ifx_shade_views(self._h, _ptr(cams), n, w, h, m, _ptr(f), _ptr(b), C.byref(st)))
        return {"film": f, "rgb8": b, "evals": int(st.evals), "ms": float(st.ms)}

    def turntable_cameras(self, n):
        """The reference's n orbit cameras (n x 10 floats)."""
        out = np.zeros((n, 10), np.float32)
        self._check(lib().ifx_turntable_cameras(self._h, int(n), _ptr(out)))
        return out

    # ---- kernel-family probes
    def dist_mat(self, pts, normals=False):
        pts = _f3(pts)
        n = pts.shape[0]
        d = np.zeros(n, np.float32)
        m = np.zeros(n, np.uint32)
        nrm = np.zeros((n, 3), np.float32) if normals else None
        self._check(lib().ifx_dist_mat(self._h, _ptr(pts), n, _ptr(d), _ptr(m), _ptr(nrm)))
        return (d, m, nrm) if normals else (d, m)

    def trace_rays(self, ro, rd, tmin, tmax):
        ro, rd = _f3(ro), _f3(rd)
        n = ro.shape[0]
        t = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        mat = np.zeros(n, np.uint32)
        hit = np.zeros(n, np.int32)
        steps = np.zeros(n, np.uint32)
        self._check(lib().ifx_trace_rays(self._h, _ptr(ro), _ptr(rd), n, tmin, tmax, _ptr(t), _ptr(nrm), _ptr(mat),
                                         _ptr(hit), _ptr(steps)))
        return hit.astype(bool), t, nrm, mat, steps

    def soft_shadows(self, ro, rd, tmin, tmax, k):
        ro, rd = _f3(ro), _f3(rd)
        n = ro.shape[0]
        tmax = np.ascontiguousarray(np.broadcast_to(np.float32(tmax), (n,)), np.float32)
        k = np.ascontiguousarray(np.broadcast_to(np.float32(k), (n,)), np.float32)
        vis = np.zeros(n, np.float32)
        steps = np.zeros(n, np.uint32)
        self._check(lib().ifx_soft_shadows(self._h, _ptr(ro), _ptr(rd), n, tmin, _ptr(tmax), _ptr(k), _ptr(vis),
                                           _ptr(steps)))
        return vis, steps

    # ---- timing
    def set_profiling(self, on):
        self._check(lib().ifx_set_profiling(self._h, 1 if on else 0))

    def kernel_time(self, name):
        n = C.c_uint64()
        ms = C.c_double()
        self._check(lib().ifx_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def reset_kernel_times(self):
        self._check(lib().ifx_reset_kernel_times(self._h))

    def set_tuning(self, key, value):
        self._check(lib().ifx_set_tuning(self._h, key.encode(), int(value)))

    def set_counting(self, on):
        """True / 1: counting instantiation (FP32 operations and primitive
        evaluations of the reference traversal); 2: the timed kernels count
        the primitive evaluations they execute; False / 0: off."""
        self._check(lib().ifx_set_counting(self._h, int(on)))

    def kernel_ops(self, name):
        n = C.c_uint64()
        self._check(lib().ifx_kernel_ops(self._h, name.encode(), C.byref(n)))
        return int(n.value)

    # ---- adjoint (no counterpart in the reference)
    def set_material_albedo(self, material, rgb):
        a = np.ascontiguousarray(rgb, np.float32).reshape(3)
        self._check(lib().ifx_set_material_albedo(self._h, int(material), _ptr(a)))

    def set_light_emit(self, light, rgb):
        a = np.ascontiguousarray(rgb, np.float32).reshape(3)
        self._check(lib().ifx_set_light_emit(self._h, int(light), _ptr(a)))

    def set_solve_tolerance(self, tol):
        self._check(lib().ifx_set_solve_tolerance(self._h, float(tol)))

    def set_adjoint(self, max_iters, rel_tol):
        self._check(lib().ifx_set_adjoint(self._h, int(max_iters), float(rel_tol)))

    def _grad_shapes(self):
        info = self.info()
        nl = info["n_quad_lights"] + info["n_sphere_lights"] + info["n_point_lights"]
        # row 0: the implicit default material; row k: the k-th material of the file
        return np.zeros((info["n_materials"] + 1, 3), np.float32), np.zeros((nl, 3), np.float32)

    def backward(self, dL_dfilm):
        """Gradients of L w.r.t. albedo ((n_materials + 1) x 3, row 0 the implicit
        default material) and emission (lights x 3) for dL/dfilm (H x W x 3) of
        the last beauty shade."""
        u = np.ascontiguousarray(dL_dfilm, np.float32)
        da, de = self._grad_shapes()
        st = BackwardStats()
        self._check(lib().ifx_backward(self._h, _ptr(u), _ptr(da), _ptr(de), C.byref(st)))
        return {"albedo": da, "emit": de, "iterations": st.iterations, "residual": st.residual, "ms": st.ms,
                "link_tiles": st.link_tiles}

    def backward_surfels(self, dL_dirr):
        """The solve's share for dL/dirr (N x 3): gradients and the adjoint state mu."""
        u = _f3(dL_dirr)
        da, de = self._grad_shapes()
        mu = np.zeros_like(u)
        st = BackwardStats()
        self._check(lib().ifx_backward_surfels(self._h, _ptr(u), _ptr(da), _ptr(de), _ptr(mu), C.byref(st)))
        return {"albedo": da, "emit": de, "mu": mu, "iterations": st.iterations, "residual": st.residual,
                "ms": st.ms, "link_tiles": st.link_tiles}

    def surfel_materials(self):
        m = np.zeros(self.num_surfels(), np.uint32)
        self._check(lib().ifx_get_surfel_materials(self._h, _ptr(m)))
        return m

    def direct_geometry(self):
        info = self.info()
        nl = info["n_quad_lights"] + info["n_sphere_lights"] + info["n_point_lights"]
        g = np.zeros((nl, self.num_surfels()), np.float32)
        self._check(lib().ifx_get_direct_geometry(self._h, _ptr(g)))
        return g


class Batch:
    """Many scenes per launch (include/interflect.h, ifx_batch_*): one frame =
    transfer + solve + beauty shade + RGB8 of every scene, each with its own
    render {} settings and camera."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.ifx_batch_create(C.byref(h), device)
        self._h = h
        if rc != 0:
            msg = L.ifx_batch_last_error(h).decode() if h.value else "create failed"
            if h.value:
                L.ifx_batch_destroy(h)
            self._h = C.c_void_p()
            raise IfxError(rc, msg)

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value and _LIB is not None:
            _LIB.ifx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def _check(self, rc):
        if rc != 0:
            raise IfxError(rc, lib().ifx_batch_last_error(self._h).decode())

    def add_scene(self, text, cluster_res=8):
        i = C.c_uint32()
        self._check(lib().ifx_batch_add_scene(self._h, text.encode(), int(cluster_res), C.byref(i)))
        return int(i.value)

    def generate_surfels(self):
        self._check(lib().ifx_batch_generate_surfels(self._h))

    def set_graph(self, on):
        self._check(lib().ifx_batch_set_graph(self._h, 1 if on else 0))

    def set_tuning(self, blocks_per_cu=7, chunk_div=8, wave_rows=0):
        self._check(lib().ifx_batch_set_tuning(self._h, int(blocks_per_cu), int(chunk_div), int(wave_rows)))

    def autotune(self):
        ms = C.c_float()
        self._check(lib().ifx_batch_autotune(self._h, C.byref(ms)))
        return float(ms.value)

    def run(self):
        ms = C.c_float()
        self._check(lib().ifx_batch_run(self._h, C.byref(ms)))
        return float(ms.value)

    def stats(self, scene):
        st = BatchSceneStats()
        self._check(lib().ifx_batch_get_stats(self._h, int(scene), C.byref(st)))
        return {k: getattr(st, k) for k, _ in BatchSceneStats._fields_}

    def image(self, scene, film=True, rgb8=True):
        st = self.stats(scene)
        w, h = st["width"], st["height"]
        f = np.zeros((h, w, 3), np.float32) if film else None
        b = np.zeros((h, w, 3), np.uint8) if rgb8 else None
        self._check(lib().ifx_batch_get_image(self._h, int(scene), _ptr(f), _ptr(b)))
        return {"film": f, "rgb8": b}

    def radiosity(self, scene):
        n = self.stats(scene)["surfels"]
        d, i, r = (np.zeros((n, 3), np.float32) for _ in range(3))
        self._check(lib().ifx_batch_get_radiosity(self._h, int(scene), _ptr(d), _ptr(i), _ptr(r)))
        return d, i, r

    def transfer_arrays(self, scene):
        st = self.stats(scene)
        n, L = st["surfels"], st["links"]
        starts = np.zeros(n + 1, np.uint32)
        idx = np.zeros(max(L, 1), np.uint32)
        w = np.zeros(max(L, 1), np.float32)
        self._check(lib().ifx_batch_get_transfer(self._h, int(scene), _ptr(starts), _ptr(idx), _ptr(w)))
        return starts, idx[:L], w[:L]
