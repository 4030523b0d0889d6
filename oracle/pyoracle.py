"""ORACLE — test infrastructure only.

ctypes binding of ``oracle/liboracle.so``, the CPU restatement of the
reference's transfer -> solve -> shade path (src/*.rs of the reference).
Only tests/, the smoke check and bench.py's CPU-baseline leg import this
module, and only as the checker: the product library never loads it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

MODES = {"beauty": 0, "direct": 1, "indirect": 2, "normals": 3, "depth": 4, "steps": 5, "reference": 6}


def lib_path():
    return os.path.join(_HERE, "liboracle.so")


def build():
    """Compile the oracle with its Makefile (g++, no FMA contraction)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.orc_create.restype = vp
    L.orc_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.orc_destroy.argtypes = [vp]
    L.orc_scene_counts.argtypes = [vp, i64p]
    L.orc_scene_cfg.argtypes = [vp, u64p, f32p]
    L.orc_set_use_bvh.argtypes = [vp, C.c_int]
    L.orc_bvh.argtypes = [vp, f32p, u32p, u32p]
    L.orc_dist_mat.argtypes = [vp, f32p, C.c_int64, f32p, u32p, C.c_int]
    L.orc_normal.argtypes = [vp, f32p, C.c_int64, f32p, C.c_int]
    L.orc_trace.argtypes = [vp, f32p, f32p, C.c_int64, C.c_float, C.c_float, f32p, f32p, u32p, i32p, u32p, C.c_int]
    L.orc_soft_shadow.argtypes = [vp, f32p, f32p, C.c_int64, C.c_float, f32p, f32p, f32p, u32p, C.c_int]
    L.orc_closure.argtypes = [vp, f32p, f32p, C.c_int64, C.c_float, C.c_uint32, f32p, u32p, C.c_int]
    L.orc_direct_diffuse.argtypes = [vp, f32p, f32p, f32p, C.c_int64, f32p, u32p, C.c_int]
    L.orc_generate.restype = C.c_int64
    L.orc_generate.argtypes = [vp, C.c_uint64]
    L.orc_num_surfels.restype = C.c_int64
    L.orc_num_surfels.argtypes = [vp]
    L.orc_get_surfels.argtypes = [vp, f32p, f32p, f32p, f32p]
    L.orc_set_surfels.argtypes = [vp, f32p, f32p, f32p, f32p, C.c_int64]
    L.orc_transfer.restype = C.c_int
    L.orc_transfer.argtypes = [vp, C.c_int32, C.c_int, C.c_int64, C.c_int64, f64p]
    L.orc_get_transfer.argtypes = [vp, u32p, u32p, f32p, f32p, f32p]
    L.orc_get_clusters.argtypes = [vp, f32p, f32p, f32p, f32p, u32p, u32p]
    L.orc_solve.restype = C.c_int
    L.orc_solve.argtypes = [vp, C.c_uint32, C.c_int, f64p]
    L.orc_light_surfels.restype = C.c_uint64
    L.orc_light_surfels.argtypes = [vp, C.c_int]
    L.orc_get_radiosity.argtypes = [vp, f32p, f32p, f32p]
    L.orc_set_radiosity.argtypes = [vp, f32p, f32p, f32p]
    L.orc_render.restype = C.c_uint64
    L.orc_render.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int, C.c_int, C.c_uint32,
                             C.c_uint32, f32p]
    L.orc_gather.argtypes = [vp, f32p, f32p, C.c_int64, C.c_float, f32p]
    L.orc_encode.argtypes = [f32p, C.c_int64, C.c_float, C.c_int, u8p]
    L.orc_prim_dist.restype = C.c_float
    L.orc_prim_dist.argtypes = [C.c_int, f32p, f32p]
    L.orc_prim_bounds.restype = C.c_int
    L.orc_prim_bounds.argtypes = [C.c_int, f32p, f32p]
    L.orc_quad_ff.restype = C.c_float
    L.orc_quad_ff.argtypes = [f32p, f32p, f32p]
    L.orc_sphere_ff.restype = C.c_float
    L.orc_sphere_ff.argtypes = [C.c_float, C.c_float]
    L.orc_srgb8.restype = C.c_uint32
    L.orc_srgb8.argtypes = [C.c_float]
    L.orc_aces.argtypes = [f32p, f32p]
    L.orc_radical_inverse.restype = C.c_float
    L.orc_radical_inverse.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_frame.argtypes = [f32p, f32p, f32p]
    L.orc_set_spp.argtypes = [vp, C.c_uint32]
    L.orc_scene_export_counts.argtypes = [vp, u32p]
    L.orc_scene_export.argtypes = [vp, f32p, f32p, f32p, f32p, u32p, f32p, f32p, f32p, f32p, f32p]
    L.orc_set_tally.argtypes = [C.c_int]
    L.orc_tallies.argtypes = [vp, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
    L.orc_fresnel.argtypes = [C.c_float, f32p, f32p]
    L.orc_d_ggx.restype = C.c_float
    L.orc_d_ggx.argtypes = [C.c_float, C.c_float]
    L.orc_v_smith.restype = C.c_float
    L.orc_v_smith.argtypes = [C.c_float, C.c_float, C.c_float]
    L.orc_f0_of.argtypes = [f32p, C.c_float, f32p]
    L.orc_direct.argtypes = [vp, f32p, f32p, f32p, f32p, C.c_float, C.c_float, f32p, u32p]
    L.orc_emitter_hit.restype = C.c_int
    L.orc_emitter_hit.argtypes = [vp, f32p, f32p, C.c_float, f32p]
    L.orc_intersect_quad.restype = C.c_int
    L.orc_intersect_quad.argtypes = [f32p, f32p, f32p, f32p]
    L.orc_rng.argtypes = [C.c_uint64, C.c_uint32, f32p]
    _LIB = L
    return L


def _f(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


class OracleError(RuntimeError):
    pass


class Scene:
    """One parsed scene plus the pipeline state (surfels, transfer, radiosity)."""

    def __init__(self, src, threads=None):
        L = lib()
        err = C.create_string_buffer(512)
        self.h = L.orc_create(src.encode(), err, 512)
        if not self.h:
            raise OracleError(err.value.decode())
        self.threads = threads or os.cpu_count() or 1
        c = np.zeros(8, np.int64)
        L.orc_scene_counts(self.h, c)
        (self.n_objects, self.n_planes, self.n_lights, self.n_materials, self.n_bvh_nodes,
         self.n_sphere_lights, self.n_point_lights) = [int(x) for x in c[:7]]
        self.blend_k = float(c[7]) / 1e6
        cfg = np.zeros(4, np.uint64)
        f = np.zeros(17, np.float32)
        L.orc_scene_cfg(self.h, cfg, f)
        self.width, self.height, self.surfels_requested, self.bounces = [int(x) for x in cfg]
        self.exposure = float(f[0])
        self.lo = f[1:4].copy()
        self.hi = f[4:7].copy()
        self.camera = f[7:17].copy()
        self.diag = float(np.float32(np.sqrt(np.float32(np.sum((self.hi - self.lo).astype(np.float32) ** 2)))))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_use_bvh(self, on):
        lib().orc_set_use_bvh(self.h, 1 if on else 0)

    def bvh(self):
        n = self.n_bvh_nodes
        lohi = np.zeros((max(n, 1), 6), np.float32)
        scr = np.zeros((max(n, 1), 3), np.uint32)
        order = np.zeros(max(self.n_objects, 1), np.uint32)
        lib().orc_bvh(self.h, lohi, scr, order)
        return lohi[:n], scr[:n], order[: self.n_objects]

    def dist_mat(self, pts):
        pts = _f(pts, (-1, 3))
        n = pts.shape[0]
        d = np.zeros(n, np.float32)
        m = np.zeros(n, np.uint32)
        lib().orc_dist_mat(self.h, pts, n, d, m, self.threads)
        return d, m

    def normal(self, pts):
        pts = _f(pts, (-1, 3))
        out = np.zeros_like(pts)
        lib().orc_normal(self.h, pts, pts.shape[0], out, self.threads)
        return out

    def trace(self, ro, rd, tmin, tmax):
        ro = _f(ro, (-1, 3))
        rd = _f(rd, (-1, 3))
        n = ro.shape[0]
        t = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        mat = np.zeros(n, np.uint32)
        hit = np.zeros(n, np.int32)
        steps = np.zeros(n, np.uint32)
        lib().orc_trace(self.h, ro, rd, n, tmin, tmax, t, nrm, mat, hit, steps, self.threads)
        return hit.astype(bool), t, nrm, mat, steps

    def soft_shadow(self, ro, rd, tmin, tmax, k):
        ro = _f(ro, (-1, 3))
        rd = _f(rd, (-1, 3))
        n = ro.shape[0]
        tmax = _f(np.broadcast_to(np.float32(tmax), (n,)))
        k = _f(np.broadcast_to(np.float32(k), (n,)))
        vis = np.zeros(n, np.float32)
        steps = np.zeros(n, np.uint32)
        lib().orc_soft_shadow(self.h, ro, rd, n, tmin, tmax, k, vis, steps, self.threads)
        return vis, steps

    def closure(self, p, n, tmax, dirs=16):
        p = _f(p, (-1, 3))
        n_ = _f(n, (-1, 3))
        sky = np.zeros(p.shape[0], np.float32)
        steps = np.zeros(p.shape[0], np.uint32)
        lib().orc_closure(self.h, p, n_, p.shape[0], tmax, dirs, sky, steps, self.threads)
        return sky, steps

    def direct_diffuse(self, p, n, albedo):
        p = _f(p, (-1, 3))
        out = np.zeros_like(p)
        steps = np.zeros(p.shape[0], np.uint32)
        lib().orc_direct_diffuse(self.h, p, _f(n, (-1, 3)), _f(albedo, (-1, 3)), p.shape[0], out, steps,
                                 self.threads)
        return out, steps

    # ---- pipeline
    def generate(self, target=None):
        return int(lib().orc_generate(self.h, int(target or self.surfels_requested)))

    def surfels(self):
        n = int(lib().orc_num_surfels(self.h))
        p = np.zeros((n, 3), np.float32)
        nn = np.zeros((n, 3), np.float32)
        a = np.zeros((n, 3), np.float32)
        ar = np.zeros(n, np.float32)
        if n:
            lib().orc_get_surfels(self.h, p, nn, a, ar)
        return p, nn, a, ar

    def set_surfels(self, p, n, albedo, area):
        p = _f(p, (-1, 3))
        N = p.shape[0]
        area = _f(np.broadcast_to(np.float32(area), (N,)) if np.ndim(area) == 0 else area)
        lib().orc_set_surfels(self.h, p, _f(n, (-1, 3)), _f(albedo, (-1, 3)), area, N)

    def transfer(self, res=8, rows=None):
        st = np.zeros(4, np.float64)
        rb, re = (-1, -1) if rows is None else rows
        rc = lib().orc_transfer(self.h, res, self.threads, rb, re, st)
        if rc != 0:
            raise OracleError("transfer failed")
        self.nc, self.links = int(st[0]), int(st[1])
        return {"clusters": int(st[0]), "links": int(st[1]), "mean_row_sum": float(st[2]), "evals": int(st[3])}

    def transfer_arrays(self):
        N = int(lib().orc_num_surfels(self.h))
        starts = np.zeros(N + 1, np.uint32)
        idx = np.zeros(max(self.links, 1), np.uint32)
        w = np.zeros(max(self.links, 1), np.float32)
        sky = np.zeros(max(N, 1), np.float32)
        rs = np.zeros(max(N, 1), np.float32)
        lib().orc_get_transfer(self.h, starts, idx, w, sky, rs)
        return starts, idx[: self.links], w[: self.links], sky[:N], rs[:N]

    def clusters(self):
        N = int(lib().orc_num_surfels(self.h))
        nc = self.nc
        p = np.zeros((max(nc, 1), 3), np.float32)
        n = np.zeros((max(nc, 1), 3), np.float32)
        a = np.zeros(max(nc, 1), np.float32)
        r = np.zeros(max(nc, 1), np.float32)
        ms = np.zeros(nc + 1, np.uint32)
        mem = np.zeros(max(N, 1), np.uint32)
        lib().orc_get_clusters(self.h, p, n, a, r, ms, mem)
        return p[:nc], n[:nc], a[:nc], r[:nc], ms, mem[: int(ms[-1])]

    def solve(self, bounces=None):
        st = np.zeros(3, np.float64)
        rc = lib().orc_solve(self.h, int(bounces or self.bounces), self.threads, st)
        if rc != 0:
            raise OracleError("solve before transfer")
        return {"evals": int(st[0]), "iterations": int(st[1]), "residual": float(st[2])}

    def light_surfels(self):
        return int(lib().orc_light_surfels(self.h, self.threads))

    def radiosity(self):
        n = int(lib().orc_num_surfels(self.h))
        d = np.zeros((n, 3), np.float32)
        i = np.zeros((n, 3), np.float32)
        r = np.zeros((n, 3), np.float32)
        lib().orc_get_radiosity(self.h, d, i, r)
        return d, i, r

    def set_radiosity(self, direct, irr, rad):
        lib().orc_set_radiosity(self.h, _f(direct, (-1, 3)), _f(irr, (-1, 3)), _f(rad, (-1, 3)))

    def set_spp(self, spp):
        """Paths per pixel of the "reference" mode (src/main.rs:81)."""
        lib().orc_set_spp(self.h, int(spp))

    def render(self, mode="beauty", width=None, height=None, camera=None, use_gi=True, rows=None):
        w = int(width or self.width)
        h = int(height or self.height)
        film = np.zeros((h * w, 3), np.float32)
        cam = None
        if camera is not None:
            camb = _f(camera)
            cam = camb.ctypes.data_as(C.c_void_p)
        y0, y1 = (0, h) if rows is None else rows
        ev = lib().orc_render(self.h, MODES[mode] if isinstance(mode, str) else int(mode), w, h, cam,
                              1 if use_gi else 0, self.threads, y0, y1, film)
        return film.reshape(h, w, 3), int(ev)

    def scene_desc(self):
        """The parsed scene (objects, planes, subtractions, BVH nodes and order,
        materials, lights, bounds, camera, render settings) as arrays in the
        layout of the product library's scene descriptor."""
        c = np.zeros(8, np.uint32)
        lib().orc_scene_export_counts(self.h, c)
        no, npl, ns, nn, nm, nq, nsl, npt = (int(x) for x in c)
        a = {"objects": np.zeros((no, 13), np.float32), "planes": np.zeros((npl, 13), np.float32),
             "subtractions": np.zeros((ns, 9), np.float32), "bvh_nodes": np.zeros((nn, 9), np.float32),
             "bvh_order": np.zeros(no, np.uint32), "materials": np.zeros((nm, 8), np.float32),
             "quad_lights": np.zeros((nq, 15), np.float32), "sphere_lights": np.zeros((nsl, 7), np.float32),
             "point_lights": np.zeros((npt, 6), np.float32), "bounds": np.zeros(6, np.float32)}
        lib().orc_scene_export(self.h, *(np.ascontiguousarray(a[k]) if a[k].size else np.zeros(16, a[k].dtype)
                                         for k in ("objects", "planes", "subtractions", "bvh_nodes", "bvh_order",
                                                   "materials", "quad_lights", "sphere_lights", "point_lights",
                                                   "bounds")))
        a["camera"] = np.asarray(self.camera, np.float32)
        a["render"] = {"width": self.width, "height": self.height, "bounces": self.bounces,
                       "surfels": self.surfels_requested, "exposure": float(self.exposure)}
        return a

    def tallies(self):
        """(ops, primitive evaluations) of the last transfer's link shadow
        marches, solve's direct light and render, under the roofline's
        counting rule, when tallying is on (set_tally)."""
        out = np.zeros(6, np.uint64)
        lib().orc_tallies(self.h, out)
        return {"visibility": (int(out[0]), int(out[1])), "direct_e": (int(out[2]), int(out[3])),
                "shade": (int(out[4]), int(out[5]))}

    def direct(self, p, n, v, albedo, roughness, metallic):
        """shade::direct (src/shade.rs:50-125) at one point: (radiance, shadow steps)."""
        out = np.zeros(3, np.float32)
        st = np.zeros(1, np.uint32)
        lib().orc_direct(self.h, _f(p), _f(n), _f(v), _f(albedo), roughness, metallic, out, st)
        return out, int(st[0])

    def emitter_hit(self, ro, rd, tmax):
        """shade::emitter_hit (src/shade.rs:168-181): emission of the nearest front-facing emitter, or None."""
        e = np.zeros(3, np.float32)
        return e if lib().orc_emitter_hit(self.h, _f(ro), _f(rd), tmax, e) else None

    def gather(self, p, n, radius):
        p = _f(p, (-1, 3))
        out = np.zeros_like(p)
        lib().orc_gather(self.h, p, _f(n, (-1, 3)), p.shape[0], radius, out)
        return out


def encode(film, exposure=1.0, tonemap=True):
    f = _f(film, (-1, 3))
    out = np.zeros(f.shape[0] * 3, np.uint8)
    lib().orc_encode(f, f.shape[0], exposure, 1 if tonemap else 0, out)
    return out.reshape(np.shape(film))


def prim_dist(kind, params, p):
    q = np.zeros(11, np.float32)
    q[: len(params)] = params
    return float(lib().orc_prim_dist(kind, q, _f(p)))


def prim_bounds(kind, params):
    q = np.zeros(11, np.float32)
    q[: len(params)] = params
    out = np.zeros(6, np.float32)
    ok = lib().orc_prim_bounds(kind, q, out)
    return (out[:3], out[3:]) if ok else None


def quad_ff(verts, p, n):
    return float(lib().orc_quad_ff(_f(verts), _f(p), _f(n)))


def set_tally(on):
    """Tally FP32 operations and primitive evaluations (the roofline's
    counting rule) in the following transfers, solves and renders."""
    lib().orc_set_tally(1 if on else 0)


def fresnel(cos_theta, f0):
    out = np.zeros(3, np.float32)
    lib().orc_fresnel(cos_theta, _f(f0), out)
    return out


def d_ggx(n_dot_h, alpha):
    return float(lib().orc_d_ggx(n_dot_h, alpha))


def v_smith(n_dot_v, n_dot_l, alpha):
    return float(lib().orc_v_smith(n_dot_v, n_dot_l, alpha))


def f0_of(albedo, metallic):
    out = np.zeros(3, np.float32)
    lib().orc_f0_of(_f(albedo), metallic, out)
    return out


def intersect_quad(verts, ro, rd):
    t = np.zeros(1, np.float32)
    return float(t[0]) if lib().orc_intersect_quad(_f(verts), _f(ro), _f(rd), t) else None


def sphere_ff(cos_t, s2):
    return float(lib().orc_sphere_ff(cos_t, s2))


def srgb8(c):
    return int(lib().orc_srgb8(c))


def aces(x):
    out = np.zeros(3, np.float32)
    lib().orc_aces(_f(x), out)
    return out


def radical_inverse(i, base):
    return float(lib().orc_radical_inverse(i, base))


def frame(n):
    t = np.zeros(3, np.float32)
    b = np.zeros(3, np.float32)
    lib().orc_frame(_f(n), t, b)
    return t, b


def rng(seed, n):
    """n floats of the path tracer's PCG32 (src/reference.rs:24-46) seeded with `seed`."""
    out = np.zeros(n, np.float32)
    lib().orc_rng(int(seed), int(n), out)
    return out
