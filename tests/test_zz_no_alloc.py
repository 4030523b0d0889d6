"""ifx_solve_noalloc: the allocation-free variant of ifx_solve.

Called through the C ABI with caller-owned device arrays (allocated here with
the HIP runtime the library itself is linked against), it must equal
ifx_solve bit for bit (direct, irradiance, radiosity, iterations, residual,
evaluation count) on two calls in a row, the second with fewer surfels in the
same arrays, and reject a problem whose scratch exceeds the static array.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import interflect_amd as ia
from conftest import scene_text

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hip():
    ia.lib()  # loads the runtime the library links
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevArrays:
    """Three device arrays of n float4 records, filled with 0xff bytes."""

    def __init__(self, h, n):
        self.h, self.n = h, n
        self.p = []
        for _ in range(3):
            q = C.c_void_p()
            assert h.hipMalloc(C.byref(q), 16 * n) == 0
            self.p.append(q.value)

    def poison(self):
        for q in self.p:
            assert self.h.hipMemset(C.c_void_p(q), 0xFF, 16 * self.n) == 0
        assert self.h.hipDeviceSynchronize() == 0

    def read(self, n):
        out = []
        for q in self.p:
            a = np.zeros((n, 4), np.float32)
            assert self.h.hipMemcpy(a.ctypes.data, C.c_void_p(q), 16 * n, 2) == 0  # device to host
            out.append(a[:, :3].copy())
        return out

    def free(self):
        for q in self.p:
            self.h.hipFree(C.c_void_p(q))
        self.p = []


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def check_case(ctx, arrs, name):
    n = ctx.num_surfels()
    ref = ctx.solve()
    d0, i0, r0 = ctx.radiosity()
    arrs.poison()
    got = ctx.solve_noalloc(*arrs.p)
    d1, i1, r1 = arrs.read(n)
    say(name, n, "surfels:", ref["iterations"], "bounces,", ref["evals"], "evals")
    assert got["iterations"] == ref["iterations"]
    assert got["evals"] == ref["evals"]
    assert np.float32(got["residual"]).view(np.uint32) == np.float32(ref["residual"]).view(np.uint32)
    for a, b in ((d1, d0), (i1, i0), (r1, r0)):
        assert np.array_equal(bits(a), bits(b))
    # the context's own radiosity is left as ifx_solve made it
    d2, i2, r2 = ctx.radiosity()
    assert np.array_equal(bits(r2), bits(r0))


def test_noalloc_solve_equals_solve_twice_then_smaller():
    h = hip()
    arrs = DevArrays(h, 4000)
    try:
        ctx = ia.Context(0)
        ctx.load_scene(scene_text("product.rad"))
        ctx.generate_surfels(4000)
        ctx.transfer(8)
        check_case(ctx, arrs, "product.rad")
        # second call: another scene, fewer surfels, the same arrays
        ctx2 = ia.Context(0)
        ctx2.load_scene(scene_text("closed_box.rad"))
        ctx2.generate_surfels(1500)
        assert ctx2.num_surfels() < ctx.num_surfels()
        ctx2.transfer(8)
        check_case(ctx2, arrs, "closed_box.rad")
        # and the first context again after the smaller call
        check_case(ctx, arrs, "product.rad again")
    finally:
        arrs.free()


def test_noalloc_solve_rejects_scratch_beyond_its_array():
    # 400 point lights x 5000 surfels x 20 B of terms exceeds the 32 MiB array
    lights = "\n".join('point_light { pos: [%.2f, 0.8, 0.0], emit: [1, 1, 1] }' % (-0.4 + 0.002 * k)
                       for k in range(400))
    ctx = ia.Context(0)
    ctx.load_scene(scene_text("closed_box.rad") + "\n" + lights + "\n")
    n = ctx.generate_surfels(5000)
    assert n * 400 * 20 > 32 << 20
    ctx.transfer(8)
    with pytest.raises(ia.IfxError) as e:
        ctx.solve_noalloc(0, 0, 0)
    assert e.value.status == -6 and "scratch" in str(e.value)
