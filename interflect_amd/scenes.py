"""Seeded workload scenes, emitted as `.rad` text.

Every generator returns scene text that both the product loader and the
oracle parse, so the two sides always see the same f32 inputs.  Random draws
use PCG32 (O'Neill's XSH-RR output on a 64-bit LCG, seeded as in the
reference's src/reference.rs:32-46) so a scene is a pure function of its seed.

Scenes that use sphere lights, point lights, `blend` or `subtract_cylinder`
rely on the loader extensions (the reference's grammar has quad lights only).
"""
import numpy as np

MASK64 = (1 << 64) - 1


class Pcg32:
    def __init__(self, seed):
        self.state = (seed * 6364136223846793005 + 1) & MASK64

    def next_u32(self):
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & MASK64
        s = self.state
        x = (((s >> 18) ^ s) >> 27) & 0xFFFFFFFF
        rot = s >> 59
        return ((x >> rot) | (x << ((32 - rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next_u32() >> 8) * (1.0 / 16777216.0)
        return float(np.float32(lo + (hi - lo) * u))

    def below(self, n):
        return min(int(self.uniform() * n), n - 1)


def f(x):
    return repr(float(np.float32(x)))


def v(*xs):
    return "[" + ", ".join(f(x) for x in xs) + "]"


def box(c, size, mat, extra=""):
    return "box { center: %s, size: %s, mat: \"%s\"%s }" % (v(*c), v(*size), mat, extra)


# --------------------------------------------------------------------------- A: cornell-unit
def cornell_unit(width=260, height=260, surfels=20000, bounces=12):
    """Unit room of box walls (thickness 0.04), two r=0.3 spheres, one
    downward quad light of emission 15."""
    L = ['render { width: %d, height: %d, surfels: %d, bounces: %d, exposure: 1.0 }' % (width, height, surfels, bounces),
         'camera { pos: [0.5, 0.5, 2.0], look: [0.5, 0.5, 0.0], up: [0, 1, 0], fov: 40 }',
         'material "grey" { albedo: [0.73, 0.73, 0.73], roughness: 1.0 }',
         'material "red" { albedo: [0.63, 0.065, 0.05], roughness: 1.0 }',
         'material "green" { albedo: [0.14, 0.45, 0.09], roughness: 1.0 }',
         box((0.5, -0.02, 0.5), (1.08, 0.04, 1.08), "grey"),
         box((0.5, 1.02, 0.5), (1.08, 0.04, 1.08), "grey"),
         box((0.5, 0.5, -0.02), (1.08, 1.08, 0.04), "grey"),
         box((-0.02, 0.5, 0.5), (0.04, 1.08, 1.08), "red"),
         box((1.02, 0.5, 0.5), (0.04, 1.08, 1.08), "green"),
         'sphere { center: [0.3, 0.3, 0.3], radius: 0.3, mat: "grey" }',
         'sphere { center: [0.7, 0.3, 0.7], radius: 0.3, mat: "grey" }',
         'light { verts: [[0.4, 0.995, 0.4], [0.6, 0.995, 0.4], [0.6, 0.995, 0.6], [0.4, 0.995, 0.6]], '
         'emit: [15, 15, 15] }']
    return "\n".join(L) + "\n"


# --------------------------------------------------------------------------- B1: literal product
def product_literal(width=640, height=480, surfels=20000, bounces=12):
    """7 SDF objects (sphere; bowl = sphere - sphere - 6 radial cylinders;
    torus R .4 r .12; capsule length .8 r .15; box floor; 2 box walls),
    2 sphere lights, 6 Lambertian materials."""
    L = ['render { width: %d, height: %d, surfels: %d, bounces: %d, exposure: 1.0 }' % (width, height, surfels, bounces),
         'camera { pos: [0, 1.6, 5.5], look: [0, 0.6, 0], up: [0, 1, 0], fov: 40 }',
         'material "floor" { albedo: [0.60, 0.58, 0.55], roughness: 1.0 }',
         'material "wall" { albedo: [0.70, 0.70, 0.68], roughness: 1.0 }',
         'material "grey" { albedo: [0.95, 0.95, 0.97], roughness: 1.0 }',
         'material "amber" { albedo: [0.82, 0.42, 0.10], roughness: 1.0 }',
         'material "teal" { albedo: [0.10, 0.45, 0.48], roughness: 1.0 }',
         'material "bone" { albedo: [0.88, 0.86, 0.80], roughness: 1.0 }',
         box((0, -0.1, 0), (6, 0.2, 6), "floor"),
         box((0, 1.5, -2.0), (6, 3, 0.2), "wall"),
         box((-3.0, 1.5, 0), (0.2, 3, 6), "wall"),
         'sphere { center: [-1.1, 0.5, 0.2], radius: 0.5, mat: "grey" }']
    c = np.array([0.2, 0.45, -0.3])
    subs = ['subtract_sphere: [%s, %s]' % (v(0.2, 0.62, -0.3), f(0.38))]
    for k in range(6):
        th = np.radians(60.0 * k)
        b = c + 0.6 * np.array([np.cos(th), 0.0, np.sin(th)])
        subs.append('subtract_cylinder: [%s, %s, %s]' % (v(*c), v(*b), f(0.06)))
    L.append('sphere { center: %s, radius: %s, mat: "amber", %s }' % (v(*c), f(0.45), ", ".join(subs)))
    L += ['torus { center: [1.2, 0.12, 0.6], major: 0.4, minor: 0.12, mat: "teal" }',
          'capsule { a: [1.3, 0.15, -0.8], b: [1.3, 0.95, -0.8], radius: 0.15, mat: "bone" }',
          'sphere_light { center: [-1.5, 2.8, 1.5], radius: 0.25, emit: [12, 12, 11] }',
          'sphere_light { center: [2.0, 2.2, 1.0], radius: 0.2, emit: [6, 7, 9] }']
    return "\n".join(L) + "\n"


# --------------------------------------------------------------------------- C / D: seeded stress scenes
def stress(seed=0x5EED, prims=256, surfels=200000, width=1920, height=1080, bounces=12):
    """Random sphere / box / torus / capsule mix, sizes U[.05,.3], centres in
    a 10x4x10 room, 25% in a smooth-union blend group (k = .05), 16 albedos
    U[.1,.9]^3, 4 point lights."""
    r = Pcg32(seed)
    L = ['render { width: %d, height: %d, surfels: %d, bounces: %d, exposure: 1.0 }' % (width, height, surfels, bounces),
         'camera { pos: [0, 2, 4.8], look: [0, 1, 0], up: [0, 1, 0], fov: 60 }']
    for m in range(16):
        L.append('material "m%d" { albedo: %s, roughness: 1.0 }' % (m, v(r.uniform(.1, .9), r.uniform(.1, .9),
                                                                         r.uniform(.1, .9))))
    L.append('material "room" { albedo: [0.7, 0.7, 0.7], roughness: 1.0 }')
    L += [box((0, -0.1, 0), (10.4, 0.2, 10.4), "room"), box((0, 4.1, 0), (10.4, 0.2, 10.4), "room"),
          box((0, 2, -5.1), (10.4, 4.4, 0.2), "room"), box((0, 2, 5.1), (10.4, 4.4, 0.2), "room"),
          box((-5.1, 2, 0), (0.2, 4.4, 10.4), "room"), box((5.1, 2, 0), (0.2, 4.4, 10.4), "room")]
    for _ in range(prims):
        kind = r.below(4)
        s = r.uniform(0.05, 0.3)
        cx, cy, cz = r.uniform(-5, 5), r.uniform(0, 4), r.uniform(-5, 5)
        blend = r.uniform() < 0.25
        mat = "m%d" % r.below(16)
        ext = ", blend: 0.05" if blend else ""
        if kind == 0:
            L.append('sphere { center: %s, radius: %s, mat: "%s"%s }' % (v(cx, cy, cz), f(s), mat, ext))
        elif kind == 1:
            hy, hz = s * r.uniform(.5, 1), s * r.uniform(.5, 1)
            L.append(box((cx, cy, cz), (2 * s, 2 * hy, 2 * hz), mat, ext))
        elif kind == 2:
            L.append('torus { center: %s, major: %s, minor: %s, mat: "%s"%s }' % (v(cx, cy, cz), f(s), f(0.3 * s),
                                                                                  mat, ext))
        else:
            L.append('capsule { a: %s, b: %s, radius: %s, mat: "%s"%s }' % (v(cx, cy - s, cz), v(cx, cy + s, cz),
                                                                            f(0.4 * s), mat, ext))
    for sx in (-2.5, 2.5):
        for sz in (-2.5, 2.5):
            L.append('point_light { pos: %s, emit: [6, 6, 6] }' % v(sx, 3.8, sz))
    return "\n".join(L) + "\n"


def batched_scene(j, surfel_div=1, side_div=1):
    """Scene j of the 64-scene batch: 8..256 primitives, 5k..100k surfels,
    square film 256 + 64 * U{0..12}; otherwise as `stress`.  The divisors
    shrink surfels and film for quick checks (same geometry)."""
    r = Pcg32(1000 + j)
    prims = 8 + r.below(249)
    surfels = 5000 + r.below(95001)
    side = 256 + 64 * r.below(13)
    return stress(seed=1000 + j, prims=prims, surfels=max(1, surfels // surfel_div), width=max(8, side // side_div),
                  height=max(8, side // side_div))


# --------------------------------------------------------------------------- E: differentiable cornell
def diff_cornell(seed, surfels=50000, width=512, height=512, bounces=12):
    """Cornell-type 4.4 room: 8 materials with albedo U[.1,.9]^3 (5 walls, 2
    occupants, 1 ledge), 3 ceiling quad lights of scalar emission U[5,20]."""
    r = Pcg32(seed)
    L = ['render { width: %d, height: %d, surfels: %d, bounces: %d, exposure: 1.0 }' % (width, height, surfels, bounces),
         'camera { pos: [0, 2.0, 7.6], look: [0, 2.0, 0], up: [0, 1, 0], fov: 42 }']
    for m in range(8):
        L.append('material "m%d" { albedo: %s, roughness: 1.0 }' % (m, v(r.uniform(.1, .9), r.uniform(.1, .9),
                                                                         r.uniform(.1, .9))))
    L += [box((0, -0.1, 0), (4.4, 0.2, 4.4), "m0"), box((0, 4.1, 0), (4.4, 0.2, 4.4), "m1"),
          box((0, 2.0, -2.1), (4.4, 4.4, 0.2), "m2"), box((-2.1, 2.0, 0), (0.2, 4.4, 4.4), "m3"),
          box((2.1, 2.0, 0), (0.2, 4.4, 4.4), "m4"),
          'sphere { center: [-0.85, 0.85, 0.45], radius: 0.85, mat: "m5" }',
          box((0.82, 1.05, -0.55), (1.25, 2.1, 1.25), "m6", ", round: 0.09"),
          box((0.0, 2.6, -1.85), (1.6, 0.1, 0.3), "m7")]
    for cx in (-1.2, 0.0, 1.2):
        e = r.uniform(5, 20)
        L.append('light { verts: [%s, %s, %s, %s], emit: %s }' % (
            v(cx - 0.35, 3.94, -0.35), v(cx + 0.35, 3.94, -0.35), v(cx + 0.35, 3.94, 0.35), v(cx - 0.35, 3.94, 0.35),
            v(e, e, e)))
    return "\n".join(L) + "\n"
