"""Pins the CPU oracle (oracle/) before anything is checked against it.

Three kinds of pin: the reference's own published stdout for
scenes/product.rad, the reference's unit tests (restated as known answers,
each citing the Rust test it follows), and closed forms that a plausible slip
(wrong sign, dropped term, transposed operand) would break.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import ROOT, scene_text
from oracle import pyoracle as po

GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(po.lib_path()):
        po.build()


# --------------------------------------------------------------------------- published product.rad
@pytest.fixture(scope="module")
def product():
    sc = po.Scene(scene_text("product.rad"))
    n = sc.generate()
    tr = sc.transfer(8)
    so = sc.solve()
    film, ev = sc.render("beauty")
    return sc, n, tr, so, film, ev


def test_product_matches_published_stdout(product):
    """README.md:125-132 / TESTING.md:28-37: every count of the published run."""
    g = json.load(open(os.path.join(GOLD, "product_published.json")))
    sc, n, tr, so, film, ev = product
    assert (sc.n_objects, sc.n_lights, sc.n_materials, sc.n_bvh_nodes) == (
        g["objects"], g["lights"], g["materials"], g["bvh_nodes"])
    assert n == g["surfels_placed"]
    assert tr["clusters"] == g["clusters"]
    assert tr["links"] == g["links"]
    assert "%.3f" % tr["mean_row_sum"] == g["row_sum_3dp"]
    assert "%.1f" % (tr["evals"] / 1e6) == g["transfer_evals_1dp_M"]
    assert so["iterations"] == g["bounces"]
    assert "%.2e" % so["residual"] == g["residual_3sig"].replace("e-5", "e-05")
    assert "%.1f" % (so["evals"] / 1e6) == g["solve_evals_1dp_M"]
    assert "%.1f" % (ev / 1e6) == g["shade_evals_1dp_M"]
    assert "%.0f" % (ev / (640 * 480)) == g["evals_per_px_0dp"]


def test_product_thread_invariance(product):
    """ci.yml:81-93: output identical for any thread count."""
    sc1 = po.Scene(scene_text("product.rad"), threads=1)
    sc1.generate()
    t1 = sc1.transfer(8)
    s1, i1, w1, _, _ = sc1.transfer_arrays()
    sc, *_ = product
    s8, i8, w8, _, _ = sc.transfer_arrays()
    assert t1["links"] == product[2]["links"]
    assert np.array_equal(s1, s8) and np.array_equal(i1, i8) and np.array_equal(w1.view(np.uint32), w8.view(np.uint32))


# --------------------------------------------------------------------------- sdf.rs tests
def test_cylinder_distances_are_correct():
    """src/sdf.rs:182-217"""
    q = [0, 0, 0, 0, 2, 0, 1.0]
    for p, want in [((3, 1, 0), 2), ((0, 5, 0), 3), ((0, -1, 0), 1), ((0, 1, 0), -1), ((1, 1, 0), 0), ((0, 2, 0), 0)]:
        assert abs(po.prim_dist(5, q, p) - want) < 1e-4
    assert abs(po.prim_dist(5, q, (2, 3, 0)) - math.sqrt(2)) < 1e-4


def test_cone_distances_are_correct():
    """src/sdf.rs:219-249"""
    q = [0, 0, 0, 0, 2, 0, 1.0, 0.0]
    assert po.prim_dist(6, q, (0, 1, 0)) < 0
    assert po.prim_dist(6, q, (0, 3, 0)) > 0.9
    assert po.prim_dist(6, q, (5, 1, 0)) > 3.5
    assert abs(po.prim_dist(6, q, (1, 0, 0))) < 1e-3
    assert po.prim_dist(6, q, (0.9, 0.1, 0)) < 0
    assert po.prim_dist(6, q, (0.9, 1.8, 0)) > 0


NEW_PRIMS = [(5, [-0.4, 0.1, 0.2, 0.5, 1.3, -0.3, 0.35]), (6, [0.2, 0.0, 0.1, -0.1, 1.6, 0.4, 0.5, 0.15])]


def test_new_primitives_never_overstate_distance():
    """src/sdf.rs:254-296 (Lipschitz safety of cylinder and cone)."""
    f = np.arange(4000, dtype=np.float32)
    for kind, q in NEW_PRIMS:
        for i in range(0, 4000, 7):
            fi = f[i]
            p = np.array([np.sin(fi * np.float32(0.71)) * 3, np.cos(fi * np.float32(1.13)) * 3,
                          np.sin(fi * np.float32(0.37)) * 3], np.float32)
            d = po.prim_dist(kind, q, p)
            assert math.isfinite(d)
            if d <= 0:
                continue
            dr = np.array([np.cos(fi * np.float32(2.1)), np.sin(fi * np.float32(0.9)), np.cos(fi * np.float32(1.7))],
                          np.float32)
            dr /= np.linalg.norm(dr)
            assert po.prim_dist(kind, q, p + dr * d) > -1e-3


def test_new_primitive_bounds_enclose_the_surface():
    """src/sdf.rs:299-335"""
    for kind, q in NEW_PRIMS:
        lo, hi = po.prim_bounds(kind, q)
        for i in range(0, 20000, 13):
            fi = np.float32(i)
            p = np.array([np.sin(fi * np.float32(0.71)) * 2.5, np.cos(fi * np.float32(1.13)) * 2.5,
                          np.sin(fi * np.float32(0.37)) * 2.5], np.float32)
            if po.prim_dist(kind, q, p) < 0:
                assert np.all(p >= lo - 1e-4) and np.all(p <= hi + 1e-4)


def test_primitive_closed_forms():
    """Exact distances: sphere |p-c|-r, plane p.n-h, box face, torus ring."""
    assert po.prim_dist(0, [1, 2, 3, 0, 0, 0, 0.5], (1, 2, 5)) == pytest.approx(1.5, abs=1e-6)
    assert po.prim_dist(2, [0, 0, 0, 0, 0, 0, 0.25, 0, 0, 1, 0], (3, 2, 1)) == pytest.approx(1.75, abs=1e-6)
    # rounded box half-size (1,2,3), round 0 -> distance to +x face
    assert po.prim_dist(1, [0, 0, 0, 1, 2, 3, 0.0], (4, 0, 0)) == pytest.approx(3.0, abs=1e-6)
    assert po.prim_dist(1, [0, 0, 0, 1, 2, 3, 0.0], (0.5, 0, 0)) == pytest.approx(-0.5, abs=1e-6)
    # torus major 1 minor 0.25: point on the ring axis circle at height 1
    assert po.prim_dist(4, [0, 0, 0, 0, 0, 0, 1.0, 0.25], (1, 1, 0)) == pytest.approx(0.75, abs=1e-6)
    # capsule from (0,0,0) to (0,2,0) r 0.5: beside the shaft, and past the cap
    assert po.prim_dist(3, [0, 0, 0, 0, 2, 0, 0.5], (2, 1, 0)) == pytest.approx(1.5, abs=1e-6)
    assert po.prim_dist(3, [0, 0, 0, 0, 2, 0, 0.5], (0, 4, 0)) == pytest.approx(1.5, abs=1e-6)


# --------------------------------------------------------------------------- bvh.rs tests
def scattered_scene(n):
    """src/bvh.rs:316-335 scattered spheres, as a .rad scene."""
    lines = ['material "m0" { albedo: [0.5,0.5,0.5] }', 'material "m1" { albedo: [0.5,0.5,0.5] }',
             'material "m2" { albedo: [0.5,0.5,0.5] }', 'material "m3" { albedo: [0.5,0.5,0.5] }']
    for i in range(n):
        f = np.float32(i)
        c = (np.sin(f * np.float32(1.7)) * np.float32(4.0), np.cos(f * np.float32(2.3)) * np.float32(3.0),
             np.sin(f * np.float32(0.9)) * np.float32(4.0))
        r = np.float32(0.2) + np.float32(i % 5) * np.float32(0.05)
        lines.append('sphere { center: [%r, %r, %r], radius: %r, mat: "m%d" }' % (
            float(c[0]), float(c[1]), float(c[2]), float(r), i % 4))
    lines.append("light { verts: [[-1,8,-1],[1,8,-1],[1,8,1],[-1,8,1]], emit: [1,1,1] }")
    return "\n".join(lines)


def test_child_indices_are_consistent_with_layout():
    """src/bvh.rs:342-377"""
    sc = po.Scene(scattered_scene(200))
    lohi, scr, _ = sc.bvh()
    for i in range(len(scr)):
        if scr[i, 1] > 0:
            continue
        l, r = i + 1, int(scr[i, 2])
        assert l < len(scr) and r < len(scr) and r > l
        for c in (l, r):
            assert np.all(lohi[c, :3] >= lohi[i, :3] - 1e-4) and np.all(lohi[c, 3:] <= lohi[i, 3:] + 1e-4)


def test_every_primitive_lands_in_exactly_one_leaf():
    """src/bvh.rs:381-399"""
    sc = po.Scene(scattered_scene(200))
    _, scr, order = sc.bvh()
    seen = np.zeros(200, int)
    for s, c, _ in scr:
        if c > 0:
            seen[order[s:s + c]] += 1
    assert np.all(seen == 1)


def test_pruning_agrees_with_linear_scan():
    """src/bvh.rs:403-428"""
    sc = po.Scene(scattered_scene(300))
    f = np.arange(3000, dtype=np.float32) * np.float32(0.037)
    pts = np.stack([np.sin(f) * 6, np.cos(f * np.float32(1.3)) * 5, np.sin(f * np.float32(0.7)) * 6], 1)
    fast, _ = sc.dist_mat(pts)
    sc.set_use_bvh(False)
    slow, _ = sc.dist_mat(pts)
    assert np.max(np.abs(fast - slow)) < 1e-5


# --------------------------------------------------------------------------- trace.rs tests
def _shadow_to_light(sc, origin):
    lo = np.array([[-0.7, 3.94, -0.7], [0.7, 3.94, -0.7], [0.7, 3.94, 0.7], [-0.7, 3.94, 0.7]], np.float32)
    c = lo.mean(0).astype(np.float32)
    to = c - np.array([0, 0, 0], np.float32)
    dist = np.float32(np.linalg.norm(to))
    radius = np.float32(math.sqrt(1.4 * 1.4 / math.pi))
    k = np.clip(dist / radius, 1.5, 96.0)
    end = max(dist - radius * np.float32(1.25), np.float32(0.05))
    return sc.soft_shadow(np.array(origin, np.float32), to / dist, 0.015, end, k)[0][0]


def test_unoccluded_floor_under_flush_light_is_not_self_shadowed():
    """src/trace.rs:223-243"""
    sc = po.Scene(scene_text("flush_ceiling_light.rad"))
    assert _shadow_to_light(sc, [0, 3e-3, 0]) > 0.9


def test_real_occluder_still_casts_shadow():
    """src/trace.rs:247-262"""
    src = scene_text("flush_ceiling_light.rad") + '\nbox { center: [0,2,0], size: [1.5,0.2,1.5], mat: "w" }\n'
    sc = po.Scene(src)
    assert _shadow_to_light(sc, [0, 3e-3, 0]) < 0.1


def test_radical_inverse_is_exact():
    """Halton radical inverse (src/trace.rs:188-199) on exactly representable values."""
    assert po.radical_inverse(1, 2) == 0.5
    assert po.radical_inverse(2, 2) == 0.25
    assert po.radical_inverse(3, 2) == 0.75
    assert po.radical_inverse(6, 2) == 0.375
    assert po.radical_inverse(1, 3) == np.float32(1.0) / np.float32(3.0)
    inv3 = np.float32(1.0) / np.float32(3.0)
    assert po.radical_inverse(4, 3) == np.float32(inv3 + np.float32(inv3 * inv3))  # 4 = 11 in base 3
    assert po.radical_inverse(0, 5) == 0.0


def test_closure_open_plane_and_sealed_box():
    """Closed forms of hemisphere_closure: a lone floor sees all sky; a sealed box none."""
    open_src = ('material "w" { albedo: [0.5,0.5,0.5] }\nbox { center: [0,-0.1,0], size: [8,0.2,8], mat: "w" }\n'
                "light { verts: [[-1,4,-1],[1,4,-1],[1,4,1],[-1,4,1]], emit: [1,1,1] }")
    sc = po.Scene(open_src)
    sky, steps = sc.closure([[0, 0, 0], [1, 0, 1]], [[0, 1, 0], [0, 1, 0]], 20.0)
    assert np.all(sky == 1.0) and np.all(steps > 0)
    box = po.Scene(scene_text("closed_box.rad"))
    sky, _ = box.closure([[0, 2, 0], [1, 0.0, 0.5]], [[0, 1, 0], [0, 1, 0]], 40.0)
    assert np.all(sky == 0.0)


def test_frame_is_orthonormal():
    """ltc::frame (src/ltc.rs:88-96): t, b, n orthonormal for sampled normals."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = rng.normal(size=3).astype(np.float32)
        n /= np.linalg.norm(n)
        t, b = po.frame(n)
        m = np.stack([t, b, n])
        assert np.allclose(m @ m.T, np.eye(3), atol=1e-5)


# --------------------------------------------------------------------------- ltc.rs tests
def ceiling(half, h):
    return np.array([[-half, h, -half], [half, h, -half], [half, h, half], [-half, h, half]], np.float32).ravel()


def test_dsl_winding_yields_positive_form_factor():
    """src/ltc.rs:155-163"""
    assert po.quad_ff(ceiling(1.0, 2.0), [0, 0, 0], [0, 1, 0]) > 0.05


def test_small_distant_light_matches_analytic_limit():
    """src/ltc.rs:168-181"""
    ff = po.quad_ff(ceiling(0.05, 5.0), [0, 0, 0], [0, 1, 0])
    expect = 0.1 * 0.1 / (math.pi * 25.0)
    assert abs(ff - expect) / expect < 0.02


def test_light_below_horizon_contributes_nothing():
    """src/ltc.rs:185-189"""
    assert po.quad_ff(ceiling(1.0, -2.0), [0, 0, 0], [0, 1, 0]) < 1e-6


def test_square_light_closed_form():
    """Point receiver under the centre of a parallel square of half-side a at
    height h: four corner terms of the differential-area-to-rectangle form
    factor (Howell's catalogue, configuration B-3)."""
    for a, h in [(1.0, 2.0), (0.5, 0.3), (2.0, 1.0)]:
        x = a / h
        # differential-area-to-rectangle (point receiver): F = (1/2pi) sum of 4 corner terms
        def point_corner(X, Y):
            return (1 / (2 * math.pi)) * (X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
                                          + Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y)))
        want = 4 * point_corner(x, x)
        got = po.quad_ff(ceiling(a, h), [0, 0, 0], [0, 1, 0])
        assert abs(got - want) < 2e-5, (a, h, got, want)


def test_sphere_form_factor_closed_forms():
    """EXT sphere light: fully above the horizon F = cos(theta) sin^2(sigma)."""
    for c, s2 in [(1.0, 0.04), (0.8, 0.1), (0.5, 0.01)]:
        assert po.sphere_ff(c, s2) == pytest.approx(c * s2, rel=1e-6)
    assert po.sphere_ff(-0.9, 0.01) == 0.0
    # straddling the horizon: continuous in cos(theta)
    s2 = 0.09
    c0 = math.sqrt(s2)
    assert po.sphere_ff(c0 + 1e-4, s2) == pytest.approx(po.sphere_ff(c0 - 1e-4, s2), abs=2e-4)
    assert 0 < po.sphere_ff(0.0, s2) < s2


# --------------------------------------------------------------------------- formfactor.rs / solve.rs tests
@pytest.fixture(scope="module")
def closed_box():
    sc = po.Scene(scene_text("closed_box.rad"))
    sc.generate(3000)
    tr = sc.transfer(8)
    return sc, tr


def test_closed_box_conserves_energy(closed_box):
    """src/formfactor.rs:313-328"""
    _, tr = closed_box
    assert 0.7 < tr["mean_row_sum"] <= 1.0001


def test_transfer_matrix_is_sparse_and_well_formed(closed_box):
    """src/formfactor.rs:332-360"""
    sc, tr = closed_box
    starts, idx, w, _, _ = sc.transfer_arrays()
    N = sc.surfels()[0].shape[0]
    assert len(starts) == N + 1 and starts[-1] == len(idx) == len(w)
    assert np.all(np.diff(starts.astype(np.int64)) >= 0)
    assert np.all(idx < tr["clusters"])
    assert tr["links"] * 3 < N * tr["clusters"]


def test_solve_converges_and_no_interior_surfel_is_unlit():
    """src/solve.rs:145-196"""
    sc = po.Scene(scene_text("closed_box.rad"))
    sc.generate(3000)
    sc.transfer(8)
    st = sc.solve(64)
    assert st["residual"] < 1e-3 and st["iterations"] < 64
    _, irr, _ = sc.radiosity()
    m = irr.max(1)
    interior = m[m > 0]
    assert len(interior) > len(m) // 4
    assert interior.min() > interior.max() * 1e-3


def fp64_operators(sc):
    """W (surfel x cluster CSR) and G (area-weighted cluster aggregation) in FP64."""
    import scipy.sparse as sp

    starts, idx, w, _, _ = sc.transfer_arrays()
    cp, cn, ca, cr, ms, mem = sc.clusters()
    _, _, alb, area = sc.surfels()
    N, C = len(area), len(ca)
    W = sp.csr_matrix((w.astype(np.float64), idx, starts.astype(np.int64)), shape=(N, C))
    rows = np.repeat(np.arange(C), np.diff(ms.astype(np.int64)))
    G = sp.csr_matrix((area[mem].astype(np.float64) / ca[rows].astype(np.float64), (rows, mem)), shape=(C, N))
    return (W @ G).tocsr(), alb.astype(np.float64)


def test_solve_matches_fp64_jacobi_and_fixed_point():
    """solve::solve (src/solve.rs:47-108) is Jacobi on B = E + rho (W G B): eight
    unconverged f32 bounces equal eight FP64 bounces from B0 = E, and the
    converged result is within the 1e-4 stopping rule's truncation of the FP64
    fixed point."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    sc = po.Scene(scene_text("closed_box.rad"))
    sc.generate(3000)
    sc.transfer(8)
    st = sc.solve(8)
    assert st["iterations"] == 8
    direct, irr, rad = sc.radiosity()
    WG, alb = fp64_operators(sc)
    E = direct.astype(np.float64)
    B = E.copy()
    for _ in range(8):
        B = E + alb * (WG @ B)
    assert np.linalg.norm(rad - B) / np.linalg.norm(B) < 1e-5

    st = sc.solve(400)
    direct, irr, rad = sc.radiosity()
    Bs = np.zeros_like(E)
    for ch in range(3):
        A = sp.eye(len(E)) - sp.diags(alb[:, ch]) @ WG
        Bs[:, ch] = spl.spsolve(A.tocsc(), E[:, ch])
    assert st["residual"] < 1e-4
    # geometric tail of a contraction with rate <= 0.7 (max albedo x max row sum)
    assert np.max(np.abs(rad - Bs)) <= st["residual"] * 0.7 / 0.3 * 1.5


def test_gather_of_constant_field_is_constant():
    """surfel::gather (src/surfel.rs:262-293) is a normalised weighted mean."""
    sc = po.Scene(scene_text("closed_box.rad"))
    sc.generate(3000)
    sc.transfer(8)
    p, n, a, _ = sc.surfels()
    c = np.array([0.3, 0.5, 0.7], np.float32)
    sc.set_radiosity(np.zeros_like(p), np.tile(c, (len(p), 1)), np.zeros_like(p))
    g = sc.gather(p[:200], n[:200], 0.5)
    ok = g.max(1) > 0
    assert ok.mean() > 0.9
    assert np.allclose(g[ok], c, rtol=1e-5)


# --------------------------------------------------------------------------- film.rs tests
def test_srgb_transfer_hits_known_anchors():
    """src/film.rs:74-88"""
    assert po.srgb8(0.0) == 0 and po.srgb8(1.0) == 255
    assert abs(po.srgb8(0.2140) - 128) <= 1
    assert po.srgb8(-5.0) == 0 and po.srgb8(9.0) == 255


def test_srgb_is_monotonic():
    """src/film.rs:90-98"""
    v = [po.srgb8(i / 1000.0) for i in range(1001)]
    assert all(b >= a for a, b in zip(v, v[1:]))


def test_aces_maps_into_unit_range():
    """src/film.rs:102-114"""
    for e in [0.0, 0.1, 0.5, 1.0, 4.0, 100.0, 1e6]:
        c = po.aces([e, e, e])
        assert np.all(c >= 0) and np.all(c <= 1)
    assert po.aces([1, 1, 1])[0] > po.aces([0.5, 0.5, 0.5])[0]
    # Narkowicz fit value at 1.0: (2.51+0.03)/(2.43+0.59+0.14)
    assert po.aces([1, 1, 1])[0] == pytest.approx(2.54 / 3.16, rel=1e-6)


def test_encode_produces_three_bytes_per_pixel():
    """src/film.rs:116-125"""
    f = np.zeros((3, 4, 3), np.float32)
    f[0, 0] = [1, 0, 0]
    b = po.encode(f, 1.0, True)
    assert b.shape == (3, 4, 3) and b[0, 0, 0] > b[0, 0, 1]


# --------------------------------------------------------------------------- scene.rs tests
def test_parses_a_minimal_scene():
    """src/scene.rs:613-621"""
    sc = po.Scene(scene_text("minimal.rad"))
    assert sc.n_objects == 1 and sc.n_lights == 1 and sc.n_materials == 1


def test_planes_are_excluded_from_the_bvh():
    """src/scene.rs:631-645"""
    sc = po.Scene(scene_text("minimal.rad") + '\nplane { normal: [0,1,0], height: 0, mat: "w" }\n')
    assert sc.n_planes == 1 and sc.n_objects == 1


MALFORMED = [("", "no `light` block"), ("wibble { }", "unknown block"), ("camera { nonsense: 5 }", "unknown camera key"),
             ("camera { fov: abc }", "expected number"), ("camera { pos: [1, 2 }", "expected number"),
             ('light { emit: [1,1,1] } sphere { mat: "nope" }', "unknown material")]
LIMITS = [("render { surfels: 999999999 }", "surfels"), ("render { width: 200000, height: 200000 }", "exceeds"),
          ("render { width: 0, height: 10 }", "at least 1"), ("render { bounces: 2000000000 }", "bounces"),
          ("render { exposure: 0 }", "exposure")]
BASE_LIGHT = "light { verts: [[-1,4,-1],[1,4,-1],[1,4,1],[-1,4,1]], emit: [9,9,9] }"


@pytest.mark.parametrize("src,msg", MALFORMED)
def test_malformed_input_errors_instead_of_panicking(src, msg):
    """src/scene.rs:649-677"""
    with pytest.raises(po.OracleError, match=msg):
        po.Scene(src)


@pytest.mark.parametrize("cfg,msg", LIMITS)
def test_absurd_render_settings_are_rejected(cfg, msg):
    """src/scene.rs:685-713"""
    with pytest.raises(po.OracleError, match=msg):
        po.Scene(cfg + "\n" + BASE_LIGHT + "\n")


def test_realistic_settings_are_accepted():
    """src/scene.rs:717-733"""
    for cfg in ["render { width: 3840, height: 2160, surfels: 200000, bounces: 64 }",
                "render { width: 1920, height: 1080 }", "render { surfels: 1 }"]:
        po.Scene(cfg + "\n" + BASE_LIGHT + "\n")


def test_bounds_enclose_geometry_without_excessive_padding():
    """src/scene.rs:737-751"""
    sc = po.Scene(scene_text("minimal.rad"))
    assert np.all(sc.lo <= [0.5, 1.5, 2.5]) and np.all(sc.hi >= [1.5, 4.0, 3.5])
    geom = np.array([1.5, 4.0, 3.5]) - np.array([0.5, 1.5, -1.0])
    assert np.linalg.norm(sc.hi - sc.lo) < np.linalg.norm(geom) * 2


def test_every_scene_fixture_renders_small():
    """ci.yml:65-76: every shipped scene renders at 200x150 (here 40x30, normals+beauty)."""
    for name in sorted(os.listdir(os.path.join(GOLD, "scenes"))):
        if name in ("TEMPLATE.rad",):
            continue
        sc = po.Scene(scene_text(name))
        sc.generate(min(sc.surfels_requested, 1500))
        sc.transfer(4)
        sc.solve(4)
        film, ev = sc.render("beauty", 40, 30)
        assert np.all(np.isfinite(film)) and ev > 0, name


# --------------------------------------------------------------------------- EXT closed forms
FLOOR = ('material "w" { albedo: [0.5, 0.5, 0.5], roughness: 1.0 }\n'
         'box { center: [0, -0.1, 0], size: [40, 0.2, 40], mat: "w" }\n')


def test_point_light_irradiance_closed_form():
    """EXT point light over an open floor: E = albedo * I * cos / (pi d^2), unshadowed."""
    sc = po.Scene(FLOOR + "point_light { pos: [0.3, 2.0, -0.4], emit: [5, 6, 7] }\n")
    p = np.array([[0, 0, 0], [1.0, 0, 0.5], [-2, 0, 1]], np.float32)
    n = np.tile([0, 1, 0], (3, 1)).astype(np.float32)
    e, steps = sc.direct_diffuse(p, n, np.full((3, 3), 0.5, np.float32))
    to = np.array([0.3, 2.0, -0.4]) - p
    d2 = (to ** 2).sum(1)
    cos = to[:, 1] / np.sqrt(d2)
    want = 0.5 * np.array([5, 6, 7])[None, :] * (cos / (np.pi * d2))[:, None]
    assert np.allclose(e, want, rtol=1e-5) and np.all(steps > 0)


def test_sphere_light_irradiance_closed_form():
    """EXT sphere light wholly above the horizon: E = albedo * L * cos * (R/d)^2."""
    sc = po.Scene(FLOOR + "sphere_light { center: [0.0, 3.0, 0.0], radius: 0.4, emit: [2, 3, 4] }\n")
    p = np.array([[0, 0, 0], [1.5, 0, 0]], np.float32)
    n = np.tile([0, 1, 0], (2, 1)).astype(np.float32)
    e, _ = sc.direct_diffuse(p, n, np.full((2, 3), 0.5, np.float32))
    to = np.array([0, 3.0, 0]) - p
    d2 = (to ** 2).sum(1)
    cos = to[:, 1] / np.sqrt(d2)
    want = 0.5 * np.array([2, 3, 4])[None, :] * (cos * 0.16 / d2)[:, None]
    assert np.allclose(e, want, rtol=1e-5)


BLEND2 = ('material "w" { albedo: [0.5, 0.5, 0.5] }\n'
          'sphere { center: [0, 0, 0], radius: 0.5, mat: "w", blend: 0.2 }\n'
          'sphere { center: [1.1, 0, 0], radius: 0.5, mat: "w", blend: 0.2 }\n'
          'light { verts: [[-1,4,-1],[1,4,-1],[1,4,1],[-1,4,1]], emit: [1,1,1] }\n')


def test_blend_group_smooth_union():
    """EXT blend: d = smin(d1, d2, k) = min - h^2 k / 4 with h = max(k - |d1-d2|, 0) / k;
    equal to the hard minimum once |d1 - d2| >= k, never above it, and the
    BVH traversal with its k-widened prune agrees with the linear scan."""
    sc = po.Scene(BLEND2)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-2, 3, (4000, 3)).astype(np.float32)
    d, _ = sc.dist_mat(pts)
    d1 = np.linalg.norm(pts - [0, 0, 0], axis=1) - 0.5
    d2 = np.linalg.norm(pts - [1.1, 0, 0], axis=1) - 0.5
    h = np.maximum(0.2 - np.abs(d1 - d2), 0) / 0.2
    want = np.minimum(d1, d2) - h * h * 0.2 * 0.25
    assert np.allclose(d, want, atol=2e-6)
    assert np.all(d <= np.minimum(d1, d2) + 1e-6)
    sc.set_use_bvh(False)
    dl, _ = sc.dist_mat(pts)
    assert np.array_equal(d.view(np.uint32), dl.view(np.uint32))


def test_blend_field_is_lipschitz():
    """The smooth union never overstates distance: stepping d along any
    direction lands no deeper than -1e-3 (cf. src/sdf.rs:254-296)."""
    sc = po.Scene(BLEND2)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-2, 3, (3000, 3)).astype(np.float32)
    d, _ = sc.dist_mat(pts)
    dirs = rng.normal(size=(3000, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ok = d > 0
    d2, _ = sc.dist_mat(pts[ok] + dirs[ok] * d[ok, None])
    assert np.all(d2 > -1e-3)


def test_cylinder_subtraction():
    """EXT subtract_cylinder: max(d, -cyl) carves a hole through a sphere."""
    src = ('material "w" { albedo: [0.5, 0.5, 0.5] }\n'
           'sphere { center: [0, 0, 0], radius: 1.0, mat: "w", subtract_cylinder: [[0, -2, 0], [0, 2, 0], 0.25] }\n'
           'light { verts: [[-1,4,-1],[1,4,-1],[1,4,1],[-1,4,1]], emit: [1,1,1] }\n')
    sc = po.Scene(src)
    d, _ = sc.dist_mat(np.array([[0, 0, 0], [0.6, 0, 0], [0, 0.9, 0]], np.float32))
    assert d[0] == pytest.approx(0.25, abs=1e-6)   # axis: inside the drilled hole
    assert d[1] == pytest.approx(-0.35, abs=1e-6)  # solid shell between hole and surface
    assert d[2] == pytest.approx(0.25, abs=1e-6)


# --------------------------------------------------------------------------- reference.rs (path tracer)
def test_path_rng_matches_an_independent_restatement():
    """src/reference.rs:24-46 against the scene generator's PCG32 (written
    separately from the same lines): identical float sequences."""
    from interflect_amd.scenes import Pcg32

    for seed in (0, 1, 12345, (7 << 32) | 3):
        p = Pcg32(seed)
        want = np.array([(p.next_u32() >> 8) * (1.0 / 16777216.0) for _ in range(64)], np.float32)
        assert np.array_equal(po.rng(seed, 64), want)


OPEN_FLOOR = """render { width: 8, height: 8, surfels: 1000, bounces: 4, exposure: 1.0 }
camera { pos: [0, 3.0, 3.0], look: [0, 0, 0], fov: 30 }
material "f" { albedo: [0.5, 0.6, 0.7], roughness: 1.0 }
box { center: [0, -0.1, 0], size: [6, 0.2, 6], mat: "f" }
light { verts: [[-0.4, 2.0, -0.4], [0.4, 2.0, -0.4], [0.4, 2.0, 0.4], [-0.4, 2.0, 0.4]], emit: [10, 10, 10] }
"""


def test_path_tracer_converges_to_the_closed_form_direct_light():
    """An open floor under one quad light: every bounce after the first
    escapes, so each pixel's expectation is the floor's diffuse direct light
    albedo * emit * F (the point-to-quad form factor of the radiosity path)."""
    sc = po.Scene(OPEN_FLOOR)
    sc.set_spp(4096)
    film, _ = sc.render("reference", 8, 8, use_gi=False)
    # camera rays of the 8 x 8 film (src/main.rs:199-217), to the floor
    pos, look = np.array([0, 3.0, 3.0], np.float32), np.zeros(3, np.float32)
    fwd = (look - pos) / np.linalg.norm(look - pos)
    right = np.cross(fwd, [0, 1, 0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = math.tan(math.radians(30) * 0.5)
    rel = []
    for y in range(8):
        for x in range(8):
            px = (2 * (x + 0.5) / 8 - 1) * th
            py = (1 - 2 * (y + 0.5) / 8) * th
            d = fwd + right * px + up * py
            d /= np.linalg.norm(d)
            t = -pos[1] / d[1]
            p = (pos + d * t).astype(np.float32)
            want, _ = sc.direct_diffuse([p], [[0, 1, 0]], [[0.5, 0.6, 0.7]])
            rel.append(np.abs(film[y, x] - want[0]) / want[0])
    rel = np.array(rel)
    assert rel.mean() < 0.01, rel.mean()
    assert rel.max() < 0.05, rel.max()


def test_path_tracer_energy_ratio_gate_on_the_bench_scenes():
    """ci.yml:152-177: the radiosity image's mean luma over the path
    tracer's (RGB8, Rec.709 weights) stays within (0.85, 1.25).  Run here at
    a reduced size; the GPU test runs the gate at its own 120 x 120, 48 spp."""
    for name in ("bench_box_only.rad", "bench_sphere_only.rad"):
        sc = po.Scene(scene_text(name))
        sc.generate()
        sc.transfer(8)
        sc.solve()
        a, _ = sc.render("beauty", 40, 40)
        sc.set_spp(16)
        g, _ = sc.render("reference", 40, 40, use_gi=False)
        la = po.encode(a, sc.exposure, True).astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
        lg = po.encode(g, sc.exposure, True).astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
        ratio = la.mean() / lg.mean()
        assert 0.85 < ratio < 1.25, (name, ratio)


# --------------------------------------------------------------------------- shade.rs BSDF + emitters
# Closed forms for the specular lobe of shade::direct (src/shade.rs:8-125) and
# the emitter test (src/shade.rs:168-181, src/trace.rs:128-148): the published
# pins count evaluations only, which do not depend on these values.
def test_fresnel_schlick_closed_forms():
    """src/shade.rs:8-14: F(1) = f0, F(0) = 1, F(1/2) = f0 + (1 - f0)/32, decreasing in cos."""
    f0 = np.array([0.04, 0.5, 0.95], np.float32)
    assert np.array_equal(po.fresnel(1.0, f0), f0)
    assert np.allclose(po.fresnel(0.0, f0), 1.0, rtol=0, atol=1.2e-7)
    assert np.allclose(po.fresnel(0.5, f0), f0 + (1 - f0) / 32.0, rtol=1e-6)
    cs = np.linspace(0, 1, 41)
    vals = np.array([po.fresnel(c, f0)[0] for c in cs])
    assert np.all(np.diff(vals) <= 0)
    # clamp: cos outside [0, 1] behaves as the nearest end
    assert np.array_equal(po.fresnel(1.5, f0), f0)


@pytest.mark.parametrize("alpha", [0.05, 0.3, 1.0])
def test_ggx_distribution_is_normalised(alpha):
    """src/shade.rs:16-22: the projected GGX distribution integrates to one over
    the hemisphere, 2 pi int_0^1 D(mu) mu dmu = pi int_0^1 D(sqrt x) dx = 1
    (quadrature of the oracle's f32 D on a grid refined towards mu = 1)."""
    s = np.concatenate([[0.0], np.geomspace(1e-9, 1.0, 6000)])  # s = 1 - x
    d = np.array([po.d_ggx(math.sqrt(1.0 - v), alpha) for v in s])
    integral = math.pi * float(np.sum(0.5 * (d[1:] + d[:-1]) * np.diff(s)))
    assert integral == pytest.approx(1.0, abs=2e-3)
    # peak and grazing values: D(1) = 1/(pi a^2), D(0) = a^2/pi
    assert po.d_ggx(1.0, alpha) == pytest.approx(1.0 / (math.pi * alpha * alpha), rel=1e-5)
    assert po.d_ggx(0.0, alpha) == pytest.approx(alpha * alpha / math.pi, rel=1e-5)


def test_smith_visibility_closed_forms():
    """src/shade.rs:24-31: height-correlated Smith over 4 nl nv.  alpha = 1:
    0.5 / (nl + nv) exactly; alpha -> 0: 1 / (4 nl nv); symmetric in (nv, nl)."""
    rng = np.random.default_rng(5)
    for nv, nl in rng.uniform(0.05, 1.0, (20, 2)).astype(np.float32):
        assert po.v_smith(nv, nl, 1.0) == float(np.float32(0.5) / (nv + nl))
        assert po.v_smith(nv, nl, 1e-4) == pytest.approx(1.0 / (4.0 * nv * nl), rel=2e-5)
        for a in (0.1, 0.5, 0.9):
            assert po.v_smith(nv, nl, a) == po.v_smith(nl, nv, a)
    # between the limits, larger roughness lowers the visibility term
    assert po.v_smith(0.5, 0.7, 0.2) > po.v_smith(0.5, 0.7, 0.8)


def test_f0_is_the_lerp_of_dielectric_and_albedo():
    """src/shade.rs:35-38: f0 = lerp(0.04, albedo, metallic)."""
    alb = np.array([0.9, 0.5, 0.2], np.float32)
    assert np.array_equal(po.f0_of(alb, 0.0), np.full(3, 0.04, np.float32))
    assert np.array_equal(po.f0_of(alb, 1.0), alb)
    assert np.allclose(po.f0_of(alb, 0.5), (0.04 + alb) / 2, rtol=1e-6)


def lit_floor(a, h, emit):
    return ("render { width: 8, height: 8 }\ncamera { pos: [0, 1, 3], look: [0, 0, 0], fov: 40 }\n"
            "material \"m\" { albedo: [0.5, 0.5, 0.5], roughness: 1.0 }\n"
            "plane { normal: [0, 1, 0], height: 0, mat: \"m\" }\n"
            "light { verts: [[%r, %r, %r], [%r, %r, %r], [%r, %r, %r], [%r, %r, %r]], emit: [%r, %r, %r] }\n"
            % (-a, h, -a, a, h, -a, a, h, a, -a, h, a, emit, emit, emit))


def howell_b3(a, h):
    """Point under the centre of a parallel square of half-side a at height h."""
    x = a / h
    c = (1 / (2 * math.pi)) * 2 * (x / math.sqrt(1 + x * x) * math.atan(x / math.sqrt(1 + x * x)))
    return 4 * c


@pytest.mark.parametrize("theta_deg", [0.0, 20.0, 35.0])
@pytest.mark.parametrize("rough,metal", [(0.3, 1.0), (1.0, 0.0), (0.5, 0.5), (0.08, 1.0)])
def test_direct_light_closed_form_on_an_open_floor(theta_deg, rough, metal):
    """shade::direct (src/shade.rs:50-125) at a floor point under the centre of
    a downward square light, viewed at angle theta in the xy plane.  Nothing
    occludes (vis = 1), the mirror ray meets the light plane at
    (-h tan theta, h, 0) inside the quad, so the representative point is that
    point, l = reflect(-v) and h = n: D = 1/(pi w^2), F = Schlick at cos theta,
    V = 0.5 / (2 c sqrt(c^2 (1 - w^2) + w^2)), energy (alpha/w)^2 with
    w = clamp(alpha + radius cos(theta) / (2h), alpha, 1), plus the diffuse
    kd L F_B3."""
    a, h, L = 1.6, 2.0, 7.0  # h tan(35 deg) = 1.40 < a
    sc = po.Scene(lit_floor(a, h, L))
    th = math.radians(theta_deg)
    v = np.array([math.sin(th), math.cos(th), 0.0], np.float32)
    v /= np.linalg.norm(v)
    alb = np.array([0.9, 0.5, 0.2])
    got, steps = sc.direct([0, 0, 0], [0, 1, 0], v, alb, rough, metal)
    assert steps > 0
    c = math.cos(th)
    alpha = max(rough * rough, 1e-3)
    radius = math.sqrt((2 * a) ** 2 / math.pi)
    w = min(max(alpha + radius / (2 * h / c), alpha), 1.0)
    energy = (alpha / w) ** 2
    D = 1.0 / (math.pi * w * w)
    V = 0.5 / (2 * c * math.sqrt(c * c * (1 - w * w) + w * w))
    f0 = 0.04 * (1 - metal) + alb * metal
    F = f0 + (1 - f0) * (1 - c) ** 5
    spec = F * (D * V * energy * c) * L
    diff = alb * (1 - metal) * L * howell_b3(a, h)
    want = diff + spec
    assert np.allclose(got, want, rtol=2e-5, atol=1e-6), (got, want)


def test_direct_light_skips_an_occluded_light_and_a_back_facing_one():
    """A slab between floor and light zeroes the term; a light facing away contributes nothing."""
    src = lit_floor(0.6, 2.0, 7.0) + "box { center: [0, 1, 0], size: [3, 0.1, 3], mat: \"m\" }\n"
    got, _ = po.Scene(src).direct([0, 0, 0], [0, 1, 0], [0, 1, 0], [0.5, 0.5, 0.5], 0.5, 0.0)
    assert np.array_equal(got, np.zeros(3, np.float32))
    up = lit_floor(0.6, 2.0, 7.0).replace("[[-0.6, 2.0, -0.6], [0.6, 2.0, -0.6], [0.6, 2.0, 0.6], [-0.6, 2.0, 0.6]]",
                                          "[[-0.6, 2.0, 0.6], [0.6, 2.0, 0.6], [0.6, 2.0, -0.6], [-0.6, 2.0, -0.6]]")
    assert up != lit_floor(0.6, 2.0, 7.0)
    got, _ = po.Scene(up).direct([0, 0, 0], [0, 1, 0], [0, 1, 0], [0.5, 0.5, 0.5], 0.5, 0.0)
    assert np.array_equal(got, np.zeros(3, np.float32))


def test_intersect_quad_closed_forms():
    """src/trace.rs:128-148: hit distance inside the quad, misses outside it,
    parallel rays and hits closer than 1e-4."""
    q = ceiling(0.5, 2.0)
    assert po.intersect_quad(q, [0, 0, 0], [0, 1, 0]) == pytest.approx(2.0, abs=1e-6)
    d = np.array([0.2, 1.0, 0.1])
    d /= np.linalg.norm(d)
    assert po.intersect_quad(q, [0, 0, 0], d) == pytest.approx(2.0 / d[1], rel=1e-6)
    assert po.intersect_quad(q, [0.6, 0, 0], [0, 1, 0]) is None  # outside
    assert po.intersect_quad(q, [0, 0, 0], [1, 0, 0]) is None    # parallel
    assert po.intersect_quad(q, [0, 2.0 - 5e-5, 0], [0, 1, 0]) is None  # t < 1e-4
    assert po.intersect_quad(q, [0, 3, 0], [0, 1, 0]) is None    # behind


def test_emitter_hit_front_face_nearest_and_tmax():
    """src/shade.rs:168-181: only the emitting face is visible, the nearest
    hit wins, and hits at or beyond tmax are ignored."""
    src = lit_floor(0.6, 2.0, 7.0) + (
        "light { verts: [[-0.3, 1.0, -0.3], [0.3, 1.0, -0.3], [0.3, 1.0, 0.3], [-0.3, 1.0, 0.3]], "
        "emit: [3, 4, 5] }\n")
    sc = po.Scene(src)
    e = sc.emitter_hit([0, 0, 0], [0, 1, 0], 100.0)
    assert e is not None and np.array_equal(e, np.array([3, 4, 5], np.float32))  # the nearer (y = 1)
    e = sc.emitter_hit([0.45, 0, 0], [0, 1, 0], 100.0)
    assert e is not None and np.array_equal(e, np.full(3, 7.0, np.float32))  # only the larger covers x = .45
    assert sc.emitter_hit([0, 0, 0], [0, 1, 0], 0.99) is None  # both beyond tmax
    assert sc.emitter_hit([0, 5, 0], [0, -1, 0], 100.0) is None  # back faces
