"""GPU backward pass under a link memory budget, and at full problem size.

With "link_budget" below what keeping the transposed links costs, every
adjoint sweep recomputes the links tile by tile from the geometry.  The
summation order is fixed by (cluster, row tile, adj_seg) segments in both
layouts, so the gradients must be bit-identical to the resident layout.  A
budget below one tile's working set is an allocation failure (IFX_E_NOMEM)
that leaves the context usable.

The full-size problem is the 8-material, 3-light Cornell-type room with
50000 surfels at 512 x 512: gradients of an L2 image loss against float64
central differences (eps 1e-3, relative error <= 1e-3) and the solve's
adjoint against a float64 Jacobi adjoint on the device's own transfer.
"""
import sys

import numpy as np
import pytest

import adjoint64 as a64
import interflect_amd as ia
from interflect_amd import scenes

pytestmark = pytest.mark.gpu

FIXED_POINT_BOUNCES = 400


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def diff_ctx(seed, surfels, w, h):
    ctx = ia.Context(0)
    ctx.load_scene(scenes.diff_cornell(seed, surfels=surfels, width=w, height=h))
    # host placement, as the other adjoint tests (bit-identical to the device's)
    ctx.set_tuning("surfels_gpu", 0)
    ctx.generate_surfels()
    tr = ctx.transfer(8)
    ctx.set_solve_tolerance(0.0)
    ctx.set_adjoint(2000, 1e-6)
    return ctx, tr


def render(ctx):
    ctx.solve(FIXED_POINT_BOUNCES)
    return ctx.shade("beauty", rgb8=False)["film"].astype(np.float64)


def scene_params(seed, surfels, w, h):
    src = scenes.diff_cornell(seed, surfels=surfels, width=w, height=h)
    alb = [[0.8, 0.8, 0.8]]
    emit = []
    for line in src.splitlines():
        if line.startswith("material"):
            alb.append([float(x) for x in line.split("albedo: [")[1].split("]")[0].split(",")])
        if line.startswith("light"):
            emit.append([float(x) for x in line.split("emit: [")[1].split("]")[0].split(",")])
    return np.array(alb, np.float64), np.array(emit, np.float64)


def same(a, b):
    assert np.array_equal(bits(a["albedo"]), bits(b["albedo"]))
    assert np.array_equal(bits(a["emit"]), bits(b["emit"]))
    assert a["iterations"] == b["iterations"]
    assert bits(a["residual"]) == bits(b["residual"])


def test_budget_tiles_give_bit_identical_gradients():
    say("link budget: 4000 surfels, 64 x 64")
    ctx, tr = diff_ctx(8, 4000, 64, 64)
    N = ctx.num_surfels()
    render(ctx)
    u = np.random.default_rng(1).normal(size=(64, 64, 3)).astype(np.float32)
    ubar = np.random.default_rng(2).normal(size=(N, 3)).astype(np.float32)
    for rows in (256, 1000):
        ctx.set_tuning("adj_tile_rows", rows)
        ctx.set_tuning("link_budget", 0)
        a = ctx.backward(u)
        sa = ctx.backward_surfels(ubar)
        assert a["link_tiles"] == 0 and sa["link_tiles"] == 0
        ctx.set_tuning("link_budget", 31 * tr["links"])  # just below the resident cost
        b = ctx.backward(u)
        sb = ctx.backward_surfels(ubar)
        tiles = -(-N // rows)
        say("rows", rows, "tiles", b["link_tiles"], "sweeps", b["iterations"], "resident %.1f ms, tiled %.1f ms"
            % (a["ms"], b["ms"]))
        assert b["link_tiles"] == tiles and sb["link_tiles"] == tiles
        same(a, b)
        same(sa, sb)
        assert np.array_equal(bits(sa["mu"]), bits(sb["mu"]))
        # launch shapes of the tiled sweep
        ctx.set_tuning("jacobi_blocks", 5)
        ctx.set_tuning("vis_blocks", 3)
        c = ctx.backward(u)
        ctx.set_tuning("jacobi_blocks", 0)
        ctx.set_tuning("vis_blocks", 0)
        same(a, c)
        assert np.abs(a["albedo"]).max() > 0
    ctx.set_tuning("link_budget", 0)
    ctx.set_tuning("adj_tile_rows", 0)
    assert ctx.backward(u)["link_tiles"] == 0


def test_budget_below_one_tile_is_an_allocation_failure():
    say("link budget below one tile")
    ctx, tr = diff_ctx(8, 1500, 32, 32)
    render(ctx)
    u = np.random.default_rng(4).normal(size=(32, 32, 3)).astype(np.float32)
    ref = ctx.backward(u)
    ctx.set_tuning("link_budget", 4096)
    with pytest.raises(ia.IfxError) as e:
        ctx.backward(u)
    assert e.value.status == -5 and "link_budget" in str(e.value)
    with pytest.raises(ia.IfxError) as e:
        ctx.set_tuning("link_budget", -1)
    assert e.value.status == -1
    # the context stays usable: a workable budget with 64-row tiles, then none
    ctx.set_tuning("adj_tile_rows", 64)
    ctx.set_tuning("link_budget", 8 * tr["links"])
    r = ctx.backward(u)
    assert r["link_tiles"] == -(-ctx.num_surfels() // 64)
    ctx.set_tuning("link_budget", 0)
    r0 = ctx.backward(u)
    assert r0["link_tiles"] == 0
    same(r, r0)
    ctx.set_tuning("adj_tile_rows", 0)
    same(ref, ctx.backward(u))


# ---------------------------------------------------------------- full size
E_SURFELS, E_W, E_H = 50000, 512, 512


@pytest.fixture(scope="module")
def config_e():
    say("full size: target seed 7, start seed 8,", E_SURFELS, "surfels,", E_W, "x", E_H)
    tgt, _ = diff_ctx(7, E_SURFELS, E_W, E_H)
    target = render(tgt)
    del tgt
    ctx, tr = diff_ctx(8, E_SURFELS, E_W, E_H)
    say("transfer: %d links, %d candidates, %.1f ms" % (tr["links"], tr["candidates"], tr["ms"]))
    film = render(ctx)
    return ctx, tr, target, film


@pytest.mark.timeout(900)
def test_full_size_budgeted_gradients_match_central_differences(config_e):
    ctx, tr, target, film = config_e
    u = (2.0 * (film - target)).astype(np.float32)
    g0 = ctx.backward(u)
    ctx.set_tuning("link_budget", 16 * tr["links"])  # half of the resident cost
    g = ctx.backward(u)
    say("backward: resident %.1f ms, %d tiles %.1f ms, %d sweeps" % (g0["ms"], g["link_tiles"], g["ms"],
                                                                    g["iterations"]))
    assert g["link_tiles"] == -(-ctx.num_surfels() // 4096)
    same(g0, g)
    alb, emit = scene_params(8, E_SURFELS, E_W, E_H)

    def loss():
        return float(np.sum((render(ctx) - target) ** 2))

    eps = 1e-3
    fd_alb = np.zeros_like(alb)
    for m in range(1, alb.shape[0]):
        for c in range(3):
            a = alb[m].copy()
            a[c] += eps
            ctx.set_material_albedo(m, a)
            lp = loss()
            a[c] -= 2 * eps
            ctx.set_material_albedo(m, a)
            lm = loss()
            ctx.set_material_albedo(m, alb[m])
            fd_alb[m, c] = (lp - lm) / (2 * eps)
        say("central differences: material", m)
    fd_emit = np.zeros(len(emit))
    for l in range(len(emit)):
        ctx.set_light_emit(l, emit[l] + eps)
        lp = loss()
        ctx.set_light_emit(l, emit[l] - eps)
        lm = loss()
        ctx.set_light_emit(l, emit[l])
        fd_emit[l] = (lp - lm) / (2 * eps)
    an_alb = g["albedo"].astype(np.float64)
    an_emit = g["emit"].astype(np.float64).sum(1)
    ra = np.abs(an_alb - fd_alb) / np.maximum(np.abs(fd_alb), 1e-2 * np.abs(fd_alb).max())
    re = np.abs(an_emit - fd_emit) / np.maximum(np.abs(fd_emit), 1e-2 * np.abs(fd_emit).max())
    say("max rel err albedo", ra.max(), "emit", re.max())
    assert ra.max() <= 1e-3
    assert re.max() <= 1e-3
    assert np.array_equal(render(ctx), film)
    ctx.set_tuning("link_budget", 0)


@pytest.mark.timeout(900)
def test_full_size_surfel_adjoint_matches_fp64_jacobi(config_e):
    ctx, tr, _, _ = config_e
    say("full size: surfel adjoint vs float64 Jacobi")
    N = ctx.num_surfels()
    render(ctx)
    starts, idx, w, _, _ = ctx.transfer_arrays()
    _, _, cl_area, _, mstart, members = ctx.clusters()
    _, _, rho, area = ctx.surfels()
    mats = ctx.surfel_materials()
    g = ctx.direct_geometry()
    alb, emit = scene_params(8, E_SURFELS, E_W, E_H)
    Wm, Pm = a64.build_operators(starts, idx, w, mstart, members, area, cl_area, N)
    D = a64.direct_factor(g, emit)
    rho64 = rho.astype(np.float64)
    irr64 = a64.forward_iterative(Wm, Pm, rho64, D, tol=1e-11)
    _, irr32, _ = ctx.radiosity()
    say("forward: max |f32 - f64| / max |f64| = %.3g" % (np.abs(irr32 - irr64).max() / np.abs(irr64).max()))
    assert np.abs(irr32 - irr64).max() <= 1e-4 * np.abs(irr64).max()
    ubar = np.random.default_rng(5).normal(size=(N, 3)).astype(np.float32)
    ctx.set_tuning("link_budget", 16 * tr["links"])
    r = ctx.backward_surfels(ubar)
    ctx.set_tuning("link_budget", 0)
    assert r["link_tiles"] > 0
    mu64, da64, de64 = a64.adjoint_iterative(Wm, Pm, rho64, D, irr64, ubar.astype(np.float64), mats, g, len(alb),
                                             tol=1e-11)
    say("adjoint sweeps", r["iterations"], "mu err %.3g" % (np.abs(r["mu"] - mu64).max() / np.abs(mu64).max()))
    assert np.abs(r["mu"] - mu64).max() <= 1e-4 * np.abs(mu64).max()
    assert np.abs(r["albedo"] - da64).max() <= 1e-3 * np.abs(da64).max()
    assert np.abs(r["emit"] - de64).max() <= 1e-3 * np.abs(de64).max()
