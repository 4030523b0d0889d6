"""GPU backward pass: implicit adjoint of the frozen-geometry pipeline.

Checks (tolerance 1e-3 relative, as stated for the gradients):
  * the solve's adjoint state and gradients against a float64 direct sparse
    solve built from the device's own transfer, clusters and light factors;
  * the full film -> albedo / emission gradients against float64 central
    differences (eps 1e-3) of a float64 loss over float32 forward renders,
    with the forward solve run to its fixed point;
  * bit-identical gradients across repeat calls and launch shapes;
  * the state and material checks of the entry point.
"""
import sys

import numpy as np
import pytest

import adjoint64 as a64
import interflect_amd as ia
from conftest import scene_text
from interflect_amd import scenes

pytestmark = pytest.mark.gpu

SURFELS, W, H = 4000, 96, 96
FIXED_POINT_BOUNCES = 400


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def diff_ctx(seed):
    ctx = ia.Context(0)
    ctx.load_scene(scenes.diff_cornell(seed, surfels=SURFELS, width=W, height=H))
    # Host placement (bit-identical to the device placement, which
    # test_gpu_surfels checks): the device placement of this 4000-surfel room
    # once ran past the test's time limit, with no cause found, and stays off
    # the GPU until one is.
    ctx.set_tuning("surfels_gpu", 0)
    ctx.generate_surfels()
    ctx.transfer(8)
    ctx.set_solve_tolerance(0.0)
    ctx.set_adjoint(2000, 1e-6)
    return ctx


def render(ctx):
    ctx.solve(FIXED_POINT_BOUNCES)
    return ctx.shade("beauty", rgb8=False)["film"].astype(np.float64)


def scene_params(ctx):
    """Albedo ((materials + 1) x 3, row 0 the implicit default material, unused
    by this scene) and emission (lights x 3) from the scene text."""
    src = scenes.diff_cornell(ctx._seed, surfels=SURFELS, width=W, height=H)
    alb = [[0.8, 0.8, 0.8]]
    emit = []
    for line in src.splitlines():
        if line.startswith("material"):
            v = line.split("albedo: [")[1].split("]")[0]
            alb.append([float(x) for x in v.split(",")])
        if line.startswith("light"):
            v = line.split("emit: [")[1].split("]")[0]
            emit.append([float(x) for x in v.split(",")])
    return np.array(alb, np.float64), np.array(emit, np.float64)


@pytest.fixture(scope="module")
def problem():
    say("adjoint: target seed 7, start seed 8,", SURFELS, "surfels,", W, "x", H)
    tgt = diff_ctx(7)
    target = render(tgt)
    ctx = diff_ctx(8)
    ctx._seed = 8
    film = render(ctx)
    return ctx, target, film


def test_surfel_adjoint_matches_fp64_direct_solve(problem):
    ctx, _, _ = problem
    say("adjoint: surfel share vs float64 direct solve")
    N = ctx.num_surfels()
    starts, idx, w, _, _ = ctx.transfer_arrays()
    _, _, cl_area, _, mstart, members = ctx.clusters()
    _, _, rho, area = ctx.surfels()
    mats = ctx.surfel_materials()
    g = ctx.direct_geometry()
    alb, emit = scene_params(ctx)
    Wm, Pm = a64.build_operators(starts, idx, w, mstart, members, area, cl_area, N)
    D = a64.direct_factor(g, emit)
    irr64 = a64.forward(Wm, Pm, rho.astype(np.float64), D)
    _, irr32, _ = ctx.radiosity()
    # the float32 forward at its fixed point agrees with the float64 solve
    assert np.abs(irr32 - irr64).max() <= 1e-4 * np.abs(irr64).max()
    ubar = np.random.default_rng(5).normal(size=(N, 3)).astype(np.float32)
    r = ctx.backward_surfels(ubar)
    mu64, da64, de64 = a64.adjoint(Wm, Pm, rho.astype(np.float64), D, irr64, ubar.astype(np.float64), mats, g,
                                   len(alb))
    say("adjoint sweeps", r["iterations"], "residual", r["residual"])
    assert r["iterations"] < 2000
    assert np.abs(r["mu"] - mu64).max() <= 1e-4 * np.abs(mu64).max()
    assert np.abs(r["albedo"] - da64).max() <= 1e-3 * np.abs(da64).max()
    assert np.abs(r["emit"] - de64).max() <= 1e-3 * np.abs(de64).max()


def test_full_backward_matches_central_differences(problem):
    ctx, target, film = problem
    say("adjoint: film gradients vs central differences")
    u = (2.0 * (film - target)).astype(np.float32)
    ctx.solve(FIXED_POINT_BOUNCES)
    f2 = ctx.shade("beauty", rgb8=False)["film"]
    assert np.array_equal(f2, film.astype(np.float32))
    g = ctx.backward(u)
    alb, emit = scene_params(ctx)

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
    fd_emit = np.zeros(len(emit))
    for l in range(len(emit)):
        ctx.set_light_emit(l, emit[l] + eps)
        lp = loss()
        ctx.set_light_emit(l, emit[l] - eps)
        lm = loss()
        ctx.set_light_emit(l, emit[l])
        fd_emit[l] = (lp - lm) / (2 * eps)
    an_alb = g["albedo"].astype(np.float64)
    assert not an_alb[0].any()  # no surface uses the default material
    an_emit = g["emit"].astype(np.float64).sum(1)  # scalar emission drives all three channels
    say("albedo grads max", np.abs(fd_alb).max(), "emit grads", fd_emit)
    ra = np.abs(an_alb - fd_alb) / np.maximum(np.abs(fd_alb), 1e-2 * np.abs(fd_alb).max())
    re = np.abs(an_emit - fd_emit) / np.maximum(np.abs(fd_emit), 1e-2 * np.abs(fd_emit).max())
    say("max rel err albedo", ra.max(), "emit", re.max())
    assert ra.max() <= 1e-3
    assert re.max() <= 1e-3
    # the restored parameters reproduce the original render bit for bit
    assert np.array_equal(render(ctx), film)


def test_backward_is_bit_identical_across_runs_and_shapes(problem):
    ctx, target, film = problem
    u = (2.0 * (film - target)).astype(np.float32)
    ctx.solve(FIXED_POINT_BOUNCES)
    ctx.shade("beauty", rgb8=False)
    a = ctx.backward(u)
    b = ctx.backward(u)
    ctx.set_tuning("jacobi_blocks", 7)
    ctx.set_tuning("lds", 0)
    c = ctx.backward(u)
    ctx.set_tuning("jacobi_blocks", 0)
    ctx.set_tuning("lds", 1)
    for x in (b, c):
        assert np.array_equal(a["albedo"].view(np.uint32), x["albedo"].view(np.uint32))
        assert np.array_equal(a["emit"].view(np.uint32), x["emit"].view(np.uint32))
        assert a["iterations"] == x["iterations"]


def test_parameter_setters_match_scene_text():
    """Setting albedo / emission through the API renders exactly like loading
    a scene whose text carries those values."""
    say("adjoint: setters vs scene text")
    a = ia.Context(0)
    a.load_scene(scenes.diff_cornell(8, surfels=2000, width=48, height=48))
    b = ia.Context(0)
    b.load_scene(scenes.diff_cornell(7, surfels=2000, width=48, height=48))
    src7 = scenes.diff_cornell(7, surfels=2000, width=48, height=48)
    mats7 = [l for l in src7.splitlines() if l.startswith("material")]
    lights7 = [l for l in src7.splitlines() if l.startswith("light")]
    for m, line in enumerate(mats7):
        v = line.split("albedo: [")[1].split("]")[0]
        a.set_material_albedo(m + 1, [float(x) for x in v.split(",")])
    for l, line in enumerate(lights7):
        v = line.split("emit: [")[1].split("]")[0]
        a.set_light_emit(l, [float(x) for x in v.split(",")])
    out = []
    for ctx in (a, b):
        ctx.set_tuning("surfels_gpu", 0)  # gradient rooms are placed on the host (see the fixture)
        ctx.generate_surfels()
        ctx.transfer(8)
        ctx.solve()
        out.append(ctx.shade("beauty")["rgb8"])
    assert np.array_equal(out[0], out[1])


def test_backward_state_and_material_checks():
    ctx = ia.Context(0)
    ctx.load_scene(scenes.diff_cornell(8, surfels=1000, width=32, height=32))
    ctx.set_tuning("surfels_gpu", 0)
    ctx.generate_surfels()
    ctx.transfer(8)
    u = np.zeros((32, 32, 3), np.float32)
    with pytest.raises(ia.IfxError) as e:
        ctx.backward(u)
    assert e.value.status == -3
    ctx.solve()
    with pytest.raises(ia.IfxError) as e:  # no beauty shade since the solve
        ctx.backward(u)
    assert e.value.status == -3
    ctx.shade("beauty")
    r = ctx.backward(u)  # zero upstream gradient: zero gradients
    assert not r["albedo"].any() and not r["emit"].any()
    # product.rad has a metallic material: not supported by the diffuse adjoint
    p = ia.Context(0)
    p.load_scene(scene_text("product.rad"))
    p.set_render(32, 32, 1000, 4)
    p.generate_surfels()
    p.transfer(8)
    p.solve()
    p.shade("beauty")
    with pytest.raises(ia.IfxError) as e:
        p.backward(np.zeros((32, 32, 3), np.float32))
    assert e.value.status == -1


def test_backward_from_a_caller_scene_and_surfels():
    """A Rust host that keeps its loader: the scene as a descriptor (the
    oracle's parse) and the surfels as arrays give the gradients of a context
    that parsed the text and placed its own surfels, bit for bit."""
    say("adjoint: descriptor scene + caller surfels")
    from oracle import pyoracle as po

    src = scenes.diff_cornell(8, surfels=1500, width=40, height=40)
    a = ia.Context(0)
    a.load_scene(src)
    a.set_tuning("surfels_gpu", 0)
    a.generate_surfels()
    b = ia.Context(0)
    b.set_scene(po.Scene(src).scene_desc())
    assert b.scene_digest() == a.scene_digest()
    b.set_surfels(*a.surfels())
    assert np.array_equal(b.surfel_materials(), a.surfel_materials())
    u = np.random.default_rng(3).normal(size=(40, 40, 3)).astype(np.float32)
    out = []
    for ctx in (a, b):
        ctx.transfer(8)
        ctx.solve(64)
        f = ctx.shade("beauty", rgb8=False)["film"]
        out.append((f, ctx.backward(u)))
    (fa, ga), (fb, gb) = out
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
    assert np.array_equal(ga["albedo"].view(np.uint32), gb["albedo"].view(np.uint32))
    assert np.array_equal(ga["emit"].view(np.uint32), gb["emit"].view(np.uint32))
    assert np.abs(ga["albedo"]).max() > 0
