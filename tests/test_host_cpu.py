"""CPU checks of the product library: the C ABI loads and exports every
declared symbol, and its host-side scene loader, BVH build and surfel
placement (a host-only context, device -1) agree bit for bit with the oracle.
No kernel runs here."""
import os
import re

import numpy as np
import pytest

import interflect_amd as ia
from conftest import ROOT, scene_text
from oracle import pyoracle as po


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "interflect.h")).read()
    declared = set(re.findall(r"^(?:ifx_status|const char\*)\s+(ifx_\w+)\(", hdr, re.M))
    assert declared == set(ia.EXPORTS)
    L = ia.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_product_library_does_not_link_the_oracle():
    import subprocess

    out = subprocess.run(["ldd", ia.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", "--defined-only", ia.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in syms


@pytest.mark.parametrize("name", ["product.rad", "cornell_bench.rad", "stress.rad", "closed_box.rad", "alcove.rad",
                                  "dog.rad", "minimal.rad"])
def test_loader_matches_oracle(name):
    src = scene_text(name)
    ctx = ia.Context(-1)
    info = ctx.load_scene(src)
    sc = po.Scene(src)
    assert info["n_objects"] == sc.n_objects and info["n_planes"] == sc.n_planes
    assert info["n_quad_lights"] == sc.n_lights and info["n_materials"] == sc.n_materials
    assert info["n_bvh_nodes"] == sc.n_bvh_nodes
    assert np.array_equal(info["lo"], sc.lo) and np.array_equal(info["hi"], sc.hi)
    assert (info["width"], info["height"], info["surfels"], info["bounces"]) == (
        sc.width, sc.height, sc.surfels_requested, sc.bounces)
    assert info["exposure"] == np.float32(sc.exposure)


@pytest.mark.parametrize("name,target", [("product.rad", None), ("cornell_bench.rad", None), ("stress.rad", None),
                                         ("closed_box.rad", 3000)])
def test_surfels_match_oracle_bitwise(name, target):
    src = scene_text(name)
    ctx = ia.Context(-1)
    ctx.load_scene(src)
    sc = po.Scene(src)
    n1 = ctx.generate_surfels(target)
    n2 = sc.generate(target)
    assert n1 == n2
    for a, b in zip(ctx.surfels(), sc.surfels()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


from test_oracle_pins import LIMITS, MALFORMED, BASE_LIGHT  # noqa: E402


@pytest.mark.parametrize("src,msg", MALFORMED)
def test_loader_rejects_malformed_input_like_the_reference(src, msg):
    ctx = ia.Context(-1)
    with pytest.raises(ia.IfxError, match=msg) as e:
        ctx.load_scene(src)
    assert e.value.status == -2


@pytest.mark.parametrize("cfg,msg", LIMITS)
def test_loader_rejects_absurd_settings(cfg, msg):
    ctx = ia.Context(-1)
    with pytest.raises(ia.IfxError, match=msg):
        ctx.load_scene(cfg + "\n" + BASE_LIGHT + "\n")


def test_host_only_context_refuses_kernels():
    ctx = ia.Context(-1)
    ctx.load_scene(scene_text("minimal.rad"))
    ctx.generate_surfels(100)
    with pytest.raises(ia.IfxError) as e:
        ctx.transfer(8)
    assert e.value.status == -3
    with pytest.raises(ia.IfxError):
        ctx.dist_mat(np.zeros((4, 3), np.float32))


def test_cli_overrides_are_revalidated():
    """src/main.rs:394-413: --surfels 999999999 is rejected after parsing."""
    ctx = ia.Context(-1)
    ctx.load_scene(scene_text("minimal.rad"))
    with pytest.raises(ia.IfxError, match="surfels"):
        ctx.set_render(800, 600, 999999999, 12)
    ctx.set_render(320, 240, 5000, 8)
    info = ctx.info()
    assert (info["width"], info["height"], info["surfels"], info["bounces"]) == (320, 240, 5000, 8)


def test_stage_order_is_enforced():
    ctx = ia.Context(-1)
    with pytest.raises(ia.IfxError) as e:
        ctx.generate_surfels(10)
    assert e.value.status == -3


def test_tuning_keys_and_ranges():
    ctx = ia.Context(-1)
    for key, v in (("link_budget", 1 << 40), ("link_budget", 0), ("adj_tile_rows", 512), ("adj_seg", 256),
                   ("side_stream", 0), ("jacobi_rows", 3), ("vis_min_idle", 1), ("vis_min_idle", 64),
                   ("closure_min_idle", 32), ("vis_chunk", 128), ("flat_query", 0), ("flat_query", 1),
                   ("vis_tail", 0), ("vis_tail", 32), ("vis_tail_items", 0), ("vis_tail_items", 1 << 30)):
        ctx.set_tuning(key, v)
    for key, v in (("link_budget", -1), ("link_budget", (1 << 50) + 1), ("adj_tile_rows", (1 << 30) + 1),
                   ("spp", 0), ("no_such_key", 1), ("vis_min_idle", 0), ("vis_min_idle", 65),
                   ("closure_min_idle", 0), ("closure_min_idle", 65), ("vis_tail", -1), ("vis_tail", 33),
                   ("vis_tail_items", -1), ("vis_tail_items", (1 << 30) + 1)):
        with pytest.raises(ia.IfxError) as e:
            ctx.set_tuning(key, v)
        assert e.value.status == -1


# --------------------------------------------------------------------------- scene descriptor
SCENE_FIXTURES = ["product.rad", "cornell_bench.rad", "stress.rad", "closed_box.rad", "alcove.rad", "dog.rad",
                  "minimal.rad", "TEMPLATE.rad"]


@pytest.mark.parametrize("name", SCENE_FIXTURES)
def test_scene_descriptor_equals_the_loader(name):
    """ifx_set_scene with the oracle's parse (the reference's Scene after its
    parser and Bvh::build) gives the device records ifx_load_scene gives for
    the text, with the caller's BVH and with the library's own."""
    src = scene_text(name)
    a = ia.Context(-1)
    ia_info = a.load_scene(src)
    desc = po.Scene(src).scene_desc()
    for with_bvh in (True, False):
        b = ia.Context(-1)
        info = b.set_scene(desc, with_bvh=with_bvh)
        assert b.scene_digest() == a.scene_digest(), with_bvh
        assert info["n_bvh_nodes"] == ia_info["n_bvh_nodes"]


def test_scene_descriptor_extension_scenes():
    from interflect_amd import scenes

    for src in (scenes.stress(prims=64), scenes.product_literal()):
        a = ia.Context(-1)
        a.load_scene(src)
        b = ia.Context(-1)
        b.set_scene(po.Scene(src).scene_desc())
        assert b.scene_digest() == a.scene_digest()


def test_scene_descriptor_rejects_malformed_input():
    base = po.Scene(scene_text("product.rad")).scene_desc()

    def bad(mutate, msg):
        d = {k: (v.copy() if isinstance(v, np.ndarray) else dict(v) if isinstance(v, dict) else v)
             for k, v in base.items()}
        mutate(d)
        c = ia.Context(-1)
        with pytest.raises(ia.IfxError) as e:
            c.set_scene(d)
        assert e.value.status == -1 and msg in str(e.value), str(e.value)

    def as_u32(a):
        return a.view(np.uint32)

    bad(lambda d: as_u32(d["objects"])[0].__setitem__(11, 99), "material index")
    bad(lambda d: as_u32(d["objects"])[0].__setitem__(0, 2), "planes belong")
    bad(lambda d: as_u32(d["objects"])[0].__setitem__(0, 9), "unknown primitive")
    bad(lambda d: as_u32(d["objects"])[0].__setitem__(10, 5), "subtraction range")
    bad(lambda d: d["bvh_order"].__setitem__(0, d["bvh_order"][1]), "BVH order")
    bad(lambda d: as_u32(d["bvh_nodes"])[1].__setitem__(7, 7), "BVH leaf")
    bad(lambda d: d.__setitem__("quad_lights", np.zeros((0, 15), np.float32)), "light")
    bad(lambda d: d["render"].__setitem__("width", 0), "width")
    bad(lambda d: d["bounds"].__setitem__(0, np.inf), "bounds")


def test_set_surfels_assigns_placement_materials():
    """ifx_set_surfels takes each surfel's material from the scene at its
    position, which is what surfel placement assigns; ifx_set_surfel_materials
    overrides it and checks the range."""
    src = scene_text("product.rad")
    a = ia.Context(-1)
    a.load_scene(src)
    a.generate_surfels(3000)
    placed = a.surfel_materials()
    b = ia.Context(-1)
    b.load_scene(src)
    b.set_surfels(*a.surfels())
    assert np.array_equal(b.surfel_materials(), placed)
    b.set_surfel_materials(np.zeros_like(placed))
    assert not b.surfel_materials().any()
    with pytest.raises(ia.IfxError):
        b.set_surfel_materials(np.full_like(placed, 1000))
    with pytest.raises(ia.IfxError):
        b.set_surfel_materials(placed[:-1])
