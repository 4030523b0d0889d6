"""Benchmark of the transfer -> solve -> shade frame on one MI355X.

One step = one frame on resident surfels: transfer (H2D of the surfels,
clusters, candidate culls, cone-traced link visibility, hemisphere closure,
CSR) + solve (direct E, Jacobi) + shade (grid build, per-pixel pass, encode,
D2H of the RGB8 image).  Surfel placement, parsing and the BVH build precede
the timed region, as in the reference's stage timers (src/main.rs:432-534).

Prints one JSON line.  `value` is Gsdf-evals/s under the reference's
counting rule (trace step +1, hit normal +4, shadow step +1) over the whole
frame; ms_per_step is the frame time.  The default workload is the literal
product scene (7 SDF objects incl. a drilled CSG bowl, 2 sphere lights,
6 Lambertian materials, 640x480, 20000 surfels requested); `--also stress`
adds the 1920x1080 seeded stress scene as a secondary measurement, and
`--also diff` the gradient workload: grads/s of the implicit adjoint
(forward solve + beauty shade + backward to 24 albedo and 3 emission
parameters) on the 4.4-unit Cornell room with 8 materials and 3 quad lights,
50000 surfels, 512x512, geometry frozen.  `--also batch` measures the
64-scene batch (scenes of 8..256 primitives, 5k..100k surfels, films 256^2
to 1024^2): frames/s of one batched frame (autotuned launch shapes, HIP
graph replay) against the same 64 frames run one scene at a time.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frame ms (transfer+solve+shade) & Gsdf-evals/s at ≤1/255 image err; grads/sec"

TRACE_KERNELS = ["visibility", "closure", "direct_e", "shade"]
ALL_KERNELS = ["cluster_keys", "seg_flags", "seg_begin", "cluster_build", "cluster_compact", "ff_count", "ff_emit",
               "visibility", "closure", "row_sum", "compact", "direct_e", "direct_sum", "jacobi", "grid_keys", "shade",
               "encode"]

PEAK_VALU_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# HBM bytes per dispatch of the product frame's kernels, read from two
# rocprofv3 --pmc runs (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0, scaled against a
# plain 256 MiB device copy).  The figures and the commands that made them
# are in profiles/product_hbm_traffic_pmc.md; they hold for the product
# workload only, so other workloads report null.
PRODUCT_TRAFFIC = {
    "visibility": {"bytes": 2.484e8, "profile": "profiles/product_hbm_traffic_pmc.md"},
    "closure": {"bytes": 1.379e7, "profile": "profiles/product_hbm_traffic_pmc.md"},
    "shade": {"bytes": 1.284e7, "profile": "profiles/product_hbm_traffic_pmc.md"},
}


def _cfg(name):
    from interflect_amd import scenes

    if name == "product":
        return {"text": scenes.product_literal(), "res": 8, "sample": None,
                "workload": "product scene: sphere, bowl = sphere - sphere - 6 drilled cylinders, torus R.4 r.12, "
                            "capsule len .8 r.15, box floor + 2 box walls; 2 sphere lights; 6 Lambertian materials; "
                            "640x480; 20000 surfels requested; cluster res 8; 12 bounces; beauty"}
    if name == "product_rad":
        return {"text": open(os.path.join(ROOT, "tests", "golden", "scenes", "product.rad")).read(), "res": 8,
                "sample": None, "workload": "scenes/product.rad of the reference, 640x480, 20000 surfels requested"}
    if name == "cornell_unit":
        return {"text": scenes.cornell_unit(), "res": 8, "sample": None,
                "workload": "unit Cornell box, 2 spheres r.3, quad light emit 15, 260x260, 20000 surfels"}
    if name == "stress":
        return {"text": scenes.stress(), "res": 8, "sample": {"rows": 200, "surfels": 4000, "pixel_rows": 4},
                "workload": "seeded stress scene (0x5EED): 256 random sphere/box/torus/capsule prims in a 10x4x10 "
                            "room, 25% smooth-union k=.05, 16 albedos, 4 point lights, 200000 surfels, 1920x1080"}
    raise SystemExit("unknown config " + name)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_frame(ctx, res, bounces):
    t = ctx.transfer(res)
    s = ctx.solve(bounces)
    r = ctx.shade("beauty", film=False, rgb8=True)
    return t, s, r


def cpu_baseline(cfg, surfels, threads, gpu_frame):
    """The oracle (C++ restatement of the reference CPU path) on the host
    cores, on the GPU run's surfels.  Small scenes: one full frame after one
    warm-up frame.  Large scenes: a bounded sample of every stage (transfer
    rows, direct-light surfels, film rows), each stage's rate applied to the
    GPU frame's evaluation counts for the frame estimate."""
    from oracle import pyoracle as po

    sc = po.Scene(cfg["text"], threads=threads)
    sc.set_surfels(*surfels)
    smp = cfg["sample"]
    if smp is None:
        sc.transfer(cfg["res"])
        sc.solve()
        t0 = time.perf_counter()
        tr = sc.transfer(cfg["res"])
        t1 = time.perf_counter()
        so = sc.solve()
        t2 = time.perf_counter()
        _, ev = sc.render("beauty")
        t3 = time.perf_counter()
        evals = tr["evals"] + so["evals"] + ev
        return {"value": evals / (t3 - t0) / 1e9, "unit": "Gsdf-evals/s", "cores": threads, "kind": "port",
                "sample": "one full frame of the same workload (transfer+solve+shade) on %d threads of %s" % (
                    threads, cpu_model()),
                "frame_ms": (t3 - t0) * 1e3,
                "stages_ms": {"transfer": (t1 - t0) * 1e3, "solve": (t2 - t1) * 1e3, "shade": (t3 - t2) * 1e3}}
    p, n, a, ar = surfels
    t0 = time.perf_counter()
    tr = sc.transfer(cfg["res"], rows=(0, smp["rows"]))
    t1 = time.perf_counter()
    k = smp["surfels"]
    _, st = sc.direct_diffuse(p[:k], n[:k], a[:k])
    t2 = time.perf_counter()
    y0 = sc.height // 2
    _, ev = sc.render("direct", rows=(y0, y0 + smp["pixel_rows"]), use_gi=False)
    t3 = time.perf_counter()
    rates = {"transfer": tr["evals"] / (t1 - t0), "solve": float(st.sum()) / (t2 - t1), "shade": ev / (t3 - t2)}
    evals = tr["evals"] + float(st.sum()) + ev
    est = {s: gpu_frame[s] / rates[s] * 1e3 for s in rates}
    return {"value": evals / (t3 - t0) / 1e9, "unit": "Gsdf-evals/s", "cores": threads, "kind": "port",
            "sample": "%d transfer rows, direct light of %d surfels, %d film rows (direct mode) of the same scene "
                      "and surfels, %.1f s on %d threads of %s" % (smp["rows"], k, smp["pixel_rows"], t3 - t0,
                                                                   threads, cpu_model()),
            "stage_rates_gevals_s": {s: r / 1e9 for s, r in rates.items()},
            "frame_ms": sum(est.values()), "frame_ms_is_estimate": True,
            "stages_ms": est}


def note(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measure(name, steps, warmup, with_cpu):
    import interflect_amd as ia

    cfg = _cfg(name)
    note(name, "placing surfels")
    ctx = ia.Context(0)
    info = ctx.load_scene(cfg["text"])
    # Surfel placement (outside the frame, as in the reference's stage
    # timers): on the device, then the host path for comparison.
    ctx.generate_surfels()  # warm-up (first launches, allocations)
    t0 = time.perf_counter()
    placed = ctx.generate_surfels()
    place_gpu = (time.perf_counter() - t0) * 1e3
    ctx.set_tuning("surfels_gpu", 0)
    t0 = time.perf_counter()
    placed_host = ctx.generate_surfels()
    place_host = (time.perf_counter() - t0) * 1e3
    ctx.set_tuning("surfels_gpu", 1)
    ctx.generate_surfels()
    if placed_host != placed:
        raise SystemExit("device and host surfel placement disagree")
    bounces = info["bounces"]
    note(name, "%d surfels placed; warm-up" % placed)
    for _ in range(warmup):
        run_frame(ctx, cfg["res"], bounces)
        note(name, "warm-up frame done")

    # One untimed instrumented frame: algorithmic FP32 operations per launch
    # of each tracing kernel (same traversal, counting stack).
    ctx.reset_kernel_times()
    ctx.set_counting(True)
    run_frame(ctx, cfg["res"], bounces)
    ctx.set_counting(False)
    ops = {k: ctx.kernel_ops(k) for k in TRACE_KERNELS}
    note(name, "instrumented frame done")

    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    stage = {"transfer": 0.0, "solve": 0.0, "shade": 0.0}
    sev = {"transfer": 0, "solve": 0, "shade": 0}
    t0 = time.perf_counter()
    for _ in range(steps):
        t, s, r = run_frame(ctx, cfg["res"], bounces)
        stage["transfer"] += t["ms"]
        stage["solve"] += s["ms"]
        stage["shade"] += r["ms"]
        sev["transfer"] += t["evals"]
        sev["solve"] += s["evals"]
        sev["shade"] += r["evals"]
        note(name, "timed frame done")
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    ms_per_step = wall * 1e3 / steps
    evals = sum(sev.values())

    ktime = {}
    for k in ALL_KERNELS:
        n, ms = ctx.kernel_time(k)
        if n:
            ktime[k] = {"launches": n, "avg_ms": ms / n, "total_ms_per_frame": ms / steps}
    dominant = max(ktime, key=lambda k: ktime[k]["total_ms_per_frame"])

    # Jacobi roofline over one solve (two launches per bounce, timed as
    # groups of 8 bounces).  Algorithmic bytes per bounce 8L + 84N + 24C:
    # idx+w per link; starts, direct, albedo, rad in/out and irr out per row
    # plus member index, rad and area for the cluster sums; Bc written and
    # read once per cluster.
    N, L, C = placed, t["links"], t["clusters"]
    jac_bytes = (8 * L + 84 * N + 24 * C) * s["iterations"]
    jk = ktime.get("jacobi")
    jac_roof = None
    if jk:
        gbs = jac_bytes / (jk["total_ms_per_frame"] * 1e-3) / 1e9
        jac_roof = {"kernel": "jacobi", "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": gbs / PEAK_HBM_GBS, "traffic": None, "bytes_per_solve": jac_bytes,
                    "ms_per_solve": jk["total_ms_per_frame"], "bounces": s["iterations"]}
    roof = None
    if dominant in TRACE_KERNELS and ops.get(dominant):
        tfl = ops[dominant] / (ktime[dominant]["avg_ms"] * 1e-3) / 1e12
        measured = PRODUCT_TRAFFIC.get(dominant) if name == "product" else None
        roof = {"kernel": dominant, "bound": "valu", "achieved": tfl, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": tfl / PEAK_VALU_TFLOPS, "traffic": measured["bytes"] if measured else None,
                "traffic_unit": "bytes per dispatch (rocprofv3 PMC)",
                "traffic_profile": measured["profile"] if measured else None, "ops_per_launch": ops[dominant]}
    elif dominant == "jacobi":
        roof = jac_roof
    trace_roofs = {}
    for k in TRACE_KERNELS:
        if k in ktime and ops.get(k):
            tfl = ops[k] / (ktime[k]["avg_ms"] * 1e-3) / 1e12
            trace_roofs[k] = {"achieved_tflops": tfl, "frac": tfl / PEAK_VALU_TFLOPS, "ops_per_launch": ops[k],
                              "avg_ms": ktime[k]["avg_ms"]}
    out = {
        "value": evals / wall / 1e9,
        "ms_per_step": ms_per_step,
        "config": {"workload": cfg["workload"], "surfels_placed": placed, "clusters": C, "links": L,
                   "candidates": t["candidates"], "bounces": s["iterations"], "frame_evals": evals / steps,
                   "width": info["width"], "height": info["height"]},
        "stages_ms": {k: v / steps for k, v in stage.items()},
        "stage_evals": {k: v / steps for k, v in sev.items()},
        "kernels": ktime,
        "dominant_kernel": dominant,
        "roofline": roof,
        "roofline_jacobi": jac_roof,
        "roofline_tracing": trace_roofs,
        "surfel_placement": {"surfels": placed, "gpu_ms": place_gpu, "host_ms": place_host,
                             "host_threads": max(1, min(16, os.cpu_count() or 1))},
    }
    if with_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        note(name, "CPU baseline on %d threads" % threads)
        cb = cpu_baseline(cfg, ctx.surfels(), threads, out["stage_evals"])
        cb["speedup_frame"] = cb["frame_ms"] / ms_per_step
        cb["speedup_throughput"] = out["value"] / cb["value"]
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    return out


def measure_grads(steps, warmup, with_cpu):
    """Gradient steps of the frozen-geometry adjoint (one step = solve to
    tolerance 1e-4 + beauty shade + dL/dfilm on the host + backward)."""
    import numpy as np

    import interflect_amd as ia
    from interflect_amd import scenes

    surfels, side, res = 50000, 512, 8
    note("diff", "target render (seed 7)")
    tgt = ia.Context(0)
    tgt.load_scene(scenes.diff_cornell(7, surfels=surfels, width=side, height=side))
    tgt.generate_surfels()
    tgt.transfer(res)
    tgt.solve(64)
    target = tgt.shade("beauty", rgb8=False)["film"]
    del tgt
    ctx = ia.Context(0)
    text = scenes.diff_cornell(8, surfels=surfels, width=side, height=side)
    info = ctx.load_scene(text)
    placed = ctx.generate_surfels()
    t = ctx.transfer(res)
    note("diff", "%d surfels, %d links; warm-up" % (placed, t["links"]))

    def step():
        s = ctx.solve(64)
        f = ctx.shade("beauty", rgb8=False)["film"]
        u = 2.0 * (f - target)
        g = ctx.backward(u)
        return s, g, float(np.sum((f.astype(np.float64) - target) ** 2))

    for _ in range(warmup):
        step()
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        s, g, loss = step()
        note("diff", "gradient step done (loss %.4g, %d forward bounces, %d adjoint sweeps)" % (
            loss, s["iterations"], g["iterations"]))
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    ktime = {}
    for k in ("direct_e", "direct_sum", "jacobi", "grid_keys", "shade", "shade_bwd", "adjoint", "adj_params"):
        n, ms = ctx.kernel_time(k)
        if n:
            ktime[k] = {"launches": n, "avg_ms": ms / n, "total_ms_per_step": ms / steps}
    N, L = placed, t["links"]
    out = {"value": steps / wall, "unit": "grads/s", "ms_per_step": wall * 1e3 / steps,
           "config": {"workload": "4.4-unit Cornell room, 8 materials (albedo U[.1,.9]^3, 24 params), 3 quad "
                                  "lights (scalar emission U[5,20], 3 params), start seed 8, target seed 7, "
                                  "50000 surfels requested, 512x512, cluster res 8, loss sum (film - target)^2 "
                                  "on the linear film",
                      "surfels_placed": N, "links": L, "clusters": t["clusters"],
                      "forward_bounces": s["iterations"], "adjoint_sweeps": g["iterations"]},
           "kernels": ktime}
    ak = ktime.get("adjoint")
    if ak:
        # Algorithmic bytes per sweep: row+weight per transposed link (8L);
        # per surfel cluster id, scale, ubar, rho, mu in/out, v out and one
        # read of v (80N); per segment its range and partial out/in (32).
        nseg = (L + 1023) // 1024 + t["clusters"]
        sweeps = max(1, g["iterations"])
        byt = (8 * L + 80 * N + 32 * nseg) * sweeps
        gbs = byt / (ak["total_ms_per_step"] * 1e-3) / 1e9
        out["roofline_adjoint"] = {"kernel": "adjoint", "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS,
                                   "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS, "traffic": None,
                                   "bytes_per_backward": byt, "ms_per_backward": ak["total_ms_per_step"],
                                   "sweeps": sweeps}
    if with_cpu:
        # The reference has no backward pass: its gradient is 2*27 + 1
        # forward evaluations (central differences).  One oracle forward
        # (light_surfels + solve + beauty render) on the same surfels is timed.
        from oracle import pyoracle as po

        threads = max(1, min(16, os.cpu_count() or 1))
        note("diff", "CPU forward on %d threads" % threads)
        sc = po.Scene(text, threads=threads)
        sc.set_surfels(*ctx.surfels())
        sc.transfer(res)
        t1 = time.perf_counter()
        sc.solve(64)
        sc.render("beauty")
        fwd = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": 1.0 / (fwd * 55), "unit": "grads/s", "cores": threads, "kind": "port",
                               "sample": "one oracle forward (light_surfels + solve + beauty render, %.2f s on %d "
                                         "threads of %s) x 55 evaluations of central differences over 27 "
                                         "parameters" % (fwd, threads, cpu_model()),
                               "forward_ms": fwd * 1e3}
        out["cpu_baseline"]["speedup"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def measure_turntable(steps, with_cpu, frames=12):
    """scenes/product.rad turntable (src/main.rs:492-520): one solve, then
    `frames` orbit views in one launch; ms per view."""
    import interflect_amd as ia

    text = open(os.path.join(ROOT, "tests", "golden", "scenes", "product.rad")).read()
    ctx = ia.Context(0)
    info = ctx.load_scene(text)
    ctx.generate_surfels()
    ctx.transfer(8)
    ctx.solve()
    cams = ctx.turntable_cameras(frames)
    ctx.shade_views(cams, "beauty", film=False)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ctx.shade_views(cams, "beauty", film=False)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    out = {"value": frames * 1e3 / ms, "unit": "views/s", "ms_per_step": ms, "ms_per_view": ms / frames,
           "evals_per_view": r["evals"] / frames,
           "config": {"workload": "scenes/product.rad, %d turntable views at %dx%d from one solve, beauty + RGB8"
                                  % (frames, info["width"], info["height"])},
           "reference_published_ms_per_view": 80.0}
    if with_cpu:
        from oracle import pyoracle as po

        threads = max(1, min(16, os.cpu_count() or 1))
        sc = po.Scene(text, threads=threads)
        sc.set_surfels(*ctx.surfels())
        sc.transfer(8)
        sc.solve()
        t1 = time.perf_counter()
        sc.render("beauty", camera=cams[1])
        cpu = (time.perf_counter() - t1) * 1e3
        out["cpu_baseline"] = {"value": 1e3 / cpu, "unit": "views/s", "cores": threads, "kind": "port",
                               "sample": "one oracle beauty view of the same turntable on %d threads of %s"
                                         % (threads, cpu_model()),
                               "ms_per_view": cpu, "speedup": cpu / (ms / frames)}
    return out


def measure_batch(steps, n_scenes=64, surfel_div=1, side_div=1):
    """Batched frames of `n_scenes` seeded scenes vs sequential single-scene
    frames (the divisors shrink surfels and films of the same scenes)."""
    import numpy as np

    import interflect_amd as ia
    from interflect_amd import scenes

    texts = [scenes.batched_scene(j, surfel_div, side_div) for j in range(n_scenes)]
    b = ia.Batch(0)
    for t in texts:
        b.add_scene(t, 8)
    note("batch", "placing surfels for %d scenes" % n_scenes)
    b.generate_surfels()
    note("batch", "first eager frame")
    first = b.run()
    note("batch", "eager frame %.0f ms; autotuning" % first)
    best = b.autotune()
    note("batch", "autotuned eager frame %.0f ms; graph capture" % best)
    b.set_graph(True)
    b.run()  # capture + first replay
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
        note("batch", "graph replay done")
    batched_ms = (time.perf_counter() - t0) * 1e3 / steps
    b.set_graph(False)
    stats = [b.stats(s) for s in range(n_scenes)]
    evals = sum(st["evals_transfer"] + st["evals_solve"] + st["evals_shade"] for st in stats)
    # Sequential: one context, one scene at a time; surfel placement is
    # outside the timed region, as in the batch.
    ctx = ia.Context(0)
    seq_ms = 0.0
    same = 0
    for j, t in enumerate(texts):
        ctx.load_scene(t)
        ctx.generate_surfels()
        ctx.transfer(8)  # warm the buffers of this size
        t1 = time.perf_counter()
        ctx.transfer(8)
        ctx.solve()
        r = ctx.shade("beauty", film=False, rgb8=True)
        seq_ms += (time.perf_counter() - t1) * 1e3
        same += int(np.array_equal(r["rgb8"], b.image(j, film=False)["rgb8"]))
        if j % 8 == 7:
            note("batch", "sequential %d/%d scenes" % (j + 1, n_scenes))
    return {"value": 1e3 / batched_ms, "unit": "batched frames/s", "ms_per_step": batched_ms,
            "sequential_ms": seq_ms, "speedup_vs_sequential": seq_ms / batched_ms,
            "first_eager_ms": first, "autotuned_eager_ms": best, "gsdf_evals_per_s": evals / batched_ms / 1e6,
            "images_identical_to_single_scene": "%d/%d" % (same, n_scenes),
            "config": {"workload": "%d seeded scenes (seeds 1000+j): 8..256 random primitives, 5k..100k surfels "
                                   "/ %d, films 256^2..1024^2 / %d^2, as the stress scene otherwise; cluster res 8; "
                                   "beauty + RGB8" % (n_scenes, surfel_div, side_div),
                       "surfels": int(sum(st["surfels"] for st in stats)),
                       "links": int(sum(st["links"] for st in stats)),
                       "candidates": int(sum(st["candidates"] for st in stats)),
                       "pixels": int(sum(st["width"] * st["height"] for st in stats))}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="product",
                    choices=["product", "product_rad", "cornell_unit", "stress", "diff", "batch", "batch_small"])
    ap.add_argument("--also", default="", help="comma-separated secondary configs (e.g. stress)")
    ap.add_argument("--also-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (HBM-counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this path has no multi-GPU mode (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    if args.config in ("diff", "batch", "batch_small"):
        # one secondary workload on its own (profiling runs)
        if args.config == "diff":
            m = measure_grads(args.steps, args.warmup, not args.no_cpu)
        else:
            m = measure_batch(args.steps, 64, *((20, 4) if args.config == "batch_small" else (1, 1)))
        print(json.dumps(dict(m, metric=METRIC, n_gpus=1, steps=args.steps, warmup=args.warmup)))
        return
    m = measure(args.config, args.steps, args.warmup, not args.no_cpu)
    line = {
        "metric": METRIC,
        "value": m["value"],
        "unit": "Gsdf-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded scene + deterministic Halton surfel placement (no dataset)",
    }
    line.update({k: v for k, v in m.items() if k not in ("value", "ms_per_step")})
    for extra in [x for x in args.also.split(",") if x]:
        if extra == "diff":
            line.setdefault("secondary", {})[extra] = measure_grads(args.also_steps, 1, not args.no_cpu)
        elif extra == "batch":
            line.setdefault("secondary", {})[extra] = measure_batch(args.also_steps)
        elif extra == "turntable":
            line.setdefault("secondary", {})[extra] = measure_turntable(max(3, args.also_steps), not args.no_cpu)
        elif extra == "batch_small":
            line.setdefault("secondary", {})[extra] = measure_batch(max(3, args.also_steps), 64, 20, 4)
        else:
            line.setdefault("secondary", {})[extra] = measure(extra, args.also_steps, 1, not args.no_cpu)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
