"""Benchmark of the transfer -> solve -> shade frame on one MI355X.

One step = one frame on resident surfels: transfer (H2D of the surfels,
clusters, candidate culls, cone-traced link visibility, closure, CSR) +
solve (direct E, Jacobi) + shade (grid build, per-pixel pass, encode, D2H of
the RGB8 image).  Surfel placement, parsing and the BVH build precede the
timed region, as in the reference's stage timers (src/main.rs:432-534).

Prints one JSON line.  `value` is Gsdf-evals/s under the reference's
counting rule (trace step +1, hit normal +4, shadow step +1) over the whole
frame; ms_per_step is the frame time.
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

# Kernel -> reference stage.  Tracing kernels are FP32-VALU bound; the Jacobi
# gather is HBM bound.
TRACE_KERNELS = ["visibility", "closure", "direct_e", "shade"]
ALL_KERNELS = ["cluster_keys", "seg_flags", "seg_begin", "cluster_build", "cluster_compact", "ff_count", "ff_emit",
               "visibility", "closure", "row_sum", "compact", "direct_e", "aggregate", "spmv", "grid_keys", "shade",
               "encode"]

PEAK_VALU_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

CONFIGS = {
    "product": {"scene": "tests/golden/scenes/product.rad", "res": 8,
                "workload": "product.rad (7 SDF objects incl. CSG, 2 quad lights, 6 materials), 640x480, "
                            "20000 surfels requested, cluster res 8, Jacobi cap 16, beauty"},
}


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


def cpu_baseline(scene_text, res, threads):
    """The oracle (C++ restatement of the reference CPU path) on the host cores:
    one full frame of the same workload, after one untimed warm-up frame."""
    from oracle import pyoracle as po

    sc = po.Scene(scene_text, threads=threads)
    sc.generate()
    sc.transfer(res)
    sc.solve()
    t0 = time.perf_counter()
    tr = sc.transfer(res)
    t1 = time.perf_counter()
    so = sc.solve()
    t2 = time.perf_counter()
    _, ev = sc.render("beauty")
    t3 = time.perf_counter()
    evals = tr["evals"] + so["evals"] + ev
    frame_ms = (t3 - t0) * 1e3
    return {"value": evals / (t3 - t0) / 1e9, "unit": "Gsdf-evals/s", "cores": threads, "kind": "port",
            "sample": "one full frame (transfer+solve+shade) of the same workload on %d threads of %s" % (
                threads, cpu_model()),
            "frame_ms": frame_ms, "stages_ms": {"transfer": (t1 - t0) * 1e3, "solve": (t2 - t1) * 1e3,
                                                "shade": (t3 - t2) * 1e3}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="product", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this path has no multi-GPU mode (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import interflect_amd as ia

    cfg = CONFIGS[args.config]
    text = open(os.path.join(ROOT, cfg["scene"])).read()
    ctx = ia.Context(0)
    info = ctx.load_scene(text)
    placed = ctx.generate_surfels()
    bounces = info["bounces"]
    for _ in range(args.warmup):
        run_frame(ctx, cfg["res"], bounces)

    # One untimed instrumented frame: algorithmic FP32 operations per launch
    # of each tracing kernel (same traversal, counting stack).
    ctx.reset_kernel_times()
    ctx.set_counting(True)
    run_frame(ctx, cfg["res"], bounces)
    ctx.set_counting(False)
    ops = {k: ctx.kernel_ops(k) for k in TRACE_KERNELS}

    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    stage = {"transfer": 0.0, "solve": 0.0, "shade": 0.0}
    evals = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t, s, r = run_frame(ctx, cfg["res"], bounces)
        stage["transfer"] += t["ms"]
        stage["solve"] += s["ms"]
        stage["shade"] += r["ms"]
        evals += t["evals"] + s["evals"] + r["evals"]
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    ms_per_step = wall * 1e3 / args.steps

    ktime = {}
    for k in ALL_KERNELS:
        n, ms = ctx.kernel_time(k)
        if n:
            ktime[k] = {"launches": n, "avg_ms": ms / n, "total_ms_per_frame": ms / args.steps}
    dominant = max(ktime, key=lambda k: ktime[k]["total_ms_per_frame"])

    # SpMV roofline: algorithmic bytes per Jacobi bounce 8L + 64N + 12C for the
    # gather (idx+w per link; starts, direct, albedo, rad in/out, irr out per
    # row; one Bc read per cluster).
    N, L, C = placed, t["links"], t["clusters"]
    spmv_bytes = 8 * L + 64 * N + 12 * C
    sp = ktime.get("spmv")
    spmv_roof = None
    if sp:
        gbs = spmv_bytes / (sp["avg_ms"] * 1e-3) / 1e9
        spmv_roof = {"kernel": "spmv", "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": gbs / PEAK_HBM_GBS, "traffic": None, "bytes_per_launch": spmv_bytes}

    roof = None
    if dominant in TRACE_KERNELS and ops.get(dominant):
        tfl = ops[dominant] / (ktime[dominant]["avg_ms"] * 1e-3) / 1e12
        roof = {"kernel": dominant, "bound": "valu", "achieved": tfl, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": tfl / PEAK_VALU_TFLOPS, "traffic": None, "ops_per_launch": ops[dominant]}
    elif dominant == "spmv":
        roof = spmv_roof
    trace_roofs = {}
    for k in TRACE_KERNELS:
        if k in ktime and ops.get(k):
            tfl = ops[k] / (ktime[k]["avg_ms"] * 1e-3) / 1e12
            trace_roofs[k] = {"achieved_tflops": tfl, "frac": tfl / PEAK_VALU_TFLOPS, "ops_per_launch": ops[k],
                              "avg_ms": ktime[k]["avg_ms"]}

    frame_evals = evals / args.steps
    out = {
        "metric": METRIC,
        "value": evals / wall / 1e9,
        "unit": "Gsdf-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: scene file + deterministic Halton surfel placement (no dataset)",
        "config": {"workload": cfg["workload"], "surfels_placed": placed, "clusters": C, "links": L,
                   "bounces": s["iterations"], "frame_evals": frame_evals},
        "stages_ms": {k: v / args.steps for k, v in stage.items()},
        "kernels": ktime,
        "dominant_kernel": dominant,
        "roofline": roof,
        "roofline_spmv": spmv_roof,
        "roofline_tracing": trace_roofs,
    }
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        cb = cpu_baseline(text, cfg["res"], threads)
        cb["speedup_frame"] = cb["frame_ms"] / ms_per_step
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
