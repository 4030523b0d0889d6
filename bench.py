"""Benchmark of the transfer -> solve -> shade frame on one MI355X.

One step = one frame on resident surfels: transfer (clusters, candidate
culls, cone-traced link visibility, hemisphere closure, CSR) + solve (direct
E, Jacobi) + shade (grid build, per-pixel pass, encode, D2H of the RGB8
image).  Surfel placement, parsing and the BVH build precede the timed
region, as in the reference's stage timers (src/main.rs:432-534); the
surfels are uploaded by the first transfer after they are placed, before
the timed region.

Prints one JSON line.  `value` is Gsdf-evals/s under the reference's
counting rule (trace step +1, hit normal +4, shadow step +1) over the whole
frame; ms_per_step is the frame time.  The default workload is the seeded
stress scene: 256 random sphere / box / torus / capsule primitives in a
10x4x10 room, 25 % in a smooth-union group, 16 albedos, 4 point lights,
200000 surfels requested, 1920x1080.  The timed frames run with per-kernel
events off; kernel times come from a separate profiled pass.  The same run
checks the frame:
  * repeat: a second frame's CSR, radiosity and RGB8 bytes against the first;
  * FP64: the solve's radiosity against a float64 Jacobi (same bounces) and
    the float64 fixed point, on the device's own links and clusters;
  * oracle (the CPU baseline leg): link lists of a window of transfer rows,
    direct light of a surfel sample, and beauty film rows rendered from the
    device's radiosity, compared with the device's; the same sample times
    the oracle on the host cores this process may use.
`--also` adds secondary workloads: product (the literal product scene),
product_rad (the reference's scenes/product.rad), cornell_unit, diff
(grads/s of the implicit adjoint), batch / batch_small (64-scene batches),
turntable.
"""
import argparse
import hashlib
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frame ms (transfer+solve+shade) & Gsdf-evals/s at ≤1/255 image err; grads/sec"

TRACE_KERNELS = ["visibility", "closure", "direct_e", "shade"]
ALL_KERNELS = ["morton", "cluster_keys", "seg_flags", "seg_begin", "cluster_build", "cluster_compact", "ff_count",
               "ff_emit", "vis_order", "visibility", "vis_scatter", "closure", "row_sum", "compact", "direct_e",
               "direct_sum", "jacobi", "grid_keys", "shade", "encode"]

PEAK_VALU_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def source_hash():
    """Hash of the kernel sources and build flags: a profiled HBM traffic
    figure is reported only for the build it was measured on."""
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "interflect_amd", "csrc")
    for name in sorted(os.listdir(csrc)):
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(name.encode() + f.read())
    for extra in (os.path.join(ROOT, "include", "interflect.h"), os.path.join(ROOT, "interflect_amd", "build.py")):
        with open(extra, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def profiled_traffic(workload):
    """HBM bytes per timed launch group from
    profiles/<workload>_hbm_traffic_pmc.md (lines `traffic <kernel> <bytes>`
    under `source <hash>`: the link visibility per transfer, both of its
    launches; the Jacobi pair per bounce), or {} when the profile is missing
    or was taken on other sources."""
    path = os.path.join(ROOT, "profiles", "%s_hbm_traffic_pmc.md" % workload)
    if not os.path.exists(path):
        return {}
    src, out = None, {}
    for line in open(path):
        f = line.strip().strip("`").split()
        if len(f) == 2 and f[0] == "source":
            src = f[1]
        elif len(f) == 3 and f[0] == "traffic":
            out[f[1]] = float(f[2])
    if src != source_hash():
        return {}
    return {k: {"bytes": v, "profile": os.path.relpath(path, ROOT), "source": src} for k, v in out.items()}


def host_cores():
    """CPUs this process may run on: affinity, capped by a cgroup CPU quota."""
    n = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = n
    quota = None
    for path in ("/sys/fs/cgroup/cpu.max",):
        try:
            q, per = open(path).read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
        except (OSError, ValueError):
            pass
    usable = aff if quota is None else max(1, min(aff, int(math.floor(quota + 1e-9))))
    return {"nproc": n, "affinity": aff, "cgroup_quota": quota, "threads": usable, "model": cpu_model()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


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
        return {"text": scenes.stress(), "res": 8,
                "sample": {"rows": 4000, "surfels": 50000, "film_rows": [135 * k for k in range(8)],
                           "rows_per_band": 3},
                "workload": "seeded stress scene (0x5EED): 256 random sphere/box/torus/capsule prims in a 10x4x10 "
                            "room, 25% smooth-union k=.05, 16 albedos, 4 point lights, 200000 surfels requested, "
                            "1920x1080, cluster res 8, 12 bounces, beauty"}
    raise SystemExit("unknown config " + name)


def run_frame(ctx, res, bounces):
    t = ctx.transfer(res)
    s = ctx.solve(bounces)
    r = ctx.shade("beauty", film=False, rgb8=True)
    return t, s, r


def image_close(a, b):
    """Fraction of pixels within 1/255 on every channel, and the worst channel difference."""
    diff = np.abs(a.astype(np.int32) - b.astype(np.int32)).max(axis=-1)
    return float(np.mean(diff <= 1)), int(diff.max()) if diff.size else 0


def frame_state(ctx):
    """Bytes of one frame's transfer (starts, idx, w, sky, row sums) and radiosity."""
    return list(ctx.transfer_arrays()) + list(ctx.radiosity())


def digest(arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).view(np.uint8).tobytes())
    return h.hexdigest()[:16]


def fp64_solve_check(ctx, iterations):
    """The device's radiosity against a float64 Jacobi on the device's own
    links W, cluster means P, direct light E and albedo rho:
    B <- E + rho (.) W (P B), from B = E (src/solve.rs:60-102), the same
    number of bounces, then to the float64 fixed point."""
    import scipy.sparse as sp

    starts, idx, w, _, _ = ctx.transfer_arrays()
    _, _, cl_area, _, mstart, members = ctx.clusters()
    p, n, rho, area = ctx.surfels()
    direct, irr, rad = ctx.radiosity()
    N, nc = len(starts) - 1, len(cl_area)
    W = sp.csr_matrix((w.astype(np.float64), idx.astype(np.int64), starts.astype(np.int64)), shape=(N, nc))
    rows = np.repeat(np.arange(nc), np.diff(mstart.astype(np.int64)))
    mem = members.astype(np.int64)
    P = sp.csr_matrix((area.astype(np.float64)[mem] / np.maximum(cl_area.astype(np.float64)[rows], 1e-12),
                       (rows, mem)), shape=(nc, N))
    E = direct.astype(np.float64)
    rho = rho.astype(np.float64)
    B = E.copy()
    for _ in range(iterations):
        B = E + rho * (W @ (P @ B))
    same = B.copy()
    for it in range(1000):
        Bn = E + rho * (W @ (P @ B))
        delta = float(np.max(np.abs(Bn - B)))
        B = Bn
        if delta < 1e-12:
            break
    ref = np.linalg.norm(same)
    return {"bounces": iterations, "rel_l2_vs_fp64_same_bounces": float(np.linalg.norm(rad - same) / ref),
            "rel_l2_vs_fp64_fixed_point": float(np.linalg.norm(rad - B) / np.linalg.norm(B)),
            "fp64_fixed_point_iterations": iterations + it + 1}


def cpu_baseline(cfg, gpu, threads, gpu_stage_evals):
    """The oracle (C++ restatement of the reference CPU path) on the host
    cores, on the device run's surfels, as the checker of that run and as
    the CPU baseline.

    Small scenes: one warm-up and one timed full frame (transfer + solve +
    beauty shade); the device's links, radiosity and image are compared
    with the oracle's.  Large scenes: a bounded sample of every stage, each
    compared with the device's result and timed: a window of transfer rows
    (link lists exact), the direct light of a surfel sample (rel-L2), and
    beauty film rows shaded from the device's radiosity (RGB8 within
    1/255); each stage's measured rate is applied to the device frame's
    evaluation counts for the frame estimate."""
    from oracle import pyoracle as po

    sc = po.Scene(cfg["text"], threads=threads)
    p, n, a, ar = gpu["surfels"]
    sc.set_surfels(p, n, a, ar)
    smp = cfg["sample"]
    sg, ig, wg, skyg, _ = gpu["csr"]
    if smp is None:
        sc.transfer(cfg["res"])
        sc.solve()
        t0 = time.perf_counter()
        tr = sc.transfer(cfg["res"])
        t1 = time.perf_counter()
        so = sc.solve()
        t2 = time.perf_counter()
        film, ev = sc.render("beauty")
        t3 = time.perf_counter()
        so_, io_, wo_, _, _ = sc.transfer_arrays()
        links_exact = bool(np.array_equal(sg, so_) and np.array_equal(ig, io_) and
                           np.array_equal(wg.view(np.uint32), wo_.view(np.uint32)))
        _, _, ro = sc.radiosity()
        rel = float(np.linalg.norm(gpu["rad"] - ro) / max(np.linalg.norm(ro), 1e-30))
        frac, worst = image_close(gpu["rgb8"], po.encode(film, sc.exposure, True))
        evals = tr["evals"] + so["evals"] + ev
        check = {"links_exact": links_exact, "rows_checked": int(len(sg) - 1), "radiosity_rel_l2": rel,
                 "image_frac_within_1": frac, "image_max_diff": worst, "pixels_checked": int(gpu["rgb8"].size // 3),
                 "iterations_equal": so["iterations"] == gpu["iterations"]}
        return {"value": evals / (t3 - t0) / 1e9, "unit": "Gsdf-evals/s", "cores": threads, "kind": "port",
                "sample": "one full frame of the same workload (transfer+solve+shade) on %d threads" % threads,
                "frame_ms": (t3 - t0) * 1e3,
                "stages_ms": {"transfer": (t1 - t0) * 1e3, "solve": (t2 - t1) * 1e3, "shade": (t3 - t2) * 1e3}}, check
    N = len(sg) - 1
    r0 = N // 2
    r1 = min(N, r0 + smp["rows"])
    t0 = time.perf_counter()
    tr = sc.transfer(cfg["res"], rows=(r0, r1))
    t1 = time.perf_counter()
    so_, io_, wo_, skyo, _ = sc.transfer_arrays()
    bad = 0
    for i in range(r0, r1):
        x, y = slice(sg[i], sg[i + 1]), slice(so_[i], so_[i + 1])
        if not (np.array_equal(ig[x], io_[y]) and np.array_equal(wg[x].view(np.uint32), wo_[y].view(np.uint32))):
            bad += 1
    sky_exact = bool(np.array_equal(skyg[r0:r1].view(np.uint32), skyo[r0:r1].view(np.uint32)))
    k = min(smp["surfels"], N)
    t2 = time.perf_counter()
    dco, st = sc.direct_diffuse(p[:k], n[:k], a[:k])
    t3 = time.perf_counter()
    drel = float(np.linalg.norm(gpu["direct"][:k] - dco) / max(np.linalg.norm(dco), 1e-30))
    sc.set_radiosity(gpu["direct"], gpu["irr"], gpu["rad"])
    H = gpu["rgb8"].shape[0]
    fracs, worst, px, ev_shade, t_shade = [], 0, 0, 0, 0.0
    for y0 in smp["film_rows"]:
        y1 = min(H, y0 + smp["rows_per_band"])
        t4 = time.perf_counter()
        film, ev = sc.render("beauty", rows=(y0, y1))
        t_shade += time.perf_counter() - t4
        ev_shade += ev
        img = po.encode(film[y0:y1], sc.exposure, True)
        f, wv = image_close(gpu["rgb8"][y0:y1], img)
        fracs.append(f * img.shape[0] * img.shape[1])
        px += img.shape[0] * img.shape[1]
        worst = max(worst, wv)
    rates = {"transfer": tr["evals"] / (t1 - t0), "solve": float(st.sum()) / (t3 - t2), "shade": ev_shade / t_shade}
    evals = tr["evals"] + float(st.sum()) + ev_shade
    est = {s: gpu_stage_evals[s] / rates[s] * 1e3 for s in rates}
    wall = (t1 - t0) + (t3 - t2) + t_shade
    check = {"links_exact": bad == 0, "rows_checked": r1 - r0, "rows_mismatched": bad, "closure_sky_exact": sky_exact,
             "direct_rel_l2": drel, "direct_surfels_checked": k, "image_frac_within_1": sum(fracs) / px,
             "image_max_diff": worst, "pixels_checked": px,
             "image_note": "film rows shaded by the oracle from the device's radiosity"}
    return {"value": evals / wall / 1e9, "unit": "Gsdf-evals/s", "cores": threads, "kind": "port",
            "sample": "transfer rows %d..%d, direct light of %d surfels, %d beauty film rows (from the device's "
                      "radiosity) of the same scene and surfels: %.1f s on %d threads" % (
                          r0, r1, k, px // gpu["rgb8"].shape[1], wall, threads),
            "stage_rates_gevals_s": {s: r / 1e9 for s, r in rates.items()},
            "frame_ms": sum(est.values()), "frame_ms_is_estimate": True, "stages_ms": est}, check


def note(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measure(name, steps, warmup, with_cpu, prof_frames=2):
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
    prims_ref = {k: ctx.kernel_ops(k + ".prims") for k in TRACE_KERNELS}
    # and one with the executed-evaluation tally of the timed traversal
    # (persistent tracing kernels: every lane of a shared traversal counted)
    ctx.reset_kernel_times()
    ctx.set_counting(2)
    run_frame(ctx, cfg["res"], bounces)
    ctx.set_counting(False)
    prims_exec = {k: ctx.kernel_ops(k + ".exec_prims") for k in ("visibility", "closure")}
    note(name, "instrumented frames done")

    # Timed frames: no per-kernel events.
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
    ms_per_step = wall * 1e3 / steps
    evals = sum(sev.values())

    # Repeat check: the last timed frame's bytes against one more frame.
    first = frame_state(ctx) + [r["rgb8"]]
    d1 = digest(first)
    t2, s2, r2 = run_frame(ctx, cfg["res"], bounces)
    second = frame_state(ctx) + [r2["rgb8"]]
    repeat = {"bit_identical": d1 == digest(second), "digest": d1,
              "arrays": "CSR starts/idx/w, closure sky, row sums, direct, irradiance, radiosity, RGB8",
              "counts_equal": (t2["links"], s2["iterations"], t2["evals"], s2["evals"], r2["evals"]) ==
                              (t["links"], s["iterations"], t["evals"], s["evals"], r["evals"])}
    note(name, "repeat frame %s" % ("identical" if repeat["bit_identical"] else "DIFFERS"))
    gpu = {"surfels": ctx.surfels(), "csr": first[:5], "direct": first[5], "irr": first[6], "rad": first[7],
           "rgb8": first[8], "iterations": s["iterations"]}
    del second
    fp64 = fp64_solve_check(ctx, s["iterations"])
    note(name, "fp64 solve check done (rel-L2 %.2e same bounces, %.2e fixed point)" % (
        fp64["rel_l2_vs_fp64_same_bounces"], fp64["rel_l2_vs_fp64_fixed_point"]))

    # Kernel times from a separate profiled pass.
    ctx.reset_kernel_times()
    ctx.set_profiling(True)
    for _ in range(prof_frames):
        run_frame(ctx, cfg["res"], bounces)
    ctx.set_profiling(False)
    ktime = {}
    for k in ALL_KERNELS:
        nl, ms = ctx.kernel_time(k)
        if nl:
            ktime[k] = {"launches": nl, "avg_ms": ms / nl, "total_ms_per_frame": ms / prof_frames}
    dominant = max(ktime, key=lambda k: ktime[k]["total_ms_per_frame"])
    note(name, "profiled pass done")

    # Jacobi roofline over one solve (two launches per bounce, timed as
    # groups of 8 bounces).  Algorithmic bytes per bounce 8L + 84N + 24C:
    # idx+w per link; starts, direct, albedo, rad in/out and irr out per row
    # plus member index, rad and area for the cluster sums; Bc written and
    # read once per cluster.
    N, L, C = placed, t["links"], t["clusters"]
    traffic = profiled_traffic(name)
    jac_bytes = (8 * L + 84 * N + 24 * C) * s["iterations"]
    jk = ktime.get("jacobi")
    jac_roof = None
    if jk:
        gbs = jac_bytes / (jk["total_ms_per_frame"] * 1e-3) / 1e9
        jt = traffic.get("jacobi")
        jac_roof = {"kernel": "jacobi", "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": gbs / PEAK_HBM_GBS, "traffic": jt["bytes"] * s["iterations"] if jt else None,
                    "traffic_profile": jt["profile"] if jt else None, "bytes_per_solve": jac_bytes,
                    "ms_per_solve": jk["total_ms_per_frame"], "bounces": s["iterations"]}
    roof = None
    if dominant in TRACE_KERNELS and ops.get(dominant):
        tfl = ops[dominant] / (ktime[dominant]["avg_ms"] * 1e-3) / 1e12
        measured = traffic.get(dominant)
        roof = {"kernel": dominant, "bound": "valu", "achieved": tfl, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": tfl / PEAK_VALU_TFLOPS, "traffic": measured["bytes"] if measured else None,
                "traffic_unit": "bytes per timed launch group (rocprofv3 PMC; the link visibility's group is "
                                "its first launch plus the parked tails' launch)",
                "traffic_profile": measured["profile"] if measured else None, "ops_per_launch": ops[dominant]}
    elif dominant == "jacobi":
        roof = jac_roof
    trace_roofs = {}
    for k in TRACE_KERNELS:
        if k in ktime and ops.get(k):
            tfl = ops[k] / (ktime[k]["avg_ms"] * 1e-3) / 1e12
            trace_roofs[k] = {"achieved_tflops": tfl, "frac": tfl / PEAK_VALU_TFLOPS, "ops_per_launch": ops[k],
                              "avg_ms": ktime[k]["avg_ms"], "prim_evals_reference": prims_ref.get(k)}
            if prims_exec.get(k):
                trace_roofs[k]["prim_evals_executed"] = prims_exec[k]
                trace_roofs[k]["executed_over_reference"] = prims_exec[k] / max(prims_ref.get(k) or 1, 1)
    evals_note = ("Gsdf-evals: scene-distance samples under the reference's counting rule (trace step +1, hit "
                  "normal +4, shadow step +1); every counted sample is executed once per lane.  Below a sample, "
                  "prim_evals_reference counts the primitive evaluations of the reference's per-lane BVH "
                  "traversal and prim_evals_executed those the timed kernel performs (its wavefront-shared "
                  "traversal evaluates each opened object on every lane).")
    out = {
        "value": evals / wall / 1e9,
        "ms_per_step": ms_per_step,
        "config": {"workload": cfg["workload"], "surfels_placed": placed, "clusters": C, "links": L,
                   "candidates": t["candidates"], "bounces": s["iterations"], "frame_evals": evals / steps,
                   "width": info["width"], "height": info["height"]},
        "stages_ms": {k: v / steps for k, v in stage.items()},
        "stage_evals": {k: v / steps for k, v in sev.items()},
        "repeat": repeat,
        "fp64_solve": fp64,
        "kernels": ktime,
        "kernels_note": "from %d profiled frames after the timed ones (per-kernel HIP events)" % prof_frames,
        "dominant_kernel": dominant,
        "roofline": roof,
        "roofline_jacobi": jac_roof,
        "roofline_tracing": trace_roofs,
        "evals_note": evals_note,
        "surfel_placement": {"surfels": placed, "gpu_ms": place_gpu, "host_ms": place_host},
    }
    if with_cpu:
        cores = host_cores()
        note(name, "CPU baseline and oracle checks on %d threads" % cores["threads"])
        cb, check = cpu_baseline(cfg, gpu, cores["threads"], out["stage_evals"])
        cb["host"] = cores
        cb["sample"] += " of %s (nproc %d, affinity %d, cgroup quota %s)" % (
            cores["model"], cores["nproc"], cores["affinity"], cores["cgroup_quota"])
        cb["speedup_frame"] = cb["frame_ms"] / ms_per_step
        cb["speedup_throughput"] = out["value"] / cb["value"]
        out["cpu_baseline"] = cb
        out["oracle_check"] = check
        out["image_ok"] = check["image_frac_within_1"] >= 0.999 and check["links_exact"]
        note(name, "oracle check: links exact %s, image %.5f within 1/255 (max %d)" % (
            check["links_exact"], check["image_frac_within_1"], check["image_max_diff"]))
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
    # Gradient rooms are placed on the host (bit-identical to the device
    # placement; one device placement of such a room once ran past its time
    # limit with no cause found), outside the timed region either way.
    tgt.set_tuning("surfels_gpu", 0)
    tgt.generate_surfels()
    tgt.transfer(res)
    tgt.solve(64)
    target = tgt.shade("beauty", rgb8=False)["film"]
    del tgt
    ctx = ia.Context(0)
    text = scenes.diff_cornell(8, surfels=surfels, width=side, height=side)
    info = ctx.load_scene(text)
    ctx.set_tuning("surfels_gpu", 0)
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

        threads = host_cores()["threads"]
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

        threads = host_cores()["threads"]
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
    ap.add_argument("--config", default="stress",
                    choices=["product", "product_rad", "cornell_unit", "stress", "diff", "batch", "batch_small"])
    ap.add_argument("--also", default="product,product_rad,cornell_unit,diff",
                    help="comma-separated secondary configs ('' for none)")
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
        "source_hash": source_hash(),
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
