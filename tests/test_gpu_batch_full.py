"""The 64-scene batch at full size (8..256 primitives, 5k..100k surfels and
256^2..1024^2 films per scene, about 3.1M surfels and 650M links in all)
against single-scene runs of the same scenes.

Property checked per scene: the batched frame reproduces, bit for bit, what
a single-scene context computes from the same scene text and surfels (CSR
starts, indices and weight bits; cluster / candidate / link counts;
evaluation counts; iterations and residual; radiosity; the linear film and
the RGB8 image), and a HIP graph replay of the batched frame reproduces the
eager one.  The smaller batches of test_gpu_batch.py check every scene
against the oracle as well; this one covers the full-size shapes (scene ids
in the sort keys, offsets past 2^28 links, LDS classes of every size).

The batched and the sequential frame times are printed for the record; the
speed-up is not asserted (a full-size scene already fills the GPU with its
link visibility, which both modes run with the same kernels).
"""
import sys
import time

import numpy as np
import pytest

import interflect_amd as ia
from interflect_amd import scenes

pytestmark = pytest.mark.gpu

N_SCENES = 64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.timeout(900)
def test_full_size_batch_equals_single_scene_runs():
    texts = [scenes.batched_scene(j) for j in range(N_SCENES)]
    b = ia.Batch(0)
    for t in texts:
        b.add_scene(t, 8)
    say("full batch: placing surfels for %d scenes" % N_SCENES)
    b.generate_surfels()
    eager = b.run()
    say("full batch: eager frame %.0f ms" % eager)
    stats = [b.stats(s) for s in range(N_SCENES)]
    links = sum(st["links"] for st in stats)
    say("full batch: %d surfels, %d links, %d pixels" % (
        sum(st["surfels"] for st in stats), links, sum(st["width"] * st["height"] for st in stats)))
    assert links > 0

    # graph replay of the same frame, every image identical to the eager one
    ref = [b.image(s, film=False)["rgb8"] for s in range(N_SCENES)]
    b.set_graph(True)
    b.run()  # capture and first replay
    t0 = time.perf_counter()
    b.run()
    batched_ms = (time.perf_counter() - t0) * 1e3
    for s in range(N_SCENES):
        assert np.array_equal(b.image(s, film=False)["rgb8"], ref[s]), s
    b.set_graph(False)
    say("full batch: graph replay %.0f ms" % batched_ms)

    ctx = ia.Context(0)
    seq_ms = 0.0
    for s, text in enumerate(texts):
        ctx.load_scene(text)
        ctx.generate_surfels()
        ctx.transfer(8)  # grow the buffers to this scene's sizes outside the timed frame
        t1 = time.perf_counter()
        tr = ctx.transfer(8)
        so = ctx.solve()
        sh = ctx.shade("beauty")
        seq_ms += (time.perf_counter() - t1) * 1e3
        st = stats[s]
        assert st["surfels"] == ctx.num_surfels(), s
        assert (st["clusters"], st["candidates"], st["links"]) == (tr["clusters"], tr["candidates"], tr["links"]), s
        assert (st["evals_transfer"], st["evals_solve"], st["evals_shade"]) == (tr["evals"], so["evals"],
                                                                               sh["evals"]), s
        assert st["iterations"] == so["iterations"], s
        assert bits(st["residual"]) == bits(so["residual"]), s
        starts, idx, w = b.transfer_arrays(s)
        s1, i1, w1, _, _ = ctx.transfer_arrays()
        assert np.array_equal(starts, s1), s
        assert np.array_equal(idx, i1), s
        assert np.array_equal(bits(w), bits(w1)), s
        del starts, idx, w, s1, i1, w1
        for a, c in zip(b.radiosity(s), ctx.radiosity()):
            assert np.array_equal(bits(a), bits(c)), s
        img = b.image(s)
        assert np.array_equal(bits(img["film"]), bits(sh["film"])), s
        assert np.array_equal(img["rgb8"], sh["rgb8"]), s
        if s % 8 == 7:
            say("full batch: %d/%d scenes equal their single-scene runs" % (s + 1, N_SCENES))
    say("full batch: batched frame %.0f ms, 64 single-scene frames %.0f ms (%.2fx)" % (
        batched_ms, seq_ms, seq_ms / batched_ms))
