"""Build step for libinterflect_hip.so (gfx950) and the oracle library.

`python -m interflect_amd.build` or `__graft_entry__.build()`.  Objects go to
interflect_amd/_build/ (git-ignored); the shared library lands in-tree next
to this file so it travels with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libinterflect_hip.so")
ARCH = os.environ.get("IFX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["engine.hip", "shade_wave_lds.hip", "shade_wave_glb.hip", "shade_lane.hip", "shade_count.hip", "probe.hip",
           "adjoint.hip", "batch.hip", "surfels.hip", "host_scene.cpp", "host_surfels.cpp", "capi.cpp"]

# Bit-parity with the reference needs IEEE f32: no contraction to FMA,
# correctly rounded division and sqrt, denormals kept.  Packed-FP32 VALU
# forms are disabled for all device code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return "hipcc"


def _stale(src, obj):
    """An object is stale when it is missing or older than its source or
    any header of the source tree."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps.append(os.path.join(ROOT, "include", "interflect.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, incremental=False):
    obj = os.path.join(OUT, src + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    else:
        cmd += ["-x", "hip"] + DEVICE
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build_lib(verbose=True, incremental=False):
    """Compile every source (incremental=True: only those whose object is
    stale) and link the library."""
    os.makedirs(OUT, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=True):
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", os.path.join(ROOT, "oracle", "liboracle.so"))


def main():
    build_lib(incremental="--incremental" in sys.argv)
    if "--no-oracle" not in sys.argv:
        build_oracle()


if __name__ == "__main__":
    main()
