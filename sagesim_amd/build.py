"""Build the native libraries in-tree.

libsagestep.so  - HIP (gfx950) kernels + C ABI (include/sagestep.h), hipcc
libsagegraph.so - host C++ graph generators / CSR builder, g++ + OpenMP

Run `python -m sagesim_amd.build` (or call build()).  Every call recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("SAGESIM_OFFLOAD_ARCH", "gfx950")

STEP_SOURCES = ["tick_kernels.hip", "aux_kernels.hip", "hashmap.hip", "runtime.hip", "noalloc.hip", "ghosts.hip", "rtc.hip"]
# Packed-FP32 instructions are kept out of every device code object.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=off", "-Wall", "-Wno-unused-result", "-Wno-unused-function"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def embed_step_runtime():
    """step_rt.h (the device runtime of hipRTC step functions) as a C++ raw
    string, so libsagestep carries it and needs no header files at run time."""
    with open(os.path.join(CSRC, "step_rt.h")) as f:
        body = f.read()
    assert ")SSRT\"" not in body
    out = os.path.join(CSRC, "step_rt_embed.inc")
    with open(out + ".tmp", "w") as f:
        f.write('R"SSRT(' + body + ')SSRT"\n')
    os.replace(out + ".tmp", out)


def build_step_lib(verbose=True):
    embed_step_runtime()
    srcs = [os.path.join(CSRC, s) for s in STEP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    out = os.path.join(HERE, "libsagestep.so")
    tmp = out + ".tmp"
    _run(["hipcc", *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs,
          "-o", tmp, "-lhiprtc"])
    os.replace(tmp, out)
    return out


def build_graph_lib():
    out = os.path.join(HERE, "libsagegraph.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-march=x86-64-v3",
          "-Wall", os.path.join(CSRC, "graphgen.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build():
    build_graph_lib()
    build_step_lib()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "graph" in what:
        build_graph_lib()
    if "all" in what or "step" in what:
        build_step_lib()
