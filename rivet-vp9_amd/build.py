"""Builds libvp9r.so in-tree: the C++ host side (parser, C ABI, frame
orchestration) and the HIP kernels for gfx950, with hipcc.

    python3 rivet-vp9_amd/build.py [-j N]

Every source is compiled on every call (the build is also the "does it
build" check); objects go to rivet-vp9_amd/_build/.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvp9r.so")

HOST_SOURCES = [
    "csrc/host/parser.cpp",
    "csrc/host/capi_parser.cpp",
]
# compiled as HIP (device code, or host code that launches kernels)
HIP_SOURCES = [
    "csrc/gpu/itx.hip",
    "csrc/gpu/mc.hip",
    "csrc/gpu/intra.hip",
    "csrc/gpu/lf.hip",
    "csrc/gpu/lf_cone.hip",
    "csrc/gpu/frame.hip",
    "csrc/gpu/pixel.hip",
    "csrc/host/recon.cpp",
    "csrc/host/capi_seam.cpp",
    "csrc/host/decoder.cpp",
    "csrc/host/server.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-missing-braces", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc")]
# gfx950 device code; packed-FP32 instructions disabled; the first 16 dwords
# of kernel arguments preloaded into SGPRs (the code keeps the loading
# preamble for a dispatcher that does not preload).
GPU = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
       "-munsafe-fp-atomics", "-mllvm", "-amdgpu-kernarg-preload-count=16"]


# Per-file flags.  lf.hip and intra.hip (whose k_wave inlines it): the loop filter keeps a 16-sample window of each
# line in a register array that the filters update in place under per-lane
# branches; promoted to one vector register, every partial update costs a
# copy of the whole window (16 moves per edge), so the promotion is off.
FILE_FLAGS = {
    "csrc/gpu/lf.hip": ["-mllvm", "-disable-promote-alloca-to-vector"],
    "csrc/gpu/intra.hip": ["-mllvm", "-disable-promote-alloca-to-vector"],  # k_wave runs the loop filter too
    "csrc/gpu/lf_cone.hip": ["-mllvm", "-disable-promote-alloca-to-vector"],
}


def compile_one(src, is_hip):
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    cmd = [HIPCC] + COMMON + (GPU + ["-x", "hip"] if is_hip else []) + FILE_FLAGS.get(src, []) + \
        ["-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 or "error" in r.stderr:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    work = [(s, False) for s in HOST_SOURCES] + [(s, True) for s in HIP_SOURCES]
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda a: compile_one(*a), work))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC"] + (GPU if HIP_SOURCES else []) + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build(a.j))
