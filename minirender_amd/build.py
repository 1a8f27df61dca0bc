"""Build recipe for libmrg.so (gfx950). Compiles in-tree: minirender_amd/lib/libmrg.so.

Flags: -ffp-contract=off everywhere (the kernels must round every product and sum separately, as the
FP32 twin does), and packed-fp32 instructions disabled for device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libmrg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MRG_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fhip-fp32-correctly-rounded-divide-sqrt"]
SOURCES = [("mrg_kernels.hip", True), ("mrg_api.cpp", False), ("mrg_bvh.cpp", False), ("mrg_energy.cpp", False), ("mrg_envmap.cpp", False)]


def _compile(src, device):
    obj = os.path.join(OUT_DIR, "obj", os.path.basename(src) + ".o")
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    cmd = [HIPCC] + COMMON + (DEVICE + ["-x", "hip"] if device else ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include"]) + \
        ["-I", os.path.join(HERE, "..", "include"), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


SESSION_CHECK = os.path.join(OUT_DIR, "session_check")


def build_session_check(verbose=False):
    """The C++ consumer of include/mrg_session.hpp used by the tests (host code only, g++)."""
    src = os.path.join(HERE, "..", "tests", "cpp", "session_check.cpp")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(HERE, "..", "include"), src,
           "-o", SESSION_CHECK, "-L", OUT_DIR, "-lmrg", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"session_check build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", SESSION_CHECK)
    return SESSION_CHECK


PROBE_LIB = os.path.join(OUT_DIR, "libmrg_probe.so")


def build_probe(verbose=False):
    """libmrg_probe.so: the cache-hierarchy read-bandwidth probe behind bench.py --workload cache
    (measurement only; the render path does not use it)."""
    src = os.path.join(CSRC, "probe", "mrg_probe.hip")
    cmd = [HIPCC] + COMMON + DEVICE + ["-shared", "-o", PROBE_LIB, "-x", "hip", src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", PROBE_LIB)
    return PROBE_LIB


RENDER_CLI = os.path.join(OUT_DIR, "mrg_render")


def build_cli(verbose=False):
    """mrg_render: the --render command line (scene JSON + OBJ in, EXR + PNG out) over libmrg."""
    host = os.path.join(CSRC, "host")
    srcs = [os.path.join(host, f) for f in ("render_cli.cpp", "scene_io.cpp", "image_io.cpp")]
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(HERE, "..", "include"), "-I", host] + srcs + \
          ["-o", RENDER_CLI, "-L", OUT_DIR, "-lmrg", "-lz", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mrg_render build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", RENDER_CLI)
    return RENDER_CLI


if __name__ == "__main__":
    build(verbose=True)
    build_session_check(verbose=True)
    build_cli(verbose=True)
    build_probe(verbose=True)
    sys.exit(0)
