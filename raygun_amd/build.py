"""Builds raygun_amd/librgrt.so from raygun_amd/csrc with hipcc for gfx950.

Each .hip source compiles to an object under raygun_amd/_build (in parallel)
and the objects link into one shared library exporting the C ABI of
include/rgrt.h. Packed fp32 instructions are disabled for all device code.
Then the C++ host side over that ABI (raygun_amd/host) is built with g++:
libraygun_host.so and the scene_demo program that drives it.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librgrt.so")
HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libraygun_host.so")
DEMO = os.path.join(HOST, "scene_demo")
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src, build_dir=BUILD, extra=()):
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + list(extra) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, out=LIB, build_dir=BUILD, extra=()):
    """Compiles every source and links librgrt.so (always rebuilds). `out`,
    `build_dir` and `extra` (more hipcc flags, e.g. -D tuning macros) make
    another build of the library elsewhere."""
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(lambda src: _compile(src, build_dir, extra), srcs))
    if verbose:
        for obj, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if out == LIB:
        build_host()
    return out


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_host():
    """The C++ host mirror (libraygun_host.so, linked to librgrt.so next to
    it) and scene_demo."""
    flags = ["-std=c++17", "-O2", "-Wall", "-Wextra"]
    tmp = HOST_LIB + ".tmp"
    _run([CXX] + flags + ["-fPIC", "-shared", os.path.join(HOST, "raygun_rt.cpp"), "-o", tmp, "-L" + HERE, "-lrgrt",
                          "-Wl,-rpath,$ORIGIN"], "host library build")
    os.replace(tmp, HOST_LIB)
    _run([CXX] + flags + [os.path.join(HOST, "scene_demo.cpp"), "-o", DEMO, "-L" + HERE, "-lraygun_host", "-lrgrt",
                          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"], "scene_demo build")
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
