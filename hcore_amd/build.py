"""Build recipe of the product library hcore_amd/lib/libhcore_amd.so.

hipcc compiles the HIP kernels and the host orchestration for gfx950 only
and links one shared library exporting the C ABI declared in include/*.h.
Packed-FP32 instruction selection is disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libhcore_amd.so")
# every translation unit of csrc/, slowest first (they compile in parallel)
_FIRST = ["k_qrr_f64.hip", "k_qrf_f64.hip", "k_qrr_f32.hip", "k_qrf_f32.hip", "k_core_reg.hip"]
SOURCES = _FIRST + sorted(f for f in os.listdir(CSRC)
                          if f.endswith((".hip", ".cpp")) and f not in _FIRST)
ARCH = "gfx950"

FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument",
         # a value-returning function that can fall off its end is an error:
         # at -O3 the host code would run into the padding after it
         "-Werror=return-type",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def _compile(src, verbose):
    obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [_hipcc()] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on " + src)
    return obj


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
