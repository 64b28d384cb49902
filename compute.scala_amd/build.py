"""Builds libcsmi.so, the gfx950 back-end library, in-tree with hipcc.

Usage: python compute.scala_amd/build.py [-v]

Every HIP translation unit is compiled for gfx950 only, with packed FP32 instructions
disabled, and linked with hiprtc into compute.scala_amd/lib/libcsmi.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libcsmi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["runtime.cpp", "fused.cpp", "reduce.hip", "rng.hip", "gemm.hip", "mathcheck.hip", "mlp.hip",
           "stream.hip"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _embed_headers():
    """Writes the device headers that generated kernels include as C++ string literals."""
    for header, name in (("precise_math.h", "kPreciseMathSource"),):
        with open(os.path.join(CSRC, header)) as f:
            text = f.read().replace("#pragma once\n", "")
        with open(os.path.join(OBJ, header.replace(".h", ".inc")), "w") as f:
            f.write('static const char %s[] = R"csmi_src(%s)csmi_src";\n' % (name, text))


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + lang + COMMON + ["-I", OBJ, "-c", path, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return src, cmd, proc


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    _embed_headers()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    failed = False
    for src, cmd, proc in results:
        if verbose or proc.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        if proc.returncode != 0:
            failed = True
    if failed:
        raise RuntimeError("libcsmi compilation failed")
    objs = [os.path.join(OBJ, s + ".o") for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
    if proc.returncode != 0:
        raise RuntimeError("libcsmi link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
