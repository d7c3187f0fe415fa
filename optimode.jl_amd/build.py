"""Build libomx.so (HIP, gfx950) in-tree with hipcc.

Usage: python optimode.jl_amd/build.py [--jobs N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "omx", "libomx.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["operator.hip", "operator3d.hip", "block.hip", "dense.hip", "fields.hip", "solver.hip", "adjoint.hip", "krylov.hip", "refine.hip", "smooth.hip", "tensor.hip", "eme.hip",
           "warmnet.hip", "noalloc.hip", "batch.cpp", "capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled on purpose: the compiler must not form
# v_pk_{add,mul,fma}_f32 from scalar or float2 code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-I", CSRC, "-I", INCLUDE, "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def compile_one(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(jobs=8, verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = False
    for src, obj, r in results:
        if r.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {src} ---\n{r.stdout}{r.stderr}\n")
        elif verbose and (r.stderr.strip()):
            sys.stderr.write(f"--- {src} (warnings) ---\n{r.stderr}\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + [o for _, o, _ in results] + ["-Wl,-soname,libomx.so"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
