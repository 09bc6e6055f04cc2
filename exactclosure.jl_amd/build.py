"""Build libexactclosure.so (gfx950) and the FP64 CPU oracle library.

    python exactclosure.jl_amd/build.py            # build the HIP library
    python exactclosure.jl_amd/build.py --oracle   # also the oracle (test infrastructure)

Each HIP translation unit is compiled with hipcc for gfx950 only, in
parallel, then linked into one shared library next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libexactclosure.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
    "-Wno-unused-value",
    # packed-FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    objs = []
    failed = False
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for src, obj, r in ex.map(compile_one, srcs):
            msg = (r.stdout + r.stderr).strip()
            # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
            msg = "\n".join(x for x in msg.splitlines() if "packed-fp32-ops" not in x and x.strip())
            if r.returncode != 0:
                failed = True
                print(f"[build] FAILED {os.path.basename(src)}\n{msg}", file=sys.stderr)
            elif verbose and msg:
                print(f"[build] {os.path.basename(src)}\n{msg}", file=sys.stderr)
            objs.append(obj)
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return os.path.join(ROOT, "oracle", "liboracle.so")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build_hip(verbose=a.v))
    if a.oracle:
        print(build_oracle())
