"""Build libbempp_hip.so (all HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python build_hip.py [--jobs N]. Objects are compiled in parallel into _build/,
then linked into bempp_amd/lib/libbempp_hip.so.
"""

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "bempp_amd", "lib", "libbempp_hip.so")
SOURCES = ["common.hip", "regular.hip", "near.hip", "singular.hip", "potential.hip", "kbatch.hip", "adjoint.hip", "matvec.hip", "mfie.hip"] + [f"regular_op{k}.hip" for k in range(9)] + ["regular_dual0.hip", "regular_dual3.hip", "regular_sym0.hip", "regular_sym3.hip", "regular_upper.hip", "regular_mfie.hip", "regular_symsl0.hip", "regular_symsl3.hip", "regular_symsl7.hip", "regular_symsl8.hip", "regular_symsl6.hip", "regular_wide.hip"]
EXTRA_SOURCES = []

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + SRC,
    "-Wno-unused-result",
]


def sources():
    return [s for s in SOURCES + EXTRA_SOURCES if os.path.exists(os.path.join(SRC, s))]


def compile_one(name):
    obj = os.path.join(BUILD, name.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, name), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed for {name}")
    return obj


def build(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, sources()))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    print(build(args.jobs))
