"""Build libhptt_amd.so in-tree with hipcc for gfx950.

    python -m hptt_amd.build          # or hptt_amd.build.build()

Device sources (*.hip) are compiled for gfx950 with FP contraction off and
packed-FP32 instruction selection disabled; host sources (*.cpp) as plain C++.
Objects go to hptt_amd/_build/, the library to hptt_amd/libhptt_amd.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhptt_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-fno-gpu-rdc",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str, build_dir: str = BUILD, defines=()) -> str:
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-D" + d for d in defines]
    if src.endswith(".hip"):
        cmd += DEVICE
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose: bool = False, defines=(), out: str = LIB) -> str:
    """Compile every source and link `out`.  `defines` (NAME or NAME=VALUE)
    are passed to every compile; objects of such builds go to their own
    _build_<defines> folder."""
    build_dir = BUILD if not defines else BUILD + "_" + "_".join(
        d.replace("=", "") for d in defines)
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")), 16)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir, defines), srcs))
    LIBOUT = out
    tmp = LIBOUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + [
        "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIBOUT)
    if verbose:
        print("built", LIBOUT)
    return LIBOUT


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--out", default=LIB)
    a = ap.parse_args()
    build(verbose=True, defines=a.defines, out=a.out)
    sys.exit(0)
