"""Builds the HIP library libstructured_gpu.so in-tree with hipcc for gfx950.

  python -m structured_amd.build      (or structured_amd.build.build())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libstructured_gpu.so")
SOURCES = ["sg_api.cpp", "sg_kernels.hip", "sg_linalg.hip", "sg_solver.cpp", "sg_mlp.hip",
           "sg_mlp_api.cpp", "sg_inversion.cpp", "sg_cilu.hip", "sg_explicit.hip",
           "sg_io.cpp", "sg_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fopenmp"]
# Per-source additions.  The chunked-ILU kernels run one work-group per chunk
# (a few dozen on the whole chip), so their occupancy does not matter and their
# level loop is a latency chain: the max-ILP scheduler issues a level's LDS
# reads together and interleaves its FMA chains, where the default
# occupancy-driven schedule serialises them.
SOURCE_FLAGS = {"sg_cilu.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]}


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC, "-x", "hip", *FLAGS, *SOURCE_FLAGS.get(src, []), "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join("---- %s\n%s" % f for f in failed)
        raise RuntimeError("hipcc failed:\n" + msg)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    shutil.rmtree(objdir, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
