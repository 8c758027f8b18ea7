"""Builds libdrape_hip.so in-tree with hipcc for gfx950 (no GPU needed to compile)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdrape_hip.so")
SOURCES = [
    "kernels/pcg.hip",
    "kernels/elements.hip",
    "kernels/twolevel.hip",
    "kernels/gnn.hip",
    "kernels/gnn_train.hip",
    "kernels/selfcontact.hip",
    "kernels/prolong.hip",
    "kernels/builder.hip",
    "host/context.cpp",
    "host/newton.cpp",
    "host/problem.cpp",
]
# -ffp-contract=off keeps FP32 arithmetic in the reference's operation order (no FMA contraction), which
# makes the PCG iterates bit-identical to the CPU emulation; divisions and square roots stay correctly
# rounded; packed FP32 instructions are disabled (some of their forms have given wrong results on this
# hardware without an error).
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles every source (in parallel, `jobs` processes; 0 = one per source up to the CPU count) and
    links the library."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(HERE, "_build", os.path.basename(src) + ".o")
        cmds.append([hipcc, *[f for f in FLAGS if f != "-shared"], "-x", "hip", "-c", path, "-o", obj])
        objs.append(obj)
    workers = jobs if jobs > 0 else max(1, min(len(cmds), os.cpu_count() or 1, 16))

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(compile_one, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
