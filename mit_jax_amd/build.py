"""Builds libmjamd.so (the gfx950 HIP library of this package) in-tree with hipcc.

    python -m mit_jax_amd.build

All sources are compiled with -ffp-contract=off (no fused multiply-add: the FP64 results are the IEEE results of
the written operation order, bitwise comparable with the CPU oracle) and without packed-FP32 instructions.
"""

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ["mj_halo.hip", "mj_cg2d.hip", "mj_gad.hip", "mj_adjoint.hip", "mj_mixed.hip", "mj_sfp.hip", "mj_implicit_r.hip"]
OUT = HERE / "libmjamd.so"
ARCH = os.environ.get("MJ_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", f"--offload-arch={ARCH}", "-Wno-unused-command-line-argument"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (Path(c).exists() or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    srcs = [str(HERE / "csrc" / s) for s in SOURCES if (HERE / "csrc" / s).exists()]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
