"""Builds libmjamd.so (the gfx950 HIP library of this package) in-tree with hipcc.

    python -m mit_jax_amd.build

All sources are compiled with -ffp-contract=off (no fused multiply-add: the FP64 results are the IEEE results of
the written operation order, bitwise comparable with the CPU oracle) and without packed-FP32 instructions.
"""

import os
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ["mj_halo.hip", "mj_cg2d.hip", "mj_gad.hip", "mj_adjoint.hip", "mj_mixed.hip", "mj_sfp.hip", "mj_implicit_r.hip",
           "mj_flux.hip"]
OUT = HERE / "libmjamd.so"
ARCH = os.environ.get("MJ_OFFLOAD_ARCH", "gfx950")
COMPILE_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                 "-packed-fp32-ops", f"--offload-arch={ARCH}", "-Wno-unused-command-line-argument"]
# per-source extra flags: the register-resident GAD_IMPLICIT_R sweep unrolls its 50 rows fully
EXTRA_FLAGS = {"mj_implicit_r.hip": ["-mllvm", "-pragma-unroll-threshold=500000"]}


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (Path(c).exists() or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")


def build_library(verbose=False):
    """Compiles each source to an object (in parallel, with its own extra flags), then links libmjamd.so."""
    from concurrent.futures import ThreadPoolExecutor
    srcs = [s for s in SOURCES if (HERE / "csrc" / s).exists()]
    with tempfile.TemporaryDirectory(prefix="mjamd_build_") as d:
        objs = [str(Path(d) / (Path(s).stem + ".o")) for s in srcs]
        cmds = [[hipcc(), *COMPILE_FLAGS, *EXTRA_FLAGS.get(s, []), "-c", "-o", o, str(HERE / "csrc" / s)]
                for s, o in zip(srcs, objs)]
        with ThreadPoolExecutor(min(len(cmds), 8)) as pool:
            list(pool.map(lambda c: _run(c, verbose), cmds))
        tmp = OUT.with_suffix(".so.tmp")
        _run([hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *objs], verbose)
        os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
