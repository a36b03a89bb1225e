"""MI355X-native (gfx950) CG2D free-surface solve, GAD DST3 tracer advection and exch2 halo exchange for the
LLC90 ocean time step of mitgcm_jax: hand-written HIP kernels behind a C ABI (include/mjamd.h, libmjamd.so),
exposed as PyTorch-ROCm ops with the reference's argument signatures.

    from mit_jax_amd import cg2d, gad, exchange
"""

from .layout import LLC90, Layout, interior_mask  # noqa: F401

__version__ = "0.2.0"


def release():
    """Destroys the library objects this package caches (halo plans, DST3 plans, device grid fields) after the
    device has finished with them. Call before the process exits when something else (a profiler, a HIP runtime
    torn down early) runs at exit: the library then holds no device objects there. Later calls rebuild the caches
    on use."""
    import sys
    import torch
    if torch.cuda.is_available() and torch.cuda.is_initialized():
        torch.cuda.synchronize()
    for name in ("exchange", "gad"):
        mod = sys.modules.get(f"{__name__}.{name}")
        if mod is not None:
            plans = list(mod._plans.values())
            mod._plans.clear()
            for p in plans:
                p.close()
    sfp = sys.modules.get(f"{__name__}.solve_for_pressure")
    if sfp is not None:
        sfp._grid_cache.clear()
