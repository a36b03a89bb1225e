"""MI355X-native (gfx950) CG2D free-surface solve, GAD DST3 tracer advection and exch2 halo exchange for the
LLC90 ocean time step of mitgcm_jax: hand-written HIP kernels behind a C ABI (include/mjamd.h, libmjamd.so),
exposed as PyTorch-ROCm ops with the reference's argument signatures.

    from mit_jax_amd import cg2d, gad, exchange
"""

from .layout import LLC90, Layout, interior_mask  # noqa: F401

__version__ = "0.1.0"
