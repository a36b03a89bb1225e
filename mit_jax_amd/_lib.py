"""ctypes binding of libmjamd.so, the HIP (gfx950) library behind every op of this package.

The library is loaded from the package directory (built in-tree by `mit_jax_amd.build`). torch is imported first so
that the library binds to the HIP runtime torch already loaded (both carry the soname libamdhip64.so.7). There is
no fallback: if the library is missing or cannot be loaded, every op raises.
"""

import ctypes
from pathlib import Path

import torch  # noqa: F401  (must precede the library load: one HIP runtime per process)

LIB_PATH = Path(__file__).resolve().parent / "libmjamd.so"

MJ_SUM_FORTRAN = 0
MJ_SUM_TREE = 1

P = ctypes.c_void_p
I32 = ctypes.c_int
I64 = ctypes.c_int64
F64 = ctypes.c_double


class MjError(RuntimeError):
    pass


class HaloGeom(ctypes.Structure):
    _fields_ = [("nouter", I64), ("outer_stride", I64), ("tile_stride", I64), ("ninner", I64),
                ("inner_stride", I64)]


class Cg2dDiag(ctypes.Structure):
    _fields_ = [("firstResidual", F64), ("lastResidual", F64), ("sumRHS", F64), ("rhsMax", F64),
                ("numIters", ctypes.c_int32), ("converged", ctypes.c_int32), ("outerIters", ctypes.c_int32),
                ("launches", ctypes.c_int32)]


class Cg2dOps(ctypes.Structure):
    _fields_ = [("aW2d", P), ("aS2d", P), ("aC2d", P), ("pW", P), ("pS", P), ("pC", P), ("member_stride", I64),
                ("op_index", P)]


class Cg2dConfig(ctypes.Structure):
    _fields_ = [("batch", I32), ("ntiles", I32), ("sNx", I32), ("sNy", I32), ("OL", I32), ("member_stride", I64),
                ("tolerance", F64), ("max_iters", I32), ("sum_order", I32), ("block_threads", I32),
                ("grid_blocks", I32), ("launch_mode", I32), ("poll_iters", I32)]


class Cg2dMixedConfig(ctypes.Structure):
    _fields_ = [("op_slots", I32), ("inner_tolerance", F64), ("inner_max_iters", I32), ("max_outer", I32)]


class GadGrid(ctypes.Structure):
    _fields_ = [(k, P) for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "recip_rA", "rA", "maskInC", "maskW",
                                 "maskS", "maskC", "dTlev", "drF", "recip_drF", "rhoFacC", "recip_rhoFacC",
                                 "deepFacC", "recip_deepFacC", "recip_deepFac2C", "recip_drC")] + [("oneSixth", F64)]


class GadConfig(ctypes.Structure):
    _fields_ = [("batch", I32), ("nr", I32), ("member_stride3d", I64), ("mask_member_stride", I64), ("nlev", P),
                ("mask_index", P), ("levels_per_block", I32), ("mask_slots", I32)]


class Dst3FluxConfig(ctypes.Structure):
    _fields_ = [("batch", I32), ("ntiles", I32), ("nr", I32), ("sNx", I32), ("sNy", I32), ("OLx", I32), ("OLy", I32),
                ("member_stride3d", I64)]


class FsParams(ctypes.Structure):
    _fields_ = [(k, F64) for k in ("deltaTMom", "deltaTFreeSurf", "freeSurfFac", "implicSurfPress", "implicDiv2DFlow",
                                   "mass2rUnit")] + [("useRealFreshWaterFlux", I32)]


class SfpGrid(ctypes.Structure):
    _fields_ = [(k, P) for k in ("dyG", "dxG", "recip_dxC", "recip_dyC", "rA", "maskInC", "kSurfC", "drF",
                                 "deepFacC", "rhoFacC", "deepFac2F")]


class SfpConfig(ctypes.Structure):
    _fields_ = [("batch", I32), ("ntiles", I32), ("sNx", I32), ("sNy", I32), ("OL", I32), ("nr", I32),
                ("member_stride3d", I64), ("member_stride2d", I64)]


class ImprGrid(ctypes.Structure):
    _fields_ = [(k, P) for k in ("maskC", "rA", "recip_rA", "dTtracerLev", "recip_drF", "recip_drC")] + \
        [("rkSign", F64)]


class ImprConfig(ctypes.Structure):
    _fields_ = [("batch", I32), ("ntiles", I32), ("sNx", I32), ("sNy", I32), ("OL", I32), ("nr", I32),
                ("member_stride3d", I64), ("mask_member_stride", I64), ("implicitAdvection", I32)]


# name -> (restype, argtypes): every entry point declared in include/mjamd.h
SIGNATURES = {
    "mj_version": (I32, []),
    "mj_last_error": (I32, [ctypes.c_char_p, ctypes.c_size_t]),
    "mj_halo_plan_create": (I32, [I32, P, P, P, P, P, I32, I32, I32, ctypes.POINTER(P)]),
    "mj_halo_plan_destroy": (I32, [P]),
    "mj_halo_plan_inplace_ok": (I32, [P]),
    "mj_halo_exchange": (I32, [P, P, P, ctypes.POINTER(HaloGeom), P]),
    "mj_halo_exchange_uv": (I32, [P, P, P, P, P, ctypes.POINTER(HaloGeom), P]),
    "mj_halo_exchange_adjoint": (I32, [P, P, P, P, P, ctypes.POINTER(HaloGeom), P]),
    "mj_cg2d_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(Cg2dConfig)]),
    "mj_cg2d_solve": (I32, [P, ctypes.POINTER(Cg2dOps), P, P, P, P, ctypes.POINTER(Cg2dConfig), P, ctypes.c_size_t,
                            P, P]),
    "mj_cg2d_solve_noalloc": (I32, [P, ctypes.POINTER(Cg2dOps), P, P, P, P, ctypes.POINTER(Cg2dConfig), P, P]),
    "mj_cg2d_matvec": (I32, [P, ctypes.POINTER(Cg2dOps), P, P, P, ctypes.POINTER(Cg2dConfig), P, P]),
    "mj_cg2d_coeff_grad": (I32, [P, P, P, P, P, P, P, ctypes.POINTER(Cg2dConfig), P, P]),
    "mj_cg2d_mixed_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(Cg2dConfig), ctypes.POINTER(Cg2dMixedConfig),
                                                        I64]),
    "mj_cg2d_solve_mixed": (I32, [P, ctypes.POINTER(Cg2dOps), P, P, P, P, ctypes.POINTER(Cg2dConfig),
                                  ctypes.POINTER(Cg2dMixedConfig), P, ctypes.c_size_t, P, P]),
    "mj_ini_cg2d_norm": (I32, [ctypes.POINTER(SfpGrid), ctypes.POINTER(SfpConfig), F64, F64, P, P, P, P]),
    "mj_update_cg2d": (I32, [P, ctypes.POINTER(SfpGrid), ctypes.POINTER(FsParams), ctypes.POINTER(SfpConfig), P, F64,
                             P, P, P, P, P, P, P, P, P, P]),
    "mj_update_cg2d_adjoint": (I32, [P, ctypes.POINTER(SfpGrid), ctypes.POINTER(FsParams), ctypes.POINTER(SfpConfig),
                                     P, P, P, P, P, P, P, P]),
    "mj_cg2d_rhs": (I32, [ctypes.POINTER(SfpGrid), ctypes.POINTER(FsParams), ctypes.POINTER(SfpConfig), P, P, P, P,
                          P, P, P, P, P, P, P]),
    "mj_cg2d_rhs_adjoint": (I32, [ctypes.POINTER(SfpGrid), ctypes.POINTER(FsParams), ctypes.POINTER(SfpConfig), P, P,
                                  P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mj_gad_implicit_r": (I32, [ctypes.POINTER(ImprGrid), ctypes.POINTER(ImprConfig), P, P, P, P, P, P]),
    "mj_gad_implicit_r_adjoint": (I32, [ctypes.POINTER(ImprGrid), ctypes.POINTER(ImprConfig), P, P, P, P, P, P, P,
                                        P, P]),
    "mj_gad_plan_create": (I32, [I32, P, P, I32, I32, I32, I32, ctypes.POINTER(P)]),
    "mj_gad_plan_destroy": (I32, [P]),
    "mj_gad_plan_describe": (I32, [P, P, P]),
    "mj_gad_tables_host": (I32, [I32, P, P, I32, I32, I32, I32, P, P]),
    "mj_gad_dst3_adjoint_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(GadConfig), I32, I32]),
    "mj_gad_dst3_adjoint": (I32, [P, ctypes.POINTER(GadGrid), ctypes.POINTER(GadConfig), P, P, P, P, P, P, P, P, P,
                                  P, P, P, P, ctypes.c_size_t, P]),
    "mj_dst3_adv_x": (I32, [ctypes.POINTER(Dst3FluxConfig), F64, P, P, P, P, P, P, P, P, P]),
    "mj_dst3_adv_y": (I32, [ctypes.POINTER(Dst3FluxConfig), F64, P, P, P, P, P, P, P, P, P]),
    "mj_gad_dst3_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(GadConfig), I32, I32]),
    "mj_gad_dst3": (I32, [P, ctypes.POINTER(GadGrid), ctypes.POINTER(GadConfig), P, P, P, P, P, P, P, P, P,
                          ctypes.c_size_t, P]),
}

_lib = None


def lib():
    """The loaded library (raises MjError if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise MjError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `python -m mit_jax_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(2048)
    lib().mj_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise MjError(f"{what} failed ({rc}): {last_error()}")


def ptr(t):
    """Device (or host) address of a tensor / ndarray, or None."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return ctypes.c_void_p(t.data_ptr())
    return t.ctypes.data_as(ctypes.c_void_p)


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
