"""Loader of libtrtt_hip.so (the gfx950 projector library) through ctypes.

The library is built in-tree by ``yoricktrtt_amd/build.py``.  There is no
CPU fallback: importing the projectors without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrtt_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "trtt_hip.h")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
lp = C.POINTER(C.c_long)
vp = C.c_void_p
L, I, D = C.c_long, C.c_int, C.c_double

TRTT_SUCCESS, TRTT_FAILURE = 0, -1
F32, F64 = 0, 1
PB3D, CB3D, PB2D, FB2D, LF2D = 0, 1, 2, 3, 4


class Tuning(C.Structure):
    """trtt_hip_tuning_t: launch shape of the axial kernels of a plan."""

    _fields_ = [("fwd_tile", C.c_int), ("bwd_tile", C.c_int), ("fwd_window_bytes", C.c_int),
                ("fwd_views_per_block", C.c_int), ("bwd_window_bytes", C.c_int),
                ("fwd_chunk_views", C.c_int), ("fwd_flush", C.c_int), ("fwd_xcds", C.c_int)]


class LF2DParameters(C.Structure):
    """lf2d_parameters_t (LongFessler2D.h:16-28)."""

    _fields_ = [("pixelSize", D), ("pixelStep", D), ("voxelSize", D), ("sourceDistance", D),
                ("detectorDistance", D), ("objectOffset1", D), ("objectOffset2", D),
                ("objectDimension1", L), ("objectDimension2", L), ("detectorDimension", L)]


_TAB3 = [dp, dp, L, L, L, D, I, dp, dp, L, D, dp, L, L, D, D, dp, dp, I]
_CUB3 = [dp, dp, L, L, L, D, I, L, L, D, D, dp, dp, I]
_TAB2 = [dp, dp, L, L, D, I, dp, dp, L, D, dp, L, D, dp, dp, I]
_CUB2 = [dp, dp, L, L, D, I, L, D, dp, dp, I]

_SIGS = {
    "trtt_PB3D": (_TAB3, I), "trtt_CB3D": (_TAB3, I),
    "trtt_PB3D_cubic": (_CUB3, I), "trtt_CB3D_cubic": (_CUB3, I),
    "trtt_PB2D": (_TAB2, I), "trtt_FB2D": (_TAB2, I),
    "trtt_PB2D_cubic": (_CUB2, I), "trtt_FB2D_cubic": (_CUB2, I),
    "trtt_PB2D_sparser": ([dp, L, L, L, L, D, I, dp, dp, L, D, dp, L, D, dp, dp], I),
    "trtt_FB2D_sparser": ([dp, L, L, L, L, D, I, dp, dp, L, D, dp, L, D, dp, dp], I),
    "lf2d_direct": ([C.POINTER(LF2DParameters), D, D, dp, dp, I], None),
    "lf2d_adjoint": ([C.POINTER(LF2DParameters), D, D, dp, dp, I], None),
    "lf2d_line_integ_direct": ([C.POINTER(LF2DParameters), D, D, dp, dp, I], None),
    "lf2d_line_integ_adjoint": ([C.POINTER(LF2DParameters), D, D, dp, dp, I], None),
    "lf2d_sparser": ([C.POINTER(LF2DParameters), D, D, L, L, dp, I], None),
    "lf2d_line_integ_sparser": ([C.POINTER(LF2DParameters), D, D, L, L, dp, I], None),
    "trtt_hip_plan_set_lf_line": ([vp, I], I),
    "trtt_hip_eval_set_data": ([vp, dp, dp], I),
    "trtt_hip_eval": ([vp, dp, dp, dp, I], I),
    "trtt_hip_set_device": ([I], I),
    "trtt_hip_set_precision": ([I], I),
    "trtt_hip_last_error": ([], C.c_char_p),
    "trtt_hip_plan_create": ([I, I, L, L, L, L, D, I, L, L, D, D, dp, dp, L, dp, L, D, D, ip, ip,
                              dp, dp], vp),
    "trtt_hip_plan_create_lf2d": ([I, C.POINTER(LF2DParameters), dp, dp, L], vp),
    "trtt_hip_forward": ([vp, vp, vp, vp], I),
    "trtt_hip_backward": ([vp, vp, vp, vp], I),
    "trtt_hip_weighted_residual": ([vp, vp, vp, vp, vp, vp], I),
    "trtt_hip_plan_error_flag": ([vp], I),
    "trtt_hip_plan_path": ([vp], I),
    "trtt_hip_plan_nviews": ([vp], L),
    "trtt_hip_plan_get_tuning": ([vp, C.POINTER(Tuning)], I),
    "trtt_hip_plan_set_tuning": ([vp, C.POINTER(Tuning)], I),
    "trtt_hip_plan_set_timing": ([vp, I], I),
    "trtt_hip_plan_kernel_time": ([vp, I, dp, ip], I),
    "trtt_hip_plan_volume_size": ([vp], L),
    "trtt_hip_plan_sino_size": ([vp], L),
    "trtt_hip_plan_destroy": ([vp], I),
    "trtt_hip_sparse_create": ([vp, I], vp),
    "trtt_hip_sparse_nnz": ([vp], L),
    "trtt_hip_sparse_coo": ([vp, dp, lp, lp], I),
    "trtt_hip_sparse_forward": ([vp, vp, vp, vp], I),
    "trtt_hip_sparse_backward": ([vp, vp, vp, vp], I),
    "trtt_hip_sparse_destroy": ([vp], I),
    "trtt_hip_tinterp_create": ([I, L, L, L, I, ip, dp], vp),
    "trtt_hip_tinterp_apply": ([vp, I, vp, vp, vp], I),
    "trtt_hip_tinterp_destroy": ([vp], I),
    "trtt_hip_dd_create": ([I, L, L, L, D, L, L, D, D, dp, dp, L], vp),
    "trtt_hip_dd_forward": ([vp, vp, vp, vp], I),
    "trtt_hip_dd_backward": ([vp, vp, vp, I, vp], I),
    "trtt_hip_dd_destroy": ([vp], I),
    "trtt_hip_tv": ([vp, vp, vp, vp, L, L, L, L, D, D, D, I, vp], I),
    "trtt_hip_bspline_filter": ([vp, vp, L, L, L, L, I, I, vp], I),
    "trtt_unet_conv_fwd": ([vp, vp, vp, vp, vp, I, I, I, I, I, I, I, I, I, vp], I),
    "trtt_unet_conv_wgrad": ([vp, vp, vp, vp, I, I, I, I, I, I, I, I, vp], I),
    "trtt_unet_prep_weights": ([vp, vp, vp, I, I, I, I, I, vp], I),
    "trtt_unet_pool": ([vp, vp, vp, vp, vp, I, I, I, I, I, I, vp], I),
    "trtt_unet_upsample": ([vp, vp, I, I, I, I, I, I, I, I, vp], I),
    "trtt_unet_channel_copy": ([vp, vp, L, I, I, I, I, vp], I),
    "trtt_unet_relu_bwd": ([vp, vp, vp, vp, L, vp], I),
    "trtt_unet_cast_bf16": ([vp, vp, L, vp], I),
    "trtt_unet_residual_mse": ([vp, vp, vp, vp, vp, L, vp, vp], I),
    "trtt_unet_adam": ([vp, vp, vp, vp, L, C.c_float, C.c_float, C.c_float, C.c_float, vp], I),
    "trtt_xform3d_new": ([], vp),
    "trtt_xform3d_delete": ([vp], None),
    "trtt_xform3d_move": ([vp, D, D, D, vp], I),
    "trtt_xform3d_scale": ([vp, D, D, D, vp], I),
    "trtt_xform3d_rotate": ([vp, D, D, D, D, vp], I),
    "trtt_xform3d_combine": ([vp, vp, vp], I),
    "trtt_xform3d_invert": ([vp, vp], I),
    "trtt_xform3d_get_coefficients": ([dp, vp], I),
    "trtt_xform3d_set_coefficients": ([vp, dp], I),
    "trtt_xform3d_apply": ([dp, vp, dp], I),
    "trtt_mvmult": ([dp, dp, L, dp, lp, lp], I),
    "trtt_DD_PB3D": ([dp, dp, L, L, L, D, L, L, D, D, dp, dp, I], I),
    "trtt_hip_DD_PB3D_adjoint": ([dp, dp, L, L, L, D, L, L, D, D, dp, dp], I),
    "trtt_hip_dropin_noalloc": ([I, dp, dp, L, L, L, D, I, dp, L, D, D, L, L, D, D, dp, dp, I], I),
}

_lib = None


def header_functions(path: str = HEADER) -> list[str]:
    """Names of every function the public header declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    text = text.replace('extern "C" {', "").replace("}", ";")
    names = set()
    for stmt in text.split(";"):
        stmt = " ".join(stmt.split())
        if "(" not in stmt or stmt.startswith("typedef") or "struct" in stmt.split("(")[0]:
            continue
        m = re.search(r"([A-Za-z_][A-Za-z0-9_]*)\s*\($", stmt.split("(")[0] + "(")
        if m:
            names.add(m.group(1))
    return sorted(names)


def load(path: str = None) -> C.CDLL:
    """Load (once) and type the library; raises when it has not been built.
    ``TRTT_HIP_LIB`` names another build of the same library (used to compare
    compiler settings)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("TRTT_HIP_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g;"
                           " g.build()'` (there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (args, res) in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().trtt_hip_last_error()
    return msg.decode() if msg else ""
