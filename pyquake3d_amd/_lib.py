"""ctypes binding of the HIP C-ABI library ``libpq3d.so``.

There is no fallback: if the library is missing, or a call returns a nonzero
status, a ``RuntimeError`` carrying ``pq3d_last_error()`` is raised.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_bool, c_double, c_float, c_int, c_int64, c_size_t, c_void_p
from pathlib import Path

import os

# PQ3D_LIB names another build of the same library (e.g. one compiled with
# different tuning macros); the default is the in-tree build
LIB_PATH = Path(os.environ.get("PQ3D_LIB") or
                Path(__file__).resolve().parent / "lib" / "libpq3d.so")

_D = POINTER(c_double)
_F = POINTER(c_float)
_L = POINTER(c_int64)
_I = POINTER(c_int)

# name -> (restype, argtypes). Kept in the same order as include/pyquake3d_amd.h.
SIGNATURES = {
    "pq3d_last_error": (ctypes.c_char_p, []),
    "pq3d_version": (c_int, []),
    # reference drop-in symbols (src/TDstressFS_C.cpp:2048, :2098)
    "TDstressFS_C": (None, [_D, _D, _D, c_size_t, _D, _D, _D, c_double, c_double, c_double,
                            c_double, c_double, c_bool, _D, _D]),
    "TDstressEachSourceAtReceiver_C": (None, [c_double, c_double, c_double, _D, _D, _D, c_size_t,
                                              c_double, c_double, c_double, c_double, c_double,
                                              c_bool, _D, _D]),
    # device-pointer TDE kernels
    "pq3d_tde_source_records": (c_int, [_D, c_int64, _D]),
    "pq3d_tde_stress_points": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_double,
                                       c_double, c_double, c_double, c_double, c_void_p, c_void_p,
                                       c_void_p]),
    "pq3d_tde_stress_sources": (c_int, [c_double, c_double, c_double, c_void_p, c_int64, c_double,
                                        c_double, c_double, c_double, c_double, c_void_p,
                                        c_void_p, c_void_p]),
    "pq3d_tde_traction_block": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p,
                                        c_int64, c_int64, c_double, c_double, c_void_p, c_int64,
                                        c_int64, c_int64, c_int64, c_void_p]),
    "pq3d_tde_traction_block_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                            c_void_p, c_int64, c_int64, c_double, c_double,
                                            c_void_p, c_int64, c_int64, c_int64, c_int64,
                                            c_void_p]),
    "pq3d_tde_hs_records": (c_int, [_D, c_int64, _D]),
    "pq3d_tde_stress_points_hs": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                          c_void_p, c_double, c_double, c_double, c_double,
                                          c_double, c_void_p, c_void_p, c_void_p]),
    "pq3d_tde_stress_sources_hs": (c_int, [c_double, c_double, c_double, c_void_p, c_void_p,
                                           c_int64, c_double, c_double, c_double, c_double,
                                           c_double, c_void_p, c_void_p, c_void_p]),
    "pq3d_tde_traction_block_hs": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                           c_int64, c_void_p, c_int64, c_int64, c_double,
                                           c_double, c_void_p, c_int64, c_int64, c_int64,
                                           c_int64, c_void_p]),
    "pq3d_tde_traction_block_hs_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                               c_int64, c_void_p, c_int64, c_int64, c_double,
                                               c_double, c_void_p, c_int64, c_int64, c_int64,
                                               c_int64, c_void_p]),
    "pq3d_tde_disp_points": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                     c_double, c_double, c_double, c_double, c_void_p,
                                     c_void_p]),
    "pq3d_tde_disp_sum": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p,
                                  c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "pq3d_tree_level": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64,
                                c_int64, c_void_p, c_void_p, c_void_p]),
    "pq3d_block_level": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_double, c_int64,
                                 c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    # H-matrix
    "pq3d_layout_constants": (c_int, [_L, c_int64]),
    "pq3d_hmat_dense_assemble": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                         c_double, c_double, c_void_p, c_int, c_void_p]),
    "pq3d_hmat_aca": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                              c_double, c_double, c_double, c_int, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p]),
    "pq3d_hmat_aca_wide": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                   c_double, c_double, c_double, c_int, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pq3d_hmat_aca_pack": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_int, c_void_p]),
    "pq3d_hmat_matvec": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p,
                                 c_void_p, c_int64, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_int64, c_void_p]),
    "pq3d_hmat_matvec_noalloc": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                                         c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int,
                                         c_int, c_void_p, c_void_p, c_int64, c_int64, c_void_p,
                                         c_void_p, c_void_p, c_int64, c_void_p]),
    "pq3d_hmat_matvec_mfma": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                                      c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int,
                                      c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int,
                                      c_void_p]),
    "pq3d_hmat_vpanel_pack": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "pq3d_hmat_matvec_mfs": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p,
                                     c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_int64, c_int, c_int, c_int, c_int, c_void_p]),
    # time step
    "pq3d_rk45_attempt": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int64,
                                  c_void_p]),
    "pq3d_rk45_commit": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "pq3d_rk45_rounds": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_int, c_double] + [c_void_p] * 19 +
                         [c_double, c_int64, c_void_p, c_void_p]),
    "pq3d_rk45_advance": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p] + [c_void_p] * 20 +
                          [c_int64, c_int64, c_void_p, c_void_p, c_double, c_void_p,
                           c_void_p]),
    "pq3d_controller_widths": (c_int, [_L, c_int64]),
    "pq3d_deficit": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p]),
    "pq3d_deficit_interleave": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_void_p]),
    "pq3d_deficit_w32": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p]),
    "pq3d_qd_rhs": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_double, c_double, c_double, c_void_p, c_void_p]),
    "pq3d_add_loading": (c_int, [c_int64, c_void_p, c_void_p, c_int, c_void_p]),
    # learned step size
    "pq3d_mlp_num_params": (c_int, []),
    "pq3d_mlp_forward_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                          c_void_p, c_void_p, c_void_p, c_void_p]),
    "pq3d_mlp_adam": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_float,
                              c_float, c_float, c_float, c_int, c_void_p, c_void_p]),
    "pq3d_features": (c_int, [c_int64, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_double, c_void_p, c_void_p]),
    "pq3d_learned_propose": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p]),
}

_lib = None


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libpq3d.so once and bind every declared symbol."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"HIP library {p} is not built; run `python -m pyquake3d_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


LAYOUT_NAMES = ("dense_block", "lowrank_block", "aca_ws", "item", "unit", "reduce", "tile",
                "recv", "rec", "hrec", "member", "vpanel", "punit")


def layout_constants() -> dict:
    """Row widths of the tables the kernels read (from the library itself)."""
    import numpy as np
    out = np.zeros(len(LAYOUT_NAMES), dtype=np.int64)
    k = load().pq3d_layout_constants(out.ctypes.data_as(_L), len(out))
    if k != len(out):
        raise RuntimeError("pq3d_layout_constants returned %d values" % k)
    return dict(zip(LAYOUT_NAMES, out.tolist()))


def check_width(name: str, table) -> None:
    """Raise unless ``table`` is a C-contiguous 2-D array/tensor whose row
    width equals the kernel's for ``name``."""
    w = layout_constants()[name]
    shape = tuple(table.shape)
    if len(shape) != 2 or shape[1] != w:
        raise ValueError(f"{name} table has shape {shape}; the kernels read rows of {w}")
    contiguous = table.is_contiguous() if hasattr(table, "is_contiguous") else table.flags["C_CONTIGUOUS"]
    if not contiguous:
        raise ValueError(f"{name} table is not contiguous")


CONTROLLER_NAMES = ("run", "lstate", "sample")


def controller_widths() -> dict:
    """Record widths of the device controller (from the library itself)."""
    import numpy as np
    out = np.zeros(len(CONTROLLER_NAMES), dtype=np.int64)
    k = load().pq3d_controller_widths(out.ctypes.data_as(_L), len(out))
    if k != len(out):
        raise RuntimeError("pq3d_controller_widths returned %d values" % k)
    return dict(zip(CONTROLLER_NAMES, out.tolist()))


def exported(lib: ctypes.CDLL) -> dict:
    return {name: hasattr(lib, name) for name in SIGNATURES}


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().pq3d_last_error()
        raise RuntimeError(f"{what or 'pyquake3d_amd'}: status {rc}: "
                           f"{msg.decode() if msg else ''}")


class DevPtr(c_void_p):
    """A device/host address. Pointer parameters only accept DevPtr (or
    None / 0 for a null pointer) and integer/float parameters reject it, so an
    argument list that is out of order fails before anything is launched."""


def _checked_args(name: str, fn, args):
    types = fn.argtypes
    if len(args) != len(types):
        raise TypeError(f"{name}: expected {len(types)} arguments, got {len(args)}")
    for k, (a, t) in enumerate(zip(args, types)):
        if t is c_void_p:
            if not (isinstance(a, DevPtr) or a is None or (isinstance(a, int) and a == 0)):
                raise TypeError(f"{name}: argument {k} must be a pointer (DevPtr/None), "
                                f"got {type(a).__name__}")
        elif t in (c_int64, c_int, c_size_t, c_double, c_float, c_bool):
            if isinstance(a, c_void_p):
                raise TypeError(f"{name}: argument {k} must be a {t.__name__}, got a pointer")
    return args


def call(name: str, *args) -> None:
    fn = getattr(load(), name)
    check(fn(*_checked_args(name, fn, args)), name)


def ptr(t):
    """Device (or host) address of a torch tensor as a DevPtr, or None."""
    return None if t is None else DevPtr(int(t.data_ptr()))


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return DevPtr(int(s.cuda_stream)) if s.cuda_stream else None
