"""ctypes binding of libfew_rocm.so (the HIP kernels behind the C ABI in include/few_rocm.h).

There is no CPU fallback: if the library is missing or cannot be loaded, every
entry point raises :class:`BackendNotInstalled`.
"""

from __future__ import annotations

import ctypes
import os

from .exceptions import BackendNotInstalled, FewRocmError

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libfew_rocm.so")

FEW_PRECISION_FP64 = 0
FEW_PRECISION_FACT64 = 1
FEW_PRECISION_MIXED = 2
FEW_PRECISION_MIXED16 = 3

# size of the static device array behind few_get_waveform_plan_noalloc
FEW_NOALLOC_SCRATCH_BYTES = 32 << 20

_lib = None

_d = ctypes.c_void_p  # every pointer travels as an untyped address


def _declare(lib: ctypes.CDLL) -> None:
    c_int, c_long, c_double, c_size = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_size_t
    sig = {
        "few_last_error": ([], ctypes.c_char_p),
        "few_abi_version": ([], c_int),
        "few_interpolate_arrays": ([_d, _d, c_int, c_int, _d, _d, _d, _d, _d], c_int),
        "few_get_waveform": ([_d, _d, _d, _d, _d, _d, _d, c_int, c_int, c_int, _d, c_double, _d,
                              c_int, _d], c_int),
        "few_get_waveform_async": ([_d, _d, _d, _d, _d, _d, _d, c_int, c_int, c_int, _d, c_double,
                                    _d, c_int, _d, c_size, _d], c_int),
        "few_waveform_workspace_bytes": ([c_int, c_int, c_int, c_int], c_size),
        "few_spline_build_async": ([_d, c_long, c_long, _d, c_long, c_long, _d, c_long, c_long,
                                    c_long, c_int, c_int, c_int, _d, _d], c_int),
        "few_prepare_modes_async": ([_d, _d, _d, c_int, c_int, _d, _d, _d, _d], c_int),
        "few_release_workspace": ([], c_int),
        "few_mode_plan_create": ([_d, _d, _d, c_int], ctypes.c_void_p),
        "few_mode_plan_info": ([_d, _d], c_int),
        "few_mode_plan_destroy": ([_d], None),
        "few_plan_workspace_bytes": ([_d, c_int, c_int, c_int], c_size),
        "few_get_waveform_plan_async": ([_d, _d, _d, _d, _d, _d, _d, _d, c_int, c_int, c_int, _d,
                                         c_double, _d, c_int, c_int, c_int, _d, c_size, _d], c_int),
        "few_batch_create": (
                             [c_int, _d, _d, _d, _d, _d, _d, _d], ctypes.c_void_p),
        "few_batch_destroy": ([_d], None),
        "few_batch_layout": ([_d, _d, _d], c_int),
        "few_batch_workspace_bytes": ([_d, c_int], c_size),
        "few_batch_run_async": ([_d, _d, _d, _d, _d, _d, _d, _d, c_int, c_int, c_int, _d, c_size,
                                 _d], c_int),
        "few_adjoint_workspace_bytes": ([c_int, c_int, c_int, c_int], c_size),
        "few_waveform_adjoint_async": ([_d, _d, _d, _d, _d, _d, _d, c_int, c_int, c_int, _d,
                                        c_double, _d, _d, _d, c_int, _d, c_size, _d], c_int),
        "few_spline_adjoint_async": ([_d, _d, c_long, c_long, c_long, c_int, c_int, _d, c_long,
                                      c_long, _d, _d], c_int),
        "few_modes_adjoint_async": ([_d, _d, _d, c_int, c_int, _d, _d, _d, _d], c_int),
        "few_plan_adjoint_workspace_bytes": ([_d, c_int, c_int, c_int], c_size),
        "few_spline_eval_async": ([_d, c_long, c_long, c_int, c_int, _d, c_long, c_long, c_long, _d,
                                   c_long, c_int, c_int, _d, _d, _d], c_int),
        "few_direct_mode_sum_async": ([_d, _d, _d, _d, _d, _d, _d, _d, _d, c_int, c_int, _d, _d],
                                      c_int),
        "few_waveform_adjoint_plan_async": ([_d, _d, _d, _d, _d, _d, _d, _d, c_int, c_int, c_int, _d,
                                             c_double, _d, _d, _d, c_int, _d, c_size, _d], c_int),
        "few_fd_workspace_bytes": ([c_int, c_int, ctypes.POINTER(c_size)], c_int),
        "few_get_waveform_generic_fd_async": ([_d, _d, _d, _d, _d, _d, _d, c_int, c_double, _d,
                                               c_int, c_int, _d, _d, _d, c_int, _d, c_int, c_int,
                                               c_int, _d, c_size, _d], c_int),
        "few_get_waveform_generic_fd": ([_d, _d, _d, _d, _d, _d, _d, c_int, c_double, _d, c_int,
                                         c_int, _d, _d, _d, c_int, _d, c_int, c_int, c_int, _d],
                                        c_int),
        "few_interp2d_workspace_bytes": ([c_int], c_size),
        "few_interp2d_async": ([_d, _d, c_int, _d, c_int, _d, c_int, c_int, _d, c_int, _d, c_int,
                                c_int, c_int, _d, c_size, _d], c_int),
        "few_interp2D_wrap": ([_d, _d, c_int, _d, c_int, _d, c_int, c_int, _d, c_int, _d, c_int,
                               c_int, c_int, _d], c_int),
        "few_aak_workspace_bytes": ([c_int], c_size),
        "few_aak_waveform_async": ([_d, _d] + [c_double] * 8 + [c_int, c_int, c_int, c_int,
                                                                 c_double, _d, _d, c_size, _d],
                                   c_int),
        "few_aak_get_waveform": ([_d, _d] + [c_double] * 8 + [c_int, c_int, c_int, c_int,
                                                               c_double, _d, _d], c_int),
        "few_get_waveform_plan_noalloc": ([_d, _d, _d, _d, _d, _d, _d, _d, c_int, c_int, c_int,
                                           _d, c_double, _d, c_int, c_int, c_int, _d], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def lib() -> ctypes.CDLL:
    """Load libfew_rocm.so once; raise BackendNotInstalled when it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BackendNotInstalled(
                f"{LIB_PATH} is not built; run `python -m fastemriwaveforms_amd._build` "
                "(hipcc, gfx950) first")
        try:
            handle = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise BackendNotInstalled(f"cannot load {LIB_PATH}: {e}") from e
        _declare(handle)
        _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().few_last_error().decode(errors="replace")
        raise FewRocmError(f"{what} failed (code {rc}): {msg}")


def exported_symbols() -> list[str]:
    return [
        "few_last_error", "few_abi_version", "few_interpolate_arrays", "few_get_waveform",
        "few_get_waveform_async", "few_waveform_workspace_bytes", "few_spline_build_async",
        "few_prepare_modes_async", "few_release_workspace", "few_mode_plan_create",
        "few_mode_plan_info", "few_mode_plan_destroy", "few_plan_workspace_bytes",
        "few_get_waveform_plan_async", "few_batch_create", "few_batch_destroy", "few_batch_layout",
        "few_batch_workspace_bytes", "few_batch_run_async", "few_adjoint_workspace_bytes",
        "few_waveform_adjoint_async", "few_spline_adjoint_async", "few_modes_adjoint_async",
        "few_plan_adjoint_workspace_bytes", "few_waveform_adjoint_plan_async",
        "few_spline_eval_async", "few_direct_mode_sum_async", "few_fd_workspace_bytes",
        "few_get_waveform_generic_fd_async", "few_get_waveform_generic_fd",
        "few_interp2d_workspace_bytes", "few_interp2d_async", "few_interp2D_wrap",
        "few_aak_workspace_bytes", "few_aak_waveform_async", "few_aak_get_waveform",
        "few_get_waveform_plan_noalloc",
    ]
