"""ctypes binding of libemd_gfx950.so (the C ABI declared in include/emd_gfx950.h).

The library is built in-tree by ``make -C emd-cuda_amd/csrc`` (``__graft_entry__.build()``).
There is no fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemd_gfx950.so")

EMD_OK = 0
EMD_ERR_VALUE = -1
EMD_ERR_RUNTIME = -2
EMD_ERR_MEMORY = -3

_c_int_p = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class EmdConfigC(ctypes.Structure):
    _fields_ = [("max_extrema", ctypes.c_int), ("nbsym", ctypes.c_int),
                ("max_iteration", ctypes.c_int), ("max_imfs", ctypes.c_int),
                ("threads_per_block", ctypes.c_int), ("tile_slots", ctypes.c_int)]


class EmdLaunchC(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int), ("tile_slots", ctypes.c_int),
                ("wgs_per_cu", ctypes.c_int), ("persistent", ctypes.c_int),
                ("traces", ctypes.c_int)]


class EmdChunkC(ctypes.Structure):
    _fields_ = [("values", ctypes.c_void_p), ("offsets", ctypes.c_void_p), ("N", ctypes.c_int),
                ("B", ctypes.c_int), ("max_n", ctypes.c_int), ("order", ctypes.c_void_p),
                ("launch", EmdLaunchC), ("imfs", ctypes.c_void_p),
                ("out_offsets", ctypes.c_void_p), ("n_imfs", ctypes.c_void_p),
                ("overflow", ctypes.c_void_p), ("debug_n", ctypes.c_void_p),
                ("debug_extno", ctypes.c_void_p), ("debug_nzm", ctypes.c_void_p),
                ("stats", ctypes.c_void_p), ("row_offsets", ctypes.c_void_p),
                ("compact", ctypes.c_void_p)]


# name -> (restype, argtypes); every symbol declared in include/emd_gfx950.h
PROTOTYPES = {
    "emd_last_error": (ctypes.c_char_p, []),
    "emd_version": (ctypes.c_int, []),
    "emd_static_lds_bytes": (ctypes.c_int, [ctypes.c_int]),
    "emd_lds_slot_bytes": (ctypes.c_int, []),
    "emd_device_info": (ctypes.c_int, [_c_int_p, _c_int_p, _c_int_p]),
    "emd_create": (ctypes.c_int, [ctypes.POINTER(EmdConfigC), ctypes.POINTER(_vp)]),
    "emd_destroy": (ctypes.c_int, [_vp]),
    "emd_run_padded": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp,
                                      _vp, _vp, _vp, _vp]),
    "emd_run_ragged": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                      ctypes.POINTER(EmdLaunchC), _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp, _vp]),
    "emd_run_ragged_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                              ctypes.POINTER(EmdLaunchC), _vp, _vp, _vp, _vp,
                                              _vp, _vp, _vp, _vp, _vp]),
    "emd_envelope_only": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                         _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "emd_run_padded_times": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            _vp]),
    "emd_run_ragged_times": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                            ctypes.POINTER(EmdLaunchC), _vp, _vp, _vp, _vp, _vp,
                                            _vp, _vp, _vp, _vp]),
    "emd_envelope_only_times": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                               _vp, _vp, _vp]),
    "emd_max_resident": (ctypes.c_int, [_vp, ctypes.POINTER(EmdLaunchC), ctypes.c_int,
                                        _c_int_p]),
    "emd_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "emd_kernel_time_ms": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "emd_compact_offsets": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "emd_compact": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                   _vp, _vp, _vp]),
    "emd_chunk_graph_create": (ctypes.c_int, [_vp, ctypes.POINTER(EmdChunkC),
                                              ctypes.POINTER(_vp)]),
    "emd_chunk_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "emd_chunk_graph_destroy": (ctypes.c_int, [_vp]),
    "emd_features": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp]),
    "emd_mlp_param_count": (ctypes.c_int, []),
    "emd_mlp_workspace_floats": (ctypes.c_int64, [ctypes.c_int]),
    "emd_mlp_forward": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "emd_mlp_backward": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "emd_mlp_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp]),
    "emd_segy_sample_bytes": (ctypes.c_int, [ctypes.c_int]),
    "emd_segy_decode": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int, _vp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "emd_hilbert": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   _vp, _vp, _vp, _vp]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the gfx950 library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C emd-cuda_amd/csrc -j8` "
            "(or __graft_entry__.build()).  There is no CPU fallback.")
    # torch ships its own HIP runtime under the same SONAME (libamdhip64.so.7).  Loading
    # torch first makes this library bind to that one runtime; loading this library first
    # would pull in the system runtime and leave torch without a usable device.
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    """Map a C status to the reference's Python exception classes."""
    if rc == EMD_OK:
        return
    msg = load().emd_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == EMD_ERR_VALUE:
        raise ValueError(msg)
    if rc == EMD_ERR_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)
