"""ctypes binding of libmmsb_amd.so, the MI355X-native MMSB tensor-method library.

The library's C-ABI (include/mmsb_stpm.h) is the product boundary. This module only declares
argument types, converts status codes into exceptions and offers small helpers for passing
torch tensors (device memory and streams) through the ABI. It never computes anything itself:
if the shared library is missing it raises ImportError instead of falling back to CPU code.

torch is imported before the library is opened so that the process holds a single HIP runtime:
torch ships its own libamdhip64.so.7 and the library's DT_NEEDED entry then resolves to it.
"""
from __future__ import annotations

import ctypes
import os

try:  # one HIP runtime per process: torch's copy must be resident first
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmmsb_amd.so")

MMSB_OK = 0
MMSB_ERR_ARG = -1
MMSB_ERR_IO = -2
MMSB_ERR_HIP = -3
MMSB_ERR_NOMEM = -4
MMSB_ERR_NUMERIC = -5
MMSB_ERR_STATE = -6


class MMSBError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mmsb status {code}: {msg}")
        self.code = code


_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/mmsb_stpm.h appears here.
SIGNATURES: dict = {
    "mmsb_last_error": (ctypes.c_char_p, []),
    "mmsb_version": (ctypes.c_char_p, []),
    "mmsb_spmm_f32": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp,
                                     ctypes.c_float, _vp, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error() -> str:
    return lib.mmsb_last_error().decode()


def check(status: int) -> int:
    if status != MMSB_OK:
        raise MMSBError(status, last_error())
    return status


def ptr(t) -> int | None:
    """Device (or host) address of a torch tensor or numpy array; None stays NULL."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_handle(stream=None) -> int | None:
    """hipStream_t of a torch stream (default: torch's current stream on the current device)."""
    if torch is None or not torch.cuda.is_available():
        return None
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def spmm_f32(rowptr, colidx, vals, B, Y, r: int, n: int, alpha: float = 1.0, stream=None):
    """Y[:m, :r] = alpha * G @ B[:, :r] through mmsb_spmm_f32 (all device tensors)."""
    m = rowptr.numel() - 1
    check(lib.mmsb_spmm_f32(stream_handle(stream), m, n, colidx.numel(), ptr(rowptr), ptr(colidx), ptr(vals),
                            float(alpha), ptr(B), B.stride(0), r, ptr(Y), Y.stride(0)))
    return Y
