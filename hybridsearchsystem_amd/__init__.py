"""MI355X (gfx950) filtered top-K search: Python face of the C ABI in ``include/frk_capi.h``.

The product is the in-tree shared library ``libfrk.so`` (HIP kernels + the C++ ``NpuRetrieval`` host
engine). This module only loads it and forwards calls through ctypes; device buffers are torch tensors
on ``cuda`` (HIP) or raw device addresses. There is no CPU fallback: every entry point raises if the
library is missing or a call fails.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FRK_LIB_PATH") or os.path.join(HERE, "libfrk.so")

FRK_DTYPE_F32, FRK_DTYPE_BF16, FRK_DTYPE_FP8E4M3, FRK_DTYPE_F16X2 = 0, 1, 2, 3
FILTER_MATCH_ALL = 1
SELECT_MAX_K = 4096

_LIB = None


class FrkError(RuntimeError):
    pass


class FilterQuery(ctypes.Structure):
    _fields_ = [("expr_off", ctypes.c_uint32), ("expr_len", ctypes.c_uint32), ("table_off", ctypes.c_uint32),
                ("n_postings", ctypes.c_uint32), ("op_num", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class DocStore(ctypes.Structure):
    _fields_ = [("tiles", ctypes.c_void_p), ("scales", ctypes.c_void_p), ("dtype", ctypes.c_uint32),
                ("dim", ctypes.c_uint32), ("n_docs", ctypes.c_uint64)]


def lib():
    """The loaded libfrk.so; raises FrkError if it was not built (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise FrkError("libfrk.so is not built: run python -m hybridsearchsystem_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    L.frk_last_error.restype = ctypes.c_char_p
    L.frk_filter_eval.restype = i32
    L.frk_filter_eval.argtypes = [vp, u32, vp, vp, u32, u32, u64, vp, u64, vp, vp, u32, vp]
    L.frk_chunk_dims.restype = u32
    L.frk_chunk_dims.argtypes = [u32]
    L.frk_store_dim.restype = u32
    L.frk_store_dim.argtypes = [u32, u32]
    L.frk_score_select.restype = i32
    L.frk_score_select.argtypes = [ctypes.POINTER(DocStore), vp, u32, vp, u32, vp, u64, u64, u64, vp, u32, vp, u32, vp]
    L.frk_select_topk.restype = i32
    L.frk_select_topk.argtypes = [vp, u32, vp, u32, u32, vp, u32, vp, u32, vp, vp, vp]
    L.frk_select_top_set.restype = i32
    L.frk_select_top_set.argtypes = [vp, u32, vp, u32, u32, vp, u32, vp, u32, vp, vp, vp]
    _bind_optional(L)
    _LIB = L
    return L


def _bind_optional(L):
    """Signatures of the handle API and the later kernel entry points."""
    from . import _capi_sigs
    _capi_sigs.bind(L)


def check(rc, what=""):
    if rc != 0:
        msg = lib().frk_last_error().decode(errors="replace")
        raise FrkError("%s failed (%d): %s" % (what or "frk call", rc, msg))


def ptr(t):
    """Device address of a torch tensor (or None / int passthrough)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
