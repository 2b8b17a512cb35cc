"""ctypes signatures of the C ABI beyond the core kernel entry points (include/frk_capi.h)."""
import ctypes

vp, u32, u64, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float


def bind(L):
    L.frk_tiles_bytes.restype = u64
    L.frk_tiles_bytes.argtypes = [u64, u32, u32]
    L.frk_tile_docs.restype = i32
    L.frk_tile_docs.argtypes = [vp, u64, u32, u32, u32, vp, vp, u32]
