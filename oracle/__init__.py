"""CPU oracle for the filtered top-K pipeline -- TEST INFRASTRUCTURE ONLY.

Python bindings (ctypes + numpy) over ``liboracle.so`` (built from ``frk_oracle.c`` by
``oracle/Makefile``) and, where it was built, over ``_ref/libfrkref*.so``: the reference's own host
sources compiled from the reference checkout. Only tests, the smoke check and the bench's CPU
baseline use this package; the product library never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = {}

TERM, AND, OR, NOT = 0, 1, 2, 3
OP_CONJ, OP_AND, OP_OR, OP_NOT = 0, 1, 2, 3

_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so (and _ref when the reference checkout is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_hash64.restype = ctypes.c_uint64
        L.orc_hash64.argtypes = [ctypes.c_char_p, ctypes.c_uint32]
        L.orc_tag_token.restype = ctypes.c_uint64
        L.orc_tag_token.argtypes = [ctypes.c_uint64]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_f16_to_f32.restype = ctypes.c_float
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_select_layout.restype = ctypes.c_int
        L.orc_select_layout.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_int]
        L.orc_posting_encode.restype = ctypes.c_int64
        L.orc_posting_encode.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_float,
                                         ctypes.c_int, _u8p, ctypes.c_uint64]
        L.orc_posting_decode.restype = ctypes.c_int
        L.orc_posting_decode.argtypes = [_u8p, ctypes.c_uint64, ctypes.c_uint32, _u64p, ctypes.c_uint32]
        L.orc_parse_filter.restype = ctypes.c_int
        L.orc_parse_filter.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int]
        L.orc_eval_filter_doc.restype = ctypes.c_int
        L.orc_eval_filter_doc.argtypes = [_i64p, ctypes.c_int, _u64p, ctypes.c_uint32]
        L.orc_filter_bitmap.restype = None
        L.orc_filter_bitmap.argtypes = [_i64p, ctypes.c_int, _u64p, ctypes.c_uint64, ctypes.c_uint32, _u64p]
        L.orc_compile_postfix.restype = ctypes.c_int
        L.orc_compile_postfix.argtypes = [_i64p, ctypes.c_int, _u32p, ctypes.c_int, _u64p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int), _u32p]
        L.orc_check_postfix.restype = ctypes.c_int
        L.orc_check_postfix.argtypes = [_u32p, ctypes.c_int, ctypes.c_uint32]
        L.orc_postfix_max_stack.restype = ctypes.c_int
        L.orc_postfix_max_stack.argtypes = [_u32p, ctypes.c_int]
        L.orc_postfix_eval.restype = ctypes.c_int
        L.orc_postfix_eval.argtypes = [_u32p, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(_u64p),
                                       ctypes.c_uint32, _u64p]
        L.orc_topk.restype = ctypes.c_int
        L.orc_topk.argtypes = [_f32p, ctypes.c_uint64, ctypes.c_uint32, _f32p, ctypes.c_uint32, _u64p,
                               ctypes.c_uint64, _i32p, ctypes.c_int32, ctypes.c_int, _i64p, _f64p, _i32p,
                               ctypes.c_int]
        _LIB = L
    return _LIB


def ref(packed=False):
    """The reference's own host code (oracle/_ref), or None when it was not built here."""
    key = bool(packed)
    if key in _REF:
        return _REF[key]
    path = os.path.join(HERE, "_ref", "libfrkref_packed.so" if packed else "libfrkref.so")
    if not os.path.exists(path):
        _REF[key] = None
        return None
    L = ctypes.CDLL(path)
    L.ref_init.argtypes = [ctypes.c_int]
    L.ref_init(1 if packed else 0)
    L.ref_hash64.restype = ctypes.c_uint64
    L.ref_hash64.argtypes = [ctypes.c_char_p, ctypes.c_uint32]
    L.ref_f32_to_f16.restype = ctypes.c_uint16
    L.ref_f32_to_f16.argtypes = [ctypes.c_float]
    L.ref_f16_to_f32.restype = ctypes.c_float
    L.ref_f16_to_f32.argtypes = [ctypes.c_uint16]
    L.ref_posting_serialize.restype = ctypes.c_int64
    L.ref_posting_serialize.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_float,
                                        _u8p, ctypes.c_uint64]
    L.ref_parse_filter.restype = ctypes.c_int
    L.ref_parse_filter.argtypes = [ctypes.c_char_p, _i64p, ctypes.c_int]
    for name in ("ref_filter_bitmap", "ref_flat_filter_bitmap"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [_i64p, ctypes.c_int, _u64p, ctypes.c_uint64, ctypes.c_uint32, _u64p]
    L.ref_cpu_topk_batch.restype = ctypes.c_int
    L.ref_cpu_topk_batch.argtypes = [_f32p, _u64p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _f32p,
                                     ctypes.c_uint32, _i64p, _i32p, _i32p, _i32p, ctypes.c_int, ctypes.c_uint64,
                                     ctypes.c_uint64, ctypes.c_int32, _i64p, _i32p]
    _REF[key] = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------------------------
# hashing / codecs
def hash64(s):
    b = s.encode() if isinstance(s, str) else bytes(s)
    return int(lib().orc_hash64(b, len(b)))


def tag_token(tag):
    return int(lib().orc_tag_token(int(tag)))


def f32_to_f16(x):
    return int(lib().orc_f32_to_f16(float(x)))


def f16_to_f32(h):
    return float(lib().orc_f16_to_f32(int(h)))


def select_layout(hits, doc_num, is_matrix=False, threshold=0.05, packed=True):
    return int(lib().orc_select_layout(hits, doc_num, int(is_matrix), threshold, int(packed)))


def posting_encode(ids, doc_num, is_matrix=False, threshold=0.05, packed=True):
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    cap = 16 + 6 * max(len(ids), 1) + doc_num + 2 * ((doc_num + 15) // 16)
    out = np.zeros(cap, dtype=np.uint8)
    n = lib().orc_posting_encode(_p(ids, _u32p), len(ids), doc_num, int(is_matrix), threshold, int(packed),
                                 _p(out, _u8p), cap)
    if n < 0:
        raise ValueError("posting_encode failed")
    return out[:n].tobytes()


def posting_decode(buf, doc_num):
    a = np.frombuffer(bytes(buf), dtype=np.uint8).copy()
    nw = (doc_num + 63) // 64
    words = np.zeros(max(nw, 1), dtype=np.uint64)
    lay = lib().orc_posting_decode(_p(a, _u8p), len(a), doc_num, _p(words, _u64p), max(nw, 1))
    if lay < 0:
        raise ValueError("malformed posting")
    return lay, words[:nw]


# ---------------------------------------------------------------------------------------------
# filters
def parse_filter(s):
    """Flat pre-order (type, arg) pairs as an int64 array of shape (n, 2); None on a parse error."""
    cap = 4 * len(s) + 8
    out = np.zeros(2 * cap, dtype=np.int64)
    n = lib().orc_parse_filter(s.encode(), _p(out, _i64p), cap)
    if n < 0:
        return None
    return out[: 2 * n].reshape(n, 2)


def filter_bitmap(ast, doc_bitmaps):
    """Per-doc EvalFilterBitmap over doc tag bitmaps (n_docs, stride) u64 -> bit words."""
    ast = np.ascontiguousarray(ast, dtype=np.int64).reshape(-1)
    doc_bitmaps = np.ascontiguousarray(doc_bitmaps, dtype=np.uint64)
    n_docs, stride = doc_bitmaps.shape
    out = np.zeros((n_docs + 63) // 64, dtype=np.uint64)
    lib().orc_filter_bitmap(_p(ast, _i64p), len(ast) // 2, _p(doc_bitmaps, _u64p), n_docs, stride, _p(out, _u64p))
    return out


def compile_postfix(ast):
    ast = np.ascontiguousarray(ast, dtype=np.int64).reshape(-1)
    n_pairs = len(ast) // 2
    cap = 4 * n_pairs + 8
    expr = np.zeros(cap, dtype=np.uint32)
    posts = np.zeros(cap, dtype=np.uint64)
    npost = ctypes.c_int(0)
    opn = ctypes.c_uint32(0)
    n = lib().orc_compile_postfix(_p(ast, _i64p), n_pairs, _p(expr, _u32p), cap, _p(posts, _u64p), cap,
                                  ctypes.byref(npost), ctypes.byref(opn))
    if n < 0:
        raise ValueError("compile_postfix failed")
    return expr[:n].copy(), posts[: npost.value].copy(), int(opn.value)


def check_postfix(expr, n_postings):
    expr = np.ascontiguousarray(expr, dtype=np.uint32)
    return lib().orc_check_postfix(_p(expr, _u32p), len(expr), n_postings) == 0


def postfix_max_stack(expr):
    expr = np.ascontiguousarray(expr, dtype=np.uint32)
    return int(lib().orc_postfix_max_stack(_p(expr, _u32p), len(expr)))


def postfix_eval(expr, op_num, operands):
    """Runs the postfix program over per-posting word arrays (all of equal length)."""
    expr = np.ascontiguousarray(expr, dtype=np.uint32)
    ops = [np.ascontiguousarray(o, dtype=np.uint64) for o in operands]
    nw = len(ops[0]) if ops else 0
    arr = (_u64p * max(len(ops), 1))(*[_p(o, _u64p) for o in ops])
    out = np.zeros(nw, dtype=np.uint64)
    rc = lib().orc_postfix_eval(_p(expr, _u32p), len(expr), op_num, arr, nw, _p(out, _u64p))
    if rc != 0:
        raise ValueError("invalid postfix program")
    return out


# ---------------------------------------------------------------------------------------------
# top-K
def topk(docs, queries, ks, masks=None, mode=0, threads=None):
    """Brute-force filtered top-K. mode 0 = float accumulation (the FP32 truth), 1 = double.

    Returns (ids [Q, kmax] int64 padded with -1, scores [Q, kmax] float64, counts [Q])."""
    docs = np.ascontiguousarray(docs, dtype=np.float32)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    n_docs, dim = docs.shape
    nq = queries.shape[0]
    ks = np.ascontiguousarray(np.broadcast_to(np.asarray(ks, dtype=np.int32), (nq,)))
    kcap = int(max(1, ks.max() if nq else 1))
    ids = np.full((nq, kcap), -1, dtype=np.int64)
    sc = np.zeros((nq, kcap), dtype=np.float64)
    cnt = np.zeros(nq, dtype=np.int32)
    if masks is not None:
        masks = np.ascontiguousarray(masks, dtype=np.uint64)
        mw = masks.shape[1]
        mp = _p(masks, _u64p)
    else:
        mw, mp = 0, None
    lib().orc_topk(_p(docs, _f32p), n_docs, dim, _p(queries, _f32p), nq, mp, mw, _p(ks, _i32p), kcap, mode,
                   _p(ids, _i64p), _p(sc, _f64p), _p(cnt, _i32p), int(threads or os.cpu_count() or 1))
    return ids, sc, cnt


def ref_cpu_topk_batch(docs, doc_bitmaps, queries, asts, ks, fp32=True, doc_offset=0, n_scan=None):
    """The reference's CpuTopKBatch (10.13 recall_ref.cpp:90-140) through oracle/_ref; id sets."""
    L = ref()
    if L is None:
        raise RuntimeError("oracle/_ref was not built (no reference checkout)")
    docs = np.ascontiguousarray(docs, dtype=np.float32)
    doc_bitmaps = np.ascontiguousarray(doc_bitmaps, dtype=np.uint64)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    n_docs, dim = docs.shape
    nq = queries.shape[0]
    flat, off, ln = [], [], []
    for a in asts:
        a = np.zeros((0, 2), np.int64) if a is None else np.asarray(a, dtype=np.int64).reshape(-1, 2)
        off.append(len(flat) // 2)
        ln.append(a.shape[0])
        flat.extend(a.reshape(-1).tolist())
    flat = np.asarray(flat + [0, 0], dtype=np.int64)
    off = np.asarray(off, dtype=np.int32)
    ln = np.asarray(ln, dtype=np.int32)
    ks = np.ascontiguousarray(np.broadcast_to(np.asarray(ks, dtype=np.int32), (nq,)))
    kcap = int(max(1, ks.max()))
    ids = np.full((nq, kcap), -1, dtype=np.int64)
    cnt = np.zeros(nq, dtype=np.int32)
    n_scan = n_docs - doc_offset if n_scan is None else n_scan
    L.ref_cpu_topk_batch(_p(docs, _f32p), _p(doc_bitmaps, _u64p), n_docs, dim, doc_bitmaps.shape[1],
                         _p(queries, _f32p), nq, _p(flat, _i64p), _p(off, _i32p), _p(ln, _i32p), _p(ks, _i32p),
                         int(fp32), doc_offset, n_scan, kcap, _p(ids, _i64p), _p(cnt, _i32p))
    return [set(ids[q, : cnt[q]].tolist()) for q in range(nq)]


def words_from_bits(bits):
    """bool array of doc flags -> u64 words, bit d%64 of word d/64."""
    bits = np.asarray(bits, dtype=bool)
    n = len(bits)
    pad = (-n) % 64
    b = np.concatenate([bits, np.zeros(pad, dtype=bool)]).reshape(-1, 64)
    return (b.astype(np.uint64) << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)


def bits_from_words(words, n):
    words = np.asarray(words, dtype=np.uint64)
    b = ((words[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool).reshape(-1)
    return b[:n]
