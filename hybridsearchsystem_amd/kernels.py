"""Thin Python wrappers over the kernel-level C ABI (include/frk_capi.h, layer 1).

Device buffers are torch tensors on the HIP device; every call is stream-ordered on torch's current
stream unless a stream is given, and raises FrkError on a failed call.
"""
import ctypes

import numpy as np

from . import DocStore, FilterQuery, check, lib, ptr, stream_ptr

FRK_DTYPE_F32, FRK_DTYPE_BF16, FRK_DTYPE_FP8E4M3, FRK_DTYPE_F16X2 = 0, 1, 2, 3


def chunk_dims(dtype):
    return int(lib().frk_chunk_dims(dtype))


def padded_dim(dim, dtype):
    """The store dim the scorer runs over: dim zero-padded to the next compiled chunk count."""
    return int(lib().frk_store_dim(dim, dtype))


def tile_docs(vectors, dtype, dim_pad=None, threads=0):
    """Row-major FP32 [n, dim] -> (tiles uint8 ndarray, scales float32 ndarray or None, dim_pad)."""
    v = np.ascontiguousarray(vectors, dtype=np.float32)
    n, dim = v.shape
    dim_pad = dim_pad or padded_dim(dim, dtype)
    nbytes = int(lib().frk_tiles_bytes(n, dim_pad, dtype))
    if nbytes == 0:
        raise ValueError("bad dim_pad %d for dtype %d" % (dim_pad, dtype))
    tiles = np.empty(nbytes, dtype=np.uint8)
    scales = np.empty(n, dtype=np.float32) if dtype in (FRK_DTYPE_FP8E4M3, FRK_DTYPE_F16X2) else None
    check(lib().frk_tile_docs(v.ctypes.data, n, dim, dtype, dim_pad, tiles.ctypes.data,
                              scales.ctypes.data if scales is not None else None, threads), "frk_tile_docs")
    return tiles, scales, dim_pad


def doc_store(tiles_dev, scales_dev, dtype, dim_pad, n_docs):
    return DocStore(ptr(tiles_dev), ptr(scales_dev), dtype, dim_pad, n_docs)


def filter_queries_array(descs):
    arr = (FilterQuery * len(descs))()
    for i, d in enumerate(descs):
        arr[i] = FilterQuery(*d)
    return np.frombuffer(bytes(arr), dtype=np.uint32).reshape(len(descs), 6)


def filter_eval(d_queries, n_queries, d_expr, d_table, n_segments, seg_docs, n_docs, d_out, out_stride, d_counts,
                d_status, max_postings=0, stream=None):
    check(lib().frk_filter_eval(ptr(d_queries), n_queries, ptr(d_expr), ptr(d_table), n_segments, seg_docs, n_docs,
                                ptr(d_out), out_stride, ptr(d_counts), ptr(d_status), max_postings,
                                stream_ptr(stream)),
          "frk_filter_eval")


def masks_to_blocks(d_masks, mask_stride, n_queries, n_docs, d_block_masks, q_stride, stream=None):
    check(lib().frk_masks_to_blocks(ptr(d_masks), mask_stride, n_queries, n_docs, ptr(d_block_masks), q_stride,
                                    stream_ptr(stream)), "frk_masks_to_blocks")


def block_masks_host(bits):
    """bool [n_queries, n_docs] -> the scorer's per-doc query words, uint32 [n_blocks, round_up(n_queries, 32)]:
    word [b, 32 g + j] bit x = bits[32 g + x, 32 b + j]."""
    bits = np.asarray(bits, dtype=bool)
    nq, n = bits.shape
    nb = (n + 31) // 32
    qs = (nq + 31) // 32 * 32
    pad = np.zeros((qs, nb * 32), dtype=bool)
    pad[:nq, :n] = bits
    t = pad.reshape(qs // 32, 32, nb, 32)  # [g, x, b, j]
    w = (t.astype(np.uint64) << np.arange(32, dtype=np.uint64)[None, :, None, None]).sum(axis=1, dtype=np.uint64)
    return np.ascontiguousarray(w.transpose(1, 0, 2).reshape(nb, qs).astype(np.uint32))  # [b, g, j]


def score_select(store, d_queries, n_queries, d_block_masks, mask_q_stride, d_floor, block_begin, block_step,
                 block_count, d_cand, cand_cap, d_counts, stream=None, counts_stride=1):
    check(lib().frk_score_select(ctypes.byref(store), ptr(d_queries), n_queries, ptr(d_block_masks), mask_q_stride,
                                 ptr(d_floor), block_begin, block_step, block_count, ptr(d_cand), cand_cap,
                                 ptr(d_counts), counts_stride, stream_ptr(stream)), "frk_score_select")


def score_matrix(store, d_queries, n_queries, d_scores, score_stride, stream=None):
    check(lib().frk_score_matrix(ctypes.byref(store), ptr(d_queries), n_queries, ptr(d_scores), score_stride,
                                 stream_ptr(stream)), "frk_score_matrix")


def aggregate(d_scores, d_bits, n_docs, d_cand, cap, d_count, stream=None):
    check(lib().frk_aggregate(ptr(d_scores), ptr(d_bits), n_docs, ptr(d_cand), cap, ptr(d_count), stream_ptr(stream)),
          "frk_aggregate")


def select_topk(d_cand, cand_cap, d_counts, n_queries, d_k, k_default, d_out, k_stride, d_out_counts, d_kth=None,
                stream=None, counts_stride=1):
    check(lib().frk_select_topk(ptr(d_cand), cand_cap, ptr(d_counts), counts_stride, n_queries, ptr(d_k), k_default,
                                ptr(d_out), k_stride, ptr(d_out_counts), ptr(d_kth), stream_ptr(stream)),
          "frk_select_topk")


def select_top_set(d_cand, cand_cap, d_counts, n_queries, d_k, k_default, d_out, k_stride, d_out_counts, d_kth=None,
                   stream=None, counts_stride=1):
    check(lib().frk_select_top_set(ptr(d_cand), cand_cap, ptr(d_counts), counts_stride, n_queries, ptr(d_k), k_default,
                                   ptr(d_out), k_stride, ptr(d_out_counts), ptr(d_kth), stream_ptr(stream)),
          "frk_select_top_set")


def decode_keys(keys):
    """uint64 candidate keys -> (scores float32, doc ids int64)."""
    keys = np.asarray(keys, dtype=np.uint64)
    hi = (keys >> np.uint64(32)).astype(np.uint32)
    lo = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    neg = (hi & np.uint32(0x80000000)) == 0
    bits = np.where(neg, ~hi, hi & np.uint32(0x7FFFFFFF)).astype(np.uint32)
    scores = bits.view(np.float32)
    docs = (np.uint32(0xFFFFFFFF) - lo).astype(np.int64)
    return scores, docs
