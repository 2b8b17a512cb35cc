"""Seeded synthetic corpora and posting arenas for the parity tests (test infrastructure)."""
import numpy as np

import oracle


def rand_filter(rng, n_tags, depth, absent_tags=0):
    """Random boolean filter string: AND/OR/NOT over tag ids (a few may be absent from the index)."""
    if depth == 0 or rng.random() < 0.3:
        if absent_tags and rng.random() < 0.05:
            return str(n_tags + int(rng.integers(0, absent_tags)))
        return str(int(rng.integers(0, n_tags)))
    r = rng.random()
    if r < 0.2:
        return "NOT " + rand_filter(rng, n_tags, depth - 1, absent_tags)
    op = " AND " if r < 0.6 else " OR "
    return "(" + op.join(rand_filter(rng, n_tags, depth - 1, absent_tags)
                         for _ in range(int(rng.integers(2, 4)))) + ")"


def tag_matrix(rng, n_docs, densities):
    return rng.random((n_docs, len(densities))) < np.asarray(densities)[None, :]


def doc_tag_words(tags):
    """bool [n_docs, n_tags] -> per-doc tag bitmaps u64 [n_docs, ceil(n_tags/64)] (HYDSET2 style)."""
    n, t = tags.shape
    stride = (t + 63) // 64
    pad = np.zeros((n, stride * 64), dtype=bool)
    pad[:, :t] = tags
    w = (pad.reshape(n, stride, 64).astype(np.uint64) << np.arange(64, dtype=np.uint64)).sum(axis=2, dtype=np.uint64)
    return np.ascontiguousarray(w)


def build_posting_arena(tags, seg_docs, layout_of_tag, threshold=0.05):
    """Encodes every (tag, segment) posting with the oracle's codec (byte-identical to the reference
    serializer) into one arena; each posting starts 8 bytes before a 16-byte boundary so its payload
    is 16-byte aligned, and is zero-padded to the next boundary.

    layout_of_tag(t) -> (is_matrix, packed). A tag with no hit in a segment has no posting there
    (the builder drops empty postings, inverted_builder.cpp:180-189): its offset is -1.
    Returns (arena uint8, offsets int64 [n_tags, n_segments], layouts int [n_tags, n_segments])."""
    n_docs, n_tags = tags.shape
    n_seg = (n_docs + seg_docs - 1) // seg_docs
    chunks = []
    pos = 8
    offs = np.full((n_tags, n_seg), -1, dtype=np.int64)
    lays = np.full((n_tags, n_seg), -1, dtype=np.int64)
    chunks.append(np.zeros(8, dtype=np.uint8))
    for t in range(n_tags):
        is_matrix, packed = layout_of_tag(t)
        for s in range(n_seg):
            lo, hi = s * seg_docs, min(n_docs, (s + 1) * seg_docs)
            ids = np.nonzero(tags[lo:hi, t])[0].astype(np.uint32)
            if len(ids) == 0:
                continue
            if is_matrix:
                ids = ids[np.random.default_rng(t * 1000 + s).permutation(len(ids))]  # id lists are unordered
            b = np.frombuffer(oracle.posting_encode(ids, hi - lo, is_matrix, threshold, packed), dtype=np.uint8)
            offs[t, s] = pos
            lays[t, s] = int(b[3]) >> 4
            padded = (len(b) - 8 + 15) // 16 * 16 + 8 + 8  # payload to 16 B, then 8 B to the next head
            buf = np.zeros(padded, dtype=np.uint8)
            buf[:len(b)] = b
            chunks.append(buf)
            pos += padded
    chunks.append(np.zeros(64, dtype=np.uint8))
    return np.concatenate(chunks), offs, lays


def mixture_corpus(rng, n_docs, dim, n_centroids, sigma=0.3):
    """L2-normalised draws from a Gaussian mixture (centroids N(0, I))."""
    cent = rng.standard_normal((n_centroids, dim)).astype(np.float32)
    z = rng.integers(0, n_centroids, n_docs)
    v = cent[z] + sigma * rng.standard_normal((n_docs, dim)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v.astype(np.float32), cent


def mixture_queries(rng, cent, n, noise=0.1, sigma=0.3):
    dim = cent.shape[1]
    z = rng.integers(0, cent.shape[0], n)
    q = cent[z] + sigma * rng.standard_normal((n, dim)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q += noise * rng.standard_normal((n, dim)).astype(np.float32)
    return q.astype(np.float32)
