"""Writes the golden vectors in this folder from the reference's own host code.

Run in a container that has the reference checkout: ``make -C oracle`` builds ``oracle/_ref`` from
the reference sources, then ``python tests/golden/make_golden.py`` writes ``ref_vectors.npz``.
The vectors are data only (inputs and the reference's outputs); the parity tests check the CPU
oracle against them wherever the reference itself is absent.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

KAT_FILTERS = [
    "47", "47 AND 48", "47 OR 48", "NOT 47", "1 OR 2 AND 3", "NOT 1 AND 2", "(1 OR 2) AND 3",
    "(1 OR 2) AND NOT (3 OR 4)", "NOT NOT 5", "NOT (1 AND 2)", "1 AND 2 AND 3 AND 4", "1 OR 2 OR 3 OR 4",
    "", "   ", "(1 OR 2", "1 OR 2)", "()", "AND 2", "1 AND", "1 2", "1 OR OR 2", "NOT", "35839",
    "0 AND 63 AND 64", "((47 AND (728 OR 31852 OR 730) AND (546 OR 11)) AND (NOT (25985 OR 15399)) AND 89)",
    "(1 AND 2) AND 3", "NOTE 3", "3 and 4", "((((7))))", "NOT (NOT (1 OR 2) AND 3)",
]


def rand_expr(rng, tags, depth):
    if depth == 0 or rng.random() < 0.3:
        return str(int(rng.integers(0, tags)))
    r = rng.random()
    if r < 0.2:
        return "NOT " + rand_expr(rng, tags, depth - 1)
    op = " AND " if r < 0.6 else " OR "
    parts = [rand_expr(rng, tags, depth - 1) for _ in range(int(rng.integers(2, 4)))]
    return "(" + op.join(parts) + ")"


def main():
    R = oracle.ref(False)
    RP = oracle.ref(True)
    if R is None:
        raise SystemExit("oracle/_ref not built: run make -C oracle with the reference checkout present")
    rng = np.random.default_rng(20261203)
    out = {}
    # Hash64 of tag strings and of raw byte strings of every length 0..40
    tags = np.arange(0, 5000, dtype=np.uint64)
    out["hash_tags"] = tags
    out["hash_tag_vals"] = np.array([R.ref_hash64(str(int(t)).encode(), len(str(int(t)))) for t in tags],
                                    dtype=np.uint64)
    raw = [bytes(rng.integers(0, 256, size=n, dtype=np.uint8)) for n in range(41)]
    out["hash_raw"] = np.frombuffer(b"".join(raw), dtype=np.uint8)
    out["hash_raw_vals"] = np.array([R.ref_hash64(b, len(b)) for b in raw], dtype=np.uint64)
    # FP16 codec
    f = np.concatenate([rng.standard_normal(2000).astype(np.float32),
                        np.array([0.0, -0.0, 1e-8, -1e-8, 6e-5, 6.1e-5, 65504, 65520, 1e6, -1e6, np.inf, -np.inf,
                                  np.nan, 1.0009765625, 1.00048828125, 2.0 ** -14, 2.0 ** -15],
                                 dtype=np.float32)])
    out["f16_in"] = f
    out["f16_out"] = np.array([R.ref_f32_to_f16(float(x)) for x in f], dtype=np.uint16)
    h = np.arange(0, 65536, 7, dtype=np.uint16)
    out["f16_h"] = h
    out["f16_back"] = np.array([R.ref_f16_to_f32(int(x)) for x in h], dtype=np.float32)
    # postings: several doc sets x doc counts x (matrix, threshold, packed)
    cases, blobs, lens = [], [], []
    for doc_num in (1, 15, 16, 17, 100, 1000, 8192, 19264):
        for dens in (0.0, 0.001, 0.02, 0.049, 0.05, 0.2, 0.9):
            n = int(round(dens * doc_num))
            ids = np.sort(rng.choice(doc_num, size=min(n, doc_num), replace=False)).astype(np.uint32)
            for is_matrix in (0, 1):
                for packed in (0, 1):
                    L = RP if packed else R
                    cap = 64 + 6 * max(len(ids), 1) + doc_num * 2
                    buf = np.zeros(cap, dtype=np.uint8)
                    m = L.ref_posting_serialize(ids.ctypes.data_as(oracle._u32p), len(ids), doc_num, is_matrix, 0.05,
                                                buf.ctypes.data_as(oracle._u8p), cap)
                    assert m > 0
                    cases.append((doc_num, is_matrix, packed, len(blobs)))
                    blobs.append((ids, buf[:m].copy()))
    out["post_cases"] = np.array(cases, dtype=np.int64)
    out["post_ids"] = np.concatenate([b[0] for b in blobs])
    out["post_ids_len"] = np.array([len(b[0]) for b in blobs], dtype=np.int64)
    out["post_bytes"] = np.concatenate([b[1] for b in blobs])
    out["post_bytes_len"] = np.array([len(b[1]) for b in blobs], dtype=np.int64)
    # filter grammar: KATs plus random expressions
    exprs = list(KAT_FILTERS) + [rand_expr(rng, 70, 4) for _ in range(300)]
    parsed, plen = [], []
    for e in exprs:
        buf = np.zeros(2 * 4096, dtype=np.int64)
        n = R.ref_parse_filter(e.encode(), buf.ctypes.data_as(oracle._i64p), 4096)
        plen.append(n)
        if n > 0:
            parsed.append(buf[: 2 * n])
    out["filter_exprs"] = np.array(exprs, dtype=object).astype(str)
    out["filter_pairs_len"] = np.array(plen, dtype=np.int64)
    out["filter_pairs"] = np.concatenate(parsed)
    # per-doc evaluation of every parsed expression over random tag bitmaps (stride 2 -> tags 0..127)
    docs_bm = rng.integers(0, 2 ** 63, size=(512, 2), dtype=np.int64).astype(np.uint64)
    docs_bm &= rng.integers(0, 2 ** 63, size=(512, 2), dtype=np.int64).astype(np.uint64)
    out["eval_docs"] = docs_bm
    ev, evf = [], []
    off = 0
    for n in plen:
        if n <= 0:
            continue
        a = out["filter_pairs"][2 * off: 2 * (off + n)].copy()
        off += n
        w = np.zeros(8, dtype=np.uint64)
        R.ref_filter_bitmap(a.ctypes.data_as(oracle._i64p), n, docs_bm.ctypes.data_as(oracle._u64p), 512, 2,
                            w.ctypes.data_as(oracle._u64p))
        wf = np.zeros(8, dtype=np.uint64)
        R.ref_flat_filter_bitmap(a.ctypes.data_as(oracle._i64p), n, docs_bm.ctypes.data_as(oracle._u64p), 512, 2,
                                 wf.ctypes.data_as(oracle._u64p))
        ev.append(w)
        evf.append(wf)
    out["eval_words"] = np.stack(ev)
    out["eval_flat_words"] = np.stack(evf)
    # CpuTopKBatch id sets (fp32 truth) on a small corpus, with tie-heavy variants
    nd, dim, nq = 2000, 16, 24
    docs = rng.standard_normal((nd, dim)).astype(np.float32)
    dbm = (rng.random((nd, 64)) < 0.3)
    bm = np.stack([oracle.words_from_bits(r) for r in dbm]).reshape(nd, 1)
    qs = rng.standard_normal((nq, dim)).astype(np.float32)
    tie_docs = np.round(docs * 2) / 2  # few value levels -> many exact ties
    fexprs = [rand_expr(rng, 64, 3) for _ in range(nq)]
    asts = [oracle.parse_filter(e) for e in fexprs]
    ks = rng.choice([1, 5, 40, 5000], size=nq).astype(np.int32)
    for name, d in (("topk", docs), ("topk_tie", tie_docs)):
        sets = oracle.ref_cpu_topk_batch(d, bm, qs, asts, ks, fp32=True)
        out[name + "_docs"] = d
        out[name + "_ids"] = np.concatenate([np.array(sorted(s), dtype=np.int64) for s in sets])
        out[name + "_lens"] = np.array([len(s) for s in sets], dtype=np.int64)
    out["topk_bm"] = bm
    out["topk_q"] = qs
    out["topk_k"] = ks
    out["topk_filters"] = np.array(fexprs).astype(str)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))


if __name__ == "__main__":
    main()
