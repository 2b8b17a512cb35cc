"""Pins the CPU oracle before anything is compared against it.

Each oracle function is checked against (a) the golden vectors written from the reference's own
compiled host code (tests/golden/ref_vectors.npz), (b) that compiled code live, when it was built
in this checkout (oracle/_ref), and (c) what the mathematics fixes independently: the published
MurmurHash3 verification value, decode-of-encode identity, brute-force numpy top-K in float64.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


# ---------------------------------------------------------------- hashing
def test_murmur3_x64_128_published_verification(oracle):
    # SMHasher's verification procedure: hash keys {0}, {0,1}, ... {0..254} with seed 256-i, hash the
    # concatenated 16-byte results with seed 0; the first 4 bytes must read 0x6384BA69.
    import ctypes
    L = oracle.lib()
    L.orc_murmur3_x64_128.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_uint64)]
    key = np.arange(256, dtype=np.uint8)
    hashes = np.zeros(256 * 2, dtype=np.uint64)
    for i in range(256):
        out = (ctypes.c_uint64 * 2)()
        L.orc_murmur3_x64_128(key.ctypes.data, i, 256 - i, out)
        hashes[2 * i], hashes[2 * i + 1] = out[0], out[1]
    final = (ctypes.c_uint64 * 2)()
    L.orc_murmur3_x64_128(hashes.ctypes.data, 256 * 16, 0, final)
    assert (final[0] & 0xFFFFFFFF) == 0x6384BA69


def test_hash64_matches_reference_vectors(gold):
    vals = [oracle.hash64(str(int(t))) for t in gold["hash_tags"]]
    assert np.array_equal(np.array(vals, dtype=np.uint64), gold["hash_tag_vals"])
    assert all(oracle.tag_token(int(t)) == int(v) for t, v in zip(gold["hash_tags"][:500], gold["hash_tag_vals"]))
    raw = gold["hash_raw"].tobytes()
    off = 0
    for n, v in enumerate(gold["hash_raw_vals"]):
        assert oracle.hash64(raw[off:off + n]) == int(v)
        off += n


def test_hash64_matches_reference_live():
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built here (no reference checkout)")
    rng = np.random.default_rng(7)
    for n in list(range(0, 70)) + [127, 128, 129, 1000]:
        b = bytes(rng.integers(0, 256, size=n, dtype=np.uint8))
        assert oracle.hash64(b) == R.ref_hash64(b, n)


# ---------------------------------------------------------------- fp16 codec
def test_f16_codec_matches_reference(gold):
    got = np.array([oracle.f32_to_f16(float(x)) for x in gold["f16_in"]], dtype=np.uint16)
    assert np.array_equal(got, gold["f16_out"])
    back = np.array([oracle.f16_to_f32(int(h)) for h in gold["f16_h"]], dtype=np.float32)
    assert np.array_equal(back.view(np.uint32), gold["f16_back"].view(np.uint32))


def test_f16_agrees_with_ieee_on_normal_range():
    # Inside the normal range the reference codec is IEEE round-to-nearest-even, which numpy implements.
    x = np.random.default_rng(3).uniform(-60000, 60000, 4000).astype(np.float32)
    x = x[np.abs(x) > 6.2e-5]
    ours = np.array([oracle.f32_to_f16(float(v)) for v in x], dtype=np.uint16)
    assert np.array_equal(ours, x.astype(np.float16).view(np.uint16))


# ---------------------------------------------------------------- posting codec
def _iter_post_cases(gold):
    ids_off = 0
    b_off = 0
    for (doc_num, is_matrix, packed, idx), nid, nb in zip(gold["post_cases"], gold["post_ids_len"],
                                                          gold["post_bytes_len"]):
        ids = gold["post_ids"][ids_off:ids_off + nid]
        blob = gold["post_bytes"][b_off:b_off + nb].tobytes()
        ids_off += nid
        b_off += nb
        yield int(doc_num), int(is_matrix), int(packed), ids, blob


def test_posting_encode_matches_reference_bytes(gold):
    n = 0
    for doc_num, is_matrix, packed, ids, blob in _iter_post_cases(gold):
        ours = oracle.posting_encode(ids, doc_num, bool(is_matrix), 0.05, bool(packed))
        layout = int.from_bytes(blob[:4], "little") >> 28
        if layout == 3:
            # SPARSE_IDLIST is written in unordered_set order: same header, same id multiset
            assert ours[:8] == blob[:8]
            assert sorted(np.frombuffer(ours[8:], np.uint32)) == sorted(np.frombuffer(blob[8:], np.uint32))
        else:
            assert ours == blob, (doc_num, is_matrix, packed, len(ids))
        n += 1
    assert n >= 200


def test_posting_decode_inverts_reference_encoding(gold):
    seen = set()
    for doc_num, is_matrix, packed, ids, blob in _iter_post_cases(gold):
        layout, words = oracle.posting_decode(blob, doc_num)
        seen.add(layout)
        want = np.zeros(doc_num, dtype=bool)
        want[ids] = True
        assert np.array_equal(oracle.bits_from_words(words, doc_num), want)
        # bits at or past doc_num are defined zero
        assert not oracle.bits_from_words(words, len(words) * 64)[doc_num:].any()
    assert seen == {0, 1, 2, 3, 4}


def test_posting_layout_threshold_boundary():
    # density < threshold is sparse; equality is dense (posting_serializer.cpp:72-79)
    assert oracle.select_layout(5, 100, False, 0.05, True) == 0
    assert oracle.select_layout(4, 100, False, 0.05, True) == 4
    assert oracle.select_layout(4, 100, False, 0.05, False) == 1
    assert oracle.select_layout(4, 100, True, 0.05, True) == 3
    assert oracle.select_layout(5, 100, True, 0.05, True) == 2
    assert oracle.select_layout(0, 100, False, -1.0, True) == 0  # the all-zero posting is dense


def test_posting_encode_matches_reference_live():
    R, RP = oracle.ref(False), oracle.ref(True)
    if R is None:
        pytest.skip("oracle/_ref not built here (no reference checkout)")
    rng = np.random.default_rng(11)
    for trial in range(60):
        doc_num = int(rng.integers(1, 140000))
        dens = float(np.exp(rng.uniform(np.log(1e-5), np.log(0.9))))
        ids = np.sort(rng.choice(doc_num, size=min(doc_num, int(dens * doc_num)), replace=False)).astype(np.uint32)
        for is_matrix in (0, 1):
            for packed, L in ((0, R), (1, RP)):
                cap = 64 + 6 * max(len(ids), 1) + 2 * doc_num
                buf = np.zeros(cap, np.uint8)
                m = L.ref_posting_serialize(ids.ctypes.data_as(oracle._u32p), len(ids), doc_num, is_matrix, 0.05,
                                            buf.ctypes.data_as(oracle._u8p), cap)
                ours = oracle.posting_encode(ids, doc_num, bool(is_matrix), 0.05, bool(packed))
                if ours[3] >> 4 == 3:
                    assert sorted(np.frombuffer(ours[8:], np.uint32)) == sorted(np.frombuffer(buf[8:m].tobytes(),
                                                                                              np.uint32))
                else:
                    assert ours == buf[:m].tobytes()


# ---------------------------------------------------------------- filter grammar + evaluation
def test_parse_filter_matches_reference(gold):
    off = 0
    for e, n in zip(gold["filter_exprs"], gold["filter_pairs_len"]):
        got = oracle.parse_filter(str(e))
        if n < 0:
            assert got is None, e
            continue
        assert got is not None, e
        assert got.shape[0] == n, e
        if n > 0:
            assert np.array_equal(got.reshape(-1), gold["filter_pairs"][2 * off:2 * (off + n)]), e
        off += max(n, 0)


def test_filter_kats_from_reference_tests():
    # harness tests (src/tests/test_filter_expr.cpp:54-116): (expression, tags present, expected)
    kats = [("47", {47}, 1), ("47", {48}, 0), ("47 AND 48", {47, 48}, 1), ("47 AND 48", {47}, 0),
            ("47 OR 48", {48}, 1), ("47 OR 48", {99}, 0), ("NOT 47", {48}, 1), ("NOT 47", {47}, 0),
            ("1 OR 2 AND 3", {1}, 1), ("1 OR 2 AND 3", {2}, 0), ("1 OR 2 AND 3", {3}, 0),
            ("1 OR 2 AND 3", {2, 3}, 1), ("NOT 1 AND 2", {2}, 1), ("NOT 1 AND 2", {1, 2}, 0),
            ("(1 OR 2) AND 3", {1, 3}, 1), ("(1 OR 2) AND 3", {1}, 0), ("(1 OR 2) AND NOT (3 OR 4)", {1}, 1),
            ("(1 OR 2) AND NOT (3 OR 4)", {1, 3}, 0), ("NOT NOT 5", {5}, 1), ("NOT NOT 5", {6}, 0),
            ("NOT (1 AND 2)", {1}, 1), ("NOT (1 AND 2)", {1, 2}, 0), ("1 AND 2 AND 3 AND 4", {1, 2, 3, 4}, 1),
            ("1 AND 2 AND 3 AND 4", {1, 2, 4}, 0), ("1 OR 2 OR 3 OR 4", {4}, 1), ("1 OR 2 OR 3 OR 4", {9}, 0),
            ("   ", set(), 1), ("35839", {35839}, 1), ("35839", {47}, 0), ("35839 AND 47", {35839, 47}, 1),
            ("0", {0}, 1), ("0 AND 63 AND 64", {0, 63, 64}, 1)]
    real = "((47 AND (728 OR 31852 OR 730) AND (546 OR 11)) AND (NOT (25985 OR 15399)) AND 89)"
    kats += [(real, {47, 728, 546, 89}, 1), (real, {47, 728, 546, 89, 25985}, 0), (real, {47, 728, 89}, 0),
             (real, {728, 546, 89}, 0)]
    words = 560
    for expr, tags, want in kats:
        ast = oracle.parse_filter(expr)
        bm = np.zeros((1, words), dtype=np.uint64)
        for t in tags:
            bm[0, t // 64] |= np.uint64(1) << np.uint64(t % 64)
        got = int(oracle.filter_bitmap(ast, bm)[0] & np.uint64(1))
        assert got == want, expr
    for bad in ["(1 OR 2", "1 OR 2)", "()", "AND 2", "1 AND", "1 2", "1 OR OR 2", "NOT"]:
        assert oracle.parse_filter(bad) is None, bad


def test_filter_eval_matches_reference(gold):
    docs = gold["eval_docs"]
    off = 0
    i = 0
    for n in gold["filter_pairs_len"]:
        if n <= 0:
            continue
        ast = gold["filter_pairs"][2 * off:2 * (off + n)]
        off += n
        got = oracle.filter_bitmap(ast, docs)
        assert np.array_equal(got, gold["eval_words"][i])
        assert np.array_equal(got, gold["eval_flat_words"][i])
        i += 1


# ---------------------------------------------------------------- postfix program
def _rand_ast_str(rng, tags, depth):
    if depth == 0 or rng.random() < 0.3:
        return str(int(rng.integers(0, tags)))
    r = rng.random()
    if r < 0.2:
        return "NOT " + _rand_ast_str(rng, tags, depth - 1)
    op = " AND " if r < 0.6 else " OR "
    return "(" + op.join(_rand_ast_str(rng, tags, depth - 1) for _ in range(int(rng.integers(2, 5)))) + ")"


def test_postfix_known_programs():
    # EmitAndOrOp (query_node_imp.cpp:8-47): non-term children first, then the term postings.
    expr, posts, ops = oracle.compile_postfix(oracle.parse_filter("1 AND (2 OR 3)"))
    assert expr.tolist() == [2, 2, 0, 1, 1, 1] and posts.tolist() == [2, 3, 1] and ops == 2
    expr, posts, ops = oracle.compile_postfix(oracle.parse_filter("NOT 5"))
    assert expr.tolist() == [3, 1, 0] and posts.tolist() == [5]
    expr, posts, ops = oracle.compile_postfix(oracle.parse_filter("NOT (5 OR 6) AND 7"))
    assert expr.tolist() == [2, 2, 0, 3, 0, 1, 1, 1, 1] and posts.tolist() == [5, 6, 7] and ops == 3
    expr, posts, ops = oracle.compile_postfix(oracle.parse_filter("9"))
    assert expr.tolist() == [2, 1, 0] and posts.tolist() == [9]
    expr, posts, ops = oracle.compile_postfix(np.zeros((0, 2), np.int64))
    assert len(expr) == 0 and ops == 0


def test_postfix_eval_equals_per_doc_eval():
    rng = np.random.default_rng(5)
    n_docs, stride = 640, 2
    docs = rng.integers(0, 2 ** 63, size=(n_docs, stride), dtype=np.int64).astype(np.uint64)
    docs &= rng.integers(0, 2 ** 63, size=(n_docs, stride), dtype=np.int64).astype(np.uint64)
    tag_bits = oracle.bits_from_words(docs.reshape(-1), n_docs * stride * 64).reshape(n_docs, stride * 64)
    for trial in range(300):
        ast = oracle.parse_filter(_rand_ast_str(rng, 140, 4))
        expr, posts, ops = oracle.compile_postfix(ast)
        assert oracle.check_postfix(expr, len(posts))
        operands = []
        for t in posts:
            col = tag_bits[:, t] if t < stride * 64 else np.zeros(n_docs, bool)
            operands.append(oracle.words_from_bits(col))
        got = oracle.postfix_eval(expr, ops, operands)
        want = oracle.filter_bitmap(ast, docs)
        assert np.array_equal(got, want)


def test_check_postfix_rejects_malformed():
    assert not oracle.check_postfix(np.array([1, 0, 0], np.uint32), 0)       # AND of nothing
    assert not oracle.check_postfix(np.array([1, 2, 0], np.uint32), 1)       # too few postings
    assert not oracle.check_postfix(np.array([3, 1, 1], np.uint32), 1)       # NOT of two operands
    assert not oracle.check_postfix(np.array([2, 1, 0, 2, 1, 0], np.uint32), 2)  # two results left
    assert not oracle.check_postfix(np.array([7, 1, 0], np.uint32), 1)       # unknown op
    assert oracle.check_postfix(np.array([0, 2, 1, 2], np.uint32), 3)        # CONJ (1) AND (2 OR 3)


def test_conjunction_semantics():
    # CONJ = AND over parts of the OR of each part's postings (filter_conjunction_op.h:86-118)
    rng = np.random.default_rng(9)
    ops = [rng.integers(0, 2 ** 63, 8, dtype=np.int64).astype(np.uint64) for _ in range(5)]
    got = oracle.postfix_eval(np.array([0, 3, 1, 2, 2], np.uint32), 1, ops)
    assert np.array_equal(got, ops[0] & (ops[1] | ops[2]) & (ops[3] | ops[4]))


# ---------------------------------------------------------------- top-K
def _numpy_topk64(docs, q, mask_bits, k):
    s = docs.astype(np.float64) @ q.astype(np.float64)
    idx = np.nonzero(mask_bits)[0]
    order = np.lexsort((idx, -s[idx]))
    sel = idx[order[:k]]
    return sel, s[sel]


def test_topk_fp64_matches_numpy_bruteforce():
    rng = np.random.default_rng(13)
    docs = rng.standard_normal((3000, 24)).astype(np.float32)
    docs[100:140] = docs[50]  # exact duplicates: ties broken by doc id
    qs = rng.standard_normal((10, 24)).astype(np.float32)
    bits = rng.random((10, 3000)) < 0.4
    masks = np.stack([oracle.words_from_bits(b) for b in bits])
    ks = np.array([1, 7, 50, 3000, 100, 5, 40, 41, 0, 1200], np.int32)
    ids, sc, cnt = oracle.topk(docs, qs, ks, masks, mode=1)
    for q in range(10):
        want_ids, want_s = _numpy_topk64(docs, qs[q], bits[q], int(ks[q]))
        assert cnt[q] == len(want_ids)
        assert ids[q, :cnt[q]].tolist() == want_ids.tolist()
        np.testing.assert_allclose(sc[q, :cnt[q]], want_s, rtol=1e-12, atol=1e-12)


def test_topk_fp32_matches_reference_vectors(gold):
    bm = gold["topk_bm"]
    qs = gold["topk_q"]
    ks = gold["topk_k"]
    asts = [oracle.parse_filter(str(e)) for e in gold["topk_filters"]]
    for name in ("topk", "topk_tie"):
        docs = gold[name + "_docs"]
        masks = np.stack([oracle.filter_bitmap(a, bm) for a in asts])
        ids, sc, cnt = oracle.topk(docs, qs, ks, masks, mode=0)
        off = 0
        for q, n in enumerate(gold[name + "_lens"]):
            want = set(gold[name + "_ids"][off:off + n].tolist())
            off += n
            assert set(ids[q, :cnt[q]].tolist()) == want, (name, q)


def test_topk_fp32_matches_reference_live():
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built here (no reference checkout)")
    rng = np.random.default_rng(17)
    nd, dim, nq = 2500, 20, 16
    docs = np.round(rng.standard_normal((nd, dim)) * 3).astype(np.float32) / 3  # heavy ties
    bm = np.stack([oracle.words_from_bits(r) for r in (rng.random((nd, 64)) < 0.25)]).reshape(nd, 1)
    qs = np.round(rng.standard_normal((nq, dim))).astype(np.float32)
    exprs = [_rand_ast_str(rng, 64, 3) for _ in range(nq)]
    asts = [oracle.parse_filter(e) for e in exprs]
    ks = rng.choice([1, 3, 30, 400, 4000], size=nq).astype(np.int32)
    want = oracle.ref_cpu_topk_batch(docs, bm, qs, asts, ks, fp32=True)
    masks = np.stack([oracle.filter_bitmap(a, bm) for a in asts])
    ids, sc, cnt = oracle.topk(docs, qs, ks, masks, mode=0)
    for q in range(nq):
        assert set(ids[q, :cnt[q]].tolist()) == want[q]
