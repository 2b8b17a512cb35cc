/*
 * frk_oracle.c -- CPU oracle for the filtered top-K pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see frk_oracle.h). Every function cites
 * the reference file:line whose behaviour it restates. Paths are relative to the reference
 * checkout.
 */
#define _POSIX_C_SOURCE 200809L
#include "frk_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------------
 * MurmurHash3_x64_128 -- Austin Appleby's public-domain hash (published algorithm, restated).
 * The reference bundles it as src/platform/murmurhash/MurmurHash3.cpp and calls it from
 * engine/core/hash/hash.cpp:10-16.
 * ---------------------------------------------------------------------------------------------- */
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

static uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

void orc_murmur3_x64_128(const void* key, int len, uint32_t seed, uint64_t out[2]) {
    const uint8_t* data = (const uint8_t*)key;
    const int nblocks = len / 16;
    uint64_t h1 = seed, h2 = seed;
    const uint64_t c1 = 0x87c37b91114253d5ULL, c2 = 0x4cf5ad432745937fULL;
    for (int i = 0; i < nblocks; i++) {
        uint64_t k1, k2;
        memcpy(&k1, data + 16 * i, 8);
        memcpy(&k2, data + 16 * i + 8, 8);
        k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1;
        h1 = rotl64(h1, 27); h1 += h2; h1 = h1 * 5 + 0x52dce729;
        k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2;
        h2 = rotl64(h2, 31); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
    }
    const uint8_t* tail = data + nblocks * 16;
    uint64_t k1 = 0, k2 = 0;
    switch (len & 15) {
        case 15: k2 ^= (uint64_t)tail[14] << 48; /* fall through */
        case 14: k2 ^= (uint64_t)tail[13] << 40; /* fall through */
        case 13: k2 ^= (uint64_t)tail[12] << 32; /* fall through */
        case 12: k2 ^= (uint64_t)tail[11] << 24; /* fall through */
        case 11: k2 ^= (uint64_t)tail[10] << 16; /* fall through */
        case 10: k2 ^= (uint64_t)tail[9] << 8;   /* fall through */
        case 9:
            k2 ^= (uint64_t)tail[8];
            k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2;
            /* fall through */
        case 8: k1 ^= (uint64_t)tail[7] << 56; /* fall through */
        case 7: k1 ^= (uint64_t)tail[6] << 48; /* fall through */
        case 6: k1 ^= (uint64_t)tail[5] << 40; /* fall through */
        case 5: k1 ^= (uint64_t)tail[4] << 32; /* fall through */
        case 4: k1 ^= (uint64_t)tail[3] << 24; /* fall through */
        case 3: k1 ^= (uint64_t)tail[2] << 16; /* fall through */
        case 2: k1 ^= (uint64_t)tail[1] << 8;  /* fall through */
        case 1:
            k1 ^= (uint64_t)tail[0];
            k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1;
    }
    h1 ^= (uint64_t)len;
    h2 ^= (uint64_t)len;
    h1 += h2;
    h2 += h1;
    h1 = fmix64(h1);
    h2 = fmix64(h2);
    h1 += h2;
    h2 += h1;
    out[0] = h1;
    out[1] = h2;
}

/* Hash64: seed 0x20160713, out[0] ^ out[1] (engine/core/hash/hash.cpp:10-16). */
uint64_t orc_hash64(const char* data, uint32_t len) {
    uint64_t out[2];
    orc_murmur3_x64_128(data, (int)len, 0x20160713u, out);
    return out[0] ^ out[1];
}

/* A tag's token id is Hash64(to_string(tag)) (harness/query/filter_setup.cpp:100;
 * indexer/builder/segment/inverted_builder.cpp:240). */
uint64_t orc_tag_token(uint64_t tag) {
    char buf[32];
    int n = 0;
    char rev[32];
    do {
        rev[n++] = (char)('0' + tag % 10);
        tag /= 10;
    } while (tag != 0);
    for (int i = 0; i < n; i++) buf[i] = rev[n - 1 - i];
    return orc_hash64(buf, (uint32_t)n);
}

/* ------------------------------------------------------------------------------------------------
 * Legacy FP16 codec with the reference's flush-to-zero rules (engine/core/number/trans_number.cpp:38-105).
 * ---------------------------------------------------------------------------------------------- */
uint16_t orc_f32_to_f16(float f) {
    uint32_t bits;
    memcpy(&bits, &f, 4);
    uint32_t sign = (bits >> 31) & 1u, exponent = (bits >> 23) & 0xFFu, mantissa = bits & 0x7FFFFFu;
    if (exponent == 0) return (uint16_t)(sign << 15);
    if (exponent == 0xFF) {
        uint16_t nm = (uint16_t)((mantissa >> 13) & 0x3FF);
        return (uint16_t)((sign << 15) | 0x7C00 | (nm != 0 ? nm : 1));
    }
    int32_t e16 = (int32_t)exponent - 127 + 15;
    if (e16 >= 0x1F) return (uint16_t)((sign << 15) | 0x7C00);
    if (e16 <= 0) return (uint16_t)(sign << 15);
    uint32_t m16 = mantissa >> 13;
    uint32_t round_bit = (mantissa >> 12) & 1u, sticky = mantissa & 0xFFFu;
    if (round_bit && ((m16 & 1u) || sticky)) {
        m16 += 1;
        if (m16 > 0x3FF) {
            m16 = 0;
            e16 += 1;
            if (e16 >= 0x1F) return (uint16_t)((sign << 15) | 0x7C00);
        }
    }
    return (uint16_t)((sign << 15) | ((uint32_t)e16 << 10) | m16);
}

float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = (h >> 15) & 1u, exponent = (h >> 10) & 0x1Fu, mantissa = h & 0x3FFu, bits;
    if (exponent == 0) {
        bits = sign << 31;
    } else if (exponent == 0x1F) {
        bits = mantissa == 0 ? ((sign << 31) | 0x7F800000u) : ((sign << 31) | 0x7FC00000u | (mantissa << 13));
    } else {
        bits = (sign << 31) | ((exponent - 15 + 127) << 23) | (mantissa << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* ------------------------------------------------------------------------------------------------
 * Posting codec. Layout codes: format/posting_layout.h:8. Selection: format/posting_serializer.cpp:65-79.
 * Writers: posting_serializer.cpp:103-211. Head word [layout:4][hitCount:28], then u32 payload length
 * (posting_serializer.cpp:15-39).
 * ---------------------------------------------------------------------------------------------- */
int orc_select_layout(uint32_t hits, uint32_t doc_num, int is_matrix, float threshold, int packed) {
    if (doc_num == 0) return 0;
    /* posting_serializer.cpp:72-73: the quotient is taken in double and stored as a float */
    float density = (float)((double)hits * 1.0 / (double)doc_num);
    if (density < threshold) {
        if (is_matrix) return 3;
        return packed ? 4 : 1;
    }
    return is_matrix ? 2 : 0;
}

static int put_u32(uint8_t* out, uint64_t cap, uint64_t* pos, uint32_t v) {
    if (*pos + 4 > cap) return -1;
    memcpy(out + *pos, &v, 4);
    *pos += 4;
    return 0;
}

static int put_u16(uint8_t* out, uint64_t cap, uint64_t* pos, uint16_t v) {
    if (*pos + 2 > cap) return -1;
    memcpy(out + *pos, &v, 2);
    *pos += 2;
    return 0;
}

/* ids: the posting's local doc ids (unique). For SPARSE_IDLIST they are written in the given order
 * (the reference writes unordered_set iteration order, posting_serializer.cpp:197-210). */
int64_t orc_posting_encode(const uint32_t* ids, uint32_t n_ids, uint32_t doc_num, int is_matrix, float threshold,
                           int packed, uint8_t* out, uint64_t cap) {
    int layout = orc_select_layout(n_ids, doc_num, is_matrix, threshold, packed);
    uint32_t units = (doc_num + 15) / 16;
    if (n_ids > 0x0FFFFFFFu) return -1;
    uint32_t head = ((uint32_t)layout << 28) | n_ids;
    uint64_t pos = 0;
    uint16_t* masks = NULL;
    if (layout != 3) {
        masks = (uint16_t*)calloc(units ? units : 1, sizeof(uint16_t));
        if (!masks) return -1;
        for (uint32_t i = 0; i < n_ids; i++) {
            if (ids[i] >= doc_num) {
                /* ids beyond docNum never land in a unit the writer emits */
                if (ids[i] / 16 >= units) continue;
            }
            masks[ids[i] / 16] |= (uint16_t)(1u << (ids[i] % 16));
        }
    }
    int rc = 0;
    if (layout == 0) {
        rc |= put_u32(out, cap, &pos, head);
        rc |= put_u32(out, cap, &pos, units * 2);
        for (uint32_t u = 0; u < units && !rc; u++) rc |= put_u16(out, cap, &pos, masks[u]);
    } else if (layout == 2) {
        rc |= put_u32(out, cap, &pos, head);
        rc |= put_u32(out, cap, &pos, doc_num);
        if (!rc && pos + doc_num > cap) rc = -1;
        for (uint32_t d = 0; d < doc_num && !rc; d++) out[pos++] = (uint8_t)((masks[d / 16] >> (d % 16)) & 1u);
    } else if (layout == 1 || layout == 4) {
        uint32_t k = 0;
        for (uint32_t u = 0; u < units; u++) k += masks[u] != 0;
        rc |= put_u32(out, cap, &pos, head);
        rc |= put_u32(out, cap, &pos, layout == 1 ? k * 6 : k * 4);
        if (layout == 1) {
            for (uint32_t u = 0; u < units && !rc; u++)
                if (masks[u]) rc |= put_u32(out, cap, &pos, u * 2);
            for (uint32_t u = 0; u < units && !rc; u++)
                if (masks[u]) rc |= put_u16(out, cap, &pos, masks[u]);
        } else {
            for (uint32_t u = 0; u < units && !rc; u++) {
                if (!masks[u]) continue;
                if (u > 0xFFFFu) { rc = -1; break; }
                rc |= put_u32(out, cap, &pos, (u << 16) | masks[u]);
            }
        }
    } else {
        rc |= put_u32(out, cap, &pos, head);
        rc |= put_u32(out, cap, &pos, n_ids * 4);
        for (uint32_t i = 0; i < n_ids && !rc; i++) rc |= put_u32(out, cap, &pos, ids[i]);
    }
    free(masks);
    return rc ? -1 : (int64_t)pos;
}

/* Decodes one posting into the segment's bit words (bit d%64 of word d/64 = local doc d). The bits at
 * or above doc_num are left zero. Returns the layout code, or -1 for a malformed posting. Semantics
 * of each layout: posting_serializer.cpp:103-211 (writers); units of a dense bitmap hold doc 16u+b at
 * bit b (posting_serializer.cpp:46-55). */
int orc_posting_decode(const uint8_t* p, uint64_t len, uint32_t doc_num, uint64_t* words, uint32_t n_words) {
    if (len < 8 || (uint64_t)n_words * 64 < doc_num) return -1;
    memset(words, 0, (size_t)n_words * 8);
    uint32_t head, plen;
    memcpy(&head, p, 4);
    memcpy(&plen, p + 4, 4);
    int layout = (int)(head >> 28);
    if (8 + (uint64_t)plen > len) return -1;
    const uint8_t* body = p + 8;
    uint32_t units = (doc_num + 15) / 16;
    if (layout == 0) {
        uint32_t n = plen / 2;
        if (n > units) n = units;
        for (uint32_t u = 0; u < n; u++) {
            uint16_t m;
            memcpy(&m, body + 2 * u, 2);
            words[u / 4] |= (uint64_t)m << (16 * (u % 4));
        }
    } else if (layout == 1 || layout == 4) {
        uint32_t k = plen / (layout == 1 ? 6 : 4);
        for (uint32_t i = 0; i < k; i++) {
            uint32_t unit;
            uint16_t m;
            if (layout == 1) {
                uint32_t off;
                memcpy(&off, body + 4 * i, 4);
                memcpy(&m, body + 4 * k + 2 * i, 2);
                unit = off / 2;
            } else {
                uint32_t w;
                memcpy(&w, body + 4 * i, 4);
                unit = w >> 16;
                m = (uint16_t)(w & 0xFFFFu);
            }
            if (unit >= units) return -1;
            words[unit / 4] |= (uint64_t)m << (16 * (unit % 4));
        }
    } else if (layout == 2) {
        uint32_t n = plen < doc_num ? plen : doc_num;
        for (uint32_t d = 0; d < n; d++)
            if (body[d]) words[d / 64] |= 1ULL << (d % 64);
    } else if (layout == 3) {
        uint32_t n = plen / 4;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t d;
            memcpy(&d, body + 4 * i, 4);
            if (d >= doc_num) return -1;
            words[d / 64] |= 1ULL << (d % 64);
        }
    } else {
        return -1;
    }
    /* define the bits past the segment's own doc count as zero */
    for (uint64_t d = doc_num; d < (uint64_t)n_words * 64; d++) words[d / 64] &= ~(1ULL << (d % 64));
    return layout;
}

/* ------------------------------------------------------------------------------------------------
 * Filter grammar: harness/query/filter_expr.h:31-207 (tokenizer :31-78, ParseAtom :82-104,
 * ParseNot :106-121, ParseAnd :123-144, ParseOr :146-167, ParseFilter :169-183). Precedence
 * NOT > AND > OR; AND/OR chains become one n-ary node.
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
    const char* s;
    size_t n, i;
} Tok;

enum { T_EOF, T_LP, T_RP, T_NUM, T_AND, T_OR, T_NOT, T_OTHER };

static void skip_ws(Tok* t) {
    while (t->i < t->n && (t->s[t->i] == ' ' || t->s[t->i] == '\t' || t->s[t->i] == '\r' || t->s[t->i] == '\n'))
        t->i++;
}

/* Returns the token kind at the cursor; *wlen = its length, *num = the number's value (u64 wrap). */
static int peek(Tok* t, uint64_t* num, size_t* wlen) {
    skip_ws(t);
    *wlen = 0;
    if (t->i >= t->n) return T_EOF;
    char c = t->s[t->i];
    if (c == '(') { *wlen = 1; return T_LP; }
    if (c == ')') { *wlen = 1; return T_RP; }
    if (c >= '0' && c <= '9') {
        size_t j = t->i;
        uint64_t v = 0;
        while (j < t->n && t->s[j] >= '0' && t->s[j] <= '9') v = v * 10 + (uint64_t)(t->s[j++] - '0');
        *num = v;
        *wlen = j - t->i;
        return T_NUM;
    }
    size_t j = t->i;
    while (j < t->n && ((t->s[j] >= 'A' && t->s[j] <= 'Z') || (t->s[j] >= 'a' && t->s[j] <= 'z'))) j++;
    *wlen = j - t->i;
    if (*wlen == 3 && !strncmp(t->s + t->i, "AND", 3)) return T_AND;
    if (*wlen == 2 && !strncmp(t->s + t->i, "OR", 2)) return T_OR;
    if (*wlen == 3 && !strncmp(t->s + t->i, "NOT", 3)) return T_NOT;
    return T_OTHER;
}

typedef struct {
    int64_t* out;
    int cap, n, err;
} Emit;

static int emit_pair(Emit* e, int64_t type, int64_t arg) {
    if (e->n >= e->cap) {
        e->err = -2;
        return -1;
    }
    e->out[2 * e->n] = type;
    e->out[2 * e->n + 1] = arg;
    return e->n++;
}

static int parse_or(Tok* t, Emit* e);

static int parse_atom(Tok* t, Emit* e) {
    uint64_t num = 0;
    size_t w;
    int k = peek(t, &num, &w);
    if (k == T_LP) {
        t->i += w;
        if (parse_or(t, e) < 0) return -1;
        if (peek(t, &num, &w) != T_RP) return -1;
        t->i += w;
        return 0;
    }
    if (k == T_NUM) {
        t->i += w;
        return emit_pair(e, ORC_TERM, (int64_t)num) < 0 ? -1 : 0;
    }
    return -1;
}

static int parse_not(Tok* t, Emit* e) {
    uint64_t num = 0;
    size_t w;
    if (peek(t, &num, &w) == T_NOT) {
        t->i += w;
        int at = emit_pair(e, ORC_NOT, 1);
        if (at < 0) return -1;
        return parse_not(t, e);
    }
    return parse_atom(t, e);
}

/* Moves pairs [from, n) up by one and writes (type, count) at `from`: the n-ary node header goes in
 * front of the children it was not known to have when they were parsed. */
static int insert_header(Emit* e, int from, int64_t type, int64_t count) {
    if (e->n >= e->cap) {
        e->err = -2;
        return -1;
    }
    memmove(e->out + 2 * (from + 1), e->out + 2 * from, sizeof(int64_t) * 2 * (size_t)(e->n - from));
    e->out[2 * from] = type;
    e->out[2 * from + 1] = count;
    e->n++;
    return 0;
}

static int parse_and(Tok* t, Emit* e) {
    int start = e->n;
    if (parse_not(t, e) < 0) return -1;
    int parts = 1;
    uint64_t num = 0;
    size_t w;
    while (peek(t, &num, &w) == T_AND) {
        t->i += w;
        if (parse_not(t, e) < 0) return -1;
        parts++;
    }
    if (parts > 1 && insert_header(e, start, ORC_AND, parts) < 0) return -1;
    return 0;
}

static int parse_or(Tok* t, Emit* e) {
    int start = e->n;
    if (parse_and(t, e) < 0) return -1;
    int parts = 1;
    uint64_t num = 0;
    size_t w;
    while (peek(t, &num, &w) == T_OR) {
        t->i += w;
        if (parse_and(t, e) < 0) return -1;
        parts++;
    }
    if (parts > 1 && insert_header(e, start, ORC_OR, parts) < 0) return -1;
    return 0;
}

int orc_parse_filter(const char* s, int64_t* out_pairs, int cap_pairs) {
    Tok t = {s, strlen(s), 0};
    Emit e = {out_pairs, cap_pairs, 0, 0};
    uint64_t num = 0;
    size_t w;
    if (peek(&t, &num, &w) == T_EOF) return 0; /* empty -> match all */
    if (parse_or(&t, &e) < 0) return e.err ? e.err : -1;
    if (peek(&t, &num, &w) != T_EOF) return -1; /* trailing garbage */
    return e.n;
}

/* ------------------------------------------------------------------------------------------------
 * Per-doc evaluation: EvalFilterBitmap, harness/query/filter_expr.h:212-232 (BitSet :203-209: a tag
 * past the doc bitmap's words is false).
 * ---------------------------------------------------------------------------------------------- */
static int subtree_end(const int64_t* ast, int at) {
    int64_t type = ast[2 * at], arg = ast[2 * at + 1];
    int next = at + 1;
    if (type == ORC_TERM) return next;
    int64_t kids = type == ORC_NOT ? 1 : arg;
    for (int64_t c = 0; c < kids; c++) next = subtree_end(ast, next);
    return next;
}

static int eval_node(const int64_t* ast, int at, const uint64_t* bm, uint32_t words) {
    int64_t type = ast[2 * at], arg = ast[2 * at + 1];
    if (type == ORC_TERM) {
        uint64_t tag = (uint64_t)arg;
        uint64_t w = tag >> 6;
        if (w >= words) return 0;
        return (int)((bm[w] >> (tag & 63u)) & 1ULL);
    }
    if (type == ORC_NOT) return !eval_node(ast, at + 1, bm, words);
    int child = at + 1;
    for (int64_t c = 0; c < arg; c++) {
        int v = eval_node(ast, child, bm, words);
        if (type == ORC_AND && !v) return 0;
        if (type == ORC_OR && v) return 1;
        child = subtree_end(ast, child);
    }
    return type == ORC_AND;
}

int orc_eval_filter_doc(const int64_t* ast, int n_pairs, const uint64_t* doc_bm, uint32_t words) {
    if (n_pairs == 0) return 1;
    return eval_node(ast, 0, doc_bm, words);
}

void orc_filter_bitmap(const int64_t* ast, int n_pairs, const uint64_t* doc_bitmaps, uint64_t n_docs,
                       uint32_t stride, uint64_t* out_words) {
    uint64_t nw = (n_docs + 63) / 64;
    memset(out_words, 0, (size_t)nw * 8);
    for (uint64_t d = 0; d < n_docs; d++)
        if (orc_eval_filter_doc(ast, n_pairs, doc_bitmaps + d * stride, stride)) out_words[d / 64] |= 1ULL << (d % 64);
}

/* ------------------------------------------------------------------------------------------------
 * Postfix compiler: retrieval/query/query_node_imp.cpp. AND/OR (EmitAndOrOp :8-47): non-term children
 * first, each pushing one stack entry, then the direct term children as posting operands, then
 * [op, postingNum, stackNum]. NOT (:65-96): the child, then [3, p, s] with p = 1 for a term child.
 * TERM (:124-160): appends one posting operand (an absent token still appends the all-zero posting,
 * index/posting/posting_field_data.cpp:487-495).
 * Root: an empty AST is match-all (searcher/stage/text_filter.cpp:443-448). A root that is a single
 * term is compiled as [OR, 1, 0] so that it filters by that term; the reference's root check
 * (GetChildrenNum() == 0, text_filter.cpp:443) would treat it as match-all.
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
    uint32_t* expr;
    int cap_expr, len;
    uint64_t* postings;
    int cap_post, np;
    uint32_t ops;
    int err;
} Prog;

static void prog_put(Prog* p, uint32_t v) {
    if (p->len >= p->cap_expr) {
        p->err = -2;
        return;
    }
    p->expr[p->len++] = v;
}

static void prog_posting(Prog* p, uint64_t tag) {
    if (p->np >= p->cap_post) {
        p->err = -2;
        return;
    }
    p->postings[p->np++] = tag;
}

static void compile_node(const int64_t* ast, int at, Prog* p) {
    int64_t type = ast[2 * at], arg = ast[2 * at + 1];
    if (type == ORC_TERM) {
        prog_posting(p, (uint64_t)arg);
        return;
    }
    if (type == ORC_NOT) {
        int child = at + 1;
        int is_term = ast[2 * child] == ORC_TERM;
        compile_node(ast, child, p);
        prog_put(p, ORC_OP_NOT);
        prog_put(p, is_term ? 1u : 0u);
        prog_put(p, is_term ? 0u : 1u);
        p->ops++;
        return;
    }
    uint32_t posting_num = 0, stack_num = 0;
    int child = at + 1;
    for (int64_t c = 0; c < arg; c++) {
        if (ast[2 * child] != ORC_TERM) {
            compile_node(ast, child, p);
            stack_num++;
        }
        child = subtree_end(ast, child);
    }
    child = at + 1;
    for (int64_t c = 0; c < arg; c++) {
        if (ast[2 * child] == ORC_TERM) {
            compile_node(ast, child, p);
            posting_num++;
        }
        child = subtree_end(ast, child);
    }
    prog_put(p, type == ORC_AND ? ORC_OP_AND : ORC_OP_OR);
    prog_put(p, posting_num);
    prog_put(p, stack_num);
    p->ops++;
}

int orc_compile_postfix(const int64_t* ast, int n_pairs, uint32_t* expr, int cap_expr, uint64_t* postings,
                        int cap_postings, int* n_postings, uint32_t* op_num) {
    Prog p = {expr, cap_expr, 0, postings, cap_postings, 0, 0, 0};
    *n_postings = 0;
    *op_num = 0;
    if (n_pairs == 0) return 0;
    if (ast[0] == ORC_TERM) {
        prog_posting(&p, (uint64_t)ast[1]);
        prog_put(&p, ORC_OP_OR);
        prog_put(&p, 1);
        prog_put(&p, 0);
        p.ops = 1;
    } else {
        compile_node(ast, 0, &p);
    }
    if (p.err) return p.err;
    *n_postings = p.np;
    *op_num = p.ops;
    return p.len;
}

/* CheckPostExpr, searcher/stage/text_filter.cpp:1220-1352: every op has enough postings and stack
 * entries, and the program ends with all postings consumed and exactly one stack entry. Returns 0
 * for a valid program. */
int orc_check_postfix(const uint32_t* expr, int len, uint32_t n_postings) {
    /* Counts are summed in 64 bits: the reference's 32-bit sums (text_filter.cpp:1230,1260,1291) would
     * let a count near 2^32 wrap past the checks; the rule itself is unchanged. */
    uint64_t in_stack = 0, remained = n_postings;
    for (int i = 0; i < len; i++) {
        uint32_t op = expr[i];
        if (op == ORC_OP_CONJ) {
            if (i + 1 >= len) return -1;
            uint64_t parts = expr[i + 1];
            if (parts == 0 || (uint64_t)i + 1 + parts >= (uint64_t)len) return -1;
            uint64_t sum = 0;
            for (uint64_t j = 0; j < parts; j++) {
                if (expr[i + 2 + j] == 0) return -1;
                sum += expr[i + 2 + j];
            }
            if (remained < sum) return -1;
            remained -= sum;
            in_stack += 1;
            i += (int)parts + 1;
        } else if (op == ORC_OP_AND || op == ORC_OP_OR || op == ORC_OP_NOT) {
            if (i + 2 >= len) return -1;
            uint64_t pn = expr[i + 1], sn = expr[i + 2];
            if (op == ORC_OP_NOT ? (pn + sn != 1) : (pn + sn == 0)) return -1;
            if (remained < pn || in_stack < sn) return -1;
            in_stack -= sn;
            remained -= pn;
            in_stack += 1;
            i += 2;
        } else {
            return -1;
        }
    }
    return (remained == 0 && in_stack == 1) ? 0 : -1;
}

/* Deepest the stack gets while a valid program runs. */
int orc_postfix_max_stack(const uint32_t* expr, int len) {
    int sp = 0, mx = 0;
    for (int i = 0; i < len; i++) {
        uint32_t op = expr[i];
        if (op == ORC_OP_CONJ) {
            i += (int)expr[i + 1] + 1;
            sp += 1;
        } else {
            sp -= (int)expr[i + 2];
            sp += 1;
            i += 2;
        }
        if (sp > mx) mx = sp;
    }
    return mx;
}

/* Postfix evaluation over one segment's words, as the device filter runs it:
 * retrieval/device/filter/kernel_bitmap_text_filter.cpp:52-119 (op dispatch and the operand cursor),
 * filter_and_op.h:9-107 (AND: all-ones start, postings then stack entries popped from the top),
 * filter_or_op.h:11-206 (OR: zero start), filter_not_op.h:9-88 (NOT: one operand),
 * filter_conjunction_op.h:9-118 (CONJ: AND over parts of the OR of each part's postings).
 * The last op's value is the result. Returns 0, or -1 for a program CheckPostExpr would reject. */
int orc_postfix_eval(const uint32_t* expr, int len, uint32_t op_num, const uint64_t* const* operands,
                     uint32_t n_words, uint64_t* out) {
    (void)op_num;
    int max_sp = orc_postfix_max_stack(expr, len);
    if (max_sp <= 0) return -1;
    uint64_t* stack = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)max_sp * n_words);
    if (!stack) return -1;
    int sp = 0;
    uint32_t pc = 0;
    for (int i = 0; i < len; i++) {
        uint32_t op = expr[i];
        uint64_t* res = (uint64_t*)malloc(sizeof(uint64_t) * n_words);
        if (op == ORC_OP_CONJ) {
            uint32_t parts = expr[i + 1];
            for (uint32_t w = 0; w < n_words; w++) res[w] = ~0ULL;
            for (uint32_t j = 0; j < parts; j++) {
                uint32_t n = expr[i + 2 + j];
                for (uint32_t w = 0; w < n_words; w++) {
                    uint64_t acc = 0;
                    for (uint32_t t = 0; t < n; t++) acc |= operands[pc + t][w];
                    res[w] &= acc;
                }
                pc += n;
            }
            i += (int)parts + 1;
        } else if (op == ORC_OP_AND || op == ORC_OP_OR) {
            uint32_t pn = expr[i + 1], sn = expr[i + 2];
            int is_and = op == ORC_OP_AND;
            for (uint32_t w = 0; w < n_words; w++) {
                uint64_t acc = is_and ? ~0ULL : 0ULL;
                for (uint32_t t = 0; t < pn; t++) acc = is_and ? (acc & operands[pc + t][w]) : (acc | operands[pc + t][w]);
                for (uint32_t t = 0; t < sn; t++) {
                    uint64_t v = stack[(size_t)(sp - 1 - (int)t) * n_words + w];
                    acc = is_and ? (acc & v) : (acc | v);
                }
                res[w] = acc;
            }
            pc += pn;
            sp -= (int)sn;
            i += 2;
        } else if (op == ORC_OP_NOT) {
            uint32_t pn = expr[i + 1];
            const uint64_t* src;
            if (pn == 1) {
                src = operands[pc++];
            } else {
                src = stack + (size_t)(sp - 1) * n_words;
                sp -= 1;
            }
            for (uint32_t w = 0; w < n_words; w++) res[w] = ~src[w];
            i += 2;
        } else {
            free(res);
            free(stack);
            return -1;
        }
        memcpy(stack + (size_t)sp * n_words, res, sizeof(uint64_t) * n_words);
        sp++;
        free(res);
    }
    memcpy(out, stack + (size_t)(sp - 1) * n_words, sizeof(uint64_t) * n_words);
    free(stack);
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Brute-force filtered top-K: CpuTopK / CpuTopKBatch, harness/recall/recall_ref.cpp:18-50 and the
 * batched form in 10.13/load/src/harness/recall/recall_ref.cpp:90-140: one pass over the docs,
 * the filter bit, the inner product accumulated in order over the dims, and a bounded heap ordered
 * by `Better` (score descending, then doc id ascending; 10.13 recall_ref.cpp:19-25).
 * mode 0: float accumulation, as the reference's --recall_fp32 truth; mode 1: double accumulation
 * (the FP64 adjudicator). Output per query: up to k (id, score) sorted best-first.
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
    double s;
    int64_t id;
} Cand;

static int better(const Cand* a, const Cand* b) {
    if (a->s != b->s) return a->s > b->s;
    return a->id < b->id;
}

/* min-heap by `better`: the root is the worst kept candidate */
static void heap_sift_down(Cand* h, int n, int i) {
    for (;;) {
        int l = 2 * i + 1, r = l + 1, w = i;
        if (l < n && better(&h[w], &h[l])) w = l;
        if (r < n && better(&h[w], &h[r])) w = r;
        if (w == i) return;
        Cand t = h[i];
        h[i] = h[w];
        h[w] = t;
        i = w;
    }
}

static void heap_sift_up(Cand* h, int i) {
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!better(&h[p], &h[i])) return;
        Cand t = h[i];
        h[i] = h[p];
        h[p] = t;
        i = p;
    }
}

static int cmp_better(const void* a, const void* b) {
    const Cand* x = (const Cand*)a;
    const Cand* y = (const Cand*)b;
    if (better(x, y)) return -1;
    if (better(y, x)) return 1;
    return 0;
}

typedef struct {
    const float* docs;
    uint64_t n_docs;
    uint32_t dim;
    const float* queries;
    const uint64_t* masks;
    uint64_t mask_words;
    const int32_t* ks;
    int32_t k_cap;
    int mode;
    int64_t* out_ids;
    double* out_scores;
    int32_t* out_counts;
    uint32_t q_begin, q_end;
} TopkJob;

static void* topk_worker(void* arg) {
    TopkJob* j = (TopkJob*)arg;
    Cand* heap = (Cand*)malloc(sizeof(Cand) * (size_t)(j->k_cap > 0 ? j->k_cap : 1));
    for (uint32_t q = j->q_begin; q < j->q_end; q++) {
        const float* qv = j->queries + (size_t)q * j->dim;
        const uint64_t* m = j->masks ? j->masks + (size_t)q * j->mask_words : NULL;
        int k = j->ks[q] < j->k_cap ? j->ks[q] : j->k_cap;
        int n = 0;
        for (uint64_t d = 0; d < j->n_docs && k > 0; d++) {
            if (m && !((m[d / 64] >> (d % 64)) & 1ULL)) continue;
            const float* v = j->docs + d * j->dim;
            Cand c;
            if (j->mode == 0) {
                float s = 0.0f;
                for (uint32_t t = 0; t < j->dim; t++) s += qv[t] * v[t];
                c.s = (double)s;
            } else {
                double s = 0.0;
                for (uint32_t t = 0; t < j->dim; t++) s += (double)qv[t] * (double)v[t];
                c.s = s;
            }
            c.id = (int64_t)d;
            if (n < k) {
                heap[n++] = c;
                heap_sift_up(heap, n - 1);
            } else if (better(&c, &heap[0])) {
                heap[0] = c;
                heap_sift_down(heap, n, 0);
            }
        }
        qsort(heap, (size_t)n, sizeof(Cand), cmp_better);
        for (int i = 0; i < n; i++) {
            j->out_ids[(size_t)q * j->k_cap + i] = heap[i].id;
            j->out_scores[(size_t)q * j->k_cap + i] = heap[i].s;
        }
        j->out_counts[q] = n;
    }
    free(heap);
    return NULL;
}

int orc_topk(const float* docs, uint64_t n_docs, uint32_t dim, const float* queries, uint32_t n_queries,
             const uint64_t* masks, uint64_t mask_words, const int32_t* ks, int32_t k_cap, int mode,
             int64_t* out_ids, double* out_scores, int32_t* out_counts, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    if ((uint32_t)n_threads > n_queries) n_threads = n_queries ? (int)n_queries : 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)n_threads);
    TopkJob* jobs = (TopkJob*)malloc(sizeof(TopkJob) * (size_t)n_threads);
    uint32_t per = (n_queries + (uint32_t)n_threads - 1) / (uint32_t)n_threads;
    for (int t = 0; t < n_threads; t++) {
        TopkJob j = {docs, n_docs, dim, queries, masks, mask_words, ks, k_cap, mode, out_ids, out_scores, out_counts,
                     0, 0};
        j.q_begin = (uint32_t)t * per;
        j.q_end = j.q_begin + per > n_queries ? n_queries : j.q_begin + per;
        if (j.q_begin > n_queries) j.q_begin = n_queries;
        jobs[t] = j;
        pthread_create(&th[t], NULL, topk_worker, &jobs[t]);
    }
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    return 0;
}
