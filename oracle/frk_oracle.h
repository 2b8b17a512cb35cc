/*
 * frk_oracle.h -- CPU oracle for the filtered top-K pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the parity tests, the smoke check and the
 * bench's CPU-baseline leg compare the HIP path against. The product library never links, loads or
 * calls it, and nothing here is on the product path.
 *
 * Each function restates the reference algorithm it names (file:line, paths relative to the
 * reference checkout). Plain C99, single translation unit, no dependencies.
 *
 * Filter ASTs cross this interface in a flat pre-order form of int64 pairs (type, arg):
 *   ORC_TERM: arg = tag id; ORC_AND / ORC_OR: arg = number of children (the children follow);
 *   ORC_NOT: arg = 1 (one child follows).
 * An empty AST (n_pairs == 0) means "match all", as in harness/query/filter_expr.h:196-205.
 */
#ifndef FRK_ORACLE_H
#define FRK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TERM = 0, ORC_AND = 1, ORC_OR = 2, ORC_NOT = 3 };

/* Postfix opcodes, engine/retrieval/device/device_common.h:8-13. */
enum { ORC_OP_CONJ = 0, ORC_OP_AND = 1, ORC_OP_OR = 2, ORC_OP_NOT = 3 };

/* ---- hashing and numeric codecs ---- */
void orc_murmur3_x64_128(const void* key, int len, uint32_t seed, uint64_t out[2]);
uint64_t orc_hash64(const char* data, uint32_t len);
uint64_t orc_tag_token(uint64_t tag);
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);

/* ---- posting codec ---- */
int orc_select_layout(uint32_t hits, uint32_t doc_num, int is_matrix, float threshold, int packed);
int64_t orc_posting_encode(const uint32_t* ids, uint32_t n_ids, uint32_t doc_num, int is_matrix, float threshold,
                           int packed, uint8_t* out, uint64_t cap);
int orc_posting_decode(const uint8_t* p, uint64_t len, uint32_t doc_num, uint64_t* words, uint32_t n_words);

/* ---- filter grammar and per-doc evaluation ---- */
int orc_parse_filter(const char* s, int64_t* out_pairs, int cap_pairs);
int orc_eval_filter_doc(const int64_t* ast, int n_pairs, const uint64_t* doc_bm, uint32_t words);
void orc_filter_bitmap(const int64_t* ast, int n_pairs, const uint64_t* doc_bitmaps, uint64_t n_docs,
                       uint32_t stride, uint64_t* out_words);

/* ---- postfix program: compile and evaluate ---- */
int orc_compile_postfix(const int64_t* ast, int n_pairs, uint32_t* expr, int cap_expr, uint64_t* postings,
                        int cap_postings, int* n_postings, uint32_t* op_num);
int orc_check_postfix(const uint32_t* expr, int len, uint32_t n_postings);
int orc_postfix_max_stack(const uint32_t* expr, int len);
int orc_postfix_eval(const uint32_t* expr, int len, uint32_t op_num, const uint64_t* const* operands,
                     uint32_t n_words, uint64_t* out);

/* ---- brute-force filtered top-K ---- */
int orc_topk(const float* docs, uint64_t n_docs, uint32_t dim, const float* queries, uint32_t n_queries,
             const uint64_t* masks, uint64_t mask_words, const int32_t* ks, int32_t k_cap, int mode,
             int64_t* out_ids, double* out_scores, int32_t* out_counts, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
