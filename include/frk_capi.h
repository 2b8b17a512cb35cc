/*
 * frk_capi.h -- the C ABI of the MI355X (gfx950) filtered top-K pipeline.
 *
 * Two layers, both plain C (no C++ or torch types):
 *
 *  1. Kernel entry points (frk_filter_eval, frk_masks_to_blocks, frk_score_select,
 *     frk_score_blockmax, frk_score_matrix, frk_aggregate, frk_kth_largest, frk_select_topk,
 *     frk_rescore, frk_mlp_*). They replace the
 *     reference's five AscendC launch stubs, which the host stages call through
 *     ACLRT_LAUNCH_KERNEL (each entry cites the kernel it replaces). Stream-ordered, no implicit
 *     synchronisation, no allocation; the caller owns every device buffer. Reentrant across streams.
 *
 *  2. Index and search handles (frk_table_*, frk_builder_*, frk_search, filter codecs), a C face of
 *     the C++ NpuRetrieval::DataTable / FullRecallSearcher classes (hybridsearchsystem_amd/csrc/
 *     engine/), for callers that cannot link C++ (ctypes, cgo, JNI).
 *
 * Every function returns 0 (FRK_OK) or a negative FRK_ERR_* code; frk_last_error() gives the text
 * of the calling thread's last failure.
 */
#ifndef FRK_CAPI_H
#define FRK_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRK_OK 0
#define FRK_ERR_INVALID (-1)     /* bad argument or malformed program/table */
#define FRK_ERR_HIP (-2)         /* a HIP runtime call failed */
#define FRK_ERR_UNSUPPORTED (-3) /* valid input outside what this build supports (see the text) */
#define FRK_ERR_CAPACITY (-4)    /* a caller buffer is too small */
#define FRK_ERR_NOMEM (-5)       /* device or host allocation failed */
#define FRK_ERR_IO (-6)          /* index file missing or malformed */

typedef void* frk_stream_t; /* a hipStream_t; NULL = the null stream */

const char* frk_last_error(void);

/* ------------------------------------------------------------------------------------------------
 * Filter: postfix program over posting operands -> per-query bit words.
 * Replaces BitmapTextFilter (engine/retrieval/device/filter/kernel_bitmap_text_filter.cpp:52-119) and
 * PostingBitListToSet (:18-43): sparse operands are scattered straight into the words, so there is no
 * separate conversion pass.
 *
 * Program words: the reference's postfix encoding (device_common.h:8-13; query_node_imp.cpp):
 *   CONJ=0: [0, parts, n_1 .. n_parts]   AND=1 / OR=2: [op, postingNum, stackNum]
 *   NOT=3: [3, p, s] with p + s == 1.
 * Operand table: for query q, table[table_off + seg * n_postings + p] is the device address of
 * posting p's 8-byte header in segment seg ([layout:4][hitCount:28][u32 payloadLen], then the
 * payload, format/posting_serializer.cpp:15-39), or 0 for an all-zero operand. The payload must be
 * 8-byte aligned and readable up to the next multiple of 8 bytes.
 * Output: out[q * out_stride + seg * (seg_docs / 64) + w], bit b of word w = local doc 64w + b of
 * that segment; bits at or past the segment's doc count are 0. counts[q] += set bits (the caller
 * zeroes counts). Bit layout is byte-identical to the reference's per-segment u16 bitsets when
 * seg_docs is a multiple of 64 (required).
 * max_postings: the largest n_postings of the batch (0 = unknown). It sizes the LDS staging of
 * operands; a query with more postings is still evaluated exactly, with its extra operands read
 * straight from HBM.
 * ------------------------------------------------------------------------------------------------ */
typedef struct {
    uint32_t expr_off;   /* first program word in the batch's expr array */
    uint32_t expr_len;   /* program length in words (0 with FRK_FILTER_MATCH_ALL) */
    uint32_t table_off;  /* first operand pointer in the batch's table */
    uint32_t n_postings; /* operands per segment */
    uint32_t op_num;     /* ops in the program */
    uint32_t flags;      /* FRK_FILTER_* */
} frk_filter_query;

#define FRK_FILTER_MATCH_ALL 1u   /* empty tree: every doc passes (text_filter.cpp:553-559) */
#define FRK_FILTER_MAX_STACK 16   /* deepest stack a program may use */
#define FRK_FILTER_MAX_POSTINGS 1024

int frk_filter_eval(const frk_filter_query* d_queries, uint32_t n_queries, const uint32_t* d_expr,
                    const uint64_t* d_table, uint32_t n_segments, uint32_t seg_docs, uint64_t n_docs,
                    uint64_t* d_out, uint64_t out_stride, uint32_t* d_counts, uint32_t* d_status,
                    uint32_t max_postings, frk_stream_t stream);

/* Host check of one program (TextFilter::CheckPostExpr, engine/retrieval/searcher/stage/text_filter.cpp:
 * 1220-1352), the same rule the filter kernel applies before it evaluates a program: every op's words
 * exist, postings are consumed in order within n_postings, only stack entries are popped, NOT takes one
 * operand, and the program ends with every posting consumed and one entry on the stack (no sum of
 * program words may wrap). FRK_OK and *max_depth (optional) = the deepest stack, or FRK_ERR_INVALID. */
int frk_check_program(const uint32_t* expr, uint32_t expr_len, uint32_t n_postings, uint32_t* max_depth);

/* ------------------------------------------------------------------------------------------------
 * Doc store: vectors in the MFMA-native tiled layout. Docs are grouped in blocks of 32; the dims in
 * chunks of 16 bytes per lane (f32: 4 dims, bf16: 8, fp8: 16) for two lane halves, so a chunk is
 * 8 / 16 / 32 dims. Element (doc n, dim d) with E dims per lane-half:
 *   block b = n / 32, j = n % 32, chunk c = d / (2E), h = (d % (2E)) / E, e = d % E
 *   offset (elements) = ((b * chunks + c) * 64 + (j + 32 h)) * E + e
 * One wave's 16-byte load per lane reads one whole (block, chunk) tile: 1 KiB, contiguous.
 * Replaces the ZN layout the reference scorer reads (vector_builder.cpp:34-73).
 * ------------------------------------------------------------------------------------------------ */
#define FRK_DTYPE_F32 0
#define FRK_DTYPE_BF16 1
#define FRK_DTYPE_FP8E4M3 2
/* FP32 values as two FP16 halves: with s the doc's power-of-two scale (the one that puts its largest
 * magnitude in [2^14, 2^15)), x s = hi + lo, hi = fp16(x s), lo = fp16(x s - hi), both rounded to
 * nearest even. Chunk c holds, per lane half h, the 8 halves of dims 16 (c / 2) + 8 h + e (e < 8): the
 * hi parts for even c, the lo parts for odd c, so a chunk pair is one 16-dim slice, 4 bytes per
 * element (the bytes of the FP32 store). Scored with three FP16 MFMAs per slice (hi hi + hi lo +
 * lo hi, FP32 accumulation): products within ~2^-21 of the FP32 ones, at the FP16 MFMA rate. */
#define FRK_DTYPE_F16X2 3

typedef struct {
    const void* tiles;      /* device: n_blocks * chunks * 1 KiB */
    const float* scales;    /* device: per-doc scale (FP8 and F16X2, else NULL) */
    uint32_t dtype;         /* FRK_DTYPE_* */
    uint32_t dim;           /* padded dim: multiple of the chunk dims */
    uint64_t n_docs;        /* valid docs; n_blocks = ceil(n_docs / 32) */
} frk_doc_store;

uint32_t frk_chunk_dims(uint32_t dtype);
/* The padded dim of a store the scorer runs over: dim rounded up, with zero dims, to the next chunk
 * count the scorer is compiled for (4, 8, 12, 16, 24, 32, 48, 64, 96 or 128 chunks). Above 128
 * chunks, the whole-chunk dim (such a store serves frk_rescore; frk_score_* refuse it). */
uint32_t frk_store_dim(uint32_t dim, uint32_t dtype);

/* Measurement aid, not on the search path: reads d_src[0 .. bytes) once with 16-byte loads from one
 * resident wave of workgroups and writes one XOR word per workgroup to d_sink (at most sink_words
 * workgroups), so the HBM read bandwidth this GPU attains can be timed beside a kernel's. */
int frk_read_probe(const void* d_src, uint64_t bytes, uint32_t* d_sink, uint32_t sink_words, frk_stream_t stream);

/* d[i] = value for i < n (a kernel, so it may sit inside a captured graph). */
int frk_fill_u32(uint32_t* d, uint64_t n, uint32_t value, frk_stream_t stream);

/* Host-side staging of a store: row-major FP32 [n_docs][dim] -> tiles (host memory of
 * frk_tiles_bytes() bytes; dims >= dim are zero-filled up to dim_pad). FP8 also writes the per-doc
 * scale max|v| / 448 and stores v / scale; F16X2 writes its per-doc scale and the hi/lo halves of v * scale.
 * No device work; n_threads 0 = all cores. */
uint64_t frk_tiles_bytes(uint64_t n_docs, uint32_t dim_pad, uint32_t dtype);
int frk_tile_docs(const float* src, uint64_t n_docs, uint32_t dim, uint32_t dtype, uint32_t dim_pad, void* dst_tiles,
                  float* dst_scales, uint32_t n_threads);

/* Device-side staging: rows [first_doc, first_doc + n_rows) of a store from row-major FP32 device
 * rows; first_doc must be a multiple of 32, and docs of the last touched block past n_rows are
 * zero-filled. FP8 and F16X2 write d_scales[first_doc + r]. Stream-ordered. */
int frk_tile_rows_device(const float* d_rows, uint64_t n_rows, uint32_t dim, uint64_t first_doc, uint32_t dtype,
                         uint32_t dim_pad, void* d_tiles, float* d_scales, frk_stream_t stream);

/* Filter output (query-major u64 words, as frk_filter_eval writes it) -> the per-doc query words the
 * scorer reads: out[b * q_stride + 32 g + j] bit x = query 32 g + x passes doc 32 b + j (each 32 x 32
 * bit tile of queries x docs transposed; bits of queries >= n_queries are 0). q_stride: a multiple of
 * 32, >= n_queries; d_block_masks 16-byte aligned. */
int frk_masks_to_blocks(const uint64_t* d_masks, uint64_t mask_stride, uint32_t n_queries, uint64_t n_docs,
                        uint32_t* d_block_masks, uint32_t q_stride, frk_stream_t stream);

/* ------------------------------------------------------------------------------------------------
 * Fused scorer: S[q][n] = <query q, doc n> on MFMA (FP32 accumulation), with the filter mask and a
 * per-query score floor fused in, and survivors compacted by wave ballot + prefix count into
 * per-query candidate lists. Replaces kernel_vector_mmad (scorer/kernel_vector_mmad_op.h:21-199)
 * and Aggregator (aggregator/kernel_aggregator.h:22-182): no score matrix is written.
 *
 * Doc blocks visited: block_begin + i * block_step for i < block_count. A doc n passes if its query
 * bit is set (d_block_masks[(n / 32) * mask_q_stride + 32 (q / 32) + n % 32], bit q % 32; NULL = every
 * doc; see frk_masks_to_blocks) and S[q][n] >= floor[q] (d_floor NULL = no floor). A
 * passing doc appends MakeKey(S, n) (score-ordered high word, reversed doc id low word) to
 * cand[q * cand_cap + slot]; the query's counter d_counts[q * counts_stride] counts every pass, so a
 * count > cand_cap means the list overflowed and was truncated. counts_stride >= 1 (u32 elements):
 * 32 gives each counter its own 128-byte line, so the counter atomics of different queries go to
 * different L2 channels instead of queueing on a few lines (at 8,000 survivors per query over 10M
 * FP8 docs and 128 queries: 1.01 -> 0.71 ms). d_queries: [n_queries][dim] FP32 row-major, 16-byte aligned.
 * ------------------------------------------------------------------------------------------------ */
int frk_score_select(const frk_doc_store* store, const float* d_queries, uint32_t n_queries,
                     const uint32_t* d_block_masks, uint32_t mask_q_stride, const float* d_floor,
                     uint64_t block_begin, uint64_t block_step, uint64_t block_count, uint64_t* d_cand,
                     uint32_t cand_cap, uint32_t* d_counts, uint32_t counts_stride, frk_stream_t stream);

/* The unfused score matrix, for callers that keep the reference's order of scoring every doc first and
 * aggregating per query afterwards (FullRecallSearcher::BatchScore / OneAggregator): d_scores[q *
 * score_stride + n] = <query q, doc n> for every n < n_docs (no filter; entries past n_docs untouched).
 * Replaces kernel_vector_mmad in its unfused use (vector_scorer_mmad.cpp:127-134,194-198), which
 * writes the [queries][extendDocNum] FP32 matrix; the same MFMA kernel as frk_score_select. */
int frk_score_matrix(const frk_doc_store* store, const float* d_queries, uint32_t n_queries, float* d_scores,
                     uint64_t score_stride, frk_stream_t stream);

/* The unfused aggregator: for one query, every doc n < n_docs whose filter bit is set (d_bits: the
 * words frk_filter_eval writes, bit n % 64 of word n / 64 for a table whose docs per segment are a
 * multiple of 64) appends MakeKey(d_scores[n], n) to d_cand in no particular order; *d_count (zeroed by
 * the caller) counts every passing doc, so a count above cap means the list was truncated. Replaces the
 * Aggregator kernel (aggregator/kernel_aggregator.h:22-182, launched at result_aggregator.cpp:1090-1092);
 * follow it with frk_select_topk. */
int frk_aggregate(const float* d_scores, const uint64_t* d_bits, uint64_t n_docs, uint64_t* d_cand, uint32_t cap,
                  uint32_t* d_count, frk_stream_t stream);

/* Probe pass of the same kernel: per (query q, sampled block i) the largest score among the block's
 * filter-passing docs, or -inf, written to d_blockmax[q * stride + i] (no candidate lists, no
 * atomics). The k-th largest of a query's block maxima is a valid score floor: the k blocks that
 * reach it each hold a distinct passing doc at least that good. */
int frk_score_blockmax(const frk_doc_store* store, const float* d_queries, uint32_t n_queries,
                       const uint32_t* d_block_masks, uint32_t mask_q_stride, uint64_t block_begin, uint64_t block_step,
                       uint64_t block_count, float* d_blockmax, uint32_t stride, frk_stream_t stream);

/* ------------------------------------------------------------------------------------------------
 * Exact top-k selection over candidate keys: radix select of the k-th largest key, then the k keys
 * sorted descending. Replaces TOPK (topk/kernel_topk.h:30-388) plus the host nth_element/sort
 * (result_aggregator.cpp:993-1000). Per query q: input cand[q * cand_cap ..], min(counts[q *
 * counts_stride], cand_cap) keys; output out[q * k_stride ..] the best min(k_q, n) keys sorted, out_counts[q] their
 * number, and kth_score[q] (optional) the score of the k-th key or -inf when fewer than k exist.
 * k_q = d_k ? d_k[q] : k_default; every k_q must be <= k_stride and <= FRK_SELECT_MAX_K.
 * ------------------------------------------------------------------------------------------------ */
#define FRK_SELECT_MAX_K 4096

/* k-th largest of the n floats vals[q * stride ..] per query (-inf when k > n or k == 0); used on
 * block maxima to turn the probe into a score floor. */
int frk_kth_largest(const float* d_vals, uint32_t stride, uint32_t n, uint32_t n_queries, const uint32_t* d_k,
                    uint32_t k_default, float* d_out, frk_stream_t stream);

int frk_select_topk(const uint64_t* d_cand, uint32_t cand_cap, const uint32_t* d_counts, uint32_t counts_stride,
                    uint32_t n_queries,
                    const uint32_t* d_k, uint32_t k_default, uint64_t* d_out, uint32_t k_stride,
                    uint32_t* d_out_counts, float* d_kth_score, frk_stream_t stream);
/* The same selection without the final sort: out[q * k_stride ..] holds the best min(k_q, n) keys in no
 * particular order (out_counts and kth_score as above). For callers that process the set further, such
 * as FP32 rescoring of the top k * m before the final sorted select; no reference counterpart (the
 * reference sorts on the host, result_aggregator.cpp:993-1000). */
int frk_select_top_set(const uint64_t* d_cand, uint32_t cand_cap, const uint32_t* d_counts, uint32_t counts_stride,
                       uint32_t n_queries, const uint32_t* d_k, uint32_t k_default, uint64_t* d_out, uint32_t k_stride,
                       uint32_t* d_out_counts, float* d_kth_score, frk_stream_t stream);

/* ------------------------------------------------------------------------------------------------
 * FP32 rescoring of candidates after a low-precision pass: keys_out[q][i] = MakeKey(<q, doc>, doc)
 * for doc = the doc of keys_in[q][i], i < min(counts[q], stride), computed from the FP32 store; a key
 * whose doc lies outside the store becomes 0 (the smallest key, which no selection prefers).
 * d_queries: [n_queries][f32_store->dim] FP32. No reference counterpart (the reference scores FP16
 * only, vector_scorer_mmad.cpp:75-211).
 * ------------------------------------------------------------------------------------------------ */
int frk_rescore(const frk_doc_store* f32_store, const float* d_queries, uint32_t n_queries, const uint64_t* d_keys_in,
                uint32_t stride, const uint32_t* d_counts, uint64_t* d_keys_out, frk_stream_t stream);
/* The same from row-major FP32 rows d_rows[doc * dim ..] (dim a multiple of 4, rows 16-byte aligned):
 * each candidate's row is one contiguous read instead of one 16-byte piece per tile. d_queries:
 * [n_queries][dim]. */
int frk_rescore_rows(const float* d_rows, uint32_t dim, uint64_t n_docs, const float* d_queries, uint32_t n_queries,
                     const uint64_t* d_keys_in, uint32_t stride, const uint32_t* d_counts, uint64_t* d_keys_out,
                     frk_stream_t stream);

/* ------------------------------------------------------------------------------------------------
 * Learned score floor: BF16 training and inference kernels of a small MLP
 * (in_dim -> hidden -> hidden -> 1, ReLU) that predicts a lower bound on a query's k-th best filtered
 * score from the query vector and filter features; the searcher uses it as a per-query floor
 * (frk_search_options.floors) and reruns unpruned any query it would cut below k survivors. No
 * reference counterpart (the reference has no learned component). Every GEMM rounds its operands to
 * BF16 (nearest even) and accumulates in FP32 on MFMA; tensors are FP32 row-major in device memory.
 * ------------------------------------------------------------------------------------------------ */
#define FRK_MLP_EPI_NONE 0      /* C = op(A) op(B) */
#define FRK_MLP_EPI_BIAS_RELU 1 /* C = op(A) op(B) + bias (pre-activation), H = max(C, 0) */
#define FRK_MLP_EPI_RELU_GRAD 2 /* C = Z > 0 ? op(A) op(B) : 0 */

/* C[M][N] = op(A)[M][K] op(B)[K][N]; op(A)(m,k) = trans_a ? A[k*lda + m] : A[m*lda + k],
 * op(B)(k,n) = trans_b ? B[n*ldb + k] : B[k*ldb + n]. */
int frk_mlp_gemm(int32_t M, int32_t N, int32_t K, const float* d_a, int32_t lda, int32_t trans_a, const float* d_b,
                 int32_t ldb, int32_t trans_b, float* d_c, int32_t ldc, int32_t epilogue, const float* d_bias,
                 const float* d_z, int32_t ldz, float* d_h, int32_t ldh, frk_stream_t stream);
/* y[r] = sum_j bf16(H[r][j]) bf16(w[j]) + bias[0] (bias may be NULL). */
int frk_mlp_head_forward(int32_t rows, int32_t dim, const float* d_h, int32_t ldh, const float* d_w,
                         const float* d_bias, float* d_y, frk_stream_t stream);
/* dZ[r][j] = Z[r][j] > 0 ? dy[r] bf16(w[j]) : 0. */
int frk_mlp_head_backward(int32_t rows, int32_t dim, const float* d_dy, const float* d_w, const float* d_z,
                          int32_t ldz, float* d_dz, int32_t lddz, frk_stream_t stream);
/* out[j] = sum_r (weights ? weights[r] : 1) x[r][j], x rounded to BF16 when round_bf16. */
int frk_mlp_column_sum(int32_t rows, int32_t cols, const float* d_x, int32_t ldx, const float* d_weights,
                       int32_t round_bf16, float* d_out, frk_stream_t stream);
/* Pinball loss at level tau of r = label - y, mean over rows: loss[r] (optional) and dy[r]. */
int frk_mlp_pinball(int32_t rows, const float* d_y, const float* d_label, float tau, float* d_dy, float* d_loss,
                    frk_stream_t stream);
/* One bias-corrected Adam step (step >= 1) over n FP32 parameters. */
int frk_mlp_adam(int64_t n, float* d_param, const float* d_grad, float* d_m, float* d_v, float lr, float beta1,
                 float beta2, float eps, int32_t step, frk_stream_t stream);

typedef struct {
    const float* w1; /* [hidden][in_dim] */
    const float* b1; /* [hidden] */
    const float* w2; /* [hidden][hidden] */
    const float* b2; /* [hidden] */
    const float* w3; /* [hidden] */
    const float* b3; /* [1] */
    int32_t in_dim, hidden;
} frk_mlp_params;

/* Inference: y[r] for the rows of x [rows][ldx]; d_work holds 4 * rows * hidden floats. */
int frk_mlp_forward(const frk_mlp_params* p, int32_t rows, const float* d_x, int32_t ldx, float* d_work, float* d_y,
                    frk_stream_t stream);

/* frk_mlp_forward without a caller workspace, for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid allocation); no counterpart in the
 * reference. The four activation tiles live in a fixed device array of 4 x FRK_MLP_NOALLOC_MAX_ROWS x
 * FRK_MLP_NOALLOC_MAX_HIDDEN floats (4 MiB, held by each GPU the library runs on until the process ends);
 * larger calls fail with FRK_ERR_CAPACITY. Every call on one GPU shares that array, so calls that overlap
 * on one GPU (different streams or threads) race; calls ordered on one stream do not. Allocates nothing. */
#define FRK_MLP_NOALLOC_MAX_ROWS 1024
#define FRK_MLP_NOALLOC_MAX_HIDDEN 256
int frk_mlp_forward_noalloc(const frk_mlp_params* p, int32_t rows, const float* d_x, int32_t ldx, float* d_y,
                            frk_stream_t stream);

/* floors[r] = y[r] + offset[r] (an offset of -inf disables the floor of that row). */
int frk_mlp_floors(int32_t rows, const float* d_y, const float* d_offset, float* d_floors, frk_stream_t stream);

/* A trained floor model as the searcher uses it: input row = [query | (features - mean) / std],
 * floor = prediction - margin (per decade of k when margin_by_k is given), no floor where fewer than
 * min_passing_ratio * k docs are expected to pass (k = the keys the main pass selects,
 * k * rescore_factor on low precision). */
#define FRK_FLOOR_FEATURES 32
typedef struct {
    frk_mlp_params mlp;       /* device weights; mlp.in_dim = query dim + FRK_FLOOR_FEATURES */
    const float* feat_mean;   /* host [FRK_FLOOR_FEATURES] */
    const float* feat_std;    /* host [FRK_FLOOR_FEATURES] */
    float margin;             /* for every k unless margin_by_k is set */
    float min_passing_ratio;
    const float* margin_by_k; /* optional host [4]: margins for k < 10, < 100, < 1000, >= 1000 */
} frk_floor_model;

/* ================================================================================================
 * Layer 2: index and searcher handles (C face of NpuRetrieval::DataTable / FullRecallSearcher).
 * ================================================================================================ */
typedef struct frk_table frk_table;
typedef struct frk_builder frk_builder;

/* DataTable::LoadData (engine/index/data_table.h:20): the index directory written by the builder
 * (.meta, id/ *.idm files, per-field .vector and .posting files). dtype_mask: extra stores to build,
 * bits (1 << FRK_DTYPE_BF16) | (1 << FRK_DTYPE_FP8E4M3) | (1 << FRK_DTYPE_F16X2); the FP32 store is
 * always built. */
int frk_table_load(int32_t device, const char* dir, uint32_t dtype_mask, frk_table** out);
void frk_table_free(frk_table* t);
int frk_table_info(const frk_table* t, uint64_t* n_docs, uint32_t* n_segments, uint32_t* docs_per_segment);

/* In-memory build (the offline builder's output without files). Posting lists per tag are CSR:
 * tags[t]'s docs are doc_ids[offsets[t] .. offsets[t+1]), strictly ascending; the token is
 * Hash64(decimal tag) (harness/query/filter_setup.cpp:100). Vectors: row-major FP32 rows
 * [first, first + n), host or device memory, first a multiple of 32. */
int frk_builder_create(int32_t device, uint32_t n_docs, uint32_t docs_per_segment, frk_builder** out);
void frk_builder_free(frk_builder* b);
int frk_builder_add_vector_field(frk_builder* b, const char* name, uint32_t dim, uint32_t dtype_mask);
int frk_builder_set_vectors(frk_builder* b, const char* name, const float* rows, uint64_t first, uint64_t n,
                            int rows_on_device);
int frk_builder_add_posting_field(frk_builder* b, const char* name, int is_matrix, float density_threshold,
                                  int packed);
int frk_builder_add_tag_postings(frk_builder* b, const char* field, uint64_t n_tags, const uint64_t* tags,
                                 const uint64_t* offsets, const uint32_t* doc_ids);
int frk_builder_set_gdocids(frk_builder* b, const uint64_t* ids, uint64_t n);
int frk_builder_finish(frk_builder* b, frk_table** out); /* the builder stays owned by the caller */

/* Offline index writer (the builder's output, indexer/builder/fr_builder.cpp and
 * builder/segment/{vector,inverted,idmapping}_builder.cpp): writes <dir>/poissonengine_<version>.meta,
 * id/ *.idm, and per field one .vector or .posting file per segment -- the directory
 * frk_table_load reads. Host only. Vectors: rows [n_docs][dim] FP32; file_type
 * FRK_VECTOR_FILE_FP16_ZN is the reference's FP16 ZN body (split_doc_num_zn docs per block),
 * FRK_VECTOR_FILE_FP32_ROWS keeps FP32 rows, FRK_VECTOR_FILE_BF16_ROWS BF16 rows (nearest even), and
 * FRK_VECTOR_FILE_FP8_ROWS a per-vector FP32 scale (the smallest power of two s with max|v| / s <= 448)
 * then E4M3 rows of v / s (nearest even, saturating); the loader widens BF16 / FP8 bodies to their
 * stored values. Postings: CSR over tokens, any id order, duplicates dropped; token t's text is
 * terms[t] when terms is non-null, else the decimal tags[t]. */
#define FRK_VECTOR_FILE_FP16_ZN 0u
#define FRK_VECTOR_FILE_FP32_ROWS 1u
#define FRK_VECTOR_FILE_BF16_ROWS 2u
#define FRK_VECTOR_FILE_FP8_ROWS 3u
typedef struct frk_index_vector_field {
    const char* name;
    uint32_t dim;
    uint32_t file_type;
    const float* rows;
} frk_index_vector_field;
typedef struct frk_index_posting_field {
    const char* name;
    int32_t is_matrix;
    float density_threshold;
    int32_t packed;
    uint64_t n_tokens;
    const uint64_t* tags;
    const char* const* terms;
    const uint64_t* offsets; /* n_tokens + 1 */
    const uint32_t* doc_ids;
} frk_index_posting_field;
int frk_write_index(const char* dir, uint32_t version, uint32_t n_docs, uint32_t docs_per_segment,
                    uint32_t split_doc_num_zn, uint32_t n_vector_fields, const frk_index_vector_field* vector_fields,
                    uint32_t n_posting_fields, const frk_index_posting_field* posting_fields,
                    const uint64_t* gdocids /* n_docs ids, or null for 0..n_docs-1 */);

/* FP32 -> FP16 as the builder stores vectors (core/number/trans_number.cpp:38-82): round to nearest
 * even, magnitudes below 2^-14 to signed zero, finite overflow to infinity, infinity and NaN to NaN. */
uint16_t frk_f32_to_f16(float f);

/* Host codecs of the index format: the token hash (core/hash/hash.cpp:10-16) and the posting
 * serializer (format/posting_serializer.cpp:81-211; SPARSE_IDLIST keeps the given id order).
 * frk_serialize_posting returns the byte count or a negative error. */
uint64_t frk_hash64(const char* data, uint32_t len);
int64_t frk_serialize_posting(const uint32_t* ids, uint32_t n_ids, uint32_t doc_num, int is_matrix, float threshold,
                              int packed, uint8_t* out, uint64_t cap);

/* The search driver's filter grammar (harness/query/filter_expr.h:164-183) as pre-order
 * (type, arg) pairs: 0 TERM(tag), 1 AND(n children), 2 OR(n), 3 NOT(1). Returns the pair count,
 * 0 for an empty (match-all) filter, or a negative error. */
int frk_parse_filter(const char* expr, int64_t* pairs, int32_t cap_pairs);

/* The postfix program a filter compiles to against a table (TextFilter::PrepareExpr,
 * text_filter.cpp:432-467). Returns the program length or a negative error. */
int frk_compile_filter(const frk_table* t, const char* field, const char* expr, uint32_t* out_expr, int32_t cap,
                       uint32_t* n_postings, uint32_t* op_num, int32_t* match_all);

typedef struct {
    uint32_t dtype;           /* scoring store: FRK_DTYPE_F32 or _F16X2 (plain: FP32-accurate scores, the
                                 k selected directly), _BF16 or _FP8E4M3 (k * m candidates rescored in FP32) */
    uint32_t rescore_factor;  /* low precision: FP32 rescoring of the top min(k * m, FRK_SELECT_MAX_K) */
    int32_t probe;            /* sampled k-th-best floor (exact pruning) */
    uint32_t probe_step;      /* 0 = automatic */
    const float* floors;      /* host, per query, or NULL: caller score floors */
    uint32_t cand_cap;        /* 0 = automatic */
    const frk_floor_model* floor_model; /* or NULL: learned floors computed on the device per batch */
    uint32_t graph;           /* nonzero: replay the device sequence from a captured hipGraph per
                                 batch shape (captured on the shape's second batch) */
} frk_search_options;

typedef struct {
    float ms_filter, ms_probe, ms_score, ms_select, ms_rescore, ms_total; /* stage times 0 on a replay */
    uint64_t candidates;
    uint32_t reruns, overflowed, probe_step, cand_cap, graph_replayed;
    /* host wall times: the searcher's preparation up to the enqueue, its result extraction after the
     * sync, and the whole call (filter parsing and output copies included) */
    float ms_host_prepare, ms_host_finish, ms_call;
} frk_search_stats;

/* The FRK_FLOOR_FEATURES floor-model features of n (filter, k) pairs against the table's posting
 * field (selectivity under tag independence from the table's per-token doc counts, filter shape,
 * query norm): out[i * FRK_FLOOR_FEATURES + f]. Host only. */
int frk_floor_features(const frk_table* t, const char* posting_field, uint32_t n, const char* const* filters,
                       const uint32_t* ks, const float* queries, uint32_t dim, float* out);

/* FullRecallSearcher::AddQuery x n + BatchSearch (full_recall_searcher.h:30-33) with filter strings
 * (NULL entry or NULL array = match all). Results per query: out_ids[q * k_stride + i] (global doc
 * ids) and out_scores, best first, out_counts[q] = min(k, passing docs); out_passing[q] = docs that
 * passed the filter (optional). */
int frk_search(const frk_table* t, const char* vector_field, const char* posting_field, uint32_t n_queries,
               const float* queries, uint32_t dim, const char* const* filters, const uint32_t* ks,
               const frk_search_options* opts, uint32_t k_stride, int64_t* out_ids, float* out_scores,
               uint32_t* out_counts, uint32_t* out_passing, frk_search_stats* stats);

/* Prepared filters: a filter string parsed and built into its QueryNode tree once (the harness builds
 * its trees when it loads the query file, harness/query/filter_setup.cpp:129-136), for any number of
 * later searches from any thread. expr NULL or empty = match all. */
typedef struct frk_filter frk_filter;
int frk_filter_prepare(const char* posting_field, const char* expr, frk_filter** out);
void frk_filter_free(frk_filter* f);
/* frk_search with prepared filters (one handle per query, shared handles allowed). */
int frk_search_prepared(const frk_table* t, const char* vector_field, uint32_t n_queries, const float* queries,
                        uint32_t dim, const frk_filter* const* filters, const uint32_t* ks,
                        const frk_search_options* opts, uint32_t k_stride, int64_t* out_ids, float* out_scores,
                        uint32_t* out_counts, uint32_t* out_passing, frk_search_stats* stats);

/* ------------------------------------------------------------------------------------------------
 * Serving: a continuous-batching request server over the searcher (extends the search driver's
 * serial --target_qps mode, harness/run/schedule_serial.cpp:23-51). Requests of any number of queries
 * (each with its own filter and k) are submitted at any time from any thread; a scheduler packs the
 * oldest pending requests into device batches of at most max_batch_queries queries, starting one as
 * soon as one of `workers` slots is free and its working set fits the HBM budget left by the batches
 * in flight, so batches overlap on the device. The index stays resident; with opts->graph each
 * batch shape's device sequence is replayed from a captured hipGraph. opts->floor_model (if any)
 * must stay valid while the server lives. Times are seconds on the server's clock.
 * ------------------------------------------------------------------------------------------------ */
typedef struct frk_server frk_server;
typedef struct {
    uint32_t max_batch_queries; /* 0 = 256 */
    uint32_t workers;           /* batches in flight; 0 = 2 */
    uint64_t hbm_budget_bytes;  /* 0 = unlimited */
    uint32_t pad_queries;       /* pad batches with inert queries to a multiple of this (graph reuse);
                                   0 = 64, 1 = no padding */
    uint32_t split_requests;    /* 1 = a request's queries may span consecutive batches, so every batch
                                   is filled to max_batch_queries under load; 0 = whole requests */
} frk_server_config;

typedef struct {
    uint64_t requests, queries, batches, graph_replays, reruns, overflows; /* reruns include overflows */
    double device_ms;            /* sum over batches of the device time of each */
    uint64_t doc_bytes_streamed; /* doc-store bytes read by the main score passes */
} frk_server_stats;

int frk_server_create(const frk_table* t, const char* vector_field, const char* posting_field,
                      const frk_search_options* opts, const frk_server_config* cfg, frk_server** out);
void frk_server_free(frk_server* s); /* finishes every submitted request first */
int frk_server_submit(frk_server* s, uint32_t n_queries, const float* queries, uint32_t dim,
                      const char* const* filters, const uint32_t* ks, uint64_t* ticket);
/* Blocks until the request is done. times[3] = submitted, started, finished (optional);
 * batch_queries = queries in the device batch that served it (optional). */
int frk_server_wait(frk_server* s, uint64_t ticket, uint32_t k_stride, int64_t* out_ids, float* out_scores,
                    uint32_t* out_counts, uint32_t* out_passing, double* times, uint32_t* batch_queries);
int frk_server_stats_get(const frk_server* s, frk_server_stats* out);
/* The device working set the scheduler charges against the HBM budget for a batch of n queries. */
uint64_t frk_server_batch_bytes(const frk_server* s, uint32_t n_queries);
double frk_server_now(const frk_server* s);

/* The searchers' workspace pool (grow-only per-batch device buffers, pooled between batches; the
 * reference's named GM pools, searcher/runtime/gm_memory_manager.cpp:9-75, and their [POOLSTAT]
 * report, gm_memory_pool.h:172-180): device bytes held now and at the peak, allocations, workspaces
 * alive and idle, and the retention limit (a server with an HBM budget sets it to the budget: a
 * workspace released while the pool holds more is freed instead of kept). */
typedef struct {
    uint64_t live_bytes, peak_bytes, mallocs, frees, retain_limit;
    uint32_t workspaces, pooled;
} frk_pool_stats;
int frk_workspace_pool_stats(frk_pool_stats* out);

#ifdef __cplusplus
}
#endif
#endif
