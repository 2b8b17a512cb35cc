/*
 * frk_capi.h -- the C ABI of the MI355X (gfx950) filtered top-K pipeline.
 *
 * Two layers, both plain C (no C++ or torch types):
 *
 *  1. Kernel entry points (frk_filter_eval, frk_score_select, frk_select_topk, frk_rescore,
 *     frk_mlp_*). They replace the reference's five AscendC launch stubs, which the host stages call
 *     through ACLRT_LAUNCH_KERNEL (each entry cites the kernel it replaces). Stream-ordered, no
 *     implicit synchronisation, no allocation; the caller owns every device buffer. Reentrant
 *     across streams.
 *
 *  2. Index and searcher handles (frk_table_*, frk_searcher_*), a C face of the C++
 *     NpuRetrieval::DataTable / FullRecallSearcher classes (hybridsearchsystem_amd/csrc/engine/),
 *     for callers that cannot link C++ (ctypes, cgo, JNI).
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
                    frk_stream_t stream);

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

typedef struct {
    const void* tiles;      /* device: n_blocks * chunks * 1 KiB */
    const float* scales;    /* device: per-doc scale (FP8 only, else NULL) */
    uint32_t dtype;         /* FRK_DTYPE_* */
    uint32_t dim;           /* padded dim: multiple of the chunk dims */
    uint64_t n_docs;        /* valid docs; n_blocks = ceil(n_docs / 32) */
} frk_doc_store;

uint32_t frk_chunk_dims(uint32_t dtype);

/* Host-side staging of a store: row-major FP32 [n_docs][dim] -> tiles (host memory of
 * frk_tiles_bytes() bytes; dims >= dim are zero-filled up to dim_pad). FP8 also writes the per-doc
 * scale max|v| / 448 and stores v / scale. No device work; n_threads 0 = all cores. */
uint64_t frk_tiles_bytes(uint64_t n_docs, uint32_t dim_pad, uint32_t dtype);
int frk_tile_docs(const float* src, uint64_t n_docs, uint32_t dim, uint32_t dtype, uint32_t dim_pad, void* dst_tiles,
                  float* dst_scales, uint32_t n_threads);

/* ------------------------------------------------------------------------------------------------
 * Fused scorer: S[q][n] = <query q, doc n> on MFMA (FP32 accumulation), with the filter mask and a
 * per-query score floor fused in, and survivors compacted by wave ballot + prefix count into
 * per-query candidate lists. Replaces kernel_vector_mmad (scorer/kernel_vector_mmad_op.h:21-199)
 * and Aggregator (aggregator/kernel_aggregator.h:22-182): no score matrix is written.
 *
 * Doc blocks visited: block_begin + i * block_step for i < block_count (the probe pass samples
 * with a step > 1). A doc n passes if mask bit n of query q is set (d_masks row q at stride
 * mask_stride words, bit n%64 of word n/64; NULL = all docs) and S[q][n] >= floor[q] (d_floor NULL
 * = no floor). A passing doc appends MakeKey(S, n) (score-ordered high word, reversed doc id low
 * word) to cand[q * cand_cap + slot]; counts[q] counts every pass, so counts[q] > cand_cap means
 * the list overflowed and was truncated. d_queries: [n_queries][dim] FP32 row-major.
 * ------------------------------------------------------------------------------------------------ */
int frk_score_select(const frk_doc_store* store, const float* d_queries, uint32_t n_queries,
                     const uint64_t* d_masks, uint64_t mask_stride, const float* d_floor,
                     uint64_t block_begin, uint64_t block_step, uint64_t block_count, uint64_t* d_cand,
                     uint32_t cand_cap, uint32_t* d_counts, frk_stream_t stream);

/* ------------------------------------------------------------------------------------------------
 * Exact top-k selection over candidate keys: radix select of the k-th largest key, then the k keys
 * sorted descending. Replaces TOPK (topk/kernel_topk.h:30-388) plus the host nth_element/sort
 * (result_aggregator.cpp:993-1000). Per query q: input cand[q * cand_cap ..], min(counts[q],
 * cand_cap) keys; output out[q * k_stride ..] the best min(k_q, n) keys sorted, out_counts[q] their
 * number, and kth_score[q] (optional) the score of the k-th key or -inf when fewer than k exist.
 * k_q = d_k ? d_k[q] : k_default; every k_q must be <= k_stride and <= FRK_SELECT_MAX_K.
 * ------------------------------------------------------------------------------------------------ */
#define FRK_SELECT_MAX_K 4096

int frk_select_topk(const uint64_t* d_cand, uint32_t cand_cap, const uint32_t* d_counts, uint32_t n_queries,
                    const uint32_t* d_k, uint32_t k_default, uint64_t* d_out, uint32_t k_stride,
                    uint32_t* d_out_counts, float* d_kth_score, frk_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
