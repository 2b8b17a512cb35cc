// frk_score_impl.h -- fused inner-product scorer on MFMA with the filter mask, a per-query score floor and
// survivor compaction in the epilogue. No score matrix is written.
//
// The reference computes S = Q x D^T on the Cube unit into a full [Q][docs] FP32 matrix
// (engine/retrieval/device/scorer/kernel_vector_mmad_op.h:21-199) and a second kernel re-reads it with
// the filter bitmap to compact survivors (device/aggregator/kernel_aggregator.h:22-182). Here one pass
// does both: a wave computes a 32-doc block against the whole query tile held in LDS, its MFMA chain
// starting from 0 for (query, doc) pairs that pass the filter and from NaN for the others, then each
// accumulator row is tested against its floor, ballot-ed, and the passing (score, doc) pairs are
// appended to per-query candidate lists at offsets from one atomic per (wave, query) and the prefix
// count of the ballot.
//
// Layout (frk_capi.h, "Doc store"): a 32-doc x 16-byte-per-lane chunk tile is exactly one wave load of
// 1 KiB and exactly the B operand of one MFMA step group:
//   f32:  4 x v_mfma_f32_32x32x2_f32   per chunk (lane halves carry dims 8c+e and 8c+4+e)
//   bf16: 1 x v_mfma_f32_32x32x16_bf16 per chunk
//   fp8:  1 x v_mfma_scale_f32_32x32x64_f8f6f4 per two chunks (scale exponents 127 = x1.0)
// The query tile uses the same tiling (rows = queries) and lives in LDS for the whole launch.
// Accumulator element r of lane l holds C[i][j] with j = l % 32 (the doc) and
// i = 8 (r / 4) + 4 (l / 32) + r % 4 (the query row).
#include <math.h>

#pragma once
#include <utility>

#include "frk_common.h"

namespace frk {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kXcds = 8;  // accelerator complex dies of an MI355X, each with its own L2
#ifndef FRK_SCORE_STAGE_A
#define FRK_SCORE_STAGE_A 1
#endif
#ifndef FRK_SCORE_WIDE_TILE
#define FRK_SCORE_WIDE_TILE 64  // query tiles above this many KiB get 8-wave workgroups
#endif
#ifndef FRK_SCORE_WIDE_TILE_FP8
#define FRK_SCORE_WIDE_TILE_FP8 16  // the same for FP8 and BF16 stores
#endif

// Workgroup size of a shape: 4 waves, or 8 where the query tile (RB x CH KiB of LDS) is so large that
// a CU holds only one workgroup; 8 waves then still give each SIMD two waves, one streaming its doc
// tiles while the other runs its MFMAs. All waves of a workgroup share its one query tile. Measured on
// 10M docs: FP8 768 dims 1.56 -> 1.51 ms at 128 queries and 2.30 -> 2.05 ms at 256, FP32 256 dims
// 5.12 -> 4.99 ms at 128; BF16 tiles above 64 KiB are kept at 4 waves, where 8 measured slower (768
// dims, 128 queries: 2.83 -> 3.41 ms; 1024 dims: 1.82 -> 2.31 ms on 5M docs).
template <int DT, int CH, int RB>
struct WgShape {
    // (FP8 at a 128 KiB tile would spill under the 256-register cap of two waves per SIMD). FP8 and
    // BF16 shapes also go wide from a 16 KiB tile on (BF16 up to 64 KiB): their register counts cap
    // them at two waves per SIMD anyway, and one 8-wave workgroup per CU with one tile measured faster
    // than two 4-wave ones with a tile each (10M x 256, 128 queries: FP8 in the pipeline 0.626 -> 0.606
    // ms, BF16 scorer alone 0.94 -> 0.91 ms).
    static constexpr int kThreads =
        (DT != FRK_DTYPE_BF16 || RB * CH <= 64) &&
                RB * CH > (DT == FRK_DTYPE_F32 || DT == FRK_DTYPE_F16X2 ? FRK_SCORE_WIDE_TILE : FRK_SCORE_WIDE_TILE_FP8) &&
                RB * CH <= (DT == FRK_DTYPE_FP8E4M3 ? 96 : 128)
            ? 512
            : 256;
    static constexpr int kWaves = kThreads / 64;
    // Query fragments staged in registers one group ahead: a 4-wave workgroup whose query tile leaves
    // room for only one workgroup per CU runs a single wave per SIMD, which has no other wave to cover
    // the latency of an LDS read issued just before its MFMA (and has the whole SIMD's registers).
    static constexpr bool kStageA = kThreads == 256 && RB * CH > FRK_SCORE_WIDE_TILE && FRK_SCORE_STAGE_A;
};

template <int DT>
struct Traits;
template <>
struct Traits<FRK_DTYPE_F32> {
    static constexpr int kElemsPerLane = 4;  // per 16-byte chunk and lane half
};
template <>
struct Traits<FRK_DTYPE_BF16> {
    static constexpr int kElemsPerLane = 8;
};
template <>
struct Traits<FRK_DTYPE_FP8E4M3> {
    static constexpr int kElemsPerLane = 16;
};
template <>
struct Traits<FRK_DTYPE_F16X2> {
    static constexpr int kElemsPerLane = 8;  // halves per lane: the hi (even chunk) or lo (odd chunk) parts
};

// Stores whose docs carry a per-doc power-of-two scale that the accumulators still hold (F16X2: the
// MFMAs multiply scaled halves, so a score is acc / (query scale * doc scale)).
template <int DT>
constexpr bool kUnscale = DT == FRK_DTYPE_F16X2;

// One chunk (or, for fp8, one chunk pair) of MFMA work for one 32-query row block.
template <int DT>
__device__ __forceinline__ void MfmaChunk(const u32x4& a, const u32x4& b, floatx16& acc);

template <>
__device__ __forceinline__ void MfmaChunk<FRK_DTYPE_F32>(const u32x4& a, const u32x4& b, floatx16& acc) {
    floatx4 av = __builtin_bit_cast(floatx4, a);
    floatx4 bv = __builtin_bit_cast(floatx4, b);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[0], bv[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[1], bv[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[2], bv[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[3], bv[3], acc, 0, 0, 0);
}

template <>
__device__ __forceinline__ void MfmaChunk<FRK_DTYPE_BF16>(const u32x4& a, const u32x4& b, floatx16& acc) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                  0, 0, 0);
}

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// One 16-dim F16X2 slice for one row block: (query hi, query lo) x (doc hi, doc lo) as three FP16 MFMAs,
// hi hi + hi lo + lo hi (lo lo is below 2^-21 of the product and dropped).
__device__ __forceinline__ f16x8 AsF16(const u32x4& v) { return __builtin_bit_cast(f16x8, v); }

// sa, sb: E8M0 scale exponents (2^(e - 127)) of this lane's query row and doc column
__device__ __forceinline__ void MfmaChunkPairFp8(const u32x4& a0, const u32x4& a1, const u32x4& b0, const u32x4& b1,
                                        floatx16& acc, int sa, int sb) {
    intx8 av = {static_cast<int>(a0[0]), static_cast<int>(a0[1]), static_cast<int>(a0[2]), static_cast<int>(a0[3]),
                static_cast<int>(a1[0]), static_cast<int>(a1[1]), static_cast<int>(a1[2]), static_cast<int>(a1[3])};
    intx8 bv = {static_cast<int>(b0[0]), static_cast<int>(b0[1]), static_cast<int>(b0[2]), static_cast<int>(b0[3]),
                static_cast<int>(b1[0]), static_cast<int>(b1[1]), static_cast<int>(b1[2]), static_cast<int>(b1[3])};
    // formats: 0 = FP8 E4M3 for both operands; op_sel 0 = byte 0 of each scale register
    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, sa, 0, sb);
}

// What a launch of the scorer produces: per-query candidate keys of the docs that pass filter and floor
// (the fused pipeline), per (query, sampled block) maxima (the probe), or the whole score matrix.
constexpr int kModeSelect = 0;
constexpr int kModeBlockMax = 1;
constexpr int kModeDense = 2;

struct ScoreArgs {
    const u32x4* tiles;
    const float* doc_scale;
    uint64_t n_docs;
    const float* queries;
    uint32_t n_queries;
    uint32_t dim;  // padded dim of the store
    uint32_t q_dim;  // row length of the query matrix (== dim)
    const uint32_t* bmask;    // per-doc filter words: bmask[block * bmask_stride + 32 g + j] bit x = query 32 g + x
                              // passes doc 32 block + j
    uint32_t bmask_stride;    // >= n_queries; 0 = no filter (bmask then points at one all-pass word)
    const float* floor;
    uint64_t block_begin, block_step, block_count;
    uint64_t* cand;
    uint32_t cap;
    uint32_t* counts;
    uint32_t counts_stride;  // u32 elements between consecutive queries' counters
    uint32_t n_groups;  // query groups of 32 RB rows (set by the launch)
};

// Writes the query tile into LDS in the doc tiling, converted to the store's element type; for fp8 and
// F16X2 the per-query power-of-two scale (Fp8Scale or SplitScale of max|q|) goes to q_scale_lds. Every workgroup of a launch
// does this, so it is cheap by construction: a query's max|q| is one wave's work (a 16-byte load per lane
// and a shuffle reduction), and each tile entry reads its E query values with 16-byte loads (query rows
// are 16-byte aligned and the padded dim is a multiple of 8).
template <int DT, int CH, int RB>
__device__ __forceinline__ void StageQueries(const ScoreArgs& A, u32x4* a_lds, float* q_scale_lds, uint32_t q0) {
    constexpr int E = Traits<DT>::kElemsPerLane;
    constexpr int kThreads = WgShape<DT, CH, RB>::kThreads;
    if constexpr (DT == FRK_DTYPE_FP8E4M3 || DT == FRK_DTYPE_F16X2) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int r = wave; r < 32 * RB; r += kThreads / 64) {
            const uint32_t q = q0 + r;
            float mx = 0.0f;
            if (q < A.n_queries) {
                const floatx4* row = reinterpret_cast<const floatx4*>(A.queries + static_cast<uint64_t>(q) * A.q_dim);
                for (uint32_t d4 = lane; d4 < A.dim / 4; d4 += 64) {
                    const floatx4 v = row[d4];
                    mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
                }
            }
            for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
            if (lane == 0) q_scale_lds[r] = DT == FRK_DTYPE_FP8E4M3 ? Fp8Scale(mx) : SplitScale(mx);
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < RB * CH * 64; idx += kThreads) {
        const int l = idx & 63;
        const int c = (idx >> 6) % CH;
        const int rb = idx / (CH * 64);
        const int i = l & 31, h = l >> 5;
        const uint32_t q = q0 + 32 * rb + i;
        // F16X2: chunk pair c / 2 is the 16-dim slice, the even chunk its hi halves and the odd its lo
        const uint32_t d0 = DT == FRK_DTYPE_F16X2 ? (c / 2) * 16 + h * 8 : c * 2 * E + h * E;
        float v[E];
        if (q < A.n_queries) {
            const floatx4* src = reinterpret_cast<const floatx4*>(A.queries + static_cast<uint64_t>(q) * A.q_dim + d0);
#pragma unroll
            for (int e4 = 0; e4 < E / 4; e4++) {
                const floatx4 x = src[e4];
#pragma unroll
                for (int e = 0; e < 4; e++) v[4 * e4 + e] = x[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = 0.0f;
        }
        u32x4 packed;
        uint32_t* pw = reinterpret_cast<uint32_t*>(&packed);
        if constexpr (DT == FRK_DTYPE_F32) {
            for (int e = 0; e < 4; e++) pw[e] = __builtin_bit_cast(uint32_t, v[e]);
        } else if constexpr (DT == FRK_DTYPE_BF16) {
            for (int e = 0; e < 4; e++)
                pw[e] = static_cast<uint32_t>(F32ToBf16(v[2 * e])) | (static_cast<uint32_t>(F32ToBf16(v[2 * e + 1])) << 16);
        } else if constexpr (DT == FRK_DTYPE_F16X2) {
            // v * scale (a power of two, exact) split into hi = fp16(v s) and lo = fp16(v s - hi), with the
            // hardware's round-to-nearest-even conversion
            const float sc = q_scale_lds[32 * rb + i];
            f16x8 part;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const float x = v[e] * sc;
                const _Float16 hi = static_cast<_Float16>(x);
                part[e] = (c & 1) ? static_cast<_Float16>(x - static_cast<float>(hi)) : hi;
            }
            packed = __builtin_bit_cast(u32x4, part);
        } else {
            // v / scale lies within [-448, 448] (the scale is a power of two chosen for that), where the
            // hardware conversion rounds exactly as F32ToE4M3 does (nearest even, subnormals included;
            // checked bit for bit on the GPU over 3.7e8 sampled inputs): bytes x of word e hold
            // element 4 e + x
            const float inv = 1.0f / q_scale_lds[32 * rb + i];
            for (int e = 0; e < 4; e++) {
                int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * e] * inv, v[4 * e + 1] * inv, 0, false);
                w = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * e + 2] * inv, v[4 * e + 3] * inv, w, true);
                pw[e] = static_cast<uint32_t>(w);
            }
        }
        a_lds[idx] = packed;
    }
}

// Doc streaming: a wave walks its doc blocks as a flat sequence of (block, group) steps, a group being
// G chunks of one block (GroupShape: 4, or 8 for FP8). A ring of kRingN register buffers (4, or 3 for
// FP8) keeps the loads of the next kRingN - 1 steps in flight while one step feeds the MFMAs. The loop
// body is unrolled over U = lcm(kRingN, groups per block) steps and always starts at a block boundary, so the ring slot and
// the group index of every step are compile-time constants: the ring stays in registers and the block
// epilogue appears once per body position that ends a block. Loads are unconditional (a step past
// the end reloads the last step's addresses), so the code between a load and its use is straight-line
// and each wait counts the loads still allowed in flight exactly.
#ifndef FRK_SCORE_RING
#define FRK_SCORE_RING 4
#endif
#ifndef FRK_SCORE_FP8_GROUP
#define FRK_SCORE_FP8_GROUP 8  // FP8 chunks per group where the doc's chunk count allows
#endif
#ifndef FRK_SCORE_FP8_RING
#define FRK_SCORE_FP8_RING 3
#endif
constexpr int kRing = FRK_SCORE_RING;
static_assert(kRing >= 2, "the ring needs a slot in use and one in flight");

// f(integral_constant<int, I>) for each I of the sequence, in order.
template <typename F, int... I>
__device__ __forceinline__ void StaticFor(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}

constexpr int Gcd(int a, int b) { return b == 0 ? a : Gcd(b, a % b); }

// G chunks per group and kRingN ring slots. FP8 docs are the smallest per chunk (1 KiB for 32 docs x 32
// dims), so their groups are 8 chunks (8 KiB, one 256-dim block) in a 3-slot ring where the chunk count
// allows: two groups, 16 KiB, in flight per wave instead of three of 4 KiB (12 KiB), for 32 more
// registers.
template <int DT, int CH>
struct GroupShape {
    static constexpr bool kWide = DT == FRK_DTYPE_FP8E4M3 && FRK_SCORE_FP8_GROUP > 4 && CH % FRK_SCORE_FP8_GROUP == 0;
    static constexpr int G = kWide ? FRK_SCORE_FP8_GROUP : (CH < 4 ? CH : 4);  // chunks per group
    static constexpr int kRingN = kWide ? FRK_SCORE_FP8_RING : kRing;
    static constexpr int kGroups = CH / G;                                       // groups per doc block
    static constexpr int kUnroll = kGroups / Gcd(kGroups, kRingN) * kRingN;     // steps per loop body (lcm)
    static_assert(kRingN >= 2, "the ring needs a slot in use and one in flight");
    static_assert(CH % G == 0, "chunks per doc must be a multiple of the group size");
    static_assert(kUnroll % kGroups == 0 && kUnroll % kRingN == 0, "loop body must hold whole blocks and rings");
};

// qexp: per row block the E8M0 exponent of this lane's query row scale; dexp: the doc's (FP8 only).
// a(rb, c): this lane's query fragment of row block rb and chunk c of the group.
template <int DT, int CH, int RB, typename FragA>
__device__ __forceinline__ void ComputeGroup(FragA&& a, const u32x4 (&buf)[GroupShape<DT, CH>::G], floatx16 (&acc)[RB],
                                    const int (&qexp)[RB], int dexp) {
    constexpr int G = GroupShape<DT, CH>::G;
    if constexpr (DT == FRK_DTYPE_FP8E4M3) {
#pragma unroll
        for (int c = 0; c < G; c += 2)
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
                MfmaChunkPairFp8(a(rb, c), a(rb, c + 1), buf[c], buf[c + 1], acc[rb], qexp[rb], dexp);
    } else if constexpr (DT == FRK_DTYPE_F16X2) {
        // the three products of a slice are issued row block by row block, so consecutive MFMAs of a
        // row block's chain are RB instructions apart
#pragma unroll
        for (int c = 0; c < G; c += 2) {
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
                acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(AsF16(a(rb, c)), AsF16(buf[c]), acc[rb], 0, 0, 0);
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
                acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(AsF16(a(rb, c)), AsF16(buf[c + 1]), acc[rb], 0, 0, 0);
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
                acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(AsF16(a(rb, c + 1)), AsF16(buf[c]), acc[rb], 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int c = 0; c < G; c++)
#pragma unroll
            for (int rb = 0; rb < RB; rb++) MfmaChunk<DT>(a(rb, c), buf[c], acc[rb]);
    }
}

// Survivor queue: a wave appends each (score, doc) that passed its row's floor and filter bit, with the
// row's index in the workgroup's query group, to a wave-private LDS queue; the queue is flushed in bulk,
// so no global memory access or atomic sits inside the MFMA loop. A flush groups its entries by row: LDS
// atomics rank each entry within its row, one global counter atomic per row present claims that row's
// span of its query's candidate list, and the entries are stored at span start + rank. A query with tens
// of thousands of survivors thus costs one counter atomic per flush instead of one per survivor (those
// atomics all hit one address and serialise there), and its keys land contiguously.
#ifndef FRK_SCORE_QUEUE_KEYS
#define FRK_SCORE_QUEUE_KEYS 256  // entries per wave (9 bytes each)
#endif
constexpr int kQueueKeys = FRK_SCORE_QUEUE_KEYS;
constexpr int kQueueRows = 128;  // rows of a query group (32 RB, RB <= 4)
// per wave: kQueueKeys x (score, doc) u32 pairs, then the rows, then a u32 per row (count, then offset)
constexpr int kQueueBytes = kQueueKeys * 9 + kQueueRows * 4;
static_assert(kQueueKeys % 64 == 0 && kQueueKeys >= 64, "the queue flushes whole waves of entries");
static_assert((kQueueKeys * 9) % 4 == 0, "the row counters are u32-aligned");

struct WaveQueue {
    uint2* ent;         // [kQueueKeys]: (score bits, doc)
    uint8_t* row;       // [kQueueKeys]: accumulator row (query index within the workgroup's query group)
    uint32_t* row_at;   // [kQueueRows]: during a flush, each row's entry count, then its span's start
    uint32_t n;         // uniform across the wave
};

__device__ __forceinline__ void WaveSync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// 1 / s for a power of two s = 2^e, e in [-126, 126]: the exponent field mirrored around the bias (exact).
__device__ __forceinline__ float InvPow2(float s) { return __uint_as_float(0x7F000000u - __float_as_uint(s)); }

// q_scale: the per-row query scales whose factor an entry's score still holds (F16X2), else null.
__device__ __forceinline__ void FlushQueue(const ScoreArgs& A, uint32_t q0, WaveQueue& Qu, const float* q_scale) {
    const int lane = threadIdx.x & 63;
    constexpr int kRounds = kQueueKeys / 64;
    WaveSync();
    for (int r = lane; r < kQueueRows; r += 64) Qu.row_at[r] = 0;
    WaveSync();
    // each entry's rank within its row
    uint32_t rank[kRounds];
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
        const uint32_t i = 64 * k + lane;
        rank[k] = i < Qu.n ? atomicAdd(&Qu.row_at[Qu.row[i]], 1u) : 0u;
    }
    WaveSync();
    // one counter atomic per row present: the start of its span in the query's candidate list
    for (int r = lane; r < kQueueRows; r += 64) {
        const uint32_t c = Qu.row_at[r];
        if (c) Qu.row_at[r] = atomicAdd(&A.counts[static_cast<uint64_t>(q0 + r) * A.counts_stride], c);
    }
    WaveSync();
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
        const uint32_t i = 64 * k + lane;
        if (i < Qu.n) {
            const uint2 e = Qu.ent[i];
            const uint32_t r = Qu.row[i];
            const uint32_t q = q0 + r;
            float sc = __uint_as_float(e.x);
            if (q_scale) sc *= InvPow2(q_scale[r]);
            const uint32_t at = Qu.row_at[r] + rank[k];
            if (at < A.cap) A.cand[static_cast<uint64_t>(q) * A.cap + at] = MakeKey(sc, e.y);
        }
    }
    WaveSync();
    Qu.n = 0;
}

// The block's filter words for this wave's row blocks: for row block rb, lane l loads the word of doc
// (l % 32) whose bit x says whether query q0 + 32 rb + x passes it (one coalesced 128-byte load per row
// block), together with the block's first group of doc tiles; for FP8 and F16X2 also the doc's scale.
template <int DT, int RB>
__device__ __forceinline__ void LoadMasks(const ScoreArgs& A, uint32_t q0, uint64_t blk, uint32_t (&m)[RB],
                                          float& dscale) {
    const int lane = threadIdx.x & 63;
    if constexpr (DT == FRK_DTYPE_FP8E4M3 || DT == FRK_DTYPE_F16X2) {
        // the doc's scale (a doc past the end reads the last doc's; its scores are NaN anyway)
        const uint64_t doc = blk * 32 + (lane & 31);
        dscale = A.doc_scale[doc < A.n_docs ? doc : A.n_docs - 1];
    }
    // bmask_stride 0: no filter, bmask is one all-pass word (no branch, so the loads stay in
    // straight-line code)
    const uint32_t* base = A.bmask + blk * A.bmask_stride;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const uint32_t r0 = q0 + 32 * rb;
        const uint32_t off = A.bmask_stride && r0 < A.n_queries ? r0 + (lane & 31) : 0u;
        const uint32_t v = base[off];
        m[rb] = r0 < A.n_queries ? v : 0u;
    }
}

// The filter folded into the accumulators: each block's MFMA chain starts from 0 where the doc passes
// the row's query and from NaN where it does not (or the doc is past the end), so a filtered score is
// NaN, which no floor comparison accepts and no max picks. Lane half h holds rows 8 g + 4 h + x of
// accumulator register 4 g + x; the all-ones word (a NaN) comes from sign-extending the doc's inverted
// filter bit, one bit-field extract per register.
template <int RB>
__device__ __forceinline__ void InitBias(const uint32_t (&m)[RB], bool doc_ok, floatx16 (&acc)[RB]) {
    const uint32_t half = (threadIdx.x & 63) >> 5;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const uint32_t nw = doc_ok ? ~(m[rb] >> (4 * half)) : 0xFFFFFFFFu;
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int x = 0; x < 4; x++)
                acc[rb][4 * g + x] = __int_as_float(-static_cast<int32_t>((nw >> (8 * g + x)) & 1u));
    }
}


// Accumulator row r of a lane in lane half `half` holds query row 8 (r / 4) + 4 half + r % 4 of its row block.
__device__ __forceinline__ uint32_t AccRow(int rb, int r, uint32_t half) {
    return 32u * rb + 8u * (r / 4) + 4u * half + (r % 4);
}

// One row block of the epilogue; rb is a template parameter so the accumulator is always indexed
// statically (a runtime row-block index would send it to scratch). With kUnscale<DT> the accumulators
// hold score * query scale * doc scale: dinv = 1 / doc scale is applied here (exactly, a power of two),
// the floors in floor_lds are already multiplied by the query scale, and the flush divides the query
// scale out of each survivor's score.
template <int DT, int RB, int rb>
__device__ __forceinline__ void EpilogueRowBlock(const ScoreArgs& A, const float* floor_lds, const float* q_scale_lds,
                                                 uint32_t q0, uint64_t doc, float dinv, const floatx16& accr,
                                                 WaveQueue& Qu) {
    const int lane = threadIdx.x & 63;
    const uint32_t half = lane >> 5;
    auto diff = [&](float a, float f) { return kUnscale<DT> ? fmaf(a, dinv, -f) : a - f; };
    auto score = [&](float a) { return kUnscale<DT> ? a * dinv : a; };
    // Does any score of the row block reach its row's floor? The largest (score - floor) over the lane's
    // 16 rows, reduced with three-operand maxima, then one ballot: filtered docs hold NaN, which the
    // maximum passes over (a row with nothing but NaN stays NaN and fails the test), and a score at or
    // above its floor leaves a difference >= 0 (a false alarm from rounding only sends the block through
    // the exact per-score compare below).
    float mg[4];  // per group of four rows
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const floatx4 f4 = *reinterpret_cast<const floatx4*>(floor_lds + 32 * rb + 8 * g + 4 * half);
        mg[g] = fmaxf(fmaxf(diff(accr[4 * g], f4[0]), diff(accr[4 * g + 1], f4[1])),
                      fmaxf(diff(accr[4 * g + 2], f4[2]), diff(accr[4 * g + 3], f4[3])));
    }
    if (__ballot(fmaxf(fmaxf(mg[0], mg[1]), fmaxf(mg[2], mg[3])) >= 0.0f) == 0) return;
    // Survivors, four rows at a time: one room check per group for at most 4 per lane that holds any
    // (never more than the queue), then the exact per-score compares with the group's floors read again
    // (keeping them in registers across the block would cost 16 registers per row block).
    const uint32_t d32 = static_cast<uint32_t>(doc);
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const uint64_t any = __ballot(mg[g] >= 0.0f);
        if (any == 0) continue;
        if (Qu.n + 4u * static_cast<uint32_t>(__popcll(any)) > static_cast<uint32_t>(kQueueKeys))
            FlushQueue(A, q0, Qu, kUnscale<DT> ? q_scale_lds : nullptr);
        const floatx4 f4 = *reinterpret_cast<const floatx4*>(floor_lds + 32 * rb + 8 * g + 4 * half);
#pragma unroll
        for (int x = 0; x < 4; x++) {
            const int r = 4 * g + x;
            const float sv = score(accr[r]);
            const bool hit = sv >= f4[x];
            const uint64_t bal = __ballot(hit);
            if (bal == 0) continue;
            if (hit) {
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(bal >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u));
                const uint32_t slot = Qu.n + below;
                Qu.ent[slot] = make_uint2(__float_as_uint(sv), d32);
                // the lane half's row offset, opaque here so that the compiler does not hoist all 16 RB
                // row indices out of the loop into registers of their own
                uint32_t h4 = 4u * half;
                asm volatile("" : "+v"(h4));
                Qu.row[slot] = static_cast<uint8_t>(AccRow(rb, r, 0) + h4);
            }
            Qu.n += static_cast<uint32_t>(__popcll(bal));
        }
    }
}

template <int DT, int RB, int... I>
__device__ __forceinline__ void EpilogueAll(const ScoreArgs& A, const float* floor_lds, const float* q_scale_lds,
                                            uint32_t q0, uint64_t doc, float dinv, const floatx16 (&acc)[RB],
                                            WaveQueue& Qu, std::integer_sequence<int, I...>) {
    (EpilogueRowBlock<DT, RB, I>(A, floor_lds, q_scale_lds, q0, doc, dinv, acc[I], Qu), ...);
}

// Epilogue of one finished 32-doc block: every accumulator row against its floor (filtered docs and
// docs past the end hold NaN); survivors go to the wave queue.
template <int DT, int RB>
__device__ __forceinline__ void Epilogue(const ScoreArgs& A, const float* floor_lds, const float* q_scale_lds, uint32_t q0,
                                         uint64_t blk, float dinv, const floatx16 (&acc)[RB], WaveQueue& Qu) {
    const uint64_t doc = blk * 32 + (threadIdx.x & 31);
    EpilogueAll<DT, RB>(A, floor_lds, q_scale_lds, q0, doc, dinv, acc, Qu, std::make_integer_sequence<int, RB>{});
}

// One stage of the block-max butterfly over lane bit B with N rows left per lane (M = array size):
// a lane keeps the lower or upper half of its rows (by its bit B) and maxes in the partner's copy of
// the same rows; once one row is left the remaining stages are plain exchanges. Recursion keeps every
// index a compile-time constant so the rows stay in registers.
template <int M, int N, int B>
__device__ __forceinline__ void BlockMaxStages(float (&v)[M], uint32_t j, uint32_t& base) {
    if constexpr (B >= 1) {
        if constexpr (N > 1) {
            const bool upper = (j & B) != 0;
#pragma unroll
            for (int i = 0; i < N / 2; i++) {
                const float send = upper ? v[i] : v[i + N / 2];
                const float keep = upper ? v[i + N / 2] : v[i];
                v[i] = fmaxf(keep, __shfl_xor(send, B, 64));
            }
            if (upper) base += N / 2;
            BlockMaxStages<M, N / 2, B / 2>(v, j, base);
        } else {
            v[0] = fmaxf(v[0], __shfl_xor(v[0], B, 64));
            BlockMaxStages<M, 1, B / 2>(v, j, base);
        }
    }
}

// Probe epilogue: per (query, block) the largest score among the block's docs that pass the filter
// (-inf if none), written to blockmax[q * A.cap + block index] -- no atomics, no compaction. Filtered
// docs hold NaN, which the maximum (IEEE maxNum) passes over; a row with no passing doc ends as NaN
// and is written as -inf.
// The max over the 32 docs (lanes j of each half) is a butterfly transpose-reduce: at each of the five
// lane-bit stages a lane keeps half of its rows and trades the other half with its partner, so the
// 16 RB rows a lane holds shrink to 16 RB / 32 after 16 RB - 1 exchanges in total (instead of five
// exchanges per row). After it, lane j of half h holds the rows of flat accumulator indices 2j, 2j+1
// (RB = 4), or the row of index j (RB = 2); with RB = 1 the last stage is a plain exchange of the
// single remaining row, index j / 2, which lanes j and j ^ 1 then share. With kUnscale<DT> the doc
// scale is divided out before the maximum and the query scale after it.
template <int DT, int RB>
__device__ __forceinline__ void EpilogueBlockMax(const ScoreArgs& A, const float* q_scale_lds, uint32_t q0,
                                                 uint64_t step_index, float dinv, const floatx16 (&acc)[RB]) {
    constexpr int N = 16 * RB;
    const int lane = threadIdx.x & 63;
    const uint32_t j = lane & 31;
    const uint32_t half = lane >> 5;
    float v[N];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int r = 0; r < 16; r++) v[16 * rb + r] = kUnscale<DT> ? acc[rb][r] * dinv : acc[rb][r];
    uint32_t base = 0;  // flat index of v[0] after the halving stages
    BlockMaxStages<N, N, 16>(v, j, base);
    constexpr int n = N >= 32 ? N / 32 : 1;  // rows left per lane
    float* out = reinterpret_cast<float*>(A.cand);
#pragma unroll
    for (int e = 0; e < n; e++) {
        // fewer than 32 rows: the last log2(32 / N) stages were plain exchanges, so lanes that differ
        // only in those low bits hold the same row; the lowest of them writes it
        if (N < 32 && (j & (32 / N - 1)) != 0) break;
        const uint32_t r = base + e;
        const uint32_t row = 32 * (r >> 4) + 8 * ((r >> 2) & 3) + 4 * half + (r & 3);
        const uint32_t q = q0 + row;
        const float m = kUnscale<DT> ? v[e] * InvPow2(q_scale_lds[row]) : v[e];
        if (q < A.n_queries) out[static_cast<uint64_t>(q) * A.cap + step_index] = m == m ? m : -INFINITY;
    }
}

// Dense epilogue: every score of the block to the matrix, out[q * A.cap + doc] for docs below n_docs and
// real queries. For each accumulator register, lanes 0-31 write 32 consecutive docs of one row and
// lanes 32-63 those of the row four below: two 128-byte runs per store instruction.
template <int DT, int RB>
__device__ __forceinline__ void EpilogueDense(const ScoreArgs& A, const float* q_scale_lds, uint32_t q0, uint64_t blk,
                                              float dinv, const floatx16 (&acc)[RB]) {
    const int lane = threadIdx.x & 63;
    const uint64_t doc = blk * 32 + (lane & 31);
    if (doc >= A.n_docs) return;
    float* out = reinterpret_cast<float*>(A.cand);
    const uint32_t half = lane >> 5;
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t row = 32 * rb + 8 * (r / 4) + 4 * half + (r % 4);
            const uint32_t q = q0 + row;
            const float sc = kUnscale<DT> ? acc[rb][r] * dinv * InvPow2(q_scale_lds[row]) : acc[rb][r];
            if (q < A.n_queries) out[static_cast<uint64_t>(q) * A.cap + doc] = sc;
        }
}

template <int DT, int CH, int RB, int kMode>
__global__ __launch_bounds__((WgShape<DT, CH, RB>::kThreads), 1) void ScoreKernel(ScoreArgs A) {
    constexpr int kThreads = WgShape<DT, CH, RB>::kThreads;
    constexpr int kWaves = WgShape<DT, CH, RB>::kWaves;
    constexpr int G = GroupShape<DT, CH>::G;
    constexpr int NG = GroupShape<DT, CH>::kGroups;
    constexpr int kRingN = GroupShape<DT, CH>::kRingN;
    static_assert(32 * RB <= kQueueRows, "a queue row index is a row of the query group");
    extern __shared__ u32x4 smem[];
    u32x4* a_lds = smem;                                               // [RB][CH][64]
    float* q_scale_lds = reinterpret_cast<float*>(smem + RB * CH * 64);  // [32 * RB]
    float* floor_lds = q_scale_lds + 32 * RB;                          // [32 * RB]
    uint8_t* queue_lds = reinterpret_cast<uint8_t*>(floor_lds + 32 * RB);  // [waves][kQueueBytes]
    // Workgroup -> (doc range x, query group): with several query groups every doc range is read once
    // per group, so the groups of one range go to the same XCD (workgroups are dealt to the 8 XCDs
    // round-robin by linear id) and run back to back there, sharing that XCD's L2 instead of each
    // fetching the range from HBM. The grid is 1-D, 8 * groups * (ranges / 8) workgroups.
    const uint32_t groups = A.n_groups;
    const uint32_t n_ranges = gridDim.x / groups;
    uint32_t range_x, group;
    if (groups == 1) {
        range_x = blockIdx.x;
        group = 0;
    } else {
        const uint32_t xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
        group = slot % groups;
        range_x = (slot / groups) * kXcds + xcd;
    }
    const uint32_t q0 = group * (32 * RB);
    StageQueries<DT, CH, RB>(A, a_lds, q_scale_lds, q0);
    for (int r = threadIdx.x; r < 32 * RB; r += kThreads) {
        const uint32_t q = q0 + r;
        // rows past the last query never pass
        const float fl = q < A.n_queries ? (A.floor ? A.floor[q] : -INFINITY) : INFINITY;
        // F16X2 compares scores that still hold the query's scale (a power of two: exact)
        floor_lds[r] = kUnscale<DT> ? fl * q_scale_lds[r] : fl;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    // the wave index is uniform over the wave: read it into a scalar register, so the block and step
    // arithmetic below (and every doc tile address derived from it) runs on the scalar unit
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    WaveQueue Qu;
    Qu.ent = reinterpret_cast<uint2*>(queue_lds + wave * kQueueBytes);
    Qu.row = queue_lds + wave * kQueueBytes + kQueueKeys * sizeof(uint2);
    Qu.row_at = reinterpret_cast<uint32_t*>(Qu.row + kQueueKeys);
    Qu.n = 0;
    const uint64_t per_wg = (A.block_count + n_ranges - 1) / n_ranges;
    const uint64_t i_lo = static_cast<uint64_t>(range_x) * per_wg;
    const uint64_t i_hi = i_lo + per_wg < A.block_count ? i_lo + per_wg : A.block_count;
    if (i_lo + wave >= i_hi) return;
    const uint64_t n_mine = (i_hi - i_lo - wave + kWaves - 1) / kWaves;  // blocks of this wave
    const uint64_t total = n_mine * NG;                                   // (block, group) steps
    constexpr int U = GroupShape<DT, CH>::kUnroll;
    static_assert(U % 2 == 0 || !WgShape<DT, CH, RB>::kStageA, "staged fragment buffers alternate by step");

    auto block_index = [&](uint64_t step) -> uint64_t { return i_lo + wave + (step / NG) * kWaves; };
    u32x4 ring[kRingN][G];
    uint32_t mring[kRingN][RB];  // filter words of the block whose first group sits in the same slot
    float sring[kRingN];         // FP8: that block's doc scale
    int qexp[RB] = {};          // FP8: E8M0 exponent of this lane's query row scale, per row block
    if constexpr (DT == FRK_DTYPE_FP8E4M3) {
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
            qexp[rb] = static_cast<int>((__float_as_uint(q_scale_lds[32 * rb + (lane & 31)]) >> 23) & 0xFFu);
    }
    int dexp = 127;
    float dinv = 1.0f;  // F16X2: 1 / the block's doc scale, per lane
    // staged query fragments (kStageA): two buffers, the step's group and the next step's
    constexpr bool kStageA = WgShape<DT, CH, RB>::kStageA;
    u32x4 afrag[kStageA ? 2 : 1][G][RB];
    auto load_a = [&](auto buf_c, auto group_c) __attribute__((always_inline)) {
        constexpr int B = decltype(buf_c)::value;
        constexpr int GL = decltype(group_c)::value;
#pragma unroll
        for (int c = 0; c < G; c++)
#pragma unroll
            for (int rb = 0; rb < RB; rb++) afrag[B][c][rb] = a_lds[(rb * CH + GL * G + c) * 64 + lane];
    };
    if constexpr (kStageA) load_a(std::integral_constant<int, 0>{}, std::integral_constant<int, 0>{});
    // load step `step` (group index GL, a compile-time constant) into ring slot SL
    auto load = [&](auto slot_c, auto group_c, uint64_t step) __attribute__((always_inline)) {
        constexpr int SL = decltype(slot_c)::value;
        constexpr int GL = decltype(group_c)::value;
        const uint64_t st = step < total ? step : total - 1;
        const uint64_t blk = A.block_begin + block_index(st) * A.block_step;
        const u32x4* p = A.tiles + (blk * CH + static_cast<uint64_t>(GL) * G) * 64 + lane;
#pragma unroll
        for (int c = 0; c < G; c++) ring[SL][c] = p[c * 64];
        if constexpr (GL == 0) LoadMasks<DT, RB>(A, q0, blk, mring[SL], sring[SL]);
    };
    // the first kRingN - 1 steps, in step order (each wait below counts on that order)
    StaticFor([&](auto s_c) __attribute__((always_inline)) {
        constexpr int st = decltype(s_c)::value;
        load(std::integral_constant<int, st>{}, std::integral_constant<int, st % NG>{}, st);
        __builtin_amdgcn_sched_barrier(0);
    }, std::make_integer_sequence<int, kRingN - 1>{});
    floatx16 acc[RB];  // set from the filter words at the start of every block
    // step s of the loop body (body starts at a block boundary `base`)
    auto step_fn = [&](auto s_c, uint64_t base) __attribute__((always_inline)) {
        constexpr int s = decltype(s_c)::value;
        constexpr int S = s % kRingN;
        constexpr int g = s % NG;
        const uint64_t step = base + s;
        load(std::integral_constant<int, (S + kRingN - 1) % kRingN>{}, std::integral_constant<int, (s + kRingN - 1) % NG>{},
             step + kRingN - 1);
        // keep the prefetch at the head of the step: the scheduler would otherwise sink the loads
        // towards their use and shrink the distance the ring is there to provide
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (g == 0) {
            const uint64_t blk = A.block_begin + block_index(step) * A.block_step;
            InitBias<RB>(mring[S], blk * 32 + (lane & 31) < A.n_docs, acc);
            if constexpr (DT == FRK_DTYPE_FP8E4M3) dexp = static_cast<int>((__float_as_uint(sring[S]) >> 23) & 0xFFu);
            if constexpr (kUnscale<DT>) dinv = InvPow2(sring[S]);
        }
        if constexpr (kStageA) {
            // the next step's fragments are read before this step's MFMAs are issued, so their LDS
            // latency runs under those MFMAs; U is even (asserted above), so buffer s % 2
            // continues across loop bodies
            load_a(std::integral_constant<int, (s + 1) % 2>{}, std::integral_constant<int, (s + 1) % NG>{});
            __builtin_amdgcn_sched_barrier(0);
            ComputeGroup<DT, CH, RB>([&](int rb, int c) -> const u32x4& { return afrag[s % 2][c][rb]; }, ring[S], acc,
                                     qexp, dexp);
        } else {
            ComputeGroup<DT, CH, RB>(
                [&](int rb, int c) -> u32x4 { return a_lds[(rb * CH + g * G + c) * 64 + lane]; }, ring[S], acc,
                qexp, dexp);
        }
        if constexpr (g == NG - 1) {
            // every row block's MFMAs are issued before the first epilogue; left alone, the compiler
            // sinks a row block's MFMAs past the epilogue branches of the row blocks before it, and the
            // MFMA and epilogue work of one wave then no longer overlap
#pragma unroll
            for (int rb = 0; rb < RB; rb++) asm volatile("" : "+v"(acc[rb]));
            const uint64_t i = block_index(step);
            const uint64_t blk = A.block_begin + i * A.block_step;
            if constexpr (kMode == kModeBlockMax)
                EpilogueBlockMax<DT, RB>(A, q_scale_lds, q0, i, dinv, acc);
            else if constexpr (kMode == kModeDense)
                EpilogueDense<DT, RB>(A, q_scale_lds, q0, blk, dinv, acc);
            else
                Epilogue<DT, RB>(A, floor_lds, q_scale_lds, q0, blk, dinv, acc, Qu);
        }
    };
    uint64_t base = 0;
    for (; base + U <= total; base += U)
        StaticFor([&](auto s_c) __attribute__((always_inline)) { step_fn(s_c, base); },
                  std::make_integer_sequence<int, U>{});
    if constexpr (U > NG) {
        // the last blocks when a body holds several (small blocks): whole blocks, one at a time
        StaticFor([&](auto s_c) __attribute__((always_inline)) {
            if (base + decltype(s_c)::value < total) step_fn(s_c, base);
        }, std::make_integer_sequence<int, U - NG>{});
    }
    if constexpr (kMode == kModeSelect) FlushQueue(A, q0, Qu, kUnscale<DT> ? q_scale_lds : nullptr);
}

using KernelFn = void (*)(ScoreArgs);

// chunk counts per doc the scorer is compiled for, ascending (the cases of Dispatch below)
constexpr uint32_t kScoreShapes[] = {4, 8, 12, 16, 24, 32, 48, 64, 96, 128};

template <int DT, int CH, int RB, int kMode>
int Launch(const ScoreArgs& A, hipStream_t stream, uint32_t grid_x) {
    constexpr int kThreads = WgShape<DT, CH, RB>::kThreads;
    constexpr int kWaves = WgShape<DT, CH, RB>::kWaves;
    const size_t lds = static_cast<size_t>(RB) * CH * 64 * sizeof(u32x4) + 2 * 32 * RB * sizeof(float) +
                       kWaves * kQueueBytes;
    const void* fn = reinterpret_cast<const void*>(&ScoreKernel<DT, CH, RB, kMode>);
    // workgroups of this shape one CU holds at once, with the LDS attribute set: once per process (a
    // function-local static is initialised exactly once even when several host threads launch together)
    struct Residency {
        hipError_t err;
        int per_cu;
    };
    static const Residency res = [&] {
        Residency r{hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)), 1};
        int occ = 0;
        if (r.err == hipSuccess) r.err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kThreads, lds);
        r.per_cu = occ > 0 ? occ : 1;
        return r;
    }();
    FRK_HIP_CHECK(res.err);
    const int resident_per_cu = res.per_cu;
    int cus = 0;
    if (int rc = CurrentDeviceCUs(&cus)) return rc;
    const uint32_t groups = (A.n_queries + 32 * RB - 1) / (32 * RB);
    // one wave of resident workgroups over the chip (each walks a contiguous block range, so the
    // query tile is staged once per workgroup), never more than the caller's work-based cap
    const uint32_t resident = static_cast<uint32_t>(cus > 0 ? cus : 1) * static_cast<uint32_t>(resident_per_cu);
    const uint32_t per_group = resident / groups > 0 ? resident / groups : 1;
    if (grid_x > per_group) grid_x = per_group;
    // whole rounds of the 8 XCDs, rounded down so ranges x groups stays within one resident wave
    if (groups > 1) grid_x = grid_x >= kXcds ? grid_x / kXcds * kXcds : kXcds;
    ScoreArgs a = A;
    a.n_groups = groups;
    hipLaunchKernelGGL((ScoreKernel<DT, CH, RB, kMode>), dim3(grid_x * groups), dim3(kThreads), lds, stream, a);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

#ifndef FRK_SCORE_MAX_RB
#define FRK_SCORE_MAX_RB 4  // most 32-query row blocks one workgroup takes (1, 2 or 4)
#endif
#ifndef FRK_SCORE_FP8_PROBE_MAX_RB
#define FRK_SCORE_FP8_PROBE_MAX_RB 2
#endif

// Instantiated shapes: CH = 16-byte chunks per doc (FP32 dims 32..1024, BF16 64..2048, FP8 128..4096
// at the listed counts; kScoreShapes lists them for frk_store_dim, which pads a store's dims up to one).
// RB = 32-query row blocks per workgroup, at most what keeps the query tile (RB * CH KiB) within
// 128 KiB of the 160 KiB LDS; 48 chunks (BF16 768 dims, FP32 384, FP8 1536) also take 3 row blocks (144 KiB, with the
// 4-wave queue 155 KiB), used for the main pass where it needs fewer query groups than 2 row blocks:
// every group re-reads the doc range (from the XCD's L2 after the first). Measured on 10M x 768 BF16
// (with the staged query fragments): 192 queries 4.02 -> 3.53 ms, 128 queries (still two groups of 64)
// 2.91 -> 2.75 ms from the staging alone.
}  // namespace frk

#include "frk_score_staged.h"

namespace frk {

template <int DT, int kMode>
int Dispatch(const ScoreArgs& A, uint32_t ch, hipStream_t stream, uint32_t grid_x) {
    const uint32_t q = A.n_queries;
    // wide docs and large batches: the staged scorer (one read of the doc range for up to 256 queries)
#define FRK_STAGED(CHV)                                                        \
    if (ch == CHV) {                                                           \
        bool done = false;                                                     \
        const int rc = TryStaged<DT, CHV, kMode>(A, stream, grid_x, &done);    \
        if (done) return rc;                                                   \
    }
    FRK_STAGED(8)
    FRK_STAGED(16)
    FRK_STAGED(24)
    FRK_STAGED(32)
    FRK_STAGED(48)
#undef FRK_STAGED
    int rb = q > 64 ? 4 : (q > 32 ? 2 : 1);
    if (rb > FRK_SCORE_MAX_RB) rb = FRK_SCORE_MAX_RB;
    // FP8 probe pass: two query groups of two row blocks (sharing an XCD's L2) measured faster
    if (DT == FRK_DTYPE_FP8E4M3 && kMode == kModeBlockMax && rb > FRK_SCORE_FP8_PROBE_MAX_RB)
        rb = FRK_SCORE_FP8_PROBE_MAX_RB;
    if (kMode == kModeDense) rb = 1;  // the unfused score matrix is written one row block per workgroup
    const bool three = rb == 4 && (q + 95) / 96 < (q + 63) / 64;
#define FRK_CASE(CHV)                                                                           \
    if (ch == CHV) {                                                                            \
        if constexpr (CHV * 4 <= 128 && kMode != kModeDense) {                                  \
            if (rb == 4) return Launch<DT, CHV, 4, kMode>(A, stream, grid_x);                   \
        }                                                                                       \
        if constexpr (CHV == 48 && kMode == kModeSelect) {                                      \
            if (three) return Launch<DT, CHV, 3, kMode>(A, stream, grid_x);                     \
        }                                                                                       \
        if constexpr (CHV * 2 <= 128 && kMode != kModeDense) {                                  \
            if (rb >= 2) return Launch<DT, CHV, 2, kMode>(A, stream, grid_x);                   \
        }                                                                                       \
        return Launch<DT, CHV, 1, kMode>(A, stream, grid_x);                                    \
    }
    FRK_CASE(4)
    FRK_CASE(8)
    FRK_CASE(12)
    FRK_CASE(16)
    FRK_CASE(24)
    FRK_CASE(32)
    FRK_CASE(48)
    FRK_CASE(64)
    FRK_CASE(96)
    FRK_CASE(128)
#undef FRK_CASE
    SetError("frk_score_select: %u chunks per doc is not an instantiated shape", ch);
    return FRK_ERR_UNSUPPORTED;
}

// Per-dtype entry points, each in its own translation unit (frk_score_<dtype>.hip).
struct ScoreLaunch {
    ScoreArgs args;
    uint32_t chunks;
    uint32_t grid_x;
    hipStream_t stream;
};
int ScoreF32(const ScoreLaunch& L, int mode);
int ScoreBf16(const ScoreLaunch& L, int mode);
int ScoreFp8(const ScoreLaunch& L, int mode);
int ScoreF16x2(const ScoreLaunch& L, int mode);
}  // namespace frk
