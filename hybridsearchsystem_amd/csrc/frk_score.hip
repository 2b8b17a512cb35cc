// frk_score.hip -- fused inner-product scorer on MFMA with the filter mask, a per-query score floor and
// survivor compaction in the epilogue. No score matrix is written.
//
// The reference computes S = Q x D^T on the Cube unit into a full [Q][docs] FP32 matrix
// (engine/retrieval/device/scorer/kernel_vector_mmad_op.h:21-199) and a second kernel re-reads it with
// the filter bitmap to compact survivors (device/aggregator/kernel_aggregator.h:22-182). Here one pass
// does both: a wave computes a 32-doc block against the whole query tile held in LDS, then each
// accumulator row is tested against (mask bit, floor), ballot-ed, and the passing (score, doc) pairs
// are appended to per-query candidate lists at offsets from one atomic per (wave, query) and the
// prefix count of the ballot.
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

#include "frk_common.h"

namespace frk {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

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

// One chunk (or, for fp8, one chunk pair) of MFMA work for one 32-query row block.
template <int DT>
__device__ inline void MfmaChunk(const u32x4& a, const u32x4& b, floatx16& acc);

template <>
__device__ inline void MfmaChunk<FRK_DTYPE_F32>(const u32x4& a, const u32x4& b, floatx16& acc) {
    floatx4 av = __builtin_bit_cast(floatx4, a);
    floatx4 bv = __builtin_bit_cast(floatx4, b);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[0], bv[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[1], bv[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[2], bv[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[3], bv[3], acc, 0, 0, 0);
}

template <>
__device__ inline void MfmaChunk<FRK_DTYPE_BF16>(const u32x4& a, const u32x4& b, floatx16& acc) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                  0, 0, 0);
}

__device__ inline void MfmaChunkPairFp8(const u32x4& a0, const u32x4& a1, const u32x4& b0, const u32x4& b1,
                                        floatx16& acc) {
    intx8 av = {static_cast<int>(a0[0]), static_cast<int>(a0[1]), static_cast<int>(a0[2]), static_cast<int>(a0[3]),
                static_cast<int>(a1[0]), static_cast<int>(a1[1]), static_cast<int>(a1[2]), static_cast<int>(a1[3])};
    intx8 bv = {static_cast<int>(b0[0]), static_cast<int>(b0[1]), static_cast<int>(b0[2]), static_cast<int>(b0[3]),
                static_cast<int>(b1[0]), static_cast<int>(b1[1]), static_cast<int>(b1[2]), static_cast<int>(b1[3])};
    // formats: 0 = FP8 E4M3 for both operands; scales 127 = 2^0
    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, 127, 0, 127);
}

struct ScoreArgs {
    const u32x4* tiles;
    const float* doc_scale;
    uint64_t n_docs;
    const float* queries;
    uint32_t n_queries;
    uint32_t dim;  // padded dim of the store
    uint32_t q_dim;  // row length of the query matrix (== dim)
    const uint64_t* masks;
    uint64_t mask_stride;
    const float* floor;
    uint64_t block_begin, block_step, block_count;
    uint64_t* cand;
    uint32_t cap;
    uint32_t* counts;
};

// Writes the query tile into LDS in the doc tiling, converted to the store's element type; for fp8 the
// per-query scale (max|q| / 448) goes to q_scale_lds.
template <int DT, int CH, int RB>
__device__ inline void StageQueries(const ScoreArgs& A, u32x4* a_lds, float* q_scale_lds, uint32_t q0) {
    constexpr int E = Traits<DT>::kElemsPerLane;
    if constexpr (DT == FRK_DTYPE_FP8E4M3) {
        for (int r = threadIdx.x; r < 32 * RB; r += kThreads) {
            uint32_t q = q0 + r;
            float mx = 0.0f;
            if (q < A.n_queries)
                for (uint32_t d = 0; d < A.dim; d++) mx = fmaxf(mx, fabsf(A.queries[static_cast<uint64_t>(q) * A.q_dim + d]));
            q_scale_lds[r] = mx > 0.0f ? mx / 448.0f : 1.0f;
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < RB * CH * 64; idx += kThreads) {
        const int l = idx & 63;
        const int c = (idx >> 6) % CH;
        const int rb = idx / (CH * 64);
        const int i = l & 31, h = l >> 5;
        const uint32_t q = q0 + 32 * rb + i;
        const uint32_t d0 = c * 2 * E + h * E;
        float v[E];
        for (int e = 0; e < E; e++) {
            uint32_t d = d0 + e;
            v[e] = (q < A.n_queries && d < A.dim) ? A.queries[static_cast<uint64_t>(q) * A.q_dim + d] : 0.0f;
        }
        u32x4 packed;
        uint32_t* pw = reinterpret_cast<uint32_t*>(&packed);
        if constexpr (DT == FRK_DTYPE_F32) {
            for (int e = 0; e < 4; e++) pw[e] = __builtin_bit_cast(uint32_t, v[e]);
        } else if constexpr (DT == FRK_DTYPE_BF16) {
            for (int e = 0; e < 4; e++)
                pw[e] = static_cast<uint32_t>(F32ToBf16(v[2 * e])) | (static_cast<uint32_t>(F32ToBf16(v[2 * e + 1])) << 16);
        } else {
            const float inv = 1.0f / q_scale_lds[32 * rb + i];
            for (int e = 0; e < 4; e++) {
                uint32_t w = 0;
                for (int x = 0; x < 4; x++) w |= static_cast<uint32_t>(F32ToE4M3(v[4 * e + x] * inv)) << (8 * x);
                pw[e] = w;
            }
        }
        a_lds[idx] = packed;
    }
}

// Chunks per pipeline group and groups in flight: a wave streams its doc blocks as a flat sequence of
// (block, group) steps through a ring of kRing register buffers, so kRing - 1 groups of loads are in
// flight while one group feeds the MFMAs.
constexpr int kRing = 3;

template <int CH>
struct GroupShape {
    static constexpr int G = CH < 8 ? CH : 8;  // chunks per group
    static constexpr int kGroups = CH / G;     // groups per doc block
    static_assert(CH % G == 0, "chunks per doc must be a multiple of the group size");
};

template <int DT, int CH, int RB>
__device__ inline void ComputeGroup(const u32x4* a_lds, int g, const u32x4 (&buf)[GroupShape<CH>::G],
                                    floatx16 (&acc)[RB]) {
    constexpr int G = GroupShape<CH>::G;
    const int lane = threadIdx.x & 63;
    if constexpr (DT == FRK_DTYPE_FP8E4M3) {
#pragma unroll
        for (int c = 0; c < G; c += 2) {
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const u32x4 a0 = a_lds[(rb * CH + g * G + c) * 64 + lane];
                const u32x4 a1 = a_lds[(rb * CH + g * G + c + 1) * 64 + lane];
                MfmaChunkPairFp8(a0, a1, buf[c], buf[c + 1], acc[rb]);
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < G; c++) {
#pragma unroll
            for (int rb = 0; rb < RB; rb++) MfmaChunk<DT>(a_lds[(rb * CH + g * G + c) * 64 + lane], buf[c], acc[rb]);
        }
    }
}

// Mask, floor and ballot compaction of one finished 32-doc block.
template <int DT, int RB>
__device__ inline void Epilogue(const ScoreArgs& A, const float* q_scale_lds, const float (&floor_v)[RB], uint32_t q0,
                                uint64_t blk, const floatx16 (&acc)[RB]) {
    const int lane = threadIdx.x & 63;
    const uint64_t doc = blk * 32 + (lane & 31);
    const bool doc_ok = doc < A.n_docs;
    float dscale = 1.0f;
    if constexpr (DT == FRK_DTYPE_FP8E4M3) dscale = doc_ok ? A.doc_scale[doc] : 0.0f;
    const uint64_t mword = doc >> 6;
    const uint32_t mshift = static_cast<uint32_t>(blk & 1) * 32;
    const uint32_t half = lane >> 5;
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        // lane l fetches the 32-doc mask word of query row (l % 32) of this row block
        const uint32_t my_q = q0 + 32 * rb + (lane & 31);
        uint32_t my_mask = 0;
        if (my_q < A.n_queries)
            my_mask = A.masks ? static_cast<uint32_t>(A.masks[static_cast<uint64_t>(my_q) * A.mask_stride + mword] >> mshift)
                              : 0xFFFFFFFFu;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i0 = 8 * (r / 4) + (r % 4);  // query row held by lane half 0; half 1 holds i0 + 4
            const uint32_t m_lo = __builtin_amdgcn_readlane(my_mask, i0);
            const uint32_t m_hi = __builtin_amdgcn_readlane(my_mask, i0 + 4);
            const int fb = __builtin_bit_cast(int, floor_v[rb]);
            const float f_lo = __builtin_bit_cast(float, __builtin_amdgcn_readlane(fb, i0));
            const float f_hi = __builtin_bit_cast(float, __builtin_amdgcn_readlane(fb, i0 + 4));
            float s = acc[rb][r];
            if constexpr (DT == FRK_DTYPE_FP8E4M3) s *= dscale * q_scale_lds[32 * rb + i0 + 4 * half];
            const uint32_t m = half ? m_hi : m_lo;
            const float f = half ? f_hi : f_lo;
            const bool pass = doc_ok && ((m >> (lane & 31)) & 1u) && s >= f;
            const uint64_t bal = __ballot(pass);
            if (bal == 0) continue;
            const uint32_t bits = half ? static_cast<uint32_t>(bal >> 32) : static_cast<uint32_t>(bal);
            const uint32_t q = q0 + 32 * rb + i0 + 4 * half;
            uint32_t base = 0;
            if ((lane & 31) == 0 && bits) base = atomicAdd(&A.counts[q], static_cast<uint32_t>(__popc(bits)));
            base = __shfl(base, lane & 32, 64);
            if (pass) {
                const uint32_t slot = base + static_cast<uint32_t>(__popc(bits & ((1u << (lane & 31)) - 1u)));
                if (slot < A.cap) A.cand[static_cast<uint64_t>(q) * A.cap + slot] = MakeKey(s, static_cast<uint32_t>(doc));
            }
        }
    }
}

template <int DT, int CH, int RB>
__global__ __launch_bounds__(kThreads, 1) void ScoreKernel(ScoreArgs A) {
    constexpr int G = GroupShape<CH>::G;
    constexpr int NG = GroupShape<CH>::kGroups;
    extern __shared__ u32x4 smem[];
    u32x4* a_lds = smem;                                               // [RB][CH][64]
    float* q_scale_lds = reinterpret_cast<float*>(smem + RB * CH * 64);  // [32 * RB]
    const uint32_t q0 = blockIdx.y * (32 * RB);
    StageQueries<DT, CH, RB>(A, a_lds, q_scale_lds, q0);
    float floor_v[RB];
    for (int rb = 0; rb < RB; rb++) {
        const uint32_t q = q0 + 32 * rb + (threadIdx.x & 31);
        floor_v[rb] = (A.floor && q < A.n_queries) ? A.floor[q] : -INFINITY;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t per_wg = (A.block_count + gridDim.x - 1) / gridDim.x;
    const uint64_t i_lo = static_cast<uint64_t>(blockIdx.x) * per_wg;
    const uint64_t i_hi = i_lo + per_wg < A.block_count ? i_lo + per_wg : A.block_count;
    if (i_lo + wave >= i_hi) return;
    const uint64_t n_mine = (i_hi - i_lo - wave + kWaves - 1) / kWaves;  // blocks of this wave
    const uint64_t total = n_mine * NG;                                   // (block, group) steps

    auto src = [&](uint64_t step) -> const u32x4* {
        const uint64_t j = step / NG;
        const uint64_t blk = A.block_begin + (i_lo + wave + j * kWaves) * A.block_step;
        return A.tiles + (blk * CH + (step % NG) * G) * 64 + lane;
    };
    u32x4 ring[kRing][G];
#pragma unroll
    for (int s = 0; s < kRing - 1; s++)
        if (static_cast<uint64_t>(s) < total) {
            const u32x4* p = src(s);
#pragma unroll
            for (int c = 0; c < G; c++) ring[s][c] = p[c * 64];
        }
    floatx16 acc[RB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
        for (int r = 0; r < 16; r++) acc[rb][r] = 0.0f;
    for (uint64_t base = 0; base < total; base += kRing) {
#pragma unroll
        for (int s = 0; s < kRing; s++) {
            const uint64_t step = base + s;
            if (step >= total) break;
            const uint64_t ahead = step + kRing - 1;
            if (ahead < total) {
                const u32x4* p = src(ahead);
#pragma unroll
                for (int c = 0; c < G; c++) ring[(s + kRing - 1) % kRing][c] = p[c * 64];
            }
            const int g = static_cast<int>(step % NG);
            ComputeGroup<DT, CH, RB>(a_lds, g, ring[s], acc);
            if (g == NG - 1) {
                const uint64_t blk = A.block_begin + (i_lo + wave + (step / NG) * kWaves) * A.block_step;
                Epilogue<DT, RB>(A, q_scale_lds, floor_v, q0, blk, acc);
#pragma unroll
                for (int rb = 0; rb < RB; rb++)
                    for (int r = 0; r < 16; r++) acc[rb][r] = 0.0f;
            }
        }
    }
}

using KernelFn = void (*)(ScoreArgs);

template <int DT, int CH, int RB>
int Launch(const ScoreArgs& A, hipStream_t stream, uint32_t grid_x) {
    const size_t lds = static_cast<size_t>(RB) * CH * 64 * sizeof(u32x4) + 32 * RB * sizeof(float);
    static bool attr_set = false;
    if (!attr_set) {
        FRK_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ScoreKernel<DT, CH, RB>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        attr_set = true;
    }
    const uint32_t groups = (A.n_queries + 32 * RB - 1) / (32 * RB);
    hipLaunchKernelGGL((ScoreKernel<DT, CH, RB>), dim3(grid_x, groups), dim3(kThreads), lds, stream, A);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

// Instantiated shapes: CH = 16-byte chunks per doc. RB = 32-query row blocks per workgroup, chosen so
// the query tile fits the 160 KiB LDS.
template <int DT>
int Dispatch(const ScoreArgs& A, uint32_t ch, hipStream_t stream, uint32_t grid_x) {
    const uint32_t q = A.n_queries;
    const int rb = q > 64 ? 4 : (q > 32 ? 2 : 1);
#define FRK_CASE(CHV)                                                              \
    if (ch == CHV) {                                                               \
        if (rb == 4 && CHV * 4 <= 128) return Launch<DT, CHV, 4>(A, stream, grid_x); \
        if (rb >= 2 && CHV * 2 <= 128) return Launch<DT, CHV, 2>(A, stream, grid_x); \
        return Launch<DT, CHV, 1>(A, stream, grid_x);                              \
    }
    FRK_CASE(4)
    FRK_CASE(8)
    FRK_CASE(16)
    FRK_CASE(32)
    FRK_CASE(48)
    FRK_CASE(64)
#undef FRK_CASE
    SetError("frk_score_select: %u chunks per doc is not an instantiated shape", ch);
    return FRK_ERR_UNSUPPORTED;
}

}  // namespace
}  // namespace frk

extern "C" uint32_t frk_chunk_dims(uint32_t dtype) {
    switch (dtype) {
        case FRK_DTYPE_F32: return 8;
        case FRK_DTYPE_BF16: return 16;
        case FRK_DTYPE_FP8E4M3: return 32;
    }
    return 0;
}

extern "C" int frk_score_select(const frk_doc_store* store, const float* d_queries, uint32_t n_queries,
                                const uint64_t* d_masks, uint64_t mask_stride, const float* d_floor,
                                uint64_t block_begin, uint64_t block_step, uint64_t block_count, uint64_t* d_cand,
                                uint32_t cand_cap, uint32_t* d_counts, frk_stream_t stream) {
    if (!store || !store->tiles || !d_queries || !d_cand || !d_counts) {
        frk::SetError("frk_score_select: null argument");
        return FRK_ERR_INVALID;
    }
    if (n_queries == 0 || block_count == 0) return FRK_OK;
    const uint32_t cd = frk_chunk_dims(store->dtype);
    if (cd == 0 || store->dim == 0 || store->dim % cd != 0) {
        frk::SetError("frk_score_select: dim %u is not a multiple of %u for dtype %u", store->dim, cd, store->dtype);
        return FRK_ERR_INVALID;
    }
    if (store->dtype == FRK_DTYPE_FP8E4M3 && !store->scales) {
        frk::SetError("frk_score_select: FP8 store needs per-doc scales");
        return FRK_ERR_INVALID;
    }
    const uint64_t n_blocks = (store->n_docs + 31) / 32;
    const uint64_t last = block_begin + (block_count - 1) * block_step;
    if (block_step == 0 || last >= n_blocks) {
        frk::SetError("frk_score_select: block range [%llu + i*%llu, i < %llu) exceeds %llu blocks",
                      (unsigned long long)block_begin, (unsigned long long)block_step,
                      (unsigned long long)block_count, (unsigned long long)n_blocks);
        return FRK_ERR_INVALID;
    }
    if (d_masks && mask_stride < (store->n_docs + 63) / 64) {
        frk::SetError("frk_score_select: mask stride %llu words < %llu", (unsigned long long)mask_stride,
                      (unsigned long long)((store->n_docs + 63) / 64));
        return FRK_ERR_CAPACITY;
    }
    if (n_queries > 65535u * 32u) {
        frk::SetError("frk_score_select: too many queries (%u)", n_queries);
        return FRK_ERR_UNSUPPORTED;
    }
    frk::ScoreArgs A;
    A.tiles = static_cast<const frk::u32x4*>(store->tiles);
    A.doc_scale = store->scales;
    A.n_docs = store->n_docs;
    A.queries = d_queries;
    A.n_queries = n_queries;
    A.dim = store->dim;
    A.q_dim = store->dim;
    A.masks = d_masks;
    A.mask_stride = mask_stride;
    A.floor = d_floor;
    A.block_begin = block_begin;
    A.block_step = block_step;
    A.block_count = block_count;
    A.cand = d_cand;
    A.cap = cand_cap;
    A.counts = d_counts;
    const uint32_t ch = store->dim / (cd / 2) / 2;  // 16-byte chunks per doc
    // enough workgroups to fill 256 CUs several times, at most one per 4 blocks (one per wave)
    uint64_t want = (block_count + 15) / 16;
    uint32_t grid_x = static_cast<uint32_t>(want < 2048 ? (want ? want : 1) : 2048);
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (store->dtype) {
        case FRK_DTYPE_F32: return frk::Dispatch<FRK_DTYPE_F32>(A, ch, s, grid_x);
        case FRK_DTYPE_BF16: return frk::Dispatch<FRK_DTYPE_BF16>(A, ch, s, grid_x);
        case FRK_DTYPE_FP8E4M3:
            if (ch % 2 != 0) {
                frk::SetError("frk_score_select: FP8 needs an even number of chunks (dim multiple of 64)");
                return FRK_ERR_UNSUPPORTED;
            }
            return frk::Dispatch<FRK_DTYPE_FP8E4M3>(A, ch, s, grid_x);
    }
    frk::SetError("frk_score_select: unknown dtype %u", store->dtype);
    return FRK_ERR_INVALID;
}
