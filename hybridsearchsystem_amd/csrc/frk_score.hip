// frk_score.hip -- C ABI of the fused scorer: argument checks, launch shape, dtype dispatch. The
// kernel is in frk_score_impl.h, instantiated per dtype in frk_score_{f32,bf16,fp8,f16x2}.hip.
#include <atomic>
#include <string>

#include "frk_score_impl.h"

namespace frk {
namespace {

// The mask word the scorer reads for every (block, query) when no filter masks are given.
__device__ uint32_t kAllPassMask[1] = {0xFFFFFFFFu};

int AllPassMask(const uint32_t** out) {
    constexpr int kMaxDevices = 64;
    static std::atomic<const uint32_t*> addr[kMaxDevices] = {};  // racing threads store the same address
    int dev = 0;
    FRK_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        SetError("device %d out of range", dev);
        return FRK_ERR_UNSUPPORTED;
    }
    const uint32_t* a = addr[dev].load(std::memory_order_relaxed);
    if (!a) {
        void* p = nullptr;
        FRK_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(kAllPassMask)));
        a = static_cast<const uint32_t*>(p);
        addr[dev].store(a, std::memory_order_relaxed);
    }
    *out = a;
    return FRK_OK;
}

int Prepare(const frk_doc_store* store, const float* d_queries, uint32_t n_queries, const uint32_t* d_bmask,
            uint32_t bmask_stride, const float* d_floor, uint64_t block_begin, uint64_t block_step,
            uint64_t block_count, uint64_t* d_out, uint32_t cap, uint32_t* d_counts, uint32_t counts_stride,
            frk_stream_t stream, const char* who, ScoreLaunch& L) {
    if (!store || !store->tiles || !d_queries || !d_out || (!d_counts && std::string(who) == "frk_score_select")) {
        SetError("%s: null argument", who);
        return FRK_ERR_INVALID;
    }
    const uint32_t cd = frk_chunk_dims(store->dtype);
    if (cd == 0 || store->dim == 0 || store->dim % cd != 0) {
        SetError("%s: dim %u is not a multiple of %u for dtype %u", who, store->dim, cd, store->dtype);
        return FRK_ERR_INVALID;
    }
    if ((reinterpret_cast<uintptr_t>(d_queries) & 15) != 0) {
        SetError("%s: the query matrix must be 16-byte aligned", who);
        return FRK_ERR_INVALID;
    }
    const uint32_t ch = store->dim / cd;
    if (store->dtype == FRK_DTYPE_FP8E4M3 && (!store->scales || ch % 2 != 0)) {
        SetError("%s: FP8 needs per-doc scales and a dim that is a multiple of 64", who);
        return FRK_ERR_INVALID;
    }
    if (store->dtype == FRK_DTYPE_F16X2 && (!store->scales || ch % 2 != 0)) {
        SetError("%s: F16X2 needs per-doc scales and a dim that is a multiple of 16", who);
        return FRK_ERR_INVALID;
    }
    const uint64_t n_blocks = (store->n_docs + 31) / 32;
    const uint64_t last = block_begin + (block_count - 1) * block_step;
    if (block_step == 0 || last >= n_blocks) {
        SetError("%s: block range [%llu + i*%llu, i < %llu) exceeds %llu blocks", who, (unsigned long long)block_begin,
                 (unsigned long long)block_step, (unsigned long long)block_count, (unsigned long long)n_blocks);
        return FRK_ERR_INVALID;
    }
    if (d_bmask && (bmask_stride < n_queries || bmask_stride % 32 != 0)) {
        SetError("%s: filter word stride %u must be a multiple of 32 and >= %u queries", who, bmask_stride, n_queries);
        return FRK_ERR_CAPACITY;
    }
    if (n_queries > 65535u * 32u) {
        SetError("%s: too many queries (%u)", who, n_queries);
        return FRK_ERR_UNSUPPORTED;
    }
    ScoreArgs& A = L.args;
    A.tiles = static_cast<const u32x4*>(store->tiles);
    A.doc_scale = store->scales;
    A.n_docs = store->n_docs;
    A.queries = d_queries;
    A.n_queries = n_queries;
    A.dim = store->dim;
    A.q_dim = store->dim;
    if (d_bmask) {
        A.bmask = d_bmask;
        A.bmask_stride = bmask_stride;
    } else {
        int rc = AllPassMask(&A.bmask);
        if (rc != FRK_OK) return rc;
        A.bmask_stride = 0;
    }
    A.floor = d_floor;
    A.block_begin = block_begin;
    A.block_step = block_step;
    A.block_count = block_count;
    A.cand = d_out;
    A.cap = cap;
    A.counts = d_counts;
    A.counts_stride = counts_stride;
    L.chunks = ch;
    // at most one workgroup per 16 blocks (4 per wave of a 4-wave workgroup; the launch lowers this
    // to the resident count)
    const uint64_t want = (block_count + 15) / 16;
    L.grid_x = static_cast<uint32_t>(want < 2048 ? (want ? want : 1) : 2048);
    L.stream = static_cast<hipStream_t>(stream);
    return FRK_OK;
}

int Run(const frk_doc_store* store, const ScoreLaunch& L, int mode) {
    switch (store->dtype) {
        case FRK_DTYPE_F32: return ScoreF32(L, mode);
        case FRK_DTYPE_BF16: return ScoreBf16(L, mode);
        case FRK_DTYPE_FP8E4M3: return ScoreFp8(L, mode);
        case FRK_DTYPE_F16X2: return ScoreF16x2(L, mode);
    }
    SetError("unknown dtype %u", store->dtype);
    return FRK_ERR_INVALID;
}

// Query-major filter words -> block-major 32-bit masks, through a 64 x 64 LDS tile so both the reads
// (along a query's words) and the writes (along a block's queries) are coalesced.
// A workgroup of 256 threads transposes the filter words of 64 doc blocks (32 u64 words per query)
// for a tile of up to 128 queries (4 groups of 32):
//   1. the tile's query rows are loaded into LDS with coalesced 8-byte loads (two 256-byte rows per
//      wave instruction);
//   2. thread (group g = t / 64, block b = t % 64) reads the 32 rows of its group for block b (a wave
//      reads one row's 256 bytes at a time: no bank conflicts) and transposes that 32 x 32 bit matrix
//      in registers (five rounds of off-diagonal block swaps: 16-, 8-, 4-, 2- and 1-bit blocks);
//   3. the 32 per-doc words go back to LDS in output order (rows padded to 132 words, so 16 lanes of
//      16-byte stores cover all banks), then the workgroup writes each block's 128 g words with
//      contiguous 16-byte stores.
constexpr int kMtbBlocks = 64;               // doc blocks per workgroup
constexpr int kMtbGroups = 4;                // query groups of 32 per workgroup
constexpr int kMtbOutPitch = 32 * kMtbGroups + 4;  // u32 words per block row of the output stage

__global__ __launch_bounds__(256) void MasksToBlocksKernel(const uint64_t* __restrict__ masks, uint64_t stride,
                                                           uint32_t n_queries, uint64_t n_blocks,
                                                           uint32_t* __restrict__ out, uint32_t q_stride) {
    // one 33 KiB buffer: the query rows [query][word] (32 KiB) in phases 1-2, the output stage
    // [block][g * 32 + j] in phase 3 (so four workgroups share a CU)
    __shared__ __align__(16) uint32_t lds[kMtbBlocks * kMtbOutPitch];
    uint64_t (*rows)[kMtbBlocks / 2] = reinterpret_cast<uint64_t (*)[kMtbBlocks / 2]>(lds);
    uint32_t* stage = lds;
    const uint32_t t = threadIdx.x;
    const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kMtbBlocks;
    const uint64_t w0 = b0 / 2;
    const uint64_t n_words = (n_blocks + 1) / 2;
    const uint32_t q0 = blockIdx.y * 32 * kMtbGroups;
    const uint32_t groups = min(static_cast<uint32_t>(kMtbGroups), (q_stride - q0) / 32);
    for (uint32_t e = t; e < 32 * groups * (kMtbBlocks / 2); e += 256) {
        const uint32_t r = e / (kMtbBlocks / 2), wi = e % (kMtbBlocks / 2);
        const uint32_t q = q0 + r;
        const uint64_t w = w0 + wi;
        rows[r][wi] = (q < n_queries && w < n_words) ? masks[static_cast<uint64_t>(q) * stride + w] : 0ull;
    }
    __syncthreads();
    const uint32_t g = t / kMtbBlocks, b = t % kMtbBlocks;
    uint32_t w[32];
    if (g < groups) {
#pragma unroll
        for (int x = 0; x < 32; x++) w[x] = static_cast<uint32_t>(rows[32 * g + x][b / 2] >> (32 * (b & 1)));
        // after the round of block size s, rows k and k + s (k with bit s clear) have traded the high
        // s-bit fields of row k for the low s-bit fields of row k + s in every 2s-bit group
#pragma unroll
        for (int s = 16; s >= 1; s >>= 1) {
            const uint32_t m = s == 16 ? 0x0000FFFFu : s == 8 ? 0x00FF00FFu : s == 4 ? 0x0F0F0F0Fu : s == 2 ? 0x33333333u : 0x55555555u;
#pragma unroll
            for (int k = 0; k < 32; k++) {
                if (k & s) continue;
                const uint32_t tt = ((w[k] >> s) ^ w[k + s]) & m;
                w[k] ^= tt << s;
                w[k + s] ^= tt;
            }
        }
    }
    __syncthreads();  // the rows are read: the buffer becomes the output stage
    if (g < groups) {
        u32x4* dst = reinterpret_cast<u32x4*>(stage + b * kMtbOutPitch + 32 * g);
#pragma unroll
        for (int i = 0; i < 8; i++) dst[i] = u32x4{w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]};
    }
    __syncthreads();
    const uint32_t pieces = 8 * groups;  // 16-byte pieces per block row
    for (uint32_t e = t; e < kMtbBlocks * pieces; e += 256) {
        const uint32_t bl = e / pieces, pc = e % pieces;
        if (b0 + bl >= n_blocks) break;
        const u32x4 v = *reinterpret_cast<const u32x4*>(stage + bl * kMtbOutPitch + 4 * pc);
        *reinterpret_cast<u32x4*>(out + (b0 + bl) * q_stride + q0 + 4 * pc) = v;
    }
}

}  // namespace
}  // namespace frk

extern "C" int frk_masks_to_blocks(const uint64_t* d_masks, uint64_t mask_stride, uint32_t n_queries, uint64_t n_docs,
                                   uint32_t* d_block_masks, uint32_t q_stride, frk_stream_t stream) {
    if (!d_masks || !d_block_masks || q_stride < n_queries || q_stride % 32 != 0 ||
        (reinterpret_cast<uintptr_t>(d_block_masks) & 15) != 0) {
        frk::SetError("frk_masks_to_blocks: bad arguments (q_stride %u must be a multiple of 32 and >= %u queries; "
                      "the output 16-byte aligned)", q_stride, n_queries);
        return FRK_ERR_INVALID;
    }
    const uint64_t n_blocks = (n_docs + 31) / 32;
    if (n_blocks == 0 || q_stride == 0) return FRK_OK;
    if (q_stride / 32 / frk::kMtbGroups > 65535u || (n_blocks + frk::kMtbBlocks - 1) / frk::kMtbBlocks > 0x7FFFFFFFu) {
        frk::SetError("frk_masks_to_blocks: %u queries exceed the grid", q_stride);
        return FRK_ERR_UNSUPPORTED;
    }
    const uint32_t tiles = (q_stride / 32 + frk::kMtbGroups - 1) / frk::kMtbGroups;
    dim3 grid(static_cast<uint32_t>((n_blocks + frk::kMtbBlocks - 1) / frk::kMtbBlocks), tiles);
    hipLaunchKernelGGL(frk::MasksToBlocksKernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), d_masks,
                       mask_stride, n_queries, n_blocks, d_block_masks, q_stride);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
extern "C" uint32_t frk_chunk_dims(uint32_t dtype) {
    switch (dtype) {
        case FRK_DTYPE_F32: return 8;
        case FRK_DTYPE_BF16: return 16;
        case FRK_DTYPE_FP8E4M3: return 32;
        case FRK_DTYPE_F16X2: return 8;  // a chunk pair (hi, lo) holds 16 dims
    }
    return 0;
}

extern "C" uint32_t frk_store_dim(uint32_t dim, uint32_t dtype) {
    const uint32_t cd = frk_chunk_dims(dtype);
    if (cd == 0 || dim == 0) return 0;
    // one FP8 MFMA step reads two chunks, and an F16X2 slice is a chunk pair
    const uint32_t step = dtype == FRK_DTYPE_FP8E4M3 || dtype == FRK_DTYPE_F16X2 ? 2 : 1;
    uint32_t ch = (dim + cd - 1) / cd;
    ch = (ch + step - 1) / step * step;
    for (uint32_t s : frk::kScoreShapes)
        if (s >= ch) return s * cd;
    return ch * cd;  // beyond the largest scorer shape: the rescorer reads such a store, the scorer refuses it
}

extern "C" int frk_score_select(const frk_doc_store* store, const float* d_queries, uint32_t n_queries,
                                const uint32_t* d_block_masks, uint32_t mask_q_stride, const float* d_floor,
                                uint64_t block_begin, uint64_t block_step, uint64_t block_count, uint64_t* d_cand,
                                uint32_t cand_cap, uint32_t* d_counts, uint32_t counts_stride,
                                frk_stream_t stream) {
    if (n_queries == 0 || block_count == 0) return FRK_OK;
    if (counts_stride == 0) {
        frk::SetError("frk_score_select: counts_stride must be >= 1");
        return FRK_ERR_INVALID;
    }
    frk::ScoreLaunch L;
    int rc = frk::Prepare(store, d_queries, n_queries, d_block_masks, mask_q_stride, d_floor, block_begin,
                          block_step, block_count, d_cand, cand_cap, d_counts, counts_stride, stream,
                          "frk_score_select", L);
    return rc ? rc : frk::Run(store, L, frk::kModeSelect);
}

extern "C" int frk_score_blockmax(const frk_doc_store* store, const float* d_queries, uint32_t n_queries,
                                  const uint32_t* d_block_masks, uint32_t mask_q_stride, uint64_t block_begin,
                                  uint64_t block_step, uint64_t block_count, float* d_blockmax, uint32_t stride,
                                  frk_stream_t stream) {
    if (n_queries == 0 || block_count == 0) return FRK_OK;
    if (stride < block_count) {
        frk::SetError("frk_score_blockmax: stride %u < %llu sampled blocks", stride, (unsigned long long)block_count);
        return FRK_ERR_CAPACITY;
    }
    frk::ScoreLaunch L;
    int rc = frk::Prepare(store, d_queries, n_queries, d_block_masks, mask_q_stride, nullptr, block_begin,
                          block_step, block_count, reinterpret_cast<uint64_t*>(d_blockmax), stride, nullptr, 1,
                          stream, "frk_score_blockmax", L);
    return rc ? rc : frk::Run(store, L, frk::kModeBlockMax);
}

extern "C" int frk_score_matrix(const frk_doc_store* store, const float* d_queries, uint32_t n_queries, float* d_scores,
                                uint64_t score_stride, frk_stream_t stream) {
    if (!store || n_queries == 0 || store->n_docs == 0) return FRK_OK;
    if (score_stride < store->n_docs || score_stride > 0xFFFFFFFFull) {
        frk::SetError("frk_score_matrix: row stride %llu must hold %llu docs and fit 32 bits",
                      (unsigned long long)score_stride, (unsigned long long)store->n_docs);
        return FRK_ERR_CAPACITY;
    }
    const uint64_t n_blocks = (store->n_docs + 31) / 32;
    frk::ScoreLaunch L;
    int rc = frk::Prepare(store, d_queries, n_queries, nullptr, 0, nullptr, 0, 1, n_blocks,
                          reinterpret_cast<uint64_t*>(d_scores), static_cast<uint32_t>(score_stride), nullptr, 1, stream,
                          "frk_score_matrix", L);
    return rc ? rc : frk::Run(store, L, frk::kModeDense);
}
