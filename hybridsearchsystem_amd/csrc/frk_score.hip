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

// 32 x 32 bit transpose of w[0..31] in registers (five rounds of off-diagonal block swaps).
__device__ __forceinline__ void Transpose32(uint32_t (&w)[32]) {
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

// Query-major filter words -> block-major 32-bit masks. Thread (query group g, block b) loads the 32-bit
// word of block b from each of its group's 32 query rows (a wave instruction reads 64 consecutive blocks
// of one row: 256 contiguous bytes; the 32 loads are independent and all in flight together) and
// transposes the 32 x 32 bit matrix in registers; the block's 32 per-doc words are one 128-byte line of
// the output. The wave's 64 lines (one per block, q_stride words apart) go out through a wave-private
// LDS buffer, half of them at a time, so that each 16-byte store instruction writes 8 whole lines (8
// lanes per line) rather than a 16-byte piece of 64 lines. Within a line the 8 pieces sit rotated by
// the line index, which keeps both the writes (a lane: its line's 8 pieces) and the reads (8 lanes: one
// line) free of LDS bank conflicts. Measured at 10M docs: 128 queries 0.085 -> 0.065 ms, 64 queries
// 0.046 -> 0.037 ms against each lane storing its own line (a 64-line buffer, twice the LDS and so
// fewer waves per SIMD: 0.067 / 0.047 ms); staging 64 blocks x 128 queries through a workgroup-wide LDS
// tile behind two barriers took 0.107 ms, and wider loads (2 or 4 blocks per thread) 0.13-0.14 ms.
__global__ __launch_bounds__(256) void MasksToBlocksKernel(const uint32_t* __restrict__ masks32, uint64_t stride32,
                                                           uint32_t n_queries, uint64_t n_blocks,
                                                           uint32_t* __restrict__ out, uint32_t q_stride) {
    __shared__ u32x4 lines[4][32 * 8];  // per wave: 32 lines x 8 pieces of 16 bytes
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63, wv = t >> 6;
    const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * 64;
    const uint64_t b = b0 + lane;
    const uint32_t g = blockIdx.y * 4 + wv;
    if (32 * g >= q_stride) return;  // uniform over the wave
    uint32_t w[32];
#pragma unroll
    for (int x = 0; x < 32; x++) {
        const uint32_t q = 32 * g + x;
        w[x] = q < n_queries && b < n_blocks ? masks32[static_cast<uint64_t>(q) * stride32 + b] : 0u;
    }
    Transpose32(w);
    u32x4* L = lines[wv];
    const uint32_t p = lane & 7;
    // two halves of 32 lines: lanes 32 h .. 32 h + 31 put their lines in the buffer, then store i of the
    // half has lanes 8 j .. 8 j + 7 write the 8 pieces of line 32 h + 8 i + j
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if ((lane >> 5) == static_cast<uint32_t>(h)) {
            const uint32_t r = lane & 31;
#pragma unroll
            for (int c = 0; c < 8; c++)
                L[r * 8 + ((c + r) & 7)] = u32x4{w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]};
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t k = 8 * i + (lane >> 3);
            const u32x4 v = L[k * 8 + ((p + k) & 7)];
            const uint64_t bk = b0 + 32 * h + k;
            if (bk < n_blocks) reinterpret_cast<u32x4*>(out + bk * q_stride + 32 * g)[p] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
    if (q_stride / 32 > 4u * 65535u || (n_blocks + 63) / 64 > 0x7FFFFFFFu) {
        frk::SetError("frk_masks_to_blocks: %u queries exceed the grid", q_stride);
        return FRK_ERR_UNSUPPORTED;
    }
    // 64 blocks x 4 query groups per workgroup
    dim3 grid(static_cast<uint32_t>((n_blocks + 63) / 64), (q_stride / 32 + 3) / 4);
    hipLaunchKernelGGL(frk::MasksToBlocksKernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const uint32_t*>(d_masks), 2 * mask_stride, n_queries, n_blocks, d_block_masks,
                       q_stride);
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
