// frk_aggregate.hip -- the unfused aggregator: survivors of one query's filter bitmap compacted out of a
// score row as candidate keys. The fused pipeline never needs it (the scorer compacts in its epilogue);
// it serves callers that keep the reference's unfused order (score matrix first, then aggregation per
// query: FullRecallSearcher::BatchScore / OneAggregator), and replaces the reference's Aggregator kernel
// (engine/retrieval/device/aggregator/kernel_aggregator.h:22-182), which gathers (score, doc) by the
// filter bits block by block per core.
//
// One workgroup per range of 64 bitmap words (4,096 docs): a first pass counts the range's passing docs
// from the bits alone (one global atomic per workgroup reserves their slots), a second pass has each
// wave take one 64-doc word at a time, ballot the passing lanes and write their keys at the range's base
// plus the wave's offset (an LDS counter) plus the lane's prefix count. Scores are read only for
// passing docs.
#include "frk_common.h"

namespace frk {
namespace {

constexpr int kAggThreads = 256;
constexpr int kAggWords = 64;  // bitmap words per workgroup

__device__ __forceinline__ uint64_t ValidBits(uint64_t w, uint64_t n_docs) {
    if (64 * (w + 1) <= n_docs) return ~0ull;
    if (64 * w < n_docs) return (1ull << (n_docs - 64 * w)) - 1;
    return 0;
}

__global__ __launch_bounds__(kAggThreads) void AggregateKernel(const float* __restrict__ scores,
                                                               const uint64_t* __restrict__ bits, uint64_t n_docs,
                                                               uint64_t* __restrict__ cand, uint32_t cap,
                                                               uint32_t* __restrict__ count) {
    __shared__ uint32_t wave_sum[kAggThreads / 64];
    __shared__ uint32_t base, fill;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint64_t n_words = (n_docs + 63) / 64;
    const uint64_t w_lo = static_cast<uint64_t>(blockIdx.x) * kAggWords;
    const uint64_t w_hi = w_lo + kAggWords < n_words ? w_lo + kAggWords : n_words;
    uint32_t c = 0;
    for (uint64_t w = w_lo + t; w < w_hi; w += kAggThreads) c += __popcll(bits[w] & ValidBits(w, n_docs));
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if (lane == 0) wave_sum[wave] = c;
    __syncthreads();
    if (t == 0) {
        uint32_t total = 0;
        for (int i = 0; i < kAggThreads / 64; i++) total += wave_sum[i];
        base = total ? atomicAdd(count, total) : 0;
        fill = 0;
    }
    __syncthreads();
    for (uint64_t w = w_lo + wave; w < w_hi; w += kAggThreads / 64) {
        const uint64_t word = bits[w] & ValidBits(w, n_docs);
        if (word == 0) continue;
        uint32_t off = 0;
        if (lane == 0) off = atomicAdd(&fill, static_cast<uint32_t>(__popcll(word)));
        off = __shfl(off, 0, 64);
        if ((word >> lane) & 1ull) {
            const uint32_t pos = base + off + static_cast<uint32_t>(__popcll(word & ((1ull << lane) - 1ull)));
            const uint64_t doc = 64 * w + lane;
            if (pos < cap) cand[pos] = MakeKey(scores[doc], static_cast<uint32_t>(doc));
        }
    }
}

}  // namespace
}  // namespace frk

extern "C" int frk_aggregate(const float* d_scores, const uint64_t* d_bits, uint64_t n_docs, uint64_t* d_cand,
                             uint32_t cap, uint32_t* d_count, frk_stream_t stream) {
    if (n_docs == 0) return FRK_OK;
    if (!d_scores || !d_bits || !d_count || (cap && !d_cand)) {
        frk::SetError("frk_aggregate: null device pointer");
        return FRK_ERR_INVALID;
    }
    if (n_docs > 0xFFFFFFFFull) {
        frk::SetError("frk_aggregate: %llu docs exceed 32-bit doc ids", (unsigned long long)n_docs);
        return FRK_ERR_UNSUPPORTED;
    }
    const uint64_t n_words = (n_docs + 63) / 64;
    const uint64_t groups = (n_words + frk::kAggWords - 1) / frk::kAggWords;
    hipLaunchKernelGGL(frk::AggregateKernel, dim3(static_cast<uint32_t>(groups)), dim3(frk::kAggThreads), 0,
                       static_cast<hipStream_t>(stream), d_scores, d_bits, n_docs, d_cand, cap, d_count);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
