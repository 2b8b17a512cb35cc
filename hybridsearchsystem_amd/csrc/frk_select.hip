// frk_select.hip -- exact top-k over candidate keys, one workgroup per query.
//
// The reference pre-filters with a single-core sampled-pivot partition (device/topk/kernel_topk.h:30-388),
// returns a superset of the top-k and finishes with a host nth_element + sort
// (searcher/stage/result_aggregator.cpp:993-1000). Here the whole selection stays on the device:
// an MSB-first radix select over the 64-bit keys (8 passes of 8-bit digits, histograms in LDS) finds
// the k-th largest key T exactly -- keys are unique because the low word is the doc id -- then the k
// keys >= T are gathered into LDS and bitonic-sorted descending. Because the key orders by score and
// then by ascending doc id, the result is the top-k under the reference CPU truth's ordering
// (10.13/load/src/harness/recall/recall_ref.cpp:19-25).
#include <math.h>

#include "frk_common.h"

namespace frk {
namespace {

constexpr int kThreads = 1024;

__device__ inline void BitonicSortDesc(uint64_t* s, uint32_t n_pow2) {
    for (uint32_t size = 2; size <= n_pow2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < n_pow2 / 2; t += kThreads) {
                const uint32_t i = 2 * t - (t & (stride - 1));
                const uint32_t j = i + stride;
                const bool desc = (i & size) == 0;
                const uint64_t a = s[i], b = s[j];
                if ((a < b) == desc) {
                    s[i] = b;
                    s[j] = a;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kThreads) void SelectKernel(const uint64_t* __restrict__ cand, uint32_t cap,
                                                         const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ ks, uint32_t k_default,
                                                         uint64_t* __restrict__ out, uint32_t k_stride,
                                                         uint32_t* __restrict__ out_counts, float* __restrict__ kth) {
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_remaining;
    __shared__ uint32_t s_fill;
    __shared__ uint64_t sorted[FRK_SELECT_MAX_K];

    const uint32_t q = blockIdx.x;
    const uint32_t n = counts[q] < cap ? counts[q] : cap;
    const uint32_t kq = ks ? (ks[q] < k_stride ? ks[q] : k_stride) : k_default;
    const uint32_t k = kq < n ? kq : n;
    const uint64_t* c = cand + static_cast<uint64_t>(q) * cap;
    if (k == 0) {
        if (threadIdx.x == 0) {
            out_counts[q] = 0;
            if (kth) kth[q] = -INFINITY;
        }
        return;
    }
    if (threadIdx.x == 0) {
        s_prefix = 0;
        s_remaining = k;
        s_fill = 0;
    }
    uint64_t prefix_mask = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += kThreads) hist[b] = 0;
        __syncthreads();
        const uint64_t prefix = s_prefix;
        for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
            const uint64_t key = c[i];
            if ((key & prefix_mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t rem = s_remaining, cum = 0;
            int d = 255;
            for (; d > 0; d--) {
                if (cum + hist[d] >= rem) break;
                cum += hist[d];
            }
            s_remaining = rem - cum;
            s_prefix = prefix | (static_cast<uint64_t>(d) << shift);
        }
        prefix_mask |= 0xFFull << shift;
        __syncthreads();
    }
    const uint64_t T = s_prefix;  // the k-th largest key
    for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
        const uint64_t key = c[i];
        if (key >= T) {
            const uint32_t at = atomicAdd(&s_fill, 1u);
            if (at < FRK_SELECT_MAX_K) sorted[at] = key;
        }
    }
    __syncthreads();
    uint32_t p2 = 1;
    while (p2 < k) p2 <<= 1;
    for (uint32_t i = k + threadIdx.x; i < p2; i += kThreads) sorted[i] = 0;
    __syncthreads();
    BitonicSortDesc(sorted, p2);
    uint64_t* o = out + static_cast<uint64_t>(q) * k_stride;
    for (uint32_t i = threadIdx.x; i < k; i += kThreads) o[i] = sorted[i];
    if (threadIdx.x == 0) {
        out_counts[q] = k;
        if (kth) kth[q] = (kq <= n) ? KeyScore(T) : -INFINITY;
    }
}

// k-th largest of n floats per query (radix select over order-preserving u32 keys, 4 passes).
__global__ __launch_bounds__(kThreads) void KthLargestKernel(const float* __restrict__ vals, uint32_t stride,
                                                             uint32_t n, const uint32_t* __restrict__ ks,
                                                             uint32_t k_default, float* __restrict__ out) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_prefix, s_remaining;
    const uint32_t q = blockIdx.x;
    const uint32_t k = ks ? ks[q] : k_default;
    const float* v = vals + static_cast<uint64_t>(q) * stride;
    if (k == 0 || k > n) {
        if (threadIdx.x == 0) out[q] = -INFINITY;
        return;
    }
    if (threadIdx.x == 0) {
        s_prefix = 0;
        s_remaining = k;
    }
    uint32_t mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += kThreads) hist[b] = 0;
        __syncthreads();
        const uint32_t prefix = s_prefix;
        for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
            const uint32_t key = OrderedFromFloat(v[i]);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t rem = s_remaining, cum = 0;
            int d = 255;
            for (; d > 0; d--) {
                if (cum + hist[d] >= rem) break;
                cum += hist[d];
            }
            s_remaining = rem - cum;
            s_prefix = prefix | (static_cast<uint32_t>(d) << shift);
        }
        mask |= 0xFFu << shift;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = FloatFromOrdered(s_prefix);
}

}  // namespace
}  // namespace frk

extern "C" int frk_kth_largest(const float* d_vals, uint32_t stride, uint32_t n, uint32_t n_queries,
                               const uint32_t* d_k, uint32_t k_default, float* d_out, frk_stream_t stream) {
    if (n_queries == 0) return FRK_OK;
    if (!d_vals || !d_out || stride < n) {
        frk::SetError("frk_kth_largest: bad arguments");
        return FRK_ERR_INVALID;
    }
    hipLaunchKernelGGL(frk::KthLargestKernel, dim3(n_queries), dim3(frk::kThreads), 0, static_cast<hipStream_t>(stream),
                       d_vals, stride, n, d_k, k_default, d_out);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_select_topk(const uint64_t* d_cand, uint32_t cand_cap, const uint32_t* d_counts,
                               uint32_t n_queries, const uint32_t* d_k, uint32_t k_default, uint64_t* d_out,
                               uint32_t k_stride, uint32_t* d_out_counts, float* d_kth_score, frk_stream_t stream) {
    if (n_queries == 0) return FRK_OK;
    if (!d_cand || !d_counts || !d_out || !d_out_counts) {
        frk::SetError("frk_select_topk: null device pointer");
        return FRK_ERR_INVALID;
    }
    if (k_stride > FRK_SELECT_MAX_K || (d_k == nullptr && k_default > k_stride)) {
        frk::SetError("frk_select_topk: k_stride %u / k %u beyond the supported %d", k_stride, k_default,
                      FRK_SELECT_MAX_K);
        return FRK_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(frk::SelectKernel, dim3(n_queries), dim3(frk::kThreads), 0, static_cast<hipStream_t>(stream),
                       d_cand, cand_cap, d_counts, d_k, k_default, d_out, k_stride, d_out_counts, d_kth_score);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
