// frk_rescore.hip -- exact FP32 rescoring of candidate keys from the FP32 doc store, used after a
// low-precision (BF16 / FP8) scoring pass: the candidates' scores are recomputed from the original
// vectors and re-keyed, so the final top-k is ordered by FP32 scores.
#include "frk_common.h"

namespace frk {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// One wave per candidate. Lane l reads the 16-byte (chunk l / 2, half l % 2) piece of the doc's
// tile column: 4 dims, so a wave covers 256 dims per step.
__global__ __launch_bounds__(256) void RescoreKernel(const floatx4* __restrict__ tiles, uint32_t chunks,
                                                     uint64_t n_docs, const float* __restrict__ queries,
                                                     uint32_t q_dim,
                                                     const uint64_t* __restrict__ keys_in, uint32_t stride,
                                                     const uint32_t* __restrict__ counts,
                                                     uint64_t* __restrict__ keys_out) {
    const uint32_t q = blockIdx.x;
    const uint32_t i = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const uint32_t n = counts[q] < stride ? counts[q] : stride;
    if (i >= n) return;
    const uint64_t key = keys_in[static_cast<uint64_t>(q) * stride + i];
    const uint32_t doc = KeyDoc(key);
    if (doc >= n_docs) {  // not a doc of this store: the smallest key, which no selection prefers
        if (lane == 0) keys_out[static_cast<uint64_t>(q) * stride + i] = 0;
        return;
    }
    const uint64_t b = doc / 32, j = doc % 32;
    const float* qv = queries + static_cast<uint64_t>(q) * q_dim;
    float acc = 0.0f;
    for (uint32_t p = lane; p < 2 * chunks; p += 64) {
        const uint32_t c = p >> 1, h = p & 1;
        const floatx4 v = tiles[(b * chunks + c) * 64 + j + 32 * h];
        const uint32_t d0 = 8 * c + 4 * h;
        acc += v[0] * qv[d0] + v[1] * qv[d0 + 1] + v[2] * qv[d0 + 2] + v[3] * qv[d0 + 3];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) keys_out[static_cast<uint64_t>(q) * stride + i] = MakeKey(acc, doc);
}

// Rescoring from row-major rows (row stride dim, a multiple of 4): one wave per four candidates of a
// query. Lane l covers dims [4 l + 256 r, +4), so a candidate's row is read with contiguous 16-byte
// loads (1 KiB per wave instruction at 256 dims), and the four candidates' loads are all issued before
// any of them is summed.
__global__ __launch_bounds__(256) void RescoreRowsKernel(const float* __restrict__ rows, uint32_t dim, uint64_t n_docs,
                                                         const float* __restrict__ queries,
                                                         const uint64_t* __restrict__ keys_in, uint32_t stride,
                                                         const uint32_t* __restrict__ counts,
                                                         uint64_t* __restrict__ keys_out) {
    constexpr int kPer = 4;  // candidates per wave
    const uint32_t q = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t i0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * kPer;
    const uint32_t n = counts[q] < stride ? counts[q] : stride;
    if (i0 >= n) return;
    const float* qv = queries + static_cast<uint64_t>(q) * dim;
    uint32_t doc[kPer];
    bool ok[kPer];
#pragma unroll
    for (int c = 0; c < kPer; c++) {
        ok[c] = i0 + c < n;
        doc[c] = ok[c] ? KeyDoc(keys_in[static_cast<uint64_t>(q) * stride + i0 + c]) : 0u;
        ok[c] = ok[c] && doc[c] < n_docs;
    }
    float acc[kPer] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t d = 4 * lane; d < dim; d += 256) {
        const floatx4 qq = *reinterpret_cast<const floatx4*>(qv + d);
        floatx4 v[kPer];
#pragma unroll
        for (int c = 0; c < kPer; c++)
            v[c] = ok[c] ? *reinterpret_cast<const floatx4*>(rows + static_cast<uint64_t>(doc[c]) * dim + d)
                         : floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < kPer; c++) acc[c] += v[c][0] * qq[0] + v[c][1] * qq[1] + v[c][2] * qq[2] + v[c][3] * qq[3];
    }
#pragma unroll
    for (int c = 0; c < kPer; c++)
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
    // a key whose doc lies outside the table becomes 0, the smallest key, which no selection prefers
    // (so the count stays valid and no stale key of an earlier call is read back)
    if (lane < kPer && i0 + lane < n)
        keys_out[static_cast<uint64_t>(q) * stride + i0 + lane] = ok[lane] ? MakeKey(acc[lane], doc[lane]) : 0ull;
}

}  // namespace
}  // namespace frk

extern "C" int frk_rescore(const frk_doc_store* f32_store, const float* d_queries, uint32_t n_queries,
                           const uint64_t* d_keys_in, uint32_t stride, const uint32_t* d_counts,
                           uint64_t* d_keys_out, frk_stream_t stream) {
    if (!f32_store || f32_store->dtype != FRK_DTYPE_F32 || !d_queries || !d_keys_in || !d_counts || !d_keys_out) {
        frk::SetError("frk_rescore: needs the FP32 store and device buffers");
        return FRK_ERR_INVALID;
    }
    if (n_queries == 0 || stride == 0) return FRK_OK;
    const uint32_t chunks = f32_store->dim / 8;
    dim3 grid(n_queries, (stride + 3) / 4);
    if (grid.y > 65535) {
        frk::SetError("frk_rescore: stride %u too large", stride);
        return FRK_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(frk::RescoreKernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const frk::floatx4*>(f32_store->tiles), chunks, f32_store->n_docs, d_queries,
                       f32_store->dim,
                       d_keys_in, stride, d_counts, d_keys_out);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_rescore_rows(const float* d_rows, uint32_t dim, uint64_t n_docs, const float* d_queries,
                                uint32_t n_queries, const uint64_t* d_keys_in, uint32_t stride, const uint32_t* d_counts,
                                uint64_t* d_keys_out, frk_stream_t stream) {
    if (!d_rows || !d_queries || !d_keys_in || !d_counts || !d_keys_out || dim == 0 || dim % 4 != 0 ||
        (reinterpret_cast<uintptr_t>(d_rows) & 15) != 0 || (reinterpret_cast<uintptr_t>(d_queries) & 15) != 0) {
        frk::SetError("frk_rescore_rows: needs 16-byte aligned rows and queries with a dim that is a multiple of 4");
        return FRK_ERR_INVALID;
    }
    if (n_queries == 0 || stride == 0) return FRK_OK;
    dim3 grid(n_queries, (stride + 15) / 16);
    if (grid.y > 65535) {
        frk::SetError("frk_rescore_rows: stride %u too large", stride);
        return FRK_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(frk::RescoreRowsKernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), d_rows, dim,
                       n_docs, d_queries, d_keys_in, stride, d_counts, d_keys_out);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
