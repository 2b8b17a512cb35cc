// frk_probe.hip -- a streaming-read kernel that measures the HBM read bandwidth this GPU attains, so a
// kernel's achieved bandwidth can be set beside what a pure read reaches on the same device (next to the
// published peak). Not part of the search path.
//
// One resident wave of workgroups walks the buffer in 16-byte loads, four independent loads per thread
// in flight per step, folds what it read with XOR and writes one word per workgroup, so no load can be
// dropped as dead.
#include "frk_common.h"

namespace frk {
namespace {

typedef unsigned int u32x4_probe __attribute__((ext_vector_type(4)));
constexpr int kProbeThreads = 512;
constexpr int kProbeUnroll = 4;

__global__ __launch_bounds__(kProbeThreads) void ReadProbeKernel(const u32x4_probe* __restrict__ src, uint64_t n16,
                                                                 uint32_t* __restrict__ sink) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kProbeThreads;
    uint64_t i = static_cast<uint64_t>(blockIdx.x) * kProbeThreads + threadIdx.x;
    u32x4_probe acc = {0u, 0u, 0u, 0u};
    for (; i + (kProbeUnroll - 1) * stride < n16; i += kProbeUnroll * stride) {
        u32x4_probe v[kProbeUnroll];
#pragma unroll
        for (int u = 0; u < kProbeUnroll; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < kProbeUnroll; u++) acc ^= v[u];
    }
    for (; i < n16; i += stride) acc ^= src[i];
    uint32_t x = acc[0] ^ acc[1] ^ acc[2] ^ acc[3];
    for (int off = 32; off > 0; off >>= 1) x ^= __shfl_xor(x, off, 64);
    __shared__ uint32_t wave_x[kProbeThreads / 64];
    if ((threadIdx.x & 63) == 0) wave_x[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0;
        for (int w = 0; w < kProbeThreads / 64; w++) r ^= wave_x[w];
        sink[blockIdx.x] = r;
    }
}

}  // namespace
}  // namespace frk

extern "C" int frk_read_probe(const void* d_src, uint64_t bytes, uint32_t* d_sink, uint32_t sink_words,
                              frk_stream_t stream) {
    if (!d_src || !d_sink || bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(d_src) & 15) != 0) {
        frk::SetError("frk_read_probe: needs a 16-byte aligned buffer of a multiple of 16 bytes and a sink");
        return FRK_ERR_INVALID;
    }
    int cus = 0;
    if (int rc = frk::CurrentDeviceCUs(&cus)) return rc;
    uint32_t grid = static_cast<uint32_t>(cus) * 4;  // 2,048 threads per CU
    if (grid > sink_words) grid = sink_words;
    if (grid == 0) return FRK_OK;
    hipLaunchKernelGGL(frk::ReadProbeKernel, dim3(grid), dim3(frk::kProbeThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<const frk::u32x4_probe*>(d_src), bytes / 16, d_sink);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
