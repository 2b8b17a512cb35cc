// frk_tile.hip -- device-side tiling of row-major FP32 vectors into a doc store (frk_capi.h, "Doc
// store"), with BF16 / FP8 E4M3 (+ per-doc scale) conversion or the F16X2 hi/lo split (+ per-doc scale). The index loader and builder stream
// rows through it in chunks, so the FP32 originals never need a second resident copy.
#include "frk_common.h"

namespace frk {
namespace {

__global__ __launch_bounds__(256) void RowScaleKernel(const float* __restrict__ rows, uint64_t n, uint32_t dim,
                                                      uint32_t dtype, float* __restrict__ scales) {
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n) return;
    float mx = 0.0f;
    for (uint32_t d = lane; d < dim; d += 64) mx = fmaxf(mx, fabsf(rows[r * dim + d]));
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if (lane == 0) scales[r] = dtype == FRK_DTYPE_F16X2 ? SplitScale(mx) : Fp8Scale(mx);
}

__global__ __launch_bounds__(256) void TileKernel(const float* __restrict__ rows, uint64_t n, uint32_t dim,
                                                  uint64_t first, uint32_t dtype, uint32_t chunks,
                                                  uint8_t* __restrict__ tiles, const float* __restrict__ scales,
                                                  uint64_t n_units) {
    const uint64_t u = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (u >= n_units) return;
    const uint32_t E = dtype == FRK_DTYPE_F32 ? 4 : (dtype == FRK_DTYPE_FP8E4M3 ? 16 : 8);
    const uint64_t b = first / 32 + u / (static_cast<uint64_t>(chunks) * 64);
    const uint32_t c = static_cast<uint32_t>((u / 64) % chunks);
    const uint32_t l = static_cast<uint32_t>(u % 64);
    const uint64_t doc = b * 32 + (l & 31);
    // F16X2: chunk pair c / 2 covers 16 dims, the hi halves in the even chunk and the lo halves in the odd
    const uint32_t d0 = dtype == FRK_DTYPE_F16X2 ? (c / 2) * 16 + (l >> 5) * 8 : c * 2 * E + (l >> 5) * E;
    float v[16];
    const bool have = doc >= first && doc < first + n;
    const float* row = rows + (doc - first) * dim;
    for (uint32_t e = 0; e < E; e++) v[e] = (have && d0 + e < dim) ? row[d0 + e] : 0.0f;
    uint32_t w[4];
    if (dtype == FRK_DTYPE_F32) {
        for (int e = 0; e < 4; e++) w[e] = __builtin_bit_cast(uint32_t, v[e]);
    } else if (dtype == FRK_DTYPE_BF16) {
        for (int e = 0; e < 4; e++)
            w[e] = static_cast<uint32_t>(F32ToBf16(v[2 * e])) | (static_cast<uint32_t>(F32ToBf16(v[2 * e + 1])) << 16);
    } else if (dtype == FRK_DTYPE_F16X2) {
        const float s = have ? scales[doc - first] : 1.0f;
        uint16_t half[8];
        for (int e = 0; e < 8; e++) {
            uint16_t hi, lo;
            SplitF16(v[e] * s, hi, lo);
            half[e] = (c & 1) ? lo : hi;
        }
        for (int e = 0; e < 4; e++) w[e] = static_cast<uint32_t>(half[2 * e]) | (static_cast<uint32_t>(half[2 * e + 1]) << 16);
    } else {
        const float inv = have ? 1.0f / scales[doc - first] : 0.0f;
        for (int e = 0; e < 4; e++) {
            uint32_t x = 0;
            for (int k = 0; k < 4; k++) x |= static_cast<uint32_t>(F32ToE4M3(v[4 * e + k] * inv)) << (8 * k);
            w[e] = x;
        }
    }
    uint4 out = {w[0], w[1], w[2], w[3]};
    reinterpret_cast<uint4*>(tiles)[(b * chunks + c) * 64 + l] = out;
}

}  // namespace
}  // namespace frk

extern "C" int frk_tile_rows_device(const float* d_rows, uint64_t n_rows, uint32_t dim, uint64_t first_doc,
                                    uint32_t dtype, uint32_t dim_pad, void* d_tiles, float* d_scales,
                                    frk_stream_t stream) {
    const uint32_t cd = frk_chunk_dims(dtype);
    if (!d_rows || !d_tiles || cd == 0 || dim == 0 || dim_pad < dim || dim_pad % cd != 0 || first_doc % 32 != 0) {
        frk::SetError("frk_tile_rows_device: bad arguments (dim %u, dim_pad %u, dtype %u, first %llu)", dim, dim_pad,
                      dtype, (unsigned long long)first_doc);
        return FRK_ERR_INVALID;
    }
    if (n_rows == 0) return FRK_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool scaled = dtype == FRK_DTYPE_FP8E4M3 || dtype == FRK_DTYPE_F16X2;
    if (scaled) {
        if (!d_scales) {
            frk::SetError("frk_tile_rows_device: FP8 and F16X2 need a scale array");
            return FRK_ERR_INVALID;
        }
        hipLaunchKernelGGL(frk::RowScaleKernel, dim3((n_rows + 3) / 4), dim3(256), 0, s, d_rows, n_rows, dim, dtype,
                           d_scales + first_doc);
        FRK_HIP_CHECK(hipGetLastError());
    }
    const uint32_t chunks = dim_pad / cd;
    const uint64_t n_blocks = (n_rows + 31) / 32;
    const uint64_t n_units = n_blocks * chunks * 64;
    hipLaunchKernelGGL(frk::TileKernel, dim3((n_units + 255) / 256), dim3(256), 0, s, d_rows, n_rows, dim, first_doc,
                       dtype, chunks, static_cast<uint8_t*>(d_tiles),
                       scaled ? d_scales + first_doc : nullptr, n_units);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

namespace frk {
namespace {
__global__ void FillU32Kernel(uint32_t* __restrict__ d, uint64_t n, uint32_t value) {
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        d[i] = value;
}
}  // namespace
}  // namespace frk

extern "C" int frk_fill_u32(uint32_t* d, uint64_t n, uint32_t value, frk_stream_t stream) {
    if (n == 0) return FRK_OK;
    if (!d) {
        frk::SetError("frk_fill_u32: null pointer");
        return FRK_ERR_INVALID;
    }
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(frk::FillU32Kernel, dim3(static_cast<uint32_t>(blocks < 1024 ? blocks : 1024)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), d, n, value);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
