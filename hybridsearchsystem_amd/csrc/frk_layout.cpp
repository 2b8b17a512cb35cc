// frk_layout.cpp -- host-side conversion of row-major FP32 vectors into the MFMA-native doc tiling
// (frk_capi.h, "Doc store"), with BF16 or FP8 E4M3 (+ per-doc FP32 scale) quantisation, or the
// F16X2 hi/lo split (+ per-doc scale). Used by the
// index loader (DataTable) and exported for callers that stage their own stores.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "frk_common.h"

namespace frk {
namespace {

uint32_t ElemsPerLane(uint32_t dtype) { return frk_chunk_dims(dtype) / 2; }

void TileRange(const float* src, uint64_t n_docs, uint32_t dim, uint32_t dtype, uint32_t dim_pad, uint8_t* dst,
               float* scales, uint64_t b_lo, uint64_t b_hi) {
    const uint32_t E = ElemsPerLane(dtype);
    const uint32_t chunks = dim_pad / (2 * E);
    std::vector<float> row(dim_pad);
    for (uint64_t b = b_lo; b < b_hi; b++) {
        for (uint32_t j = 0; j < 32; j++) {
            const uint64_t n = b * 32 + j;
            std::fill(row.begin(), row.end(), 0.0f);
            if (n < n_docs) memcpy(row.data(), src + n * dim, sizeof(float) * dim);
            float inv = 1.0f;
            if (dtype == FRK_DTYPE_FP8E4M3 || dtype == FRK_DTYPE_F16X2) {
                float mx = 0.0f;
                for (uint32_t d = 0; d < dim; d++) mx = std::max(mx, std::fabs(row[d]));
                const float sc = dtype == FRK_DTYPE_FP8E4M3 ? frk::Fp8Scale(mx) : frk::SplitScale(mx);
                if (n < n_docs) scales[n] = sc;
                inv = 1.0f / sc;
            }
            if (dtype == FRK_DTYPE_F16X2) {
                // chunk pair p: the hi halves (chunk 2p) and lo halves (chunk 2p + 1) of dims 16 p + 8 h + e
                const float s = 1.0f / inv;
                for (uint32_t c = 0; c < chunks; c++) {
                    for (uint32_t h = 0; h < 2; h++) {
                        uint16_t* o = reinterpret_cast<uint16_t*>(dst + ((b * chunks + c) * 64 + (j + 32 * h)) * 16);
                        const float* v = row.data() + (c / 2) * 16 + h * 8;
                        for (uint32_t e = 0; e < 8; e++) {
                            uint16_t hi, lo;
                            SplitF16(v[e] * s, hi, lo);
                            o[e] = (c % 2) ? lo : hi;
                        }
                    }
                }
                continue;
            }
            for (uint32_t c = 0; c < chunks; c++) {
                for (uint32_t h = 0; h < 2; h++) {
                    uint8_t* lane = dst + ((b * chunks + c) * 64 + (j + 32 * h)) * 16;
                    const float* v = row.data() + c * 2 * E + h * E;
                    if (dtype == FRK_DTYPE_F32) {
                        memcpy(lane, v, 16);
                    } else if (dtype == FRK_DTYPE_BF16) {
                        uint16_t* o = reinterpret_cast<uint16_t*>(lane);
                        for (uint32_t e = 0; e < 8; e++) o[e] = F32ToBf16(v[e]);
                    } else {
                        for (uint32_t e = 0; e < 16; e++) lane[e] = F32ToE4M3(v[e] * inv);
                    }
                }
            }
        }
    }
}

}  // namespace
}  // namespace frk

extern "C" uint64_t frk_tiles_bytes(uint64_t n_docs, uint32_t dim_pad, uint32_t dtype) {
    const uint32_t cd = frk_chunk_dims(dtype);
    if (cd == 0 || dim_pad % cd != 0) return 0;
    return ((n_docs + 31) / 32) * (dim_pad / cd) * 1024ull;
}

extern "C" int frk_tile_docs(const float* src, uint64_t n_docs, uint32_t dim, uint32_t dtype, uint32_t dim_pad,
                             void* dst_tiles, float* dst_scales, uint32_t n_threads) {
    const uint32_t cd = frk_chunk_dims(dtype);
    if (!src || !dst_tiles || cd == 0 || dim == 0 || dim_pad < dim || dim_pad % cd != 0) {
        frk::SetError("frk_tile_docs: bad arguments (dim %u, dim_pad %u, dtype %u)", dim, dim_pad, dtype);
        return FRK_ERR_INVALID;
    }
    if ((dtype == FRK_DTYPE_FP8E4M3 || dtype == FRK_DTYPE_F16X2) && !dst_scales) {
        frk::SetError("frk_tile_docs: FP8 and F16X2 need a scale array");
        return FRK_ERR_INVALID;
    }
    const uint64_t n_blocks = (n_docs + 31) / 32;
    uint32_t T = n_threads ? n_threads : std::max(1u, std::thread::hardware_concurrency());
    if (T > n_blocks) T = static_cast<uint32_t>(n_blocks ? n_blocks : 1);
    std::vector<std::thread> pool;
    const uint64_t per = (n_blocks + T - 1) / T;
    for (uint32_t t = 0; t < T; t++) {
        const uint64_t lo = t * per, hi = std::min(n_blocks, lo + per);
        if (lo >= hi) break;
        pool.emplace_back(frk::TileRange, src, n_docs, dim, dtype, dim_pad, static_cast<uint8_t*>(dst_tiles),
                          dst_scales, lo, hi);
    }
    for (auto& th : pool) th.join();
    return FRK_OK;
}
