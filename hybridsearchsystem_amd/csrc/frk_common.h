// frk_common.h -- shared definitions for the HIP kernels and their C-ABI wrappers (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/frk_capi.h"

namespace frk {

// Thread-local last-error text, returned by frk_last_error().
void SetError(const char* fmt, ...);

// Compute units of the calling thread's current device, queried once per device and cached.
int CurrentDeviceCUs(int* cus);

// Maps a failed HIP call to FRK_ERR_HIP with the call site in the error text.
#define FRK_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::frk::SetError("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),        \
                            __FILE__, __LINE__);                                         \
            return FRK_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

// Sortable candidate key: the high word orders by score (larger score -> larger key) and the low
// word by doc id reversed (smaller id -> larger key), so the k largest keys are the top-k under
// "score descending, then doc id ascending".
__host__ __device__ inline uint32_t OrderedFromFloat(float f) {
    if (f == 0.0f) f = 0.0f;  // -0.0 and +0.0 compare equal as scores
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__host__ __device__ inline float FloatFromOrdered(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __builtin_bit_cast(float, u);
}

__host__ __device__ inline uint64_t MakeKey(float score, uint32_t doc) {
    return (static_cast<uint64_t>(OrderedFromFloat(score)) << 32) | static_cast<uint64_t>(0xFFFFFFFFu - doc);
}

__host__ __device__ inline uint32_t KeyDoc(uint64_t key) { return 0xFFFFFFFFu - static_cast<uint32_t>(key); }

__host__ __device__ inline float KeyScore(uint64_t key) { return FloatFromOrdered(static_cast<uint32_t>(key >> 32)); }

// Round-to-nearest-even BF16 (NaN stays NaN).
__host__ __device__ inline uint16_t F32ToBf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7F800000u) == 0x7F800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}

// OCP FP8 E4M3 (e4m3fn: bias 7, max 448, no infinities), round to nearest even, saturating.
__host__ __device__ inline uint8_t F32ToE4M3(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    uint8_t sign = static_cast<uint8_t>((u >> 24) & 0x80u);
    float a = fabsf(f);
    if (!(a == a)) return static_cast<uint8_t>(sign | 0x7F);
    if (a >= 448.0f) return static_cast<uint8_t>(sign | 0x7E);
    if (a <= 0.0009765625f) return sign;  // at or below half the smallest subnormal (2^-10): ties to even 0
    int e;
    frexpf(a, &e);  // a = m * 2^e, m in [0.5, 1)
    int exp_unb = e - 1;  // a = 1.x * 2^exp_unb
    if (exp_unb < -6) {
        // subnormal: value = k * 2^-9, k in 0..7
        float k = a * 512.0f;
        float r = rintf(k);  // nearest even
        int ki = static_cast<int>(r);
        if (ki >= 8) return static_cast<uint8_t>(sign | 0x08);  // rounds up to the smallest normal
        return static_cast<uint8_t>(sign | ki);
    }
    float scaled = ldexpf(a, -exp_unb);  // [1, 2)
    float mant = rintf((scaled - 1.0f) * 8.0f);  // 3 mantissa bits, nearest even
    int m = static_cast<int>(mant);
    if (m == 8) {
        m = 0;
        exp_unb += 1;
    }
    int be = exp_unb + 7;
    if (be > 15 || (be == 15 && m > 6)) return static_cast<uint8_t>(sign | 0x7E);
    return static_cast<uint8_t>(sign | (be << 3) | m);
}

// BF16 bits -> FP32 (exact).
__host__ __device__ inline float Bf16ToF32(uint16_t h) { return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16); }

// OCP FP8 E4M3 (e4m3fn) bits -> FP32 (exact): bias 7, subnormals k * 2^-9, 0x7F / 0xFF NaN.
__host__ __device__ inline float E4M3ToF32(uint8_t b) {
    const float sign = (b & 0x80u) ? -1.0f : 1.0f;
    const int e = (b >> 3) & 0xF, m = b & 0x7;
    if (e == 15 && m == 7) return __builtin_nanf("");
    if (e == 0) return sign * ldexpf(static_cast<float>(m), -9);
    return sign * ldexpf(1.0f + m / 8.0f, e - 7);
}

// FP8 scale of a vector whose largest magnitude is mx: the smallest power of two s with mx / s <= 448
// (E4M3's largest finite value), clamped to [2^-126, 2^127]; 1 for a zero vector. A power of two is an
// E8M0 exponent, which the scaled FP8 MFMA applies per row and column by itself.
__host__ __device__ inline float Fp8Scale(float mx) {
    if (!(mx > 0.0f)) return 1.0f;
    if (!(mx <= 3.402823466e38f)) return ldexpf(1.0f, 127);
    int E;
    const float m = frexpf(mx, &E);  // mx = m * 2^E, m in [0.5, 1); 448 = 0.875 * 2^9
    int e = m <= 0.875f ? E - 9 : E - 8;
    e = e < -126 ? -126 : (e > 127 ? 127 : e);
    return ldexpf(1.0f, e);
}

// IEEE binary16, round to nearest even, subnormals kept, overflow to infinity, NaN stays NaN. One
// software definition for the host and the device, so stores tiled on either side are bit-identical.
__host__ __device__ inline uint16_t F32ToF16(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    const uint16_t sign = static_cast<uint16_t>((u >> 16) & 0x8000u);
    const uint32_t a = u & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return static_cast<uint16_t>(sign | 0x7C00u | (a > 0x7F800000u ? 0x200u : 0u));
    if (a >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);  // rounds to at least 65520: infinity
    if (a < 0x38800000u) {
        // below 2^-14: a subnormal half, k * 2^-24 with k = round(|f| * 2^24), ties to even
        if (a < 0x33000000u) return sign;  // at most 2^-25: ties to even 0
        const uint32_t m = (a & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - (a >> 23);  // 14..24 for exponents 2^-25 .. 2^-15
        const uint32_t k = m >> shift, rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        const uint32_t r = k + (rem > half || (rem == half && (k & 1u)));
        return static_cast<uint16_t>(sign | r);
    }
    // normal: rebias the exponent (127 -> 15), keep 10 mantissa bits, round to nearest even
    const uint32_t b = a - 0x38000000u;
    const uint32_t r = (b + 0xFFFu + ((b >> 13) & 1u)) >> 13;
    return static_cast<uint16_t>(sign | r);
}

// IEEE binary16 bits -> FP32 (exact).
__host__ __device__ inline float F16ToF32(uint16_t h) {
    const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    if (e == 0) {
        const float v = ldexpf(static_cast<float>(m), -24);
        return sign ? -v : v;
    }
    if (e == 31) return __builtin_bit_cast(float, sign | 0x7F800000u | (m << 13));
    return __builtin_bit_cast(float, sign | ((e + 112u) << 23) | (m << 13));
}

// F16X2 scale of a vector whose largest magnitude is mx: the power of two s = 2^e with mx * s in
// [2^14, 2^15), so hi and lo of every element keep FP16's 11-bit precision down to 2^-17 of mx (below
// that the halves are subnormal, with an absolute error under 2^-25 / s); e clamped to [-126, 126];
// 1 for a zero vector.
__host__ __device__ inline float SplitScale(float mx) {
    if (!(mx > 0.0f)) return 1.0f;
    if (!(mx <= 3.402823466e38f)) return ldexpf(1.0f, -126);
    int E;
    (void)frexpf(mx, &E);  // mx = m * 2^E, m in [0.5, 1): mx * 2^(15 - E) in [2^14, 2^15)
    int e = 15 - E;
    e = e < -126 ? -126 : (e > 126 ? 126 : e);
    return ldexpf(1.0f, e);
}

// The hi and lo halves of v (already scaled): v = F16ToF32(hi) + F16ToF32(lo) within 2^-11 |v - hi|.
__host__ __device__ inline void SplitF16(float v, uint16_t& hi, uint16_t& lo) {
    hi = F32ToF16(v);
    lo = F32ToF16(v - F16ToF32(hi));
}

}  // namespace frk
