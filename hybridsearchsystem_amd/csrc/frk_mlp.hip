// frk_mlp.hip -- BF16 training and inference kernels of the learned score floor: a 3-layer MLP
// (in -> 256 -> 256 -> 1, ReLU) that predicts a lower bound on a query's k-th best filtered score.
//
// Arithmetic: every GEMM rounds both operands to BF16 (round to nearest even) as it stages them into
// LDS and multiplies on MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation; activations, gradients
// and the FP32 master weights live in FP32 memory. The output head (one unit) and the reductions run
// on the vector ALUs with the same BF16 rounding of their inputs. No reference counterpart: the
// reference has no learned component.
#include <stdarg.h>
#include <stdio.h>

#include "frk_common.h"

namespace frk {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

constexpr int kTile = 64;   // C tile per workgroup (4 waves of 32 x 32)
constexpr int kKStep = 16;  // K per MFMA
constexpr int kKChunk = 128;  // K staged in LDS per round (a multiple of kKStep)
static_assert(kKChunk % kKStep == 0 && kTile * kKChunk % 256 == 0, "GEMM staging shape");

__device__ __forceinline__ float Bf16Round(float f) {
    return __builtin_bit_cast(float, static_cast<uint32_t>(F32ToBf16(f)) << 16);
}

// C[M][N] = op(A)[M][K] * op(B)[K][N] with the epilogue
//   FRK_MLP_EPI_NONE       C = acc
//   FRK_MLP_EPI_BIAS_RELU  C = acc + bias[n] (pre-activation), H = max(C, 0)
//   FRK_MLP_EPI_RELU_GRAD  C = Z[m][n] > 0 ? acc : 0
// op(A)(m, k) = TA ? A[k * lda + m] : A[m * lda + k]; op(B)(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n].
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void GemmKernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                  const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                                  int epi, const float* __restrict__ bias, const float* __restrict__ Z,
                                                  int ldz, float* __restrict__ H, int ldh) {
    // K is staged kKChunk at a time: every load of a round is issued before the first is awaited, so
    // a layer costs ceil(K / kKChunk) global round trips instead of one per MFMA step (the inference
    // GEMMs have few workgroups, and their time is those round trips). Rows are padded by 16 bytes so
    // the 32 rows one MFMA operand read touches start in different LDS banks.
    __shared__ __attribute__((aligned(16))) uint16_t As[kTile][kKChunk + 8];  // [row][k]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile][kKChunk + 8];  // [col][k]
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    constexpr int kPerThread = kTile * kKChunk / 256;
    floatx16_t acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += kKChunk) {
        float va[kPerThread], vb[kPerThread];
#pragma unroll
        for (int i = 0; i < kPerThread; i++) {
            const int idx = t + 256 * i;
            int r, kk;
            if (TA) { kk = idx / kTile; r = idx % kTile; } else { r = idx / kKChunk; kk = idx % kKChunk; }
            const int m = m0 + r, k = k0 + kk;
            va[i] = (m < M && k < K) ? (TA ? A[static_cast<int64_t>(k) * lda + m] : A[static_cast<int64_t>(m) * lda + k])
                                     : 0.0f;
            int c;
            if (TB) { c = idx / kKChunk; kk = idx % kKChunk; } else { kk = idx / kTile; c = idx % kTile; }
            const int n = n0 + c, kb = k0 + kk;
            vb[i] = (n < N && kb < K) ? (TB ? B[static_cast<int64_t>(n) * ldb + kb] : B[static_cast<int64_t>(kb) * ldb + n])
                                      : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < kPerThread; i++) {
            const int idx = t + 256 * i;
            int r, kk;
            if (TA) { kk = idx / kTile; r = idx % kTile; } else { r = idx / kKChunk; kk = idx % kKChunk; }
            As[r][kk] = F32ToBf16(va[i]);
            int c;
            if (TB) { c = idx / kKChunk; kk = idx % kKChunk; } else { kk = idx / kTile; c = idx % kTile; }
            Bs[c][kk] = F32ToBf16(vb[i]);
        }
        __syncthreads();
        // the same MFMA steps, in the same order, as one step of kKStep per round: K rounded up to kKStep
        for (int kk = 0; kk < kKChunk && k0 + kk < K; kk += kKStep) {
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&As[wm + (lane & 31)][kk + 8 * (lane >> 5)]);
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&Bs[wn + (lane & 31)][kk + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator element r of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = n0 + wn + (lane & 31);
        if (m >= M || n >= N) continue;
        float v = acc[r];
        if (epi == FRK_MLP_EPI_BIAS_RELU) {
            v += bias ? bias[n] : 0.0f;
            H[static_cast<int64_t>(m) * ldh + n] = v > 0.0f ? v : 0.0f;
        } else if (epi == FRK_MLP_EPI_RELU_GRAD) {
            v = Z[static_cast<int64_t>(m) * ldz + n] > 0.0f ? v : 0.0f;
        }
        C[static_cast<int64_t>(m) * ldc + n] = v;
    }
}

// y[b] = sum_j bf16(H[b][j]) * bf16(w[j]) + bias: one wave per row.
__global__ __launch_bounds__(256) void HeadForwardKernel(int rows, int D, const float* __restrict__ H, int ldh,
                                                         const float* __restrict__ w, const float* __restrict__ bias,
                                                         float* __restrict__ y) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    float s = 0.0f;
    for (int j = lane; j < D; j += 64) s += Bf16Round(H[static_cast<int64_t>(row) * ldh + j]) * Bf16Round(w[j]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[row] = s + (bias ? bias[0] : 0.0f);
}

// dZ[b][j] = Z[b][j] > 0 ? dy[b] * bf16(w[j]) : 0 (the head's input gradient through the ReLU).
__global__ void HeadBackwardKernel(int rows, int D, const float* __restrict__ dy, const float* __restrict__ w,
                                   const float* __restrict__ Z, int ldz, float* __restrict__ dZ, int lddz) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<int64_t>(rows) * D) return;
    const int b = static_cast<int>(i / D), j = static_cast<int>(i % D);
    const float z = Z[static_cast<int64_t>(b) * ldz + j];
    dZ[static_cast<int64_t>(b) * lddz + j] = z > 0.0f ? dy[b] * Bf16Round(w[j]) : 0.0f;
}

// out[j] = sum_b (wt ? wt[b] : 1) * X[b][j] (X rounded to BF16 when round_x), one thread per column,
// rows summed in order (deterministic).
__global__ void ColumnSumKernel(int rows, int N, const float* __restrict__ X, int ldx, const float* __restrict__ wt,
                                int round_x, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int b = 0; b < rows; b++) {
        float x = X[static_cast<int64_t>(b) * ldx + j];
        if (round_x) x = Bf16Round(x);
        s += (wt ? wt[b] : 1.0f) * x;
    }
    out[j] = s;
}

// Pinball (quantile) loss of the residual r = label - y at level tau, averaged over rows:
// loss = tau r (r >= 0) or (tau - 1) r; dy = (r > 0 ? -tau : 1 - tau) / rows.
__global__ void PinballKernel(int rows, const float* __restrict__ y, const float* __restrict__ label, float tau,
                              float* __restrict__ dy, float* __restrict__ loss) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= rows) return;
    const float r = label[b] - y[b];
    if (loss) loss[b] = r >= 0.0f ? tau * r : (tau - 1.0f) * r;
    dy[b] = (r > 0.0f ? -tau : 1.0f - tau) / static_cast<float>(rows);
}

__global__ void FloorsKernel(int rows, const float* __restrict__ y, const float* __restrict__ offset,
                             float* __restrict__ floors) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) floors[r] = y[r] + offset[r];
}

// Adam (bias-corrected) on FP32 master parameters.
__global__ void AdamKernel(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, float lr, float beta1, float beta2, float eps, float c1, float c2) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace
}  // namespace frk

extern "C" int frk_mlp_gemm(int32_t M, int32_t N, int32_t K, const float* d_a, int32_t lda, int32_t trans_a,
                            const float* d_b, int32_t ldb, int32_t trans_b, float* d_c, int32_t ldc, int32_t epilogue,
                            const float* d_bias, const float* d_z, int32_t ldz, float* d_h, int32_t ldh,
                            frk_stream_t stream) {
    if (M < 0 || N < 0 || K < 0) {
        frk::SetError("frk_mlp_gemm: negative size");
        return FRK_ERR_INVALID;
    }
    if (M == 0 || N == 0) return FRK_OK;
    if (!d_a || !d_b || !d_c) {
        frk::SetError("frk_mlp_gemm: null operand");
        return FRK_ERR_INVALID;
    }
    if ((trans_a ? lda < M : lda < K) || (trans_b ? ldb < K : ldb < N) || ldc < N) {
        frk::SetError("frk_mlp_gemm: leading dimension too small");
        return FRK_ERR_INVALID;
    }
    if (epilogue == FRK_MLP_EPI_BIAS_RELU && (!d_h || ldh < N)) {
        frk::SetError("frk_mlp_gemm: the bias+relu epilogue needs an activation output");
        return FRK_ERR_INVALID;
    }
    if (epilogue == FRK_MLP_EPI_RELU_GRAD && (!d_z || ldz < N)) {
        frk::SetError("frk_mlp_gemm: the relu-gradient epilogue needs the pre-activations");
        return FRK_ERR_INVALID;
    }
    if (epilogue < FRK_MLP_EPI_NONE || epilogue > FRK_MLP_EPI_RELU_GRAD) {
        frk::SetError("frk_mlp_gemm: unknown epilogue %d", epilogue);
        return FRK_ERR_INVALID;
    }
    dim3 grid((N + frk::kTile - 1) / frk::kTile, (M + frk::kTile - 1) / frk::kTile);
    hipStream_t s = static_cast<hipStream_t>(stream);
#define FRK_GEMM(TA, TB)                                                                                       \
    hipLaunchKernelGGL((frk::GemmKernel<TA, TB>), grid, dim3(256), 0, s, M, N, K, d_a, lda, d_b, ldb, d_c, ldc, \
                       epilogue, d_bias, d_z, ldz, d_h, ldh)
    if (trans_a && trans_b) FRK_GEMM(true, true);
    else if (trans_a) FRK_GEMM(true, false);
    else if (trans_b) FRK_GEMM(false, true);
    else FRK_GEMM(false, false);
#undef FRK_GEMM
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_mlp_head_forward(int32_t rows, int32_t dim, const float* d_h, int32_t ldh, const float* d_w,
                                    const float* d_bias, float* d_y, frk_stream_t stream) {
    if (rows <= 0) return FRK_OK;
    if (!d_h || !d_w || !d_y || ldh < dim) {
        frk::SetError("frk_mlp_head_forward: bad argument");
        return FRK_ERR_INVALID;
    }
    hipLaunchKernelGGL(frk::HeadForwardKernel, dim3((rows + 3) / 4), dim3(256), 0, static_cast<hipStream_t>(stream),
                       rows, dim, d_h, ldh, d_w, d_bias, d_y);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_mlp_head_backward(int32_t rows, int32_t dim, const float* d_dy, const float* d_w, const float* d_z,
                                     int32_t ldz, float* d_dz, int32_t lddz, frk_stream_t stream) {
    if (rows <= 0) return FRK_OK;
    if (!d_dy || !d_w || !d_z || !d_dz || ldz < dim || lddz < dim) {
        frk::SetError("frk_mlp_head_backward: bad argument");
        return FRK_ERR_INVALID;
    }
    const int64_t n = static_cast<int64_t>(rows) * dim;
    hipLaunchKernelGGL(frk::HeadBackwardKernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), rows, dim, d_dy, d_w, d_z, ldz, d_dz, lddz);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_mlp_column_sum(int32_t rows, int32_t cols, const float* d_x, int32_t ldx, const float* d_weights,
                                  int32_t round_bf16, float* d_out, frk_stream_t stream) {
    if (cols <= 0) return FRK_OK;
    if (!d_out || (rows > 0 && (!d_x || ldx < cols))) {
        frk::SetError("frk_mlp_column_sum: bad argument");
        return FRK_ERR_INVALID;
    }
    hipLaunchKernelGGL(frk::ColumnSumKernel, dim3((cols + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       rows, cols, d_x, ldx, d_weights, round_bf16, d_out);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_mlp_pinball(int32_t rows, const float* d_y, const float* d_label, float tau, float* d_dy,
                               float* d_loss, frk_stream_t stream) {
    if (rows <= 0) return FRK_OK;
    if (!d_y || !d_label || !d_dy || !(tau > 0.0f && tau < 1.0f)) {
        frk::SetError("frk_mlp_pinball: bad argument");
        return FRK_ERR_INVALID;
    }
    hipLaunchKernelGGL(frk::PinballKernel, dim3((rows + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       rows, d_y, d_label, tau, d_dy, d_loss);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_mlp_adam(int64_t n, float* d_param, const float* d_grad, float* d_m, float* d_v, float lr,
                            float beta1, float beta2, float eps, int32_t step, frk_stream_t stream) {
    if (n <= 0) return FRK_OK;
    if (!d_param || !d_grad || !d_m || !d_v || step < 1) {
        frk::SetError("frk_mlp_adam: bad argument");
        return FRK_ERR_INVALID;
    }
    const float c1 = 1.0f - powf(beta1, static_cast<float>(step));
    const float c2 = 1.0f - powf(beta2, static_cast<float>(step));
    hipLaunchKernelGGL(frk::AdamKernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), n, d_param, d_grad, d_m, d_v, lr, beta1, beta2, eps, c1, c2);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_mlp_forward(const frk_mlp_params* p, int32_t rows, const float* d_x, int32_t ldx,
                               float* d_work, float* d_y, frk_stream_t stream) {
    if (!p || !p->w1 || !p->w2 || !p->w3 || p->in_dim <= 0 || p->hidden <= 0) {
        frk::SetError("frk_mlp_forward: bad parameters");
        return FRK_ERR_INVALID;
    }
    if (rows <= 0) return FRK_OK;
    if (!d_x || !d_work || !d_y || ldx < p->in_dim) {
        frk::SetError("frk_mlp_forward: bad argument");
        return FRK_ERR_INVALID;
    }
    const int32_t Hd = p->hidden;
    float* z1 = d_work;
    float* h1 = z1 + static_cast<int64_t>(rows) * Hd;
    float* z2 = h1 + static_cast<int64_t>(rows) * Hd;
    float* h2 = z2 + static_cast<int64_t>(rows) * Hd;
    int rc = frk_mlp_gemm(rows, Hd, p->in_dim, d_x, ldx, 0, p->w1, p->in_dim, 1, z1, Hd, FRK_MLP_EPI_BIAS_RELU, p->b1,
                          nullptr, 0, h1, Hd, stream);
    if (rc == FRK_OK)
        rc = frk_mlp_gemm(rows, Hd, Hd, h1, Hd, 0, p->w2, Hd, 1, z2, Hd, FRK_MLP_EPI_BIAS_RELU, p->b2, nullptr, 0, h2, Hd,
                          stream);
    if (rc == FRK_OK) rc = frk_mlp_head_forward(rows, Hd, h2, Hd, p->w3, p->b3, d_y, stream);
    return rc;
}

extern "C" int frk_mlp_floors(int32_t rows, const float* d_y, const float* d_offset, float* d_floors,
                              frk_stream_t stream) {
    if (rows <= 0) return FRK_OK;
    if (!d_y || !d_offset || !d_floors) {
        frk::SetError("frk_mlp_floors: null argument");
        return FRK_ERR_INVALID;
    }
    hipLaunchKernelGGL(frk::FloorsKernel, dim3((rows + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       rows, d_y, d_offset, d_floors);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
