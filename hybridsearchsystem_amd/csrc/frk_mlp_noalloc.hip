// frk_mlp_noalloc.hip -- the floor model's inference entry point for callers that cannot allocate
// device memory (code under graph capture, tight serving loops): frk_mlp_forward with its workspace in
// a fixed device array instead of a caller buffer.
#include <mutex>

#include "frk_common.h"

// ------------------------------------------------------------------------------------------------
// Inference without a caller workspace: the four activation tiles live in a fixed device array of
// FRK_MLP_NOALLOC_MAX_ROWS x FRK_MLP_NOALLOC_MAX_HIDDEN floats each (4 MiB per GPU, held until the
// process ends). Host code never passes the array's name to a kernel: its device address is looked
// up once per device with that device current and kept per device.
namespace frk {
namespace {
constexpr int64_t kMlpScratchFloats =
    4ll * FRK_MLP_NOALLOC_MAX_ROWS * static_cast<int64_t>(FRK_MLP_NOALLOC_MAX_HIDDEN);
__device__ __align__(256) float scratch[kMlpScratchFloats];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxScratchDevices] = {};  // guarded by g_scratch_mutex

// The calling thread's current device's copy of `scratch`.
int ScratchForCurrentDevice(float** out) {
    int dev = 0;
    FRK_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) {
        SetError("device %d out of range", dev);
        return FRK_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        FRK_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return FRK_OK;
}
}  // namespace
}  // namespace frk

extern "C" int frk_mlp_forward_noalloc(const frk_mlp_params* p, int32_t rows, const float* d_x, int32_t ldx,
                                       float* d_y, frk_stream_t stream) {
    if (!p || p->hidden <= 0) {
        frk::SetError("frk_mlp_forward_noalloc: bad parameters");
        return FRK_ERR_INVALID;
    }
    if (rows > FRK_MLP_NOALLOC_MAX_ROWS || p->hidden > FRK_MLP_NOALLOC_MAX_HIDDEN) {
        frk::SetError("frk_mlp_forward_noalloc: %d rows x %d hidden units exceed the fixed workspace (%d x %d)", rows,
                      p->hidden, FRK_MLP_NOALLOC_MAX_ROWS, FRK_MLP_NOALLOC_MAX_HIDDEN);
        return FRK_ERR_CAPACITY;
    }
    if (rows <= 0) return FRK_OK;
    float* work = nullptr;
    if (int rc = frk::ScratchForCurrentDevice(&work)) return rc;
    // the kernels overwrite every activation they read, so nothing left by an earlier call is seen
    return frk_mlp_forward(p, rows, d_x, ldx, work, d_y, stream);
}
