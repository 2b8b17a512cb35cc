// frk_score_fp8.hip -- the fused scorer's FP8 instantiations (kernel body in frk_score_impl.h).
#include "frk_score_impl.h"

namespace frk {
int ScoreFp8(const ScoreLaunch& L, int mode) {
    switch (mode) {
        case kModeBlockMax: return Dispatch<FRK_DTYPE_FP8E4M3, kModeBlockMax>(L.args, L.chunks, L.stream, L.grid_x);
        case kModeDense: return Dispatch<FRK_DTYPE_FP8E4M3, kModeDense>(L.args, L.chunks, L.stream, L.grid_x);
        default: return Dispatch<FRK_DTYPE_FP8E4M3, kModeSelect>(L.args, L.chunks, L.stream, L.grid_x);
    }
}
}  // namespace frk
