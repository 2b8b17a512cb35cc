// frk_score_f16x2.hip -- the fused scorer's F16X2 instantiations (kernel body in frk_score_impl.h).
#include "frk_score_impl.h"

namespace frk {
int ScoreF16x2(const ScoreLaunch& L, int mode) {
    switch (mode) {
        case kModeBlockMax: return Dispatch<FRK_DTYPE_F16X2, kModeBlockMax>(L.args, L.chunks, L.stream, L.grid_x);
        case kModeDense: return Dispatch<FRK_DTYPE_F16X2, kModeDense>(L.args, L.chunks, L.stream, L.grid_x);
        default: return Dispatch<FRK_DTYPE_F16X2, kModeSelect>(L.args, L.chunks, L.stream, L.grid_x);
    }
}
}  // namespace frk
