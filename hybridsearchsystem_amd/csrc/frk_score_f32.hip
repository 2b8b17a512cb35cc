// frk_score_f32.hip -- the fused scorer's F32 instantiations (kernel body in frk_score_impl.h).
#include "frk_score_impl.h"

namespace frk {
int ScoreF32(const ScoreLaunch& L, int mode) {
    switch (mode) {
        case kModeBlockMax: return Dispatch<FRK_DTYPE_F32, kModeBlockMax>(L.args, L.chunks, L.stream, L.grid_x);
        case kModeDense: return Dispatch<FRK_DTYPE_F32, kModeDense>(L.args, L.chunks, L.stream, L.grid_x);
        default: return Dispatch<FRK_DTYPE_F32, kModeSelect>(L.args, L.chunks, L.stream, L.grid_x);
    }
}
}  // namespace frk
