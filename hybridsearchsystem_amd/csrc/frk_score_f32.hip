// frk_score_f32.hip -- the fused scorer's F32 instantiations (kernel body in frk_score_impl.h).
#include "frk_score_impl.h"

namespace frk {
int ScoreF32(const ScoreLaunch& L, bool block_max) {
    return block_max ? Dispatch<FRK_DTYPE_F32, true>(L.args, L.chunks, L.stream, L.grid_x)
                     : Dispatch<FRK_DTYPE_F32, false>(L.args, L.chunks, L.stream, L.grid_x);
}
}  // namespace frk
