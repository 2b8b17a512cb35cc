// frk_score_fp8.hip -- the fused scorer's FP8 instantiations (kernel body in frk_score_impl.h).
#include "frk_score_impl.h"

namespace frk {
int ScoreFp8(const ScoreLaunch& L, bool block_max) {
    return block_max ? Dispatch<FRK_DTYPE_FP8E4M3, true>(L.args, L.chunks, L.stream, L.grid_x)
                     : Dispatch<FRK_DTYPE_FP8E4M3, false>(L.args, L.chunks, L.stream, L.grid_x);
}
}  // namespace frk
