// frk_score_bf16.hip -- the fused scorer's BF16 instantiations (kernel body in frk_score_impl.h).
#include "frk_score_impl.h"

namespace frk {
int ScoreBf16(const ScoreLaunch& L, bool block_max) {
    return block_max ? Dispatch<FRK_DTYPE_BF16, true>(L.args, L.chunks, L.stream, L.grid_x)
                     : Dispatch<FRK_DTYPE_BF16, false>(L.args, L.chunks, L.stream, L.grid_x);
}
}  // namespace frk
