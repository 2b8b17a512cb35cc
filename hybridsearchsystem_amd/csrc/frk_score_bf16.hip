// frk_score_bf16.hip -- the fused scorer's BF16 instantiations (kernel body in frk_score_impl.h).
#include "frk_score_impl.h"

namespace frk {
int ScoreBf16(const ScoreLaunch& L, int mode) {
    switch (mode) {
        case kModeBlockMax: return Dispatch<FRK_DTYPE_BF16, kModeBlockMax>(L.args, L.chunks, L.stream, L.grid_x);
        case kModeDense: return Dispatch<FRK_DTYPE_BF16, kModeDense>(L.args, L.chunks, L.stream, L.grid_x);
        default: return Dispatch<FRK_DTYPE_BF16, kModeSelect>(L.args, L.chunks, L.stream, L.grid_x);
    }
}
}  // namespace frk
