// score_pair.hip -- score_kernel_pair instantiations (paired packed SW
// queries), packed int16 and packed fp16 windows.
#define SSA_SCORE_INSTANTIATE
#include "score_kernel.h"

namespace ssa {
template hipError_t launch_pair_mode<false>(int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template hipError_t launch_pair_mode<true>(int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_pair_mode<false>(int, int);
template int resident_pair_mode<true>(int, int);
}  // namespace ssa
