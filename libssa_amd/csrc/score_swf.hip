// score_swf.hip -- score_kernel instantiations of mode 3 (the SW first tier
// in packed fp16 windows).
#define SSA_SCORE_INSTANTIATE
#include "score_kernel.h"

namespace ssa {
template hipError_t launch_mode<3, false>(int, int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_mode<3, false>(int, int, int);
}  // namespace ssa
