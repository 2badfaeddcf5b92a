// score_swv.hip -- score_kernel instantiations of modes 0 and 3 (SW, packed
// int16 and fp16 windows) with the query profile in VGPRs.
#define SSA_SCORE_INSTANTIATE
#include "score_kernel.h"

namespace ssa {
template hipError_t launch_mode<0, true>(int, int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_mode<0, true>(int, int, int);
template hipError_t launch_mode<3, true>(int, int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_mode<3, true>(int, int, int);
}  // namespace ssa
