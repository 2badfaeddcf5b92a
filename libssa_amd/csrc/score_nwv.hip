// score_nwv.hip -- score_kernel instantiations of modes 2 and 4 (NW without
// the overflow trackers, packed int16 and fp16) with the query profile in
// VGPRs.
#define SSA_SCORE_INSTANTIATE
#include "score_kernel.h"

namespace ssa {
template hipError_t launch_mode<2, true>(int, int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_mode<2, true>(int, int, int);
template hipError_t launch_mode<4, true>(int, int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_mode<4, true>(int, int, int);
}  // namespace ssa
