// score_nwf.hip -- score_kernel instantiations of mode 4 (NW without the
// overflow trackers, packed fp16).
#define SSA_SCORE_INSTANTIATE
#include "score_kernel.h"

namespace ssa {
template hipError_t launch_mode<4, false>(int, int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_mode<4, false>(int, int, int);
}  // namespace ssa
