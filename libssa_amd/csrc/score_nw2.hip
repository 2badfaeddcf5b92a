// score_nw2.hip -- score_kernel instantiations of mode 2 (NW without the
// overflow trackers, packed int16).
#define SSA_SCORE_INSTANTIATE
#include "score_kernel.h"

namespace ssa {
template hipError_t launch_mode<2, false>(int, int, int, hipStream_t, const ScoreLaunch&, const ScoreParams&, int);
template int resident_mode<2, false>(int, int, int);
}  // namespace ssa
