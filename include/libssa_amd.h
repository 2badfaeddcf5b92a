/*
 * libssa_amd.h -- extensions of the MI355X library beyond the reference API.
 *
 * None of these has a counterpart in the reference (src/libssa.h); they expose
 * the GPU runtime's counters and tuning knobs.
 */
#ifndef LIBSSA_AMD_H_
#define LIBSSA_AMD_H_

#include <stddef.h>
#include <stdint.h>

#include "libssa.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* the reference's per-search counters (src/libssa_datatypes.h:40-47) */
    size_t chunk_count;
    size_t seq_count;
    size_t overflow_8_bit_count;
    size_t overflow_16_bit_count;
    /* work and time of the last sw_align / nw_align call */
    uint64_t cells;           /* sum over (query, DB entry) pairs of |q| * |d| in the first tier */
    uint64_t cells_computed;  /* cells the GPU computed in all tiers, padding included */
    int lanes_per_stream;     /* score kernel shape of the first tier: lanes per stream pair */
    int rows_per_lane;        /* and query rows per lane */
    double ms_gpu_first;      /* first-tier kernel time (HIP events) */
    double ms_gpu_requeue;    /* re-queued tiers (16-bit pass, 32-bit exact pass) */
    double ms_replay;         /* host replay of the reference's lane schedule and heap */
    double ms_traceback;      /* traceback kernel, including its copies */
    double ms_total;          /* whole call, host clock */
} ssa_amd_stats_t;

/* Counters of the last search. */
void ssa_amd_last_stats(ssa_amd_stats_t* out);

/* Shape of the score kernel: lanes per stream pair (8, 16, 32 or 64) and
 * query rows per lane (4, 8 or 16); 0, 0 chooses from the query length. */
void ssa_amd_set_kernel_shape(int lanes_per_stream, int rows_per_lane);

/* Number of waves the DB's pair streams are cut into (0 = the default, 10240).
 * Streams never get shorter than 16 blocks per lane of a group. */
void ssa_amd_set_stream_waves(int waves);

/* The current 32x32 score matrix as the search uses it ([db code * 32 + query code]). */
void ssa_amd_get_matrix(int64_t* out);

/* Name of the GPU architecture the library runs on (e.g. "gfx950:sramecc+:xnack-"). */
const char* ssa_amd_device_arch(void);

#ifdef __cplusplus
}
#endif

#endif /* LIBSSA_AMD_H_ */
