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
    uint64_t cells_scored;    /* cells of the query buffers the first tier scored (a both-strand
                                 search whose reverse strand is read from the forward strand's
                                 reverse-complement twins scores half of `cells`) */
    int lanes_per_stream;     /* score kernel shape of the first tier: lanes per stream pair */
    int rows_per_lane;        /* and query rows per lane */
    int overflow_tracking;    /* 1 when the first tier's kernel tracked overflow (every 8-bit
                                 tier and SW; NW at 16 bit only where some value could reach
                                 either end of the int16 range), 0 when it provably cannot */
    double ms_gpu_first;      /* first-tier kernel time (HIP events) */
    double ms_gpu_requeue;    /* re-queued tiers (16-bit pass, 32-bit exact pass) */
    double ms_replay;         /* host replay of the reference's lane schedule and heap */
    double ms_traceback;      /* traceback kernel, including its copies */
    double ms_total;          /* whole call, host clock */
    uint64_t d2h_bytes;       /* tier results copied to the host: 12 bytes per (query buffer,
                                 entry) pair where the host decodes every pair, else the
                                 per-chunk largest inserts the GPU selects (chunk_top_kernel)
                                 and the pairs of the chunks whose lane schedules the replay runs */
    uint64_t replayed_chunks; /* chunks whose lane schedules the host replay ran */
    double ms_graph;          /* batches: capturing the first tiers' graph and updating (or
                                 instantiating) its executable, host clock */
    int graph_reused;         /* 1 when the batch ran its context's previous executable graph,
                                 updated to the batch's arguments */
    int columns_per_step;     /* first tier's column tile: 4 or 8 DB columns per step */
    int profile_placement;    /* first tier's query profile: 1 = LDS, 2 = VGPRs */
    int cell_arithmetic;      /* first tier's cell arithmetic: 1 = packed saturating int16,
                                 2 = packed fp16 windows (SW; exact, see
                                 ssa_amd_set_cell_arithmetic) */
    double ms_setup;          /* host work from the call to the first tier's launch (query
                                 buffers, profiles, uploads; for a batch, the whole batch's) */
} ssa_amd_stats_t;

/* Counters of the last search. */
void ssa_amd_last_stats(ssa_amd_stats_t* out);

/* Restores every setting made through libssa.h -- symbol type, strands,
 * genetic codes, gap penalties, chunk size, SIMD mode, thread count, output
 * mode -- to its initial value; the score matrix and the DB stay as they are.
 * The reference's init_symbol_translation tests the symbol type in force
 * before it changes it (src/libssa.c:146-151) and so exits once TRANS_DB or
 * TRANS_BOTH is set; this is how a long-lived process leaves those types. */
void ssa_amd_reset_settings(void);

/* Shape of the score kernel: lanes per stream pair (8, 16, 32 or 64) and
 * query rows per lane (4, 8, 16 or 32), at most 512 rows per pass (lanes x
 * rows); NW searches run a 32-row shape with 16 rows; 0, 0 chooses from the
 * query length. */
void ssa_amd_set_kernel_shape(int lanes_per_stream, int rows_per_lane);

/* Variant of the score kernel: the column tile -- DB columns a lane scores
 * per step, 4 (one block of the reference's 4-column unit) or 8 (two blocks
 * per step, shapes of at most 16 rows per lane in groups of 8 or 16 lanes) --
 * and where the query profile sits: 1 = LDS (each lane reads its rows'
 * scores for a column's DB codes), 2 = VGPRs (each lane holds its rows'
 * scores for A, C, G and T and picks a column's pair with v_perm: nucleotide
 * databases of A, C, G and T only, 4-column tiles of at most 16 rows per
 * lane, SW and NW without overflow trackers).  0 leaves the choice to the
 * cost model (LDS) or the autotuner; a forced variant a search or shape lacks
 * is left at its default.  Results do not depend on the variant. */
void ssa_amd_set_kernel_variant(int columns_per_step, int profile_placement);

/* Cell arithmetic of the score kernel's SW first tier (the tier that records
 * the 8-bit and 16-bit overflow blocks and scores below 65,535).  0 (the
 * default): packed fp16, each lane's values held relative to an offset of
 * its own (a window kept within +-2047, where fp16 integers are exact) --
 * the three-input packed maximum takes a SW cell from 9 packed operations to
 * 7.5 -- wherever the matrix and gaps keep a lane's values within the window
 * (BLOSUM62 with gaps 10/1 does), packed saturating int16 elsewhere.  1:
 * packed int16 always.  NW and the other tiers compute in packed int16.
 * Results do not depend on it. */
void ssa_amd_set_cell_arithmetic(int arithmetic);

/* Number of waves the DB's pair streams are cut into (0 = the default, 10240).
 * Streams never get shorter than 16 blocks per lane of a group. */
void ssa_amd_set_stream_waves(int waves);

/* ---- autotuning ---- */

/* Measures every score-kernel shape (lanes per stream pair x query rows per
 * lane) on the loaded DB, for each of the n lengths with `copies` random
 * queries of that length in one launch (1 = as single searches run; more =
 * as batches run), with the current scoring and search type (0 = SW, 1 = NW).
 * From then on a query's shape is the one with the least measured time at the
 * tuned length nearest to its own (on a log scale), scaled by the number of
 * passes the shape needs for the query's length.  lanes_out, rows_out, ms_out (NULL
 * allowed) receive the chosen shape and its time per query for each length.
 * Returns the number of lengths tuned, -1 for invalid arguments, -2 when
 * scoring is not initialised.  Results do not depend on the shape. */
int ssa_amd_autotune(int search_type, const size_t* lengths, size_t n, int copies, int* lanes_out, int* rows_out,
                     double* ms_out);

/* The choice of the last ssa_amd_autotune for its i-th length: lanes per
 * stream pair, rows per lane, columns per step (4 or 8), where the query
 * profile sits (1 = LDS, 2 = VGPRs: measured for nucleotide DBs of A, C, G
 * and T only) and the time per query (NULL allowed for each).  Returns 0, or
 * -1 for i past the last tuned length. */
int ssa_amd_autotune_choice(size_t i, int* lanes_per_stream, int* rows_per_lane, int* columns_per_step,
                            int* profile_placement, double* ms);

/* Forgets the measured shapes (the cost model chooses again). */
void ssa_amd_autotune_reset(void);

/* ---- task trace (diagnostics) ---- */

/* No counterpart in the reference: a timeline of the score kernels for
 * callers tuning or diagnosing launches.  With ssa_amd_set_task_trace(1), the
 * first-tier score launches of the following searches and batches record
 * every task (a workgroup's 4 waves over all passes of their streams): its
 * start and end in ns from the first task's start (the GPU's constant-rate
 * clock), the workgroup that ran it, its 4-column steps per wave and the
 * launch of the batch it belongs to.  Each workgroup runs its first task by
 * its index and takes later ones from a queue, costliest first; workgroups 0
 * .. CUs - 1 hold the first slot of each CU.  Tracing costs two stores per
 * task. */
typedef struct {
    uint64_t start_ns;
    uint64_t end_ns;
    int32_t workgroup;
    int32_t steps;
    int32_t launch;
    int32_t pad;
} ssa_amd_task_record_t;

void ssa_amd_set_task_trace(int on);

/* Copies up to max records of the last traced search (or batch) to out and
 * returns how many there are. */
size_t ssa_amd_task_trace(ssa_amd_task_record_t* out, size_t max);

/* ---- batched searches ---- */

/* One search of a batch: the arguments of sw_align (search_type 0) or
 * nw_align (search_type 1). */
typedef struct {
    int search_type;
    p_query query;
    size_t hitcount;
    int bit_width;
    int align_type;
} ssa_amd_request_t;

/* Runs n searches as one batch with the current settings: the first-tier
 * launches of all of them go to the GPU as one captured HIP graph, searches
 * with the same type, tier and kernel shape share one kernel launch, and the
 * host replays run in parallel.  out[i] receives exactly the list that
 * sw_align / nw_align returns for reqs[i] alone (free each with
 * free_alignment, before freeing its query); stats (NULL allowed) receives one
 * record per search.  Returns 0; -1 for an invalid request, -2 when scoring is
 * not initialised (nothing runs in either case).  Not re-entrant, like the
 * libssa.h entry points. */
int ssa_amd_batch_align(const ssa_amd_request_t* reqs, size_t n, p_alignment_list* out, ssa_amd_stats_t* stats);

/* ---- serving: continuous batching over the resident DB ---- */

typedef struct ssa_amd_server ssa_amd_server;

typedef struct {
    size_t batches;         /* batches run so far */
    size_t requests;        /* requests taken into batches */
    size_t max_batch_seen;  /* largest batch */
    size_t queued;          /* requests waiting now */
    size_t device_budget;   /* device bytes the admission works with */
    size_t deferrals;       /* times a free worker waited instead of starting a batch */
    size_t device_in_use;   /* device bytes of the batches on the GPU now */
} ssa_amd_server_counters_t;

/* Starts a request loop with `slots` workers (1-8), each with its own device
 * stream and buffers; a free worker takes waiting requests as one batch, by
 * the policy of ssa_amd_admission_plan: at most `max_batch` of them, within
 * the cell budget (ssa_amd_server_tune), while their device footprint fits
 * what `device_bytes` (0 = half the free device memory) leaves beside the
 * batches on the GPU, and beside running batches only if the batch's
 * first-tier workgroups fit the resident slots they leave (or once the
 * oldest waiting request has waited 50 ms), so below saturation requests
 * run alone.  Each request's footprint comes from the kernel shape its search
 * would choose over the resident DB.  Makes the DB resident; uses the
 * settings current at the time (scoring, gaps, symbol type, DB), which must
 * not change while the server runs.  Returns NULL if the arguments or the
 * settings are invalid. */
ssa_amd_server* ssa_amd_server_start(int slots, size_t max_batch, size_t device_bytes);

/* Batch formation: a free worker takes the oldest waiting request first once
 * it has waited more than aging_ms (default 750), then the cheapest waiting
 * requests (estimated cells, shortest first) while the batch stays within
 * batch_cells (default 2.5e11) and the other limits.  Values <= 0 keep the
 * current setting. */
void ssa_amd_server_tune(ssa_amd_server* server, double batch_cells, double aging_ms);

/* Queues one search; returns its ticket (>= 0) or -1 for an invalid request.
 * The query must stay alive until the result has been taken and freed. */
int64_t ssa_amd_server_submit(ssa_amd_server* server, const ssa_amd_request_t* request);

/* Waits for a ticket's result: the list sw_align / nw_align returns for the
 * request alone (the caller frees it with free_alignment).  stats, latency_ms
 * (submit to result) and queue_ms (submit to batch start) may be NULL.
 * Returns NULL for an unknown ticket, and for a ticket another thread is
 * already waiting for (one waiter takes each result). */
p_alignment_list ssa_amd_server_wait(ssa_amd_server* server, int64_t ticket, ssa_amd_stats_t* stats,
                                     double* latency_ms, double* queue_ms);

void ssa_amd_server_counters(ssa_amd_server* server, ssa_amd_server_counters_t* out);

/* Finishes every queued request, then stops the workers; results not taken are freed. */
void ssa_amd_server_stop(ssa_amd_server* server);

/* The server's admission policy as a pure function, for tests and for
 * callers with their own request loop.  No counterpart in the reference. */
typedef struct {
    double wait_ms;       /* time the request has waited */
    double cells;         /* its GPU work: first-tier cells */
    size_t device_bytes;  /* device memory its buffers hold while it runs */
    double occupancy;     /* share (0-1) of the chip's resident workgroup slots its first-tier
                             grid fills: the kernel shape's LDS and VGPR use set how many
                             workgroups a CU holds */
} ssa_amd_admission_request_t;

typedef struct {
    size_t max_batch;
    double batch_cells;       /* cell budget of a batch */
    double aging_ms;          /* the oldest request goes first once it waited longer */
    double assist_ms;         /* a batch may queue behind running ones once the oldest waited longer */
    size_t device_budget;     /* device bytes all batches may hold together */
    size_t device_in_use;     /* device bytes of the batches on the GPU now */
    double occupancy_in_use;  /* their summed occupancy */
    int running;              /* batches on the GPU now */
} ssa_amd_admission_state_t;

/* Which of the n waiting requests (queue[0] the oldest) a free worker takes
 * as its next batch: candidates are the oldest first once it has waited past
 * aging_ms, then the cheapest by cells; each joins while the batch stays
 * within max_batch and batch_cells and its device bytes fit device_budget -
 * device_in_use (a request larger than the whole budget runs alone when
 * nothing runs); with batches running, the batch starts only if its summed
 * occupancy fits 1 - occupancy_in_use or the oldest request has waited
 * assist_ms.  Writes the chosen indices to chosen (room for n) and returns
 * how many; 0 means wait. */
size_t ssa_amd_admission_plan(const ssa_amd_admission_request_t* queue, size_t n,
                              const ssa_amd_admission_state_t* state, size_t* chosen);

/* ---- a search that allocates no device memory ---- */

/* One hit of ssa_amd_align_noalloc: the fields of alignment_t that a
 * score-only search fills, with the same values (db_strand and db_frame as
 * alignment_t.db_seq.strand and .frame, which the reference fills swapped,
 * src/algo/aligner.c:76-77). */
typedef struct {
    size_t db_id;
    int db_strand;
    int db_frame;
    int query_strand;
    int query_frame;
    long score;
} ssa_amd_hit_t;

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the search sw_align (search_type 0) or nw_align (1)
 * runs with COMPUTE_SCORE, and its hits -- the same hits, in the same order,
 * as that call's list -- are written to hits[0 .. *nhits), which holds
 * hitcount entries.  The call allocates and frees no device or pinned host
 * memory: every per-call device buffer (pair results, profiles, task lists,
 * pass boundaries, counters) is a piece of one file-scope device array of
 * 128 MiB that each GPU the library runs on holds until the process ends; its
 * contents persist between calls, and the kernels clear what must start at
 * zero.  Host working memory comes from the process heap.  The DB must already
 * be resident on the GPU with the current settings, with the stream layout of
 * at least one kernel shape: one sw_align or nw_align over the loaded DB
 * leaves both.  Queries longer than one pass of the kernel run on as many
 * workgroups as the array holds pass boundaries for.  Calls are serialised
 * within the process (they share the array; calls that overlapped on one GPU
 * would race).  Returns 0; -1 for invalid arguments; -2 when scoring is not
 * initialised; -3 when the DB or a stream layout is not resident (run sw_align
 * or nw_align once first); -4 when the search needs more than the array holds
 * (an error message says which). */
int ssa_amd_align_noalloc(int search_type, p_query query, size_t hitcount, int bit_width, ssa_amd_hit_t* hits,
                          size_t* nhits);

/* Writes the DB the plugin currently serves to a pre-packed file: the
 * sequences, their codes mapped for the symbol type in force (nucleotide DBs
 * with both strands' entries), the entry table and the entries in length
 * order.  init_db on that file (built-in provider) then serves the same
 * sequences and IDs and uploads the codes without mapping them again.  Packing
 * a translated symbol type is refused.  Returns 0, or -1 after an error
 * message.  No counterpart in the reference: it re-maps the DB on every search
 * (src/db_adapter.c:47-110). */
int ssa_amd_db_pack(const char* path);

/* The current 32x32 score matrix as the search uses it ([db code * 32 + query code]). */
void ssa_amd_get_matrix(int64_t* out);

/* Name of the GPU architecture the library runs on (e.g. "gfx950:sramecc+:xnack-"). */
const char* ssa_amd_device_arch(void);

/* Compute units of the GPU the library runs on (a score launch's workgroups
 * 0 .. CUs - 1 hold the first slot of each CU, as in the task trace). */
int ssa_amd_device_cus(void);

/* Self-check of the host replay of the reference's lane schedule and heap
 * (host only, no GPU call): `cases` random sets of per-pair tier results
 * (overflows at every width, heavy score ties, 1-6 query buffers, hit counts
 * 1-5000), replayed once skipping the chunks that cannot change the heap and
 * once in full.  Returns the number of cases whose hits (scores, IDs, strands,
 * frames, query buffers, order) or overflow counts differ: 0 is a pass.  No
 * counterpart in the reference. */
int ssa_amd_selfcheck_replay(uint64_t seed, int cases);

#ifdef __cplusplus
}
#endif

#endif /* LIBSSA_AMD_H_ */
