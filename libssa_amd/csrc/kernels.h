// kernels.h -- data layouts shared by the host runtime and the HIP kernels.
#pragma once

#include <cstdint>

namespace ssa {

// ---------------------------------------------------------------------------
// Score kernels (score_kernels.hip)
//
// A stream is a run of (query, DB entry) pairs laid end to end, each padded to
// whole 4-column blocks.  A lane group of G lanes scores two streams at once
// (the lo and hi int16 halves of every lane): lane g of the group owns query
// rows [p*G*R + g*R, +R) of pass p and runs one block behind lane g-1, which
// hands it the H and F of its bottom row.  A wave holds 64/G lane groups; a
// workgroup holds kWavesPerWG waves of the same query, which share the pass's
// score profile in LDS.
//
// Residue layout of a wave: uint2 res[nblocks][64/G], .x = 4 codes of the
// group's lo stream for that block, .y = the hi stream.  Byte k of a word
// holds column k's code times two in bits 0-5.  Flag bits: byte 0 bit 7 = a
// pair starts at this block, byte 1 bit 7 = a pair ends in this block, byte 2
// bits 6-7 = column of the pair's last residue in that block.
// ---------------------------------------------------------------------------
constexpr uint32_t kResStart = 0x80u;
constexpr uint32_t kResLast = 0x8000u;
constexpr int kResZShift = 22;  // (word >> 22) & 3

constexpr int kWavesPerWG = 4;
constexpr int kNoFlag = 0x7fffffff;

struct ScoreWave {
    int32_t nblocks;    // blocks of the wave's longest stream (0 = an idle filler wave)
    int32_t slot_base;  // stream (group k, half h): entries slot_first[slot_base + 2k + h] .. of slot_entries
    int64_t res_off;    // in uint2 units
};

// One workgroup task: a query against kWavesPerWG consecutive waves of the layout.
struct ScoreTask {
    int32_t query;
    int32_t wave0;
};

struct QueryDesc {
    int32_t qlen;
    int32_t qp_off;    // score kernel: int16 units, profile [pass][32 codes][G*R rows];
                       // exact tier: int32 units, profile [qlen][32 codes]
    int32_t lb_off;    // in uint32 units: NW left boundary [pass*G*R rows][2] packed (H, E)
    int32_t pair_off;  // result index of (this query, entry position 0)
    int32_t seg_off;   // packed query: its lanes' segments in ScoreLaunch::segs; -1 otherwise
    int32_t seg_off_hi;  // paired packed queries: the hi half's segments (score_kernel_pair)
};

// A packed query (SW): the rows of several queries laid end to end, each
// query starting at a lane boundary and padded to whole lanes, scored as one
// query of many passes.  Lane g of pass p takes its segment from
// segs[seg_off + p * G + g]: the result index of its query's entry position 0
// and the flags below.  A segment-top lane (its first row is its query's
// first row) takes the matrix border instead of the row above; a
// segment-bottom lane (the last lane of its query in the pass, or the pass's
// last lane) issues the pair's score; a lane past the last query writes
// nothing.
constexpr int32_t kSegTop = 1, kSegBottom = 2, kSegNone = 4;
struct SegLane {
    int32_t pair_off;
    int32_t flags;
};

// Per-launch constants in the kernel's value representation: every 16-bit
// value is packed twice into a uint32.  The 16-bit tier holds the reference's
// int16 lane values unchanged; the 8-bit tier holds the reference's int8 lane
// values shifted left by 8, so int16 saturation at -32768 is int8 saturation
// at -128 and the first value >= 127<<8 marks the int8 top saturation.
struct ScoreParams {
    uint32_t qr;          // gap open + extend
    uint32_t r;           // gap extend
    uint32_t zero;        // SW zero (biased), start value of H/E/F
    int32_t thr_a;        // SW: first flag threshold (S >= thr_a); NW: score_min (hmin < thr_a)
    int32_t thr_b;        // SW: second flag threshold; NW: hmax >= thr_b
    uint32_t hb_init[4];  // NW: top boundary H of the first block's four columns
    uint32_t fb_init[4];  // NW: top boundary F of the first block's four columns
    int32_t shift;        // the gap extend g (<= 0) of the shifted fp16 kernels, else 0: SW in fp16
                          // windows (mode 3) holds the values of DB column j less j g, NW in fp16
                          // (mode 4) those of cell (i, j) less (i + j) g
    uint32_t e4;          // mode 3: 4 (-g) as packed fp16 (a block's rise of the threshold frame)
    uint32_t thr_start;   // mode 3: packed fp16 fire threshold of a half that starts a pair
                          // (min(thr_a, thr_b) - 1 - 4 g, +inf past 2047)
};

// result per pair: score in representation (SW: maximum cell, NW: last cell),
// first flagged block (kNoFlag if none), second SW flag block.  The SW fields
// are combined over lanes and passes with atomic max / min.
struct PairResult {
    int32_t score;
    int32_t flag_a;
    int32_t flag_b;
};

// Arguments of one score-kernel launch.  The grid is persistent: each
// workgroup takes tasks from *counter (zeroed before the launch) until they
// run out, and keeps the pass boundaries of its 4 waves in its own slice of
// bnd (bnd_per_wave uint4 per wave).
// one task of a score launch, recorded when the caller traces the launches:
// start and end on the GPU's constant-rate wall clock, the workgroup that ran
// it and its 4-column steps per wave (passes x pass length)
struct TaskTrace {
    uint64_t start;
    uint64_t end;
    int32_t wg;
    int32_t steps;
};

struct ScoreLaunch {
    const ScoreWave* waves;
    const ScoreTask* tasks;
    int32_t ntasks;
    int32_t prio_steps;  // wave priority unit: a task's remaining 4-column steps in these units set it (0: off)
    int32_t* counter;
    const QueryDesc* queries;
    const int16_t* qp;
    const uint32_t* lb;
    const uint2* res;
    uint4* bnd;
    int64_t bnd_per_wave;
    int64_t bnd_rows;  // uint4 of a wave's slice that hold bottom rows; the SW fp16 windows'
                       // offsets (one uint32 per block and lane group) follow them
    const int32_t* slot_first;
    const int32_t* slot_entries;
    PairResult* results;
    const SegLane* segs;
    TaskTrace* trace;  // per-task records, or null (no tracing)
};

// one (stream, entry) of the residue interleaving kernel
struct PackEntry {
    int64_t code_off;  // byte offset of the entry's codes (padded to 4) in the DB code buffer
    int64_t dst;       // uint32 index of the entry's first word in the residue array
    int32_t nblocks;   // blocks of the entry
    int32_t last_col;  // (length - 1) % 4
    int32_t stride;    // uint32 distance between the words of consecutive blocks
    int32_t pad;
};

// exact tier (score_kernels.hip): one wave per pair, int32 or int64
constexpr int kExactR = 8;                    // query rows per lane
constexpr int kExactRows = 64 * kExactR;      // rows per pass of a wave
struct ExactPair {
    int64_t res_off;  // byte offset of the entry's codes in the DB code buffer
    int64_t bnd_off;  // element offset of the pair's pass boundaries (2 per DB column; queries of
                      // more than kExactRows rows only)
    int32_t dlen;
    int32_t query;    // index into QueryDesc (qp_off: byte offset of the query's codes)
    int32_t out;      // index into the int64 score output
    int32_t pad;
};

// ---------------------------------------------------------------------------
// Per-chunk top inserts (select_kernels.hip): the order-free first pass of
// the host's replay of the reference's heap, on the GPU.  One wave per chunk
// of DB entries decodes every (query buffer, entry) pair's tier results the
// way the host does, counts the cascade's overflows and keeps the chunk's m
// largest inserts, duplicates counted (an entry that overflowed 8 bit for m8
// query buffers is inserted m8 times at 16 bit).
// ---------------------------------------------------------------------------
constexpr int kChunkTopMax = 16;
constexpr int32_t kTopMode16 = 0;        // 16-bit first tier: res16[q] holds it
constexpr int32_t kTopModeSwSingle = 1;  // SW from one int16 pass: the 8-bit tier from flag_a, 16-bit from flag_b
constexpr int32_t kTopModeAll8 = 2;      // every pair overflowed 8 bit in its first block; res16 = the 16-bit re-run

struct TopSource {
    const PairResult* res16[6];  // per query buffer: entry e's result at [e ^ (flip >> q & 1)]
    uint32_t flip;               // bit q: query q reads the entry's reverse-complement twin (e ^ 1)
    int32_t nq;
    int32_t mode;                // kTopMode*
    int32_t type;                // 0 SW, 1 NW
    int32_t m;                   // inserts kept per chunk (<= kChunkTopMax)
};

struct ChunkTop {
    int64_t ov8, ov16;     // the chunk's overflow counts (src/libssa_datatypes.h:40-47)
    int32_t ntop;          // inserts kept
    int32_t need_exact;    // some pair that reaches the 16-bit tier overflowed it (the exact tier runs)
    int32_t top[kChunkTopMax];  // the largest inserts, descending
};

// one range of pair results to copy into a compact buffer (gather_kernel)
struct GatherSeg {
    const PairResult* src;
    int64_t dst;  // index into the compact buffer
    int32_t n;
    int32_t pad;
};

// ---------------------------------------------------------------------------
// Traceback (traceback_kernels.hip): one task per hit.
// ---------------------------------------------------------------------------
struct TraceTask {
    int64_t q_off;     // byte offset of the query codes
    int64_t d_off;     // byte offset of the DB codes
    int32_t qlen, dlen;
    int64_t dir_off;   // byte offset of this task's direction matrix scratch
    int64_t hh_off;    // int64 offset of this task's HH/EE scratch
    int64_t ops_off;   // offset into the op output (pairs of int32: op char, count)
    int32_t ops_cap;
    int32_t type;      // 0 SW, 1 NW
};

struct TraceResult {
    int64_t a_begin, a_end, b_begin, b_end;  // region (query = a, DB = b)
    int32_t nops;    // runs written (back to front)
    int32_t status;  // 0 ok, 1 region not found, 2 op buffer too small
};

}  // namespace ssa
