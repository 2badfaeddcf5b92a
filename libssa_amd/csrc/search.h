// search.h -- one database search (sw_align / nw_align) on the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "common.h"

extern "C" {
#include "libssa_amd.h"
}

namespace ssa {

// a query buffer as s_create_searchdata builds them (src/algo/searcher.c:42-90)
struct QueryBuf {
    const uint8_t* seq;
    size_t len;
    int strand;
    int frame;
};

// a DB entry as the adapter makes them (src/db_adapter.c:47-110)
struct DbEntry {
    uint64_t id;
    uint32_t len;
    uint8_t strand;
    uint8_t frame;
    int64_t code_off;  // offset of the codes in the DB code buffer (padded to 4)
};

// bumped whenever the DB plugin is (re)initialised or closed
void db_generation_bump();

// 4-column blocks a score kernel spends on an entry: an empty entry (a
// translated frame without a codon) takes one block of padding, as a lane of
// the reference's kernels does (src/algo/8/search_8_util.h:60-74)
inline int32_t entry_blocks(uint32_t len) { return len ? (int32_t)((len + 3) / 4) : 1; }

p_alignment_list run_search(SearchType type, p_query query, size_t hitcount, int bit_width, int align_type);

// one sw_align / nw_align call of a batch
struct SearchRequest {
    SearchType type;
    p_query query;
    size_t hitcount;
    int bit_width;
    int align_type;
};

struct SearchResult {
    p_alignment_list list = nullptr;
    ssa_amd_stats_t stats{};
    std::vector<ssa_amd_hit_t> hits;  // the no-allocation path: the hits instead of a list
};

// sw_align / nw_align with COMPUTE_SCORE on the no-allocation context: every
// per-call device buffer is a piece of the device's scratch array.  Returns
// 0, -3 (the DB or a stream layout is not resident) or -4 (the search needs
// more than the scratch holds).
int run_search_noalloc(SearchType type, p_query query, size_t hitcount, int bit_width,
                       std::vector<ssa_amd_hit_t>& hits);

struct AutotuneResult {
    size_t qlen;
    int lanes, rows, cols, place;
    double ms;
};
// shape measurements on the loaded DB; the fastest shape per (type, length)
// becomes the choice for queries of the nearest tuned length
int autotune(int type, const std::vector<size_t>& lengths, int copies, std::vector<AutotuneResult>& out);
void autotune_reset();

// Self-check of the host replay: random per-pair tier results (overflows at
// every width, heavy score ties, 1-6 query buffers, hit counts 1-5000),
// replayed with the chunk filter and in full; returns the number of cases
// whose hits or overflow counts differ.  Host only.
int replay_selfcheck(uint64_t seed, int cases);

class Exec;
// Runs the searches as one batch on the context's stream: every result is
// what the search returns alone.  one_graph: the first-tier launches of all
// searches go into one captured HIP graph.
void run_batch(Exec& ex, const std::vector<SearchRequest>& reqs, std::vector<SearchResult>& results, bool one_graph);
void release_device_db();

// The plugin DB mapped to search entries on the host (src/db_adapter.c:47-110,
// 212-239): per non-empty sequence one entry (amino acids), one or two
// (nucleotides: forward, and the reverse complement when strands has bit 2),
// or three or six translated frames (TRANS_DB / TRANS_BOTH, with the genetic
// code); codes padded to a multiple of 4 per entry, 4 bytes for an empty frame.
struct HostDb {
    size_t count = 0;                    // sequences the provider returned
    std::vector<DbEntry> entries;
    std::vector<uint32_t> first_entry;   // per sequence index (count + 1 values)
    std::vector<uint8_t> codes;
    std::vector<uint32_t> by_length;     // entries longest first, ties in entry order
    std::vector<std::pair<uint64_t, uint64_t>> unknown;  // (sequence index, symbols mapped to 0)
};
void map_provider_db(int symtype, int strands, int gencode, HostDb& out);

// The device DB for the current settings, built if needed.
void ensure_device_db();

// What one request takes on the device (the resident DB required): its
// first-tier cells, the HBM its buffers hold while it runs, the share of the
// chip's resident workgroup slots its first-tier grid fills (set by the
// kernel shape's LDS and VGPR use) and that shape.
struct RequestPlan {
    double cells = 0;
    size_t hbm_bytes = 0;
    double occupancy = 0;
    int lanes = 0, rows = 0;
};
RequestPlan plan_request(const SearchRequest& rq, const std::vector<size_t>& query_lengths);

ssa_amd_stats_t& last_stats();
void set_shape_override(int lanes, int rows);  // 0, 0 = chosen per query length
void set_cols_override(int cols);             // 0 = chosen per search, else 4 or 8
void set_placement_override(int placement);   // 0 = chosen per search, 1 LDS, 2 VGPRs
void set_arith_override(int arith);           // 0 = fp16 windows for the SW first tier, 1 = int16 everywhere
void set_target_waves(int waves);             // 0 = default
void set_task_trace(bool on);                 // record the first-tier score launches' tasks
std::vector<ssa_amd_task_record_t> last_task_trace();

}  // namespace ssa
