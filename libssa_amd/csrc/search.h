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

p_alignment_list run_search(SearchType type, p_query query, size_t hitcount, int bit_width, int align_type);
void release_device_db();

ssa_amd_stats_t& last_stats();
void set_shape_override(int lanes, int rows);  // 0, 0 = chosen per query length
void set_target_waves(int waves);             // 0 = default

}  // namespace ssa
