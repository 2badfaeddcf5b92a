// api.cpp -- the libssa.h entry points (reference: src/libssa.c:41-271).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"
#include "dbpack.h"
#include "gpu.h"
#include "search.h"

extern "C" {
#include "libssa_amd.h"
}

namespace ssa {
p_query query_from_string(const char* s);
p_query query_from_file(const char* path);
}  // namespace ssa

using namespace ssa;

// the choices of the last ssa_amd_autotune (ssa_amd_autotune_choice)
static std::vector<AutotuneResult>& last_autotune() {
    static std::vector<AutotuneResult> r;
    return r;
}

extern "C" {

// src/libssa.c:41-43
void set_output_mode(int mode) { config().output_mode = mode; }

// src/libssa.c:45-47 (set_max_compute_capability, src/cpu_config.c:29-33)
void set_simd_compute_mode(int mode) { config().simd_mode = mode; }

// src/libssa.c:49-56
void set_chunk_size(size_t size) {
    if (size == 0) {
        print_error("Only non zero chunk sizes are allowed. Using the default size of 1000 sequences.");
        size = kDefaultChunkSize;
    }
    config().chunk_size = size;
}

// src/libssa.c:58-60
void set_thread_count(size_t count) { config().thread_count = count; }

// src/libssa.c:81-94
void init_score_matrix(int mode, const char* m) {
    if (mode == READ_FROM_FILE) matrix_init_file(m);
    else if (mode == READ_FROM_STRING) matrix_init_text(m);
    else if (mode == MATRIX_BUILDIN) matrix_init_builtin(m);
    else fatal("Unknown mode for reading score matrices: %d", mode);
}

// src/libssa.c:102-105
void init_gap_penalties(const int8_t gap_open, const int8_t gap_extend) {
    config().gap_open = gap_open;
    config().gap_extend = gap_extend;
}

// src/libssa.c:115-117: the first argument is used as the match score
void init_constant_scores(const int8_t p, const int8_t m) { matrix_init_constant(p, m); }

// src/libssa.c:126-132
void init_db(const char* db_file) {
    ssa_db_close();
    ssa_db_init(db_file);
    db_generation_bump();
    print_info("DB read %lu sequences\n", (unsigned long)ssa_db_get_sequence_count());
}

// src/libssa.c:139-157.  The symbol-type and strand checks test the setting in
// force before the call, as the reference's do.
void init_symbol_translation(int type, int strands, int d_gencode, int q_gencode) {
    Config& c = config();
    if (!gencode_valid(q_gencode)) fatal("Illegal query genetic code specified.");
    if (!gencode_valid(d_gencode)) fatal("Illegal database genetic code specified.");
    if (c.symtype < 0 || c.symtype > 4) fatal("Illegal symbol type specified.");
    if (c.symtype < 0 || c.symtype > 2) fatal("Illegal strands specified.");
    c.symtype = type;
    c.strands = strands;
    c.q_gencode = q_gencode;
    c.d_gencode = d_gencode;
}

// src/libssa.c:165-182
p_query init_sequence_fasta(int mode, const char* fasta) {
    if (mode == READ_FROM_FILE) return query_from_file(fasta);
    if (mode == READ_FROM_STRING) return query_from_string(fasta);
    fatal("Unknown mode for reading query sequences: %d", mode);
}

// src/libssa.c:192-194
void free_sequence(p_query p) { delete p; }

// src/libssa.c:196-219
static void test_configuration(p_query query) {
    const Config& c = config();
    if (!c.gap_open && !c.gap_extend) print_warning("Gap opening and gap extension cost set to zero. Possible error.");
    if (!matrix().valid) fatal("Scoring not initialized.");
    if (!query) fatal("Query not initialized.");
    if (ssa_db_get_sequence_count() == 0) print_warning("Database contains zero sequences. Possible error.");
    if (!matrix().constant && c.symtype == NUCLEOTIDE)
        fatal("Nucleotide sequences can only be aligned using constant scores.");
}

// src/libssa.c:229-235
p_alignment_list sw_align(p_query p, size_t hitcount, int bit_width, int align_type) {
    test_configuration(p);
    return run_search(kSW, p, hitcount, bit_width, align_type);
}

// src/libssa.c:245-251
p_alignment_list nw_align(p_query p, size_t hitcount, int bit_width, int align_type) {
    test_configuration(p);
    return run_search(kNW, p, hitcount, bit_width, align_type);
}

// src/libssa.c:261-263, src/algo/aligner.c:101-143
void free_alignment(p_alignment_list alist) {
    if (!alist) return;
    if (alist->alignments) {
        for (size_t i = 0; i < alist->len; ++i) {
            p_alignment a = alist->alignments[i];
            if (!a) continue;
            free(a->db_seq.seq);
            free(a->alignment);
            free(a);
        }
        free(alist->alignments);
    }
    free(alist);
}

// src/libssa.c:266-271
void ssa_exit(void) {
    matrix_free();
    ssa_db_close();
    release_device_db();
}

// ---- libssa_amd.h ----
void ssa_amd_last_stats(ssa_amd_stats_t* out) {
    if (out) *out = last_stats();
}

void ssa_amd_reset_settings(void) { config() = Config(); }

void ssa_amd_set_kernel_shape(int lanes_per_stream, int rows_per_lane) {
    const bool autosel = lanes_per_stream == 0 && rows_per_lane == 0;
    const bool lanes_ok = lanes_per_stream == 8 || lanes_per_stream == 16 || lanes_per_stream == 32 ||
                          lanes_per_stream == 64;
    const bool rows_ok = (rows_per_lane == 4 || rows_per_lane == 8 || rows_per_lane == 16 || rows_per_lane == 32) &&
                         lanes_per_stream * rows_per_lane <= 512;
    if (!autosel && !(lanes_ok && rows_ok))
        fatal("libssa_amd: kernel shape must be 0, 0 or lanes in {8, 16, 32, 64} and rows in {4, 8, 16, 32} "
              "with at most 512 rows per pass (got %d, %d)",
              lanes_per_stream, rows_per_lane);
    set_shape_override(lanes_per_stream, rows_per_lane);
}

void ssa_amd_set_cell_arithmetic(int arithmetic) {
    if (arithmetic < 0 || arithmetic > 1)
        fatal("libssa_amd: cell arithmetic must be 0 (chosen) or 1 (int16) (got %d)", arithmetic);
    set_arith_override(arithmetic);
}

void ssa_amd_set_kernel_variant(int columns_per_step, int profile_placement) {
    if (columns_per_step != 0 && columns_per_step != 4 && columns_per_step != 8)
        fatal("libssa_amd: columns per step must be 0, 4 or 8 (got %d)", columns_per_step);
    if (profile_placement < 0 || profile_placement > 2)
        fatal("libssa_amd: profile placement must be 0, 1 (LDS) or 2 (VGPRs) (got %d)", profile_placement);
    set_cols_override(columns_per_step);
    set_placement_override(profile_placement);
}

void ssa_amd_set_stream_waves(int waves) {
    if (waves < 0) fatal("libssa_amd: stream waves must be >= 0 (got %d)", waves);
    set_target_waves(waves);
}

int ssa_amd_batch_align(const ssa_amd_request_t* reqs, size_t n, p_alignment_list* out, ssa_amd_stats_t* stats) {
    if (n == 0) return 0;
    if (!reqs || !out) return -1;
    const Config& c = config();
    if (!matrix().valid) return -2;
    if (!matrix().constant && c.symtype == NUCLEOTIDE) return -2;
    std::vector<SearchRequest> rq(n);
    for (size_t i = 0; i < n; ++i) {
        const ssa_amd_request_t& r = reqs[i];
        if (r.search_type != 0 && r.search_type != 1) return -1;
        if (!r.query) return -1;
        if (r.bit_width != BIT_WIDTH_8 && r.bit_width != BIT_WIDTH_16 && r.bit_width != BIT_WIDTH_64) return -1;
        if (r.align_type != COMPUTE_SCORE && r.align_type != COMPUTE_ALIGNMENT) return -1;
        if (r.bit_width == BIT_WIDTH_8 && c.simd_mode < COMPUTE_ON_SSE41) return -1;
        rq[i] = SearchRequest{r.search_type == 0 ? kSW : kNW, r.query, r.hitcount, r.bit_width, r.align_type};
    }
    std::vector<SearchResult> res;
    run_batch(Gpu::get().exec(), rq, res, true);
    for (size_t i = 0; i < n; ++i) {
        out[i] = res[i].list;
        if (stats) stats[i] = res[i].stats;
    }
    last_stats() = res[n - 1].stats;
    return 0;
}

int ssa_amd_align_noalloc(int search_type, p_query query, size_t hitcount, int bit_width, ssa_amd_hit_t* hits,
                          size_t* nhits) {
    if (nhits) *nhits = 0;
    if (search_type != 0 && search_type != 1) return -1;
    if (!query || !nhits || (hitcount && !hits)) return -1;
    if (bit_width != BIT_WIDTH_8 && bit_width != BIT_WIDTH_16 && bit_width != BIT_WIDTH_64) return -1;
    const Config& c = config();
    if (bit_width == BIT_WIDTH_8 && c.simd_mode < COMPUTE_ON_SSE41) return -1;
    if (!matrix().valid) return -2;
    if (!matrix().constant && c.symtype == NUCLEOTIDE) return -2;
    std::vector<ssa_amd_hit_t> got;
    const int rc = run_search_noalloc(search_type == 0 ? kSW : kNW, query, hitcount, bit_width, got);
    if (rc != 0) return rc;
    const size_t n = std::min(got.size(), hitcount);
    if (n) memcpy(hits, got.data(), n * sizeof(ssa_amd_hit_t));
    *nhits = n;
    return 0;
}

int ssa_amd_autotune(int search_type, const size_t* lengths, size_t n, int copies, int* lanes_out, int* rows_out,
                     double* ms_out) {
    if (search_type != 0 && search_type != 1) return -1;
    if (n && !lengths) return -1;
    if (!matrix().valid) return -2;
    if (!matrix().constant && config().symtype == NUCLEOTIDE) return -2;
    std::vector<size_t> L(lengths, lengths + n);
    std::vector<AutotuneResult> res;
    if (copies < 1) return -1;
    const int k = autotune(search_type, L, copies, res);
    for (size_t i = 0; i < res.size(); ++i) {
        if (lanes_out) lanes_out[i] = res[i].lanes;
        if (rows_out) rows_out[i] = res[i].rows;
        if (ms_out) ms_out[i] = res[i].ms;
    }
    last_autotune() = res;
    return k;
}

void ssa_amd_set_task_trace(int on) { set_task_trace(on != 0); }

size_t ssa_amd_task_trace(ssa_amd_task_record_t* out, size_t max) {
    const std::vector<ssa_amd_task_record_t> r = last_task_trace();
    if (out) std::copy(r.begin(), r.begin() + std::min(max, r.size()), out);
    return r.size();
}

int ssa_amd_autotune_choice(size_t i, int* lanes_per_stream, int* rows_per_lane, int* columns_per_step,
                            int* profile_placement, double* ms) {
    const std::vector<AutotuneResult>& r = last_autotune();
    if (i >= r.size()) return -1;
    if (lanes_per_stream) *lanes_per_stream = r[i].lanes;
    if (rows_per_lane) *rows_per_lane = r[i].rows;
    if (columns_per_step) *columns_per_step = r[i].cols;
    if (profile_placement) *profile_placement = r[i].place;
    if (ms) *ms = r[i].ms;
    return 0;
}

void ssa_amd_autotune_reset(void) { autotune_reset(); }

int ssa_amd_selfcheck_replay(uint64_t seed, int cases) { return cases > 0 ? replay_selfcheck(seed, cases) : 0; }

int ssa_amd_db_pack(const char* path) { return write_packed_db(path); }

void ssa_amd_get_matrix(int64_t* out) { memcpy(out, matrix().s, sizeof(matrix().s)); }

const char* ssa_amd_device_arch(void) {
    static std::string arch;
    arch = Gpu::get().arch();
    return arch.c_str();
}

int ssa_amd_device_cus(void) { return Gpu::get().cu_count(); }

}  // extern "C"
