// common.h -- shared definitions of the libssa_amd host runtime.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

extern "C" {
#include "libssa.h"
#include "libssa_extern_db.h"
}

namespace ssa {

// search types (the reference's SMITH_WATERMAN / NEEDLEMAN_WUNSCH, src/util/util.h:36-37)
enum SearchType { kSW = 0, kNW = 1 };

// DB sequences per chunk by default (src/util/util.h:50)
constexpr size_t kDefaultChunkSize = 1000;
// columns per lane block of the reference's kernels (src/algo/8/search_8_util.h:25)
constexpr int kCDepth = 4;
constexpr int kDim = 32;

// ---- messages and fatal errors (src/util/util.c:36-89) ----
[[noreturn]] void fatal(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void print_info(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void print_warning(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void print_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

// ---- process-wide configuration set through libssa.h ----
struct Config {
    int output_mode = OUTPUT_WARNING;
    int simd_mode = COMPUTE_ON_AVX2;  // selects the reference lane counts replayed
    size_t chunk_size = kDefaultChunkSize;
    size_t thread_count = (size_t)-1;
    int8_t gap_open = 0, gap_extend = 0;
    int symtype = AMINOACID;
    int strands = FORWARD_STRAND;
    int q_gencode = 1, d_gencode = 1;
};
Config& config();

// Runs fn(i) for every i in [0, n) on the library's worker threads and the
// calling thread, and returns when all have run.  The workers are started
// once and shared by concurrent callers; the caller works through its own
// items too, so a call always finishes even when every worker is busy.
void parallel_for(size_t n, const std::function<void(size_t)>& fn);
// threads a parallel_for spreads over (workers + caller, at most 16)
size_t parallel_width();

// Large host blocks for per-search arrays that are written before they are
// read.  Released blocks stay mapped in a process-wide cache (up to 4 GiB) and
// serve later requests of up to their size, so repeated searches neither
// fault in nor zero fresh pages.
void* host_block_acquire(size_t bytes, size_t* capacity);
void host_block_release(void* p, size_t capacity);

// An array of trivially copyable T over a cached host block; elements start
// unset.
template <class T>
class HostArray {
public:
    HostArray() = default;
    HostArray(const HostArray&) = delete;
    HostArray& operator=(const HostArray&) = delete;
    HostArray(HostArray&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) { o.p_ = nullptr, o.n_ = o.cap_ = 0; }
    HostArray& operator=(HostArray&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_, n_ = o.n_, cap_ = o.cap_;
            o.p_ = nullptr, o.n_ = o.cap_ = 0;
        }
        return *this;
    }
    ~HostArray() { release(); }
    // n unset elements (previous contents are not kept)
    void reset(size_t n) {
        if (n * sizeof(T) > cap_) {
            release();
            p_ = (T*)host_block_acquire(n * sizeof(T), &cap_);
        }
        n_ = n;
    }
    void fill_zero() {
        if (n_) memset((void*)p_, 0, n_ * sizeof(T));
    }
    T& operator[](size_t i) { return p_[i]; }
    const T& operator[](size_t i) const { return p_[i]; }
    T* data() { return p_; }
    const T* data() const { return p_; }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }

private:
    void release() {
        if (p_) host_block_release(p_, cap_);
        p_ = nullptr, n_ = cap_ = 0;
    }
    T* p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
};

// lanes of the reference's 8-bit and 16-bit kernels for the configured ISA
// (src/algo/8/search_8_util.h:26-35, src/algo/16/search_16_util.h:27-37)
int lanes8();
int lanes16();

// ---- score matrix (src/matrices.c) ----
struct Matrix {
    int64_t s[kDim * kDim];  // [db code * 32 + query code]
    bool constant = false;
    bool valid = false;
    int64_t at(int d, int q) const { return s[d * kDim + q]; }
};
Matrix& matrix();
void matrix_init_constant(int match, int mismatch);
void matrix_init_text(const char* text);
void matrix_init_file(const char* path);
void matrix_init_builtin(const char* name);
void matrix_free();

// ---- symbols (src/util/util_sequence.c) ----
int map_aa(int c);
int map_nt(int c);
int complement_nt(int code);
void revcomp(const uint8_t* in, size_t len, uint8_t* out);
bool gencode_valid(int code);
// DB sequences are nucleotides for NUCLEOTIDE, TRANS_DB, TRANS_BOTH; queries
// for NUCLEOTIDE, TRANS_QUERY, TRANS_BOTH (src/query.c:48-55, src/db_adapter.c:47-110)
inline bool db_is_translated(int symtype) { return symtype == TRANS_DB || symtype == TRANS_BOTH; }
inline bool query_is_translated(int symtype) { return symtype == TRANS_QUERY || symtype == TRANS_BOTH; }
// the codon table of a genetic code: [256 a + 16 b + c] = amino-acid code of
// the nt16 codon (a, b, c) (src/util/util_sequence.c:200-253)
const int8_t* translation_table(int gencode);
// us_translate_sequence (src/util/util_sequence.c:312-353): the protein of
// strand 0 or, for any other strand value, of the reverse complement, from
// frame 0..2.  A sequence shorter than its frame exits as the reference does.
void translate(const int8_t* table, const uint8_t* dna, size_t len, int strand, int frame, std::vector<uint8_t>& out);

}  // namespace ssa

// the opaque query handle of libssa.h (fields as src/libssa_datatypes.h:95-107)
struct _query {
    std::vector<uint8_t> nt[2];  // mapped nucleotide strands
    std::vector<uint8_t> aa[6];  // mapped amino-acid frames
    bool has_nt[2] = {false, false};
    bool has_aa[6] = {false, false, false, false, false, false};
    std::string header;
};
