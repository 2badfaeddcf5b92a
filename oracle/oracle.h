/*
 * oracle.h -- CPU restatement of libssa's inter-sequence SW/NW search path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker.
 *
 * Every function states the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  The code is written
 * from the behaviour, as scalar C over explicit per-lane state; it is not a
 * translation of the reference's SIMD source.
 *
 * Parity status of each part is recorded in oracle.c's header comment.
 */
#ifndef LIBSSA_ORACLE_H
#define LIBSSA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DIM 32

/* search types (src/util/util.h:36-37) */
#define ORC_SW 0
#define ORC_NW 1

/* symbol types (src/libssa.h:44-48) */
#define ORC_NUCLEOTIDE 0
#define ORC_AMINOACID 1
#define ORC_TRANS_QUERY 2
#define ORC_TRANS_DB 3
#define ORC_TRANS_BOTH 4

typedef struct {
    int64_t s[ORC_DIM * ORC_DIM]; /* [db_code * 32 + query_code], as SCORE_MATRIX_64 */
    int constant;
} orc_matrix;

/* ---- score matrices (src/matrices.c:335-460) ---- */
void orc_matrix_constant(orc_matrix* m, int match, int mismatch);
/* Parses the text table format read by mat_init_from_string; returns 0 on success. */
int orc_matrix_from_text(orc_matrix* m, const char* text);

/* ---- symbol maps (src/util/util_sequence.c:44-88, 298-318, 427-437) ---- */
int orc_map_aa(int c);             /* -1 for unknown */
int orc_map_nt(int c);             /* -1 for unknown */
int orc_complement(int code);      /* nt16 complement */
/* DB mapping: unknown -> 0 (us_map_sequence). */
void orc_map_db(int symtype, const char* raw, size_t len, uint8_t* out);
/* Query mapping: unknown symbols removed (src/query.c:102-130). Returns new length. */
size_t orc_map_query(int symtype, const char* raw, size_t len, uint8_t* out);
void orc_revcomp(const uint8_t* in, size_t len, uint8_t* out);

/* ---- translation (src/util/util_sequence.c:190-353) ---- */
/* 1 if the NCBI genetic code number is one the reference names
 * (src/util/util_sequence.c:95-123, checked by src/libssa.c:139-145). */
int orc_gencode_valid(int gencode);
/* table[256*a + 16*b + c] = amino-acid code of the nt16 codon (a, b, c); each
 * nt16 code is an IUPAC base set, and a codon of several bases translates to
 * the one amino acid all its codons give, B for a mix of D and N, Z for a mix
 * of E and Q, X otherwise (init_translate_table, :200-253). */
void orc_translate_table(int gencode, int8_t* table4096);
/* us_translate_sequence (:312-353): the protein of strand 0 (forward) or, for
 * any other strand value, the reverse complement, read from frame 0..2.
 * Writes (len - frame) / 3 codes and returns that length; returns (size_t)-1
 * when len < frame (the reference then fails to allocate and exits). */
size_t orc_translate(const int8_t* table, const uint8_t* dna, size_t len, int strand, int frame,
                     uint8_t* out);

/* ---- 64-bit tier (src/algo/64/smith_waterman_63.c:32-98, needleman_wunsch_64.c:32-98) ---- */
int64_t orc_full_sw(const uint8_t* d, size_t dlen, const uint8_t* q, size_t qlen,
                    const orc_matrix* m, int gap_open, int gap_extend);
int64_t orc_full_nw(const uint8_t* d, size_t dlen, const uint8_t* q, size_t qlen,
                    const orc_matrix* m, int gap_open, int gap_extend);

/* ---- one width kernel over a list of items with a fixed lane count ----
 * Restates search_{8,16}_{avx2,sse*}_{sw,nw} (src/algo/simd/search_simd_sw.c:279-441,
 * src/algo/simd/search_simd_nw.c:314-516) with scalar saturating arithmetic.
 * Writes events in the order the reference emits them.  flag_block[i] receives
 * the 0-based block at which item i was flagged as overflowed, or -1. */
typedef struct {
    int32_t item;   /* index into the item list */
    int32_t kind;   /* 0 = heap insert, 1 = appended to the overflow chunk */
    int64_t score;  /* valid for kind 0 */
} orc_event;

size_t orc_lane_kernel(int type, int width, int lanes,
                       const uint8_t* const* seqs, const size_t* lens, size_t n,
                       const uint8_t* q, size_t qlen, const orc_matrix* m,
                       int gap_open, int gap_extend,
                       orc_event* events, int32_t* flag_block);

/* ---- bounded min-heap (src/util/minheap.c:30-106, src/util/util.c:108-124) ---- */
typedef struct {
    uint64_t db_id;
    uint8_t db_frame;
    uint8_t db_strand;
    uint8_t query_id;
    int64_t score;
} orc_elem;

typedef struct {
    size_t alloc, count;
    orc_elem* a;
} orc_heap;

void orc_heap_init(orc_heap* h, orc_elem* storage, size_t alloc);
void orc_heap_add(orc_heap* h, const orc_elem* e);
/* Stable sort by (score desc, db_id desc), as minheap_sort under glibc's merge sort. */
void orc_heap_sort(orc_heap* h);

/* ---- full search (src/algo/manager.c:120-185 with one thread) ---- */
typedef struct {
    int type;          /* ORC_SW / ORC_NW */
    int bit_width;     /* 8, 16, 64 */
    int symtype;       /* ORC_NUCLEOTIDE / ORC_AMINOACID */
    int strands;       /* 1 forward, 2 complementary, 3 both (query_strands) */
    int lanes8;        /* 32 (AVX2) or 16 (SSE4.1) */
    int lanes16;       /* 16 (AVX2) or 8 (SSE2) */
    size_t chunk_size; /* DB sequences per chunk (set_chunk_size) */
    size_t hit_count;
    int gap_open, gap_extend;
    const orc_matrix* matrix;
    /* queries, already mapped: q_count <= 6, each with strand/frame tags */
    int q_count;
    const uint8_t* q_seq[6];
    size_t q_len[6];
    /* DB: raw (unmapped) sequences in ID order */
    size_t db_count;
    const char* const* db_raw;
    const size_t* db_len;
    /* TRANS_DB / TRANS_BOTH: genetic code of the DB sequences */
    int d_gencode;
} orc_search_params;

typedef struct {
    size_t chunk_count, seq_count;
    size_t overflow_8, overflow_16;
} orc_stats;

/* Returns the number of hits written to out (<= hit_count), sorted as the
 * reference's final list, or (size_t)-1 when a DB sequence is shorter than a
 * frame it is translated in (the reference exits there). */
size_t orc_search(const orc_search_params* p, orc_elem* out, orc_stats* stats);

/* ---- region, directions, CIGAR (src/algo/align.c:39-150, src/algo/cigar.c:48-346) ---- */
typedef struct {
    int64_t a_begin, a_end, b_begin, b_end;
} orc_region;

/* a = query, b = database sequence (as align_sequences passes them). */
int orc_region_local(const uint8_t* a, size_t alen, const uint8_t* b, size_t blen,
                     const orc_matrix* m, int gap_open, int gap_extend, orc_region* r);
void orc_region_global(size_t alen, size_t blen, orc_region* r);
/* Writes the NUL-terminated CIGAR into out (capacity cap); returns its length. */
size_t orc_cigar(int type, const uint8_t* a, size_t alen, const uint8_t* b, size_t blen,
                 const orc_region* r, const orc_matrix* m, int gap_open, int gap_extend,
                 char* out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif
