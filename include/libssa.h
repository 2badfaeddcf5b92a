/*
 * libssa.h -- public C API of the MI355X implementation of libssa's
 * inter-sequence Smith-Waterman / Needleman-Wunsch database search.
 *
 * Drop-in replacement for the reference header src/libssa.h: the same
 * constants, struct layouts and function signatures (src/libssa.h:29-263),
 * so code written against the reference compiles and links unchanged.  Each
 * declaration names the reference declaration it replaces.
 *
 * Searches run on the GPU (HIP, gfx950).  Results -- scores, hit selection and
 * order, CIGAR strings -- are those of the reference run with one thread.
 */
#ifndef LIBSSA_H_
#define LIBSSA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (src/libssa.h:29-72) ---- */
#define VERBOSE_ON 1
#define VERBOSE_OFF 0

#define SIMD_ON 1
#define SIMD_OFF 0

#define BLOSUM45 "blosum45"
#define BLOSUM50 "blosum50"
#define BLOSUM62 "blosum62"
#define BLOSUM80 "blosum80"
#define BLOSUM90 "blosum90"
#define PAM30 "pam30"
#define PAM70 "pam70"
#define PAM250 "pam250"

#define NUCLEOTIDE 0
#define AMINOACID 1
#define TRANS_QUERY 2
#define TRANS_DB 3
#define TRANS_BOTH 4

#define FORWARD_STRAND 1
#define COMPLEMENTARY_STRAND 2
#define BOTH_STRANDS 3

#define BIT_WIDTH_8 8
#define BIT_WIDTH_16 16
#define BIT_WIDTH_64 64

#define OUTPUT_SILENT 0
#define OUTPUT_ERROR 1
#define OUTPUT_WARNING 2
#define OUTPUT_INFO 3

#define COMPUTE_SCORE 0
#define COMPUTE_ALIGNMENT 1

#define READ_FROM_FILE 0
#define READ_FROM_STRING 1
#define MATRIX_BUILDIN 2

#define COMPUTE_ON_SSE2 0
#define COMPUTE_ON_SSE41 1
#define COMPUTE_ON_AVX2 2

/* ---- data types (src/libssa.h:77-117) ---- */
struct _query;
typedef struct _query* p_query; /* opaque (src/libssa.h:79-80) */

typedef struct {
    char* seq;
    size_t len;
    size_t ID;
    int strand;
    int frame;
} db_seq_t;

typedef struct {
    char* seq;
    size_t len;
    int strand;
    int frame;
} q_seq_t;

typedef struct {
    db_seq_t db_seq;
    q_seq_t query;
    char* alignment;
    size_t alignment_len;
    long score;
    size_t align_q_start;
    size_t align_q_end;
    size_t align_d_start;
    size_t align_d_end;
} alignment_t;
typedef alignment_t* p_alignment;

typedef struct {
    p_alignment* alignments;
    size_t len;
} alignment_list_t;
typedef alignment_list_t* p_alignment_list;

/* ---- technical initialisation (src/libssa.h:122-128) ---- */
/* src/libssa.h:122 */
void set_output_mode(int mode);
/* src/libssa.h:124.  Selects the reference lane width whose schedule the hit
 * selection reproduces: COMPUTE_ON_AVX2 (32/16 lanes) or the SSE widths (16/8). */
void set_simd_compute_mode(int mode);
/* src/libssa.h:126.  DB sequences per chunk; shapes tie handling as in the reference. */
void set_chunk_size(size_t size);
/* src/libssa.h:128.  Accepted for compatibility; the GPU replaces the thread pool and
 * results are always those of one reference thread. */
void set_thread_count(size_t count);

/* ---- initialisation (src/libssa.h:153-213) ---- */
/* src/libssa.h:153 */
void init_score_matrix(int mode, const char* matrix);
/* src/libssa.h:164.  The first argument is the match score, the second the mismatch
 * score, as the reference implementation uses them (src/libssa.c:115-117). */
void init_constant_scores(const int8_t p, const int8_t m);
/* src/libssa.h:172 */
void init_gap_penalties(const int8_t gapO, const int8_t gapE);
/* src/libssa.h:195 */
void init_symbol_translation(int type, int strands, int db_gencode, int q_gencode);
/* src/libssa.h:209 */
void init_db(const char* db_file);
/* src/libssa.h:217 */
p_query init_sequence_fasta(int mode, const char* fasta_seq_file);
/* src/libssa.h:227 */
void free_sequence(p_query p);

/* ---- alignment (src/libssa.h:236-263) ---- */
/* src/libssa.h:240 */
p_alignment_list sw_align(p_query p, size_t hitcount, int bit_width, int align_type);
/* src/libssa.h:250 */
p_alignment_list nw_align(p_query p, size_t hitcount, int bit_width, int align_type);
/* src/libssa.h:261 */
void free_alignment(p_alignment_list alist);
/* src/libssa.h:263 */
void ssa_exit(void);

#ifdef __cplusplus
}
#endif

#endif /* LIBSSA_H_ */
