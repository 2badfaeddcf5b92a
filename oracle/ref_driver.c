/*
 * ref_driver.c -- calls into the reference's own compiled sources (built by
 * oracle/Makefile into oracle/_ref/libssaref.so) so tests can check the oracle
 * against the reference itself.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference's full library cannot be linked here: its DB adapter and the
 * per-width cascade drivers (src/db_adapter.c, src/algo/{8,16}/search_{8,16}.c,
 * src/algo/64/search_64.c) need the un-vendored libsdb.  This driver therefore
 * calls the reference's width kernels, 64-bit functions, heap, matrices and
 * traceback directly on chunks it builds itself.  Heap insertions made by the
 * kernels are recorded by wrapping add_to_minheap at link time; the wrapped
 * function still runs.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "libssa_datatypes.h"
#include "matrices.h"
#include "util/minheap.h"
#include "util/util.h"
#include "util/util_sequence.h"
#include "algo/8/search_8.h"
#include "algo/16/search_16.h"
#include "algo/64/search_64.h"
#include "algo/align.h"

/* The reference defines these two globals in src/libssa.c, which is not part
 * of this partial build; the driver hosts them. */
int8_t gapO = 0;
int8_t gapE = 0;

/* kernel entry points (declared with opaque pointers; defined in
 * src/algo/simd/search_simd_{sw,nw}.c and src/algo/{8,16}/search_*_util.c) */
p_s8info search_8_avx2_init(p_search_data sdp);
p_s8info search_8_sse41_init(p_search_data sdp);
p_s16info search_16_avx2_init(p_search_data sdp);
p_s16info search_16_sse2_init(p_search_data sdp);
void search_8_avx2_sw(p_s8info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
void search_8_avx2_nw(p_s8info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
void search_8_sse41_sw(p_s8info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
void search_8_sse41_nw(p_s8info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
void search_16_avx2_sw(p_s16info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
void search_16_avx2_nw(p_s16info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
void search_16_sse2_sw(p_s16info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
void search_16_sse2_nw(p_s16info, p_db_chunk, p_minheap, p_db_chunk, uint8_t);
region_t find_region_for_local(sequence_t a_seq, sequence_t b_seq);
cigar_p compute_cigar_string(int search_type, sequence_t a_seq, sequence_t b_seq, region_t region);

/* ---- recording of heap insertions ---- */
typedef struct { int64_t item; int32_t query; int64_t score; } ref_insert;
static ref_insert* g_rec = 0;
static size_t g_rec_n = 0, g_rec_cap = 0;

void __real_add_to_minheap(p_minheap heap, uint8_t query_id, p_sdb_sequence db_seq, long score);
void __wrap_add_to_minheap(p_minheap heap, uint8_t query_id, p_sdb_sequence db_seq, long score) {
    if (g_rec && g_rec_n < g_rec_cap) {
        g_rec[g_rec_n].item = (int64_t)db_seq->ID;
        g_rec[g_rec_n].query = query_id;
        g_rec[g_rec_n].score = score;
        g_rec_n++;
    }
    __real_add_to_minheap(heap, query_id, db_seq, score);
}

/* ---- matrices ---- */
void ref_matrix_constant(int match, int mismatch) { mat_init_constant_scoring((int8_t)match, (int8_t)mismatch); }
void ref_matrix_builtin(const char* name) { mat_init_buildin(name); }
void ref_matrix_string(const char* text) { mat_init_from_string(text); }
void ref_matrix_get(int64_t* out64, int8_t* out8, int16_t* out16) {
    memcpy(out64, score_matrix_64, 32 * 32 * sizeof(int64_t));
    memcpy(out8, score_matrix_8, 32 * 32);
    memcpy(out16, score_matrix_16, 32 * 32 * sizeof(int16_t));
}
void ref_set_gaps(int go, int ge) { gapO = (int8_t)go; gapE = (int8_t)ge; }

/* ---- symbol maps ---- */
void ref_maps(int8_t* aa, int8_t* nt) {
    memcpy(aa, map_ncbi_aa, 256);
    memcpy(nt, map_ncbi_nt16, 256);
}
void ref_revcomp(const char* in, size_t len, char* out) {
    sequence_t a = { (char*)in, len };
    sequence_t b = { out, len };
    us_revcompl(a, b);
}

/* ---- translation: the reference's tables and us_translate_sequence ---- */
void ref_translate(int db, int q_gencode, int d_gencode, const char* dna, size_t len, int strand, int frame,
                   char* out, size_t* out_len) {
    us_init_translation(q_gencode, d_gencode);
    sequence_t d = { (char*)dna, len };
    sequence_t prot = { NULL, 0 };
    us_translate_sequence(db, d, strand, frame, &prot);
    memcpy(out, prot.seq, prot.len);
    *out_len = prot.len;
    free(prot.seq);
}

/* ---- 64-bit ---- */
int64_t ref_full(int type, const char* d, size_t dlen, const char* q, size_t qlen) {
    sequence_t ds = { (char*)d, dlen };
    sequence_t qs = { (char*)q, qlen };
    int64_t* he = (int64_t*)calloc(2 * qlen + 2, sizeof(int64_t));
    int64_t r = type == 0 ? full_sw(&ds, &qs, he) : full_nw(&ds, &qs, he);
    free(he);
    return r;
}

/* ---- one width kernel over a list of items, for one query ----
 * isa: 1 = AVX2, 0 = SSE (SSE4.1 for 8 bit, SSE2 for 16 bit).
 * Writes the insertions (item, score) in order and the overflow list in order. */
size_t ref_lane_kernel(int type, int width, int isa, const char* const* seqs, const size_t* lens, size_t n,
                       const char* q, size_t qlen, int64_t* ins_item, int64_t* ins_score,
                       int64_t* ovf_item, size_t* n_ovf) {
    search_data_t sd;
    memset(&sd, 0, sizeof sd);
    sd.maxqlen = qlen;
    sd.q_count = 1;
    sd.queries[0].seq.seq = (char*)q;
    sd.queries[0].seq.len = qlen;

    db_chunk_t chunk, ovf;
    chunk.size = n;
    chunk.fill_pointer = n;
    chunk.seq = (p_sdb_sequence*)calloc(n + 1, sizeof(p_sdb_sequence));
    sdb_sequence_t* store = (sdb_sequence_t*)calloc(n + 1, sizeof(sdb_sequence_t));
    for (size_t i = 0; i < n; i++) {
        store[i].ID = i;
        store[i].seq.seq = (char*)seqs[i];
        store[i].seq.len = lens[i];
        chunk.seq[i] = &store[i];
    }
    ovf.size = n;
    ovf.fill_pointer = 0;
    ovf.seq = (p_sdb_sequence*)calloc(n + 1, sizeof(p_sdb_sequence));

    ref_insert* rec = (ref_insert*)calloc(n + 1, sizeof(ref_insert));
    g_rec = rec;
    g_rec_n = 0;
    g_rec_cap = n;
    p_minheap heap = minheap_init(n + 1);

    if (width == 8) {
        p_s8info s = isa ? search_8_avx2_init(&sd) : search_8_sse41_init(&sd);
        if (type == 0) (isa ? search_8_avx2_sw : search_8_sse41_sw)(s, &chunk, heap, &ovf, 0);
        else (isa ? search_8_avx2_nw : search_8_sse41_nw)(s, &chunk, heap, &ovf, 0);
    } else {
        p_s16info s = isa ? search_16_avx2_init(&sd) : search_16_sse2_init(&sd);
        if (type == 0) (isa ? search_16_avx2_sw : search_16_sse2_sw)(s, &chunk, heap, &ovf, 0);
        else (isa ? search_16_avx2_nw : search_16_sse2_nw)(s, &chunk, heap, &ovf, 0);
    }
    size_t ni = g_rec_n;
    for (size_t i = 0; i < ni; i++) {
        ins_item[i] = rec[i].item;
        ins_score[i] = rec[i].score;
    }
    for (size_t i = 0; i < ovf.fill_pointer; i++) ovf_item[i] = (int64_t)ovf.seq[i]->ID;
    *n_ovf = ovf.fill_pointer;
    g_rec = 0;
    minheap_exit(heap);
    free(rec);
    free(chunk.seq);
    free(ovf.seq);
    free(store);
    return ni;
}

/* ---- min-heap: feed elements in order, return the sorted array ---- */
size_t ref_heap_run(size_t k, const int64_t* ids, const int64_t* scores, const int32_t* qids, size_t n,
                    int64_t* out_ids, int64_t* out_scores, int32_t* out_q) {
    p_minheap h = minheap_init(k);
    for (size_t i = 0; i < n; i++) {
        elem_t e;
        memset(&e, 0, sizeof e);
        e.db_id = (size_t)ids[i];
        e.query_id = (uint8_t)qids[i];
        e.score = scores[i];
        minheap_add(h, &e);
    }
    minheap_sort(h);
    size_t c = h->count;
    for (size_t i = 0; i < c; i++) {
        out_ids[i] = (int64_t)h->array[i].db_id;
        out_scores[i] = h->array[i].score;
        out_q[i] = h->array[i].query_id;
    }
    minheap_exit(h);
    return c;
}

/* ---- region + CIGAR (a = query, b = db) ---- */
int ref_align(int type, const char* a, size_t alen, const char* b, size_t blen, int64_t* region4,
              char* cigar, size_t cap) {
    sequence_t as = { (char*)a, alen };
    sequence_t bs = { (char*)b, blen };
    region_t r;
    if (type == 0) r = find_region_for_local(as, bs);
    else r = init_region_for_global(as, bs);
    region4[0] = (int64_t)r.a_begin;
    region4[1] = (int64_t)r.a_end;
    region4[2] = (int64_t)r.b_begin;
    region4[3] = (int64_t)r.b_end;
    cigar_p c = compute_cigar_string(type, as, bs, r);
    size_t n = c->len < cap - 1 ? c->len : cap - 1;
    memcpy(cigar, c->cigar, n);
    cigar[n] = 0;
    free(c->cigar);
    free(c);
    return (int)n;
}

/* ---- CPU baseline: the reference's AVX2 kernels over a whole DB ----
 * T worker threads pull chunks of `chunk` sequences from a shared counter,
 * run the 8-bit kernel for the query, the 16-bit kernel on the chunk's
 * overflow and the 64-bit function on what overflows again, each into its own
 * heap; the heaps are merged in thread order and sorted.  This is the
 * structure of src/algo/manager.c:120-185, src/algo/searcher.c:155-174 and
 * src/algo/8/search_8.c:94-124 around the reference's own kernels.  Only the
 * SW search at 8 bit (the benchmark configuration) is driven.
 * Returns the wall time of the search in seconds; writes the hits. */
#include <pthread.h>
#include <time.h>

typedef struct {
    const char* codes;
    const int64_t* offs;
    const int64_t* lens;
    size_t n;
    const char* q;
    size_t qlen;
    size_t k;
    size_t chunk;
    size_t next;
    pthread_mutex_t mu;
} bench_ctx;

static void* bench_worker(void* arg) {
    bench_ctx* c = (bench_ctx*)arg;
    search_data_t sd;
    memset(&sd, 0, sizeof sd);
    sd.maxqlen = c->qlen;
    sd.q_count = 1;
    sd.queries[0].seq.seq = (char*)c->q;
    sd.queries[0].seq.len = c->qlen;
    p_s8info s8 = search_8_avx2_init(&sd);
    p_s16info s16 = search_16_avx2_init(&sd);
    p_minheap heap = minheap_init(c->k);
    sdb_sequence_t* store = (sdb_sequence_t*)calloc(c->chunk, sizeof(sdb_sequence_t));
    db_chunk_t chunk, ovf, ovf2;
    chunk.seq = (p_sdb_sequence*)calloc(c->chunk, sizeof(p_sdb_sequence));
    ovf.seq = (p_sdb_sequence*)calloc(c->chunk, sizeof(p_sdb_sequence));
    ovf2.seq = (p_sdb_sequence*)calloc(c->chunk, sizeof(p_sdb_sequence));
    chunk.size = ovf.size = ovf2.size = c->chunk;
    int64_t* he = (int64_t*)calloc(2 * c->qlen + 2, sizeof(int64_t));
    sequence_t qs = { (char*)c->q, c->qlen };
    for (;;) {
        pthread_mutex_lock(&c->mu);
        size_t start = c->next;
        c->next += c->chunk;
        pthread_mutex_unlock(&c->mu);
        if (start >= c->n) break;
        size_t stop = start + c->chunk < c->n ? start + c->chunk : c->n;
        chunk.fill_pointer = 0;
        for (size_t i = start; i < stop; i++) {
            if (c->lens[i] == 0) continue;
            sdb_sequence_t* s = &store[chunk.fill_pointer];
            s->ID = i;
            s->seq.seq = (char*)c->codes + c->offs[i];
            s->seq.len = (size_t)c->lens[i];
            s->strand = s->frame = 0;
            chunk.seq[chunk.fill_pointer++] = s;
        }
        ovf.fill_pointer = 0;
        search_8_avx2_sw(s8, &chunk, heap, &ovf, 0);
        if (ovf.fill_pointer) {
            ovf2.fill_pointer = 0;
            search_16_avx2_sw(s16, &ovf, heap, &ovf2, 0);
            for (size_t i = 0; i < ovf2.fill_pointer; i++) {
                long sc = (long)full_sw(&ovf2.seq[i]->seq, &qs, he);
                add_to_minheap(heap, 0, ovf2.seq[i], sc);
            }
        }
    }
    free(he);
    free(store);
    free(chunk.seq);
    free(ovf.seq);
    free(ovf2.seq);
    return heap;
}

double ref_bench_sw8(const char* codes, const int64_t* offs, const int64_t* lens, size_t n, const char* q,
                     size_t qlen, size_t k, size_t chunk, int threads, int64_t* out_ids, int64_t* out_scores,
                     size_t* out_n) {
    bench_ctx c;
    c.codes = codes;
    c.offs = offs;
    c.lens = lens;
    c.n = n;
    c.q = q;
    c.qlen = qlen;
    c.k = k;
    c.chunk = chunk;
    c.next = 0;
    pthread_mutex_init(&c.mu, NULL);
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, bench_worker, &c);
    p_minheap all = minheap_init(k);
    for (int t = 0; t < threads; t++) {
        void* r;
        pthread_join(th[t], &r);
        p_minheap h = (p_minheap)r;
        for (size_t j = 0; j < h->count; j++) minheap_add(all, &h->array[j]);
        minheap_exit(h);
    }
    minheap_sort(all);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    *out_n = all->count;
    for (size_t i = 0; i < all->count; i++) {
        out_ids[i] = (int64_t)all->array[i].db_id;
        out_scores[i] = all->array[i].score;
    }
    minheap_exit(all);
    free(th);
    pthread_mutex_destroy(&c.mu);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
