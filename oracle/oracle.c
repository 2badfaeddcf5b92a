/*
 * oracle.c -- CPU restatement of libssa's inter-sequence SW/NW search path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Not linked into the product.
 *
 * Parity status (how each part is pinned; see tests/test_oracle_*.py):
 *   orc_matrix_*        pinned: reference test spot values (tests/test_matrices.c:46-111)
 *                       and entry-by-entry against the compiled reference (oracle/_ref).
 *   orc_map_* / revcomp pinned: symbol tables cross-checked against oracle/_ref; reverse
 *                       complement involution + IUPAC closed form.
 *   orc_full_sw/nw      pinned: known answers (tests/algo/64/test_*.c), brute-force Gotoh
 *                       enumeration on tiny inputs, oracle/_ref on random inputs.
 *   orc_lane_kernel     pinned: oracle/_ref (the reference's own AVX2/SSE kernels) on
 *                       random inputs including saturating and padding corner cases;
 *                       overflow counts of tests/algo/8/test_8_simd_*.c.
 *   orc_heap_*          pinned: tests/util/test_minheap.c:33-90, oracle/_ref heap.
 *   orc_search          pinned: tests/algo/test_searcher.c, tests/test_bigger_databases.c:70-92,
 *                       tests/algo/64/test_search_64.c, oracle/_ref driver; translated
 *                       queries by the TRANS_QUERY goldens of tests/algo/test_searcher.c:248-330.
 *   orc_translate*      pinned: every table entry of every genetic code and random
 *                       sequences against oracle/_ref (the reference's util_sequence.c);
 *                       textbook codons and IUPAC ambiguity cases.
 *   orc_region/cigar    pinned: tests/algo/test_cigar.c, test_align.c, test_aligner.c,
 *                       tests/test_libssa.c:48-93 CIGAR strings, oracle/_ref.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* ------------------------------------------------------------------------ */
/* score matrices                                                            */
/* ------------------------------------------------------------------------ */

/* src/matrices.c:335-345: every entry starts at -1. */
static void matrix_reset(orc_matrix* m) {
    for (int i = 0; i < ORC_DIM * ORC_DIM; i++) m->s[i] = -1;
    m->constant = 0;
}

/* src/matrices.c:447-460: only codes >= 1 are filled; code 0 keeps -1. */
void orc_matrix_constant(orc_matrix* m, int match, int mismatch) {
    matrix_reset(m);
    m->constant = 1;
    for (int a = 1; a < ORC_DIM; a++)
        for (int b = 1; b < ORC_DIM; b++)
            m->s[a * ORC_DIM + b] = (a == b) ? match : mismatch;
}

/* src/matrices.c:388-437 (read_line) driven line by line as mat_init_from_string
 * (src/matrices.c:497-531) does.  Column order comes from a line starting with
 * blank or tab; rows start with their symbol; '#' and empty lines are skipped.
 * The column counter accumulates across header lines, as the reference's does. */
int orc_matrix_from_text(orc_matrix* m, const char* text) {
    matrix_reset(m);
    int order[2048];
    int symbols = 0;
    const char* s = text;
    while (*s) {
        const char* nl = strchr(s, '\n');
        size_t len = nl ? (size_t)(nl - s) : strlen(s);
        char line[2048];
        if (len >= sizeof line) return -1;
        memcpy(line, s, len);
        line[len] = 0;
        const char* p = line;
        char c = *p++;
        if (c == '#' || c == '\n') {
            /* comment */
        } else if (c == ' ' || c == '\t') {
            int k = 0;
            char x;
            while ((x = *p++)) {
                if (x != ' ' && x != '\t' && x != '\n') {
                    order[k++] = orc_map_aa((unsigned char)x);
                    symbols++;
                }
            }
        } else {
            int a = orc_map_aa((unsigned char)c);
            for (int i = 0; i < symbols; i++) {
                char* end;
                long v = strtol(p, &end, 10);
                if (end == p) return -2; /* the reference calls fatal() here */
                p = end;
                int b = order[i];
                if (a >= 0 && b >= 0 && a < ORC_DIM && b < ORC_DIM) m->s[a * ORC_DIM + b] = v;
            }
        }
        s = nl ? nl + 1 : s + len;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* symbols                                                                   */
/* ------------------------------------------------------------------------ */

/* src/util/util_sequence.c:40-64: code = position in "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";
 * '-' itself is not accepted for amino acids, '*' is. */
int orc_map_aa(int c) {
    static const char* sym = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";
    if (c == '*') return 25;
    if (c >= 'a' && c <= 'z') c -= 32;
    if (c < 'A' || c > 'Z') return -1;
    const char* f = strchr(sym + 1, c);
    return f ? (int)(f - sym) : -1;
}

/* src/util/util_sequence.c:66-88: code = position in "-ACMGRSVTWYHKDBN", U as T.
 * The code is the IUPAC bit set (A=1, C=2, G=4, T=8). */
int orc_map_nt(int c) {
    static const char* sym = "-ACMGRSVTWYHKDBN";
    if (c == '-') return 0;
    if (c >= 'a' && c <= 'z') c -= 32;
    if (c == 'U') return 8;
    if (c < 'A' || c > 'Z') return -1;
    const char* f = strchr(sym + 1, c);
    return f ? (int)(f - sym) : -1;
}

/* src/util/util_sequence.c:174: complement swaps A<->T and C<->G bits. */
int orc_complement(int code) {
    return ((code & 1) << 3) | ((code & 2) << 1) | ((code & 4) >> 1) | ((code & 8) >> 3);
}

/* DB sequences are nucleotides for NUCLEOTIDE, TRANS_DB and TRANS_BOTH
 * (src/db_adapter.c:47-110); queries for NUCLEOTIDE, TRANS_QUERY and
 * TRANS_BOTH (src/query.c:48-55). */
static int db_is_nt(int symtype) {
    return symtype == ORC_NUCLEOTIDE || symtype == ORC_TRANS_DB || symtype == ORC_TRANS_BOTH;
}
static int query_is_nt(int symtype) { return symtype != ORC_AMINOACID && symtype != ORC_TRANS_DB; }

/* src/util/util_sequence.c:298-318 */
void orc_map_db(int symtype, const char* raw, size_t len, uint8_t* out) {
    for (size_t i = 0; i < len; i++) {
        int v = db_is_nt(symtype) ? orc_map_nt((unsigned char)raw[i]) : orc_map_aa((unsigned char)raw[i]);
        out[i] = (uint8_t)(v < 0 ? 0 : v);
    }
}

/* src/query.c:102-130 */
size_t orc_map_query(int symtype, const char* raw, size_t len, uint8_t* out) {
    size_t k = 0;
    for (size_t i = 0; i < len; i++) {
        int v = query_is_nt(symtype) ? orc_map_nt((unsigned char)raw[i]) : orc_map_aa((unsigned char)raw[i]);
        if (v >= 0) out[k++] = (uint8_t)v;
    }
    return k;
}

/* src/util/util_sequence.c:427-437 */
void orc_revcomp(const uint8_t* in, size_t len, uint8_t* out) {
    for (size_t i = 0; i < len; i++) out[i] = (uint8_t)orc_complement(in[len - 1 - i]);
}

/* ------------------------------------------------------------------------ */
/* translation                                                               */
/* ------------------------------------------------------------------------ */

/* The NCBI genetic codes the reference names (src/util/util_sequence.c:95-123),
 * each as the 64 amino acids of the codons in TCAG order with the first base
 * slowest (TTT TTC TTA TTG TCT ... GGG), from NCBI's published translation
 * tables.  Index = code number - 1; NULL where NCBI has no such code. */
static const char* const kGencodes[23] = {
    /*  1 */ "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /*  2 */ "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNKKSS**VVVVAAAADDEEGGGG",
    /*  3 */ "FFLLSSSSYY**CCWWTTTTPPPPHHQQRRRRIIMMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /*  4 */ "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /*  5 */ "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNKKSSSSVVVVAAAADDEEGGGG",
    /*  6 */ "FFLLSSSSYYQQCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /*  7 */ NULL,
    /*  8 */ NULL,
    /*  9 */ "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIIMTTTTNNNKSSSSVVVVAAAADDEEGGGG",
    /* 10 */ "FFLLSSSSYY**CCCWLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /* 11 */ "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /* 12 */ "FFLLSSSSYY**CC*WLLLSPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /* 13 */ "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNKKSSGGVVVVAAAADDEEGGGG",
    /* 14 */ "FFLLSSSSYYY*CCWWLLLLPPPPHHQQRRRRIIIMTTTTNNNKSSSSVVVVAAAADDEEGGGG",
    /* 15 */ "FFLLSSSSYY*QCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /* 16 */ "FFLLSSSSYY*LCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /* 17 */ NULL,
    /* 18 */ NULL,
    /* 19 */ NULL,
    /* 20 */ NULL,
    /* 21 */ "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNNKSSSSVVVVAAAADDEEGGGG",
    /* 22 */ "FFLLSS*SYY*LCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
    /* 23 */ "FF*LSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",
};

int orc_gencode_valid(int gencode) { return gencode >= 1 && gencode <= 23 && kGencodes[gencode - 1] != NULL; }

/* src/util/util_sequence.c:200-253.  The nt16 bit of a base: A=1, C=2, G=4,
 * T=8; its TCAG position: T=0, C=1, A=2, G=3. */
void orc_translate_table(int gencode, int8_t* table) {
    static const int tcag[4] = {2, 1, 3, 0}; /* bit index 0..3 (A, C, G, T) -> TCAG position */
    const char* code = kGencodes[gencode - 1];
    for (int a = 0; a < 16; a++)
        for (int b = 0; b < 16; b++)
            for (int c = 0; c < 16; c++) {
                /* the set of amino acids over every codon the three base sets allow */
                int seen[128] = {0};
                int nseen = 0;
                for (int i = 0; i < 4; i++)
                    for (int j = 0; j < 4; j++)
                        for (int k = 0; k < 4; k++)
                            if ((a >> i & 1) && (b >> j & 1) && (c >> k & 1)) {
                                int x = (unsigned char)code[16 * tcag[i] + 4 * tcag[j] + tcag[k]];
                                if (!seen[x]) { seen[x] = 1; nseen++; }
                            }
                char aa = 'X';
                if (nseen == 1) {
                    for (int x = 0; x < 128; x++) if (seen[x]) aa = (char)x;
                } else if (nseen == 2 && seen['D'] && seen['N']) {
                    aa = 'B';
                } else if (nseen == 2 && seen['E'] && seen['Q']) {
                    aa = 'Z';
                }
                table[256 * a + 16 * b + c] = (int8_t)orc_map_aa(aa);
            }
}

/* src/util/util_sequence.c:312-353 */
size_t orc_translate(const int8_t* table, const uint8_t* dna, size_t len, int strand, int frame, uint8_t* out) {
    if (len < (size_t)frame) return (size_t)-1;
    size_t plen = (len - (size_t)frame) / 3;
    for (size_t p = 0; p < plen; p++) {
        int c[3];
        for (int k = 0; k < 3; k++) {
            size_t i = (size_t)frame + 3 * p + (size_t)k; /* position on the strand read */
            c[k] = strand == 0 ? dna[i] : orc_complement(dna[len - 1 - i]);
        }
        out[p] = (uint8_t)table[256 * c[0] + 16 * c[1] + c[2]];
    }
    return plen;
}

/* ------------------------------------------------------------------------ */
/* 64-bit Gotoh                                                              */
/* ------------------------------------------------------------------------ */

#define SC(m, d, q) ((m)->s[(int)(d) * ORC_DIM + (int)(q)])

/* src/algo/64/smith_waterman_63.c:32-98: columns are DB positions (outer loop),
 * rows query positions; E/F start at 0, H floors at 0, score = max H. */
int64_t orc_full_sw(const uint8_t* d, size_t dlen, const uint8_t* q, size_t qlen,
                    const orc_matrix* m, int go, int ge) {
    int64_t* hh = (int64_t*)calloc(qlen + 1, sizeof(int64_t));
    int64_t* ee = (int64_t*)calloc(qlen + 1, sizeof(int64_t));
    int64_t best = 0;
    for (size_t j = 0; j < dlen; j++) {
        int64_t f = 0, diag = 0;
        for (size_t i = 0; i < qlen; i++) {
            int64_t up_left = hh[i];
            int64_t e = ee[i];
            int64_t h = diag + SC(m, d[j], q[i]);
            if (e > h) h = e;
            if (f > h) h = f;
            if (h < 0) h = 0;
            if (h > best) best = h;
            hh[i] = h;
            e += ge;
            f += ge;
            int64_t open = h + go + ge;
            if (open > e) e = open;
            if (open > f) f = open;
            ee[i] = e;
            diag = up_left;
        }
    }
    free(hh);
    free(ee);
    return best;
}

/* src/algo/64/needleman_wunsch_64.c:32-98: left column H(i,-1) = go+(i+1)ge,
 * E(i,0) = 2go+(i+2)ge, top F(0,j) = 2go+(j+2)ge, corner diag (j==0)?0:go+j*ge;
 * the score is H(qlen-1, dlen-1). */
int64_t orc_full_nw(const uint8_t* d, size_t dlen, const uint8_t* q, size_t qlen,
                    const orc_matrix* m, int go, int ge) {
    int64_t* hh = (int64_t*)calloc(qlen + 1, sizeof(int64_t));
    int64_t* ee = (int64_t*)calloc(qlen + 1, sizeof(int64_t));
    for (size_t i = 0; i < qlen; i++) {
        hh[i] = go + (int64_t)(i + 1) * ge;
        ee[i] = 2 * (int64_t)go + (int64_t)(i + 2) * ge;
    }
    for (size_t j = 0; j < dlen; j++) {
        int64_t f = 2 * (int64_t)go + (int64_t)(j + 2) * ge;
        int64_t diag = (j == 0) ? 0 : go + (int64_t)j * ge;
        for (size_t i = 0; i < qlen; i++) {
            int64_t up_left = hh[i];
            int64_t e = ee[i];
            int64_t h = diag + SC(m, d[j], q[i]);
            if (f > h) h = f;
            if (e > h) h = e;
            hh[i] = h;
            e += ge;
            f += ge;
            int64_t open = h + go + ge;
            if (f < open) f = open;
            if (e < open) e = open;
            ee[i] = e;
            diag = up_left;
        }
    }
    int64_t r = qlen ? hh[qlen - 1] : 0;
    free(hh);
    free(ee);
    return r;
}

/* ------------------------------------------------------------------------ */
/* width kernels with a fixed number of lanes                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int width;
    int64_t imin, imax, uimax;
} wspec;

static inline int64_t sat(const wspec* w, int64_t v) { return v < w->imin ? w->imin : (v > w->imax ? w->imax : v); }
/* conversion of an int to the lane type (wraps, as the reference's set1/stores do) */
static inline int64_t wrap(const wspec* w, int64_t v) { return w->width == 8 ? (int64_t)(int8_t)v : (int64_t)(int16_t)v; }
static inline int64_t mx(int64_t a, int64_t b) { return a > b ? a : b; }
static inline int64_t mn(int64_t a, int64_t b) { return a < b ? a : b; }

#define CDEPTH 4

typedef struct {
    int64_t *H, *E;       /* per-row H of the previous column and E entering the next */
    int item;             /* -1 = empty lane */
    size_t pos, end;      /* position in the item's sequence */
    int blocks;           /* blocks computed for the current item */
    int ovf;
    int newly;            /* lane was (re)assigned in this change iteration */
    int64_t S;            /* SW: running maximum */
    int64_t Sm[4];        /* NW: last-row H of the four columns of the block */
    int64_t Hb[4], Fb[4]; /* NW: top boundary H and F for the four columns */
    int64_t hmin, hmax;   /* NW: overflow trackers */
    uint8_t win[CDEPTH];
} lane_t;

/* move_db_sequence_window_{8,16} (src/algo/8/search_8_util.h:60-74): take the
 * next four residues (0 past the end); returns 1 when the sequence is used up. */
static int move_window(lane_t* L, const uint8_t* seq) {
    for (int k = 0; k < CDEPTH; k++) L->win[k] = (L->pos < L->end) ? seq[L->pos++] : 0;
    return L->pos == L->end;
}

/* One ALIGNCORE step (src/algo/simd/search_simd_sw.c:172-182 and
 * src/algo/simd/search_simd_nw.c:180-191); returns the new H of the cell. */
static inline int64_t core(const wspec* w, int64_t diag, int64_t* E, int64_t* F, int64_t V, int64_t QR, int64_t R) {
    int64_t h = sat(w, diag + V);
    h = mx(h, *F);
    h = mx(h, *E);
    int64_t t = sat(w, h + QR);
    *F = mx(sat(w, *F + R), t);
    *E = mx(sat(w, *E + R), t);
    return h;
}

/* aligncolumns_first/_rest of the SW template (src/algo/simd/search_simd_sw.c:184-277) */
static void sw_block(const wspec* w, lane_t* L, int first, const uint8_t* q, size_t qlen,
                     const orc_matrix* m, int64_t QR, int64_t R) {
    int64_t h[4], f[4];
    for (int k = 0; k < 4; k++) { h[k] = w->imin; f[k] = w->imin; }
    int64_t mask = L->newly ? w->imin : w->imax;
    for (size_t i = 0; i < qlen; i++) {
        int64_t h4 = L->H[i], e = L->E[i];
        if (first) {
            h4 = mn(h4, mask);
            e = mn(e, mask);
        }
        int64_t n[4];
        for (int k = 0; k < 4; k++) {
            int64_t V = wrap(w, SC(m, L->win[k], q[i]));
            n[k] = core(w, h[k], &e, &f[k], V, QR, R);
            L->S = mx(L->S, n[k]);
        }
        L->H[i] = n[3];
        L->E[i] = e;
        h[0] = h4; h[1] = n[0]; h[2] = n[1]; h[3] = n[2];
    }
}

/* aligncolumns_first/_rest of the NW template (src/algo/simd/search_simd_nw.c:193-312) */
static void nw_block(const wspec* w, lane_t* L, int first, const uint8_t* q, size_t qlen,
                     const orc_matrix* m, int64_t QR, int64_t R) {
    if (first) { L->hmin = 0; L->hmax = 0; }
    int64_t h[4], f[4];
    for (int k = 0; k < 4; k++) { h[k] = L->Hb[k]; f[k] = sat(w, L->Fb[k] + QR); }
    int64_t mqr = (first && L->newly) ? QR : 0;
    int64_t mr = (first && L->newly) ? R : 0;
    int64_t mext = mqr;
    for (size_t i = 0; i < qlen; i++) {
        int64_t h4 = L->H[i], e = L->E[i];
        if (first) {
            if (L->newly) { h4 = 0; e = 0; } /* unsigned saturating subtract of all-ones */
            h4 = sat(w, h4 + mext);
            e = sat(w, e + mext);
            e = sat(w, e + mqr);
            mext = sat(w, mext + mr);
        }
        int64_t n[4];
        for (int k = 0; k < 4; k++) {
            int64_t V = wrap(w, SC(m, L->win[k], q[i]));
            n[k] = core(w, h[k], &e, &f[k], V, QR, R);
            L->hmin = mn(L->hmin, n[k]);
            L->hmax = mx(L->hmax, n[k]);
        }
        L->H[i] = n[3];
        L->E[i] = e;
        h[0] = h4; h[1] = n[0]; h[2] = n[1]; h[3] = n[2];
    }
    if (qlen) {
        L->Sm[0] = h[1]; L->Sm[1] = h[2]; L->Sm[2] = h[3]; L->Sm[3] = L->H[qlen - 1];
    }
}

size_t orc_lane_kernel(int type, int width, int lanes,
                       const uint8_t* const* seqs, const size_t* lens, size_t n,
                       const uint8_t* q, size_t qlen, const orc_matrix* m,
                       int go, int ge, orc_event* ev, int32_t* flag_block) {
    wspec w;
    w.width = width;
    w.imin = width == 8 ? -128 : -32768;
    w.imax = width == 8 ? 127 : 32767;
    w.uimax = width == 8 ? 255 : 65535;
    /* gap vectors (src/algo/simd/search_simd_sw.c:324-325): the lane type holds
     * go+ge after conversion */
    const int64_t QR = wrap(&w, go + ge);
    const int64_t R = wrap(&w, ge);
    const int64_t score_min = wrap(&w, w.imin - go - ge - 1); /* search_simd_nw.c:372 */

    lane_t* L = (lane_t*)calloc((size_t)lanes, sizeof(lane_t));
    for (int c = 0; c < lanes; c++) {
        L[c].H = (int64_t*)calloc(qlen + 1, sizeof(int64_t));
        L[c].E = (int64_t*)calloc(qlen + 1, sizeof(int64_t));
        L[c].item = -1;
    }
    for (size_t i = 0; i < n; i++) flag_block[i] = -1;

    size_t next = 0, done = 0, nev = 0;
    int change = 1;
    for (;;) {
        int first;
        if (!change) {
            for (int c = 0; c < lanes; c++) {
                L[c].newly = 0;
                if (L[c].item >= 0) change |= move_window(&L[c], seqs[L[c].item]);
            }
            first = 0;
        } else {
            change = 0;
            for (int c = 0; c < lanes; c++) {
                lane_t* l = &L[c];
                if (!l->ovf && l->pos < l->end) {
                    l->newly = 0;
                    change |= move_window(l, seqs[l->item]);
                    continue;
                }
                l->newly = 1;
                if (l->item >= 0) {
                    int accept;
                    int64_t score;
                    if (type == ORC_SW) {
                        score = l->S - w.imin;
                        accept = !l->ovf && score < w.uimax;
                    } else {
                        size_t z = (lens[l->item] + 3) % 4;
                        score = l->Sm[z];
                        accept = !l->ovf && score > w.imin && score < w.imax;
                    }
                    ev[nev].item = l->item;
                    ev[nev].kind = accept ? 0 : 1;
                    ev[nev].score = accept ? score : 0;
                    nev++;
                    done++;
                }
                if (type == ORC_SW) l->S = w.imin;
                l->ovf = 0;
                if (next < n) {
                    l->item = (int)next++;
                    l->pos = 0;
                    l->end = lens[l->item];
                    l->blocks = 0;
                    if (type == ORC_NW) {
                        /* search_simd_nw.c:445-453: per-lane boundary constants,
                         * converted to the lane type */
                        l->Hb[0] = 0;
                        for (int k = 1; k < 4; k++) l->Hb[k] = wrap(&w, go + k * ge);
                        for (int k = 0; k < 4; k++) l->Fb[k] = wrap(&w, go + (k + 1) * ge);
                    }
                    change |= move_window(l, seqs[l->item]);
                } else {
                    l->item = -1;
                    l->pos = l->end = 0;
                    for (int k = 0; k < CDEPTH; k++) l->win[k] = 0;
                }
            }
            if (done == n) break;
            first = 1;
        }
        int any_ovf = 0;
        for (int c = 0; c < lanes; c++) {
            lane_t* l = &L[c];
            if (type == ORC_SW) {
                sw_block(&w, l, first, q, qlen, m, QR, R);
                l->ovf = (l->S == w.imax);
            } else {
                nw_block(&w, l, first, q, qlen, m, QR, R);
                l->ovf = (l->hmin < score_min) || (l->hmax == w.imax);
                /* search_simd_nw.c:494-502: advance the top boundary by four columns */
                l->Fb[0] = sat(&w, l->Fb[3] + R);
                for (int k = 1; k < 4; k++) l->Fb[k] = sat(&w, l->Fb[k - 1] + R);
                l->Hb[0] = sat(&w, l->Hb[3] + R);
                for (int k = 1; k < 4; k++) l->Hb[k] = sat(&w, l->Hb[k - 1] + R);
            }
            if (l->item >= 0) {
                if (l->ovf && flag_block[l->item] < 0) flag_block[l->item] = l->blocks;
                l->blocks++;
            }
            any_ovf |= l->ovf;
        }
        if (any_ovf) change = 1;
    }
    for (int c = 0; c < lanes; c++) {
        free(L[c].H);
        free(L[c].E);
    }
    free(L);
    return nev;
}

/* ------------------------------------------------------------------------ */
/* bounded min-heap                                                          */
/* ------------------------------------------------------------------------ */

void orc_heap_init(orc_heap* h, orc_elem* storage, size_t alloc) {
    h->a = storage;
    h->alloc = alloc;
    h->count = 0;
}

/* src/util/minheap.c:60-91: sift up while strictly smaller; when full,
 * replace the root only if it is strictly smaller than the new element. */
void orc_heap_add(orc_heap* h, const orc_elem* e) {
    if (h->count < h->alloc) {
        size_t i = h->count++;
        while (i > 0) {
            size_t p = (i - 1) / 2;
            if (!(e->score < h->a[p].score)) break;
            h->a[i] = h->a[p];
            i = p;
        }
        h->a[i] = *e;
    } else if (h->count > 0 && h->a[0].score < e->score) {
        orc_elem tmp = *e;
        size_t p = 0, c = 1;
        while (c < h->count) {
            if (c + 1 < h->count && h->a[c + 1].score < h->a[c].score) c++;
            if (h->a[c].score < tmp.score) h->a[p] = h->a[c];
            else break;
            p = c;
            c = 2 * p + 1;
        }
        h->a[p] = tmp;
    }
}

static int elem_before(const orc_elem* x, const orc_elem* y) {
    /* CMP_ASC on score then db_id (src/util/util.h:31, src/util/minheap.c:93-106) */
    if (x->score != y->score) return x->score > y->score;
    return x->db_id > y->db_id;
}

void orc_heap_sort(orc_heap* h) {
    /* stable insertion sort: equal keys keep their heap-array order */
    for (size_t i = 1; i < h->count; i++) {
        orc_elem t = h->a[i];
        size_t j = i;
        while (j > 0 && elem_before(&t, &h->a[j - 1])) {
            h->a[j] = h->a[j - 1];
            j--;
        }
        h->a[j] = t;
    }
}

/* ------------------------------------------------------------------------ */
/* search driver                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t id;
    uint8_t strand, frame;
    uint8_t* seq;
    size_t len;
} entry_t;

static void insert(orc_heap* h, const entry_t* e, int qid, int64_t score) {
    orc_elem x;
    x.db_id = e->id;
    x.db_frame = e->frame;
    x.db_strand = e->strand;
    x.query_id = (uint8_t)qid;
    x.score = score;
    orc_heap_add(h, &x);
}

static int64_t full64(int type, const entry_t* e, const uint8_t* q, size_t ql, const orc_search_params* p) {
    return type == ORC_SW ? orc_full_sw(e->seq, e->len, q, ql, p->matrix, p->gap_open, p->gap_extend)
                          : orc_full_nw(e->seq, e->len, q, ql, p->matrix, p->gap_open, p->gap_extend);
}

/* search_64_chunk (src/algo/64/search_64.c:44-56) */
static void chunk64(const orc_search_params* p, orc_heap* h, entry_t** items, size_t n) {
    for (int qi = 0; qi < p->q_count; qi++)
        for (size_t i = 0; i < n; i++)
            insert(h, items[i], qi, full64(p->type, items[i], p->q_seq[qi], p->q_len[qi], p));
}

/* search_16_chunk (src/algo/16/search_16.c:92-114); returns the overflow count */
static size_t chunk16(const orc_search_params* p, orc_heap* h, entry_t** items, size_t n) {
    entry_t** ovf = (entry_t**)malloc((n * p->q_count + 1) * sizeof(entry_t*));
    size_t novf = 0;
    const uint8_t** seqs = (const uint8_t**)malloc((n + 1) * sizeof(uint8_t*));
    size_t* lens = (size_t*)malloc((n + 1) * sizeof(size_t));
    orc_event* ev = (orc_event*)malloc((n + 1) * sizeof(orc_event));
    int32_t* fb = (int32_t*)malloc((n + 1) * sizeof(int32_t));
    for (size_t i = 0; i < n; i++) { seqs[i] = items[i]->seq; lens[i] = items[i]->len; }
    for (int qi = 0; qi < p->q_count; qi++) {
        size_t ne = orc_lane_kernel(p->type, 16, p->lanes16, seqs, lens, n, p->q_seq[qi], p->q_len[qi],
                                    p->matrix, p->gap_open, p->gap_extend, ev, fb);
        for (size_t k = 0; k < ne; k++) {
            if (ev[k].kind == 0) insert(h, items[ev[k].item], qi, ev[k].score);
            else ovf[novf++] = items[ev[k].item];
        }
    }
    if (novf) chunk64(p, h, ovf, novf);
    free(ovf); free(seqs); free(lens); free(ev); free(fb);
    return novf;
}

/* search_8_chunk (src/algo/8/search_8.c:94-124) */
static void chunk8(const orc_search_params* p, orc_heap* h, entry_t** items, size_t n, orc_stats* st) {
    entry_t** ovf = (entry_t**)malloc((n * p->q_count + 1) * sizeof(entry_t*));
    size_t novf = 0;
    const uint8_t** seqs = (const uint8_t**)malloc((n + 1) * sizeof(uint8_t*));
    size_t* lens = (size_t*)malloc((n + 1) * sizeof(size_t));
    orc_event* ev = (orc_event*)malloc((n + 1) * sizeof(orc_event));
    int32_t* fb = (int32_t*)malloc((n + 1) * sizeof(int32_t));
    for (size_t i = 0; i < n; i++) { seqs[i] = items[i]->seq; lens[i] = items[i]->len; }
    for (int qi = 0; qi < p->q_count; qi++) {
        size_t ne = orc_lane_kernel(p->type, 8, p->lanes8, seqs, lens, n, p->q_seq[qi], p->q_len[qi],
                                    p->matrix, p->gap_open, p->gap_extend, ev, fb);
        for (size_t k = 0; k < ne; k++) {
            if (ev[k].kind == 0) insert(h, items[ev[k].item], qi, ev[k].score);
            else ovf[novf++] = items[ev[k].item];
        }
    }
    if (novf) {
        st->overflow_8 += novf;
        st->overflow_16 += chunk16(p, h, ovf, novf);
    }
    free(ovf); free(seqs); free(lens); free(ev); free(fb);
}

size_t orc_search(const orc_search_params* p, orc_elem* out, orc_stats* st) {
    orc_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof *st);

    /* DB entries as the adapter makes them (src/db_adapter.c:47-110, 212-239) */
    const int trans_db = p->symtype == ORC_TRANS_DB || p->symtype == ORC_TRANS_BOTH;
    int per_seq = (p->symtype == ORC_NUCLEOTIDE && (p->strands & 2)) ? 2 : 1;
    if (trans_db) per_seq = p->strands == 3 ? 6 : 3;
    int8_t dtab[4096];
    if (trans_db) orc_translate_table(p->d_gencode, dtab);
    entry_t* entries = (entry_t*)calloc(p->db_count * per_seq + 1, sizeof(entry_t));
    size_t* first_entry = (size_t*)calloc(p->db_count + 1, sizeof(size_t));
    size_t ne = 0;
    for (size_t id = 0; id < p->db_count; id++) {
        first_entry[id] = ne;
        size_t len = p->db_len[id];
        if (len == 0) continue;
        if (trans_db) {
            /* set_translated_sequences (src/db_adapter.c:72-104): frames 0..2 of
             * both strands (strand field 0/1), or of the one strand selected,
             * whose strand field is then the strand setting itself (1 or 2) */
            uint8_t* nt = (uint8_t*)malloc(len + 1);
            orc_map_db(p->symtype, p->db_raw[id], len, nt);
            int bad = 0;
            for (int k = 0; k < per_seq; k++) {
                int s = per_seq == 6 ? k / 3 : p->strands - 1, f = k % 3;
                entry_t* e = &entries[ne++];
                e->id = id;
                e->strand = (uint8_t)(per_seq == 6 ? s : p->strands);
                e->frame = (uint8_t)f;
                e->seq = (uint8_t*)malloc(len / 3 + 1);
                e->len = orc_translate(dtab, nt, len, s, f, e->seq);
                if (e->len == (size_t)-1) { e->len = 0; bad = 1; }
            }
            free(nt);
            if (bad) {
                for (size_t i = 0; i < ne; i++) free(entries[i].seq);
                free(entries);
                free(first_entry);
                return (size_t)-1;
            }
            continue;
        }
        entry_t* e = &entries[ne++];
        e->id = id;
        e->len = len;
        e->seq = (uint8_t*)malloc(len + 1);
        orc_map_db(p->symtype, p->db_raw[id], len, e->seq);
        if (per_seq == 2) {
            entry_t* r = &entries[ne++];
            r->id = id;
            r->len = len;
            r->strand = 1;
            r->seq = (uint8_t*)malloc(len + 1);
            orc_revcomp(e->seq, len, r->seq);
        }
    }
    first_entry[p->db_count] = ne;

    orc_heap heap;
    orc_elem* storage = (orc_elem*)calloc(p->hit_count + 1, sizeof(orc_elem));
    orc_heap_init(&heap, storage, p->hit_count);

    entry_t** items = (entry_t**)malloc((p->chunk_size * per_seq + 1) * sizeof(entry_t*));
    for (size_t start = 0; start < p->db_count; start += p->chunk_size) {
        size_t stop = start + p->chunk_size;
        if (stop > p->db_count) stop = p->db_count;
        size_t n = 0;
        for (size_t k = first_entry[start]; k < first_entry[stop]; k++) items[n++] = &entries[k];
        /* search_{8,16,64} loop while the chunk is not empty */
        if (n == 0) break;
        st->chunk_count++;
        st->seq_count += n;
        if (p->bit_width == 8) chunk8(p, &heap, items, n, st);
        else if (p->bit_width == 16) st->overflow_16 += chunk16(p, &heap, items, n);
        else chunk64(p, &heap, items, n);
    }
    free(items);

    orc_heap_sort(&heap);
    size_t count = heap.count;
    memcpy(out, heap.a, count * sizeof(orc_elem));
    free(storage);
    for (size_t i = 0; i < ne; i++) free(entries[i].seq);
    free(entries);
    free(first_entry);
    return count;
}

/* ------------------------------------------------------------------------ */
/* region, directions, CIGAR                                                 */
/* ------------------------------------------------------------------------ */

#define MX(a, b) ((a) > (b) ? (a) : (b))

/* src/algo/align.c:39-139.  Forward pass over DB rows (outer) and query
 * columns (inner) with scores S(query, db); end = first strict maximum.
 * Reverse pass from the end cell; its arrays are reset only for indices up to
 * b_end, so entries above b_end keep forward-pass values.  Returns -1 when the
 * forward pass finds no positive cell (the reference's end cell is then
 * uninitialised) or when the reverse pass does not reach the score. */
int orc_region_local(const uint8_t* a, size_t alen, const uint8_t* b, size_t blen,
                     const orc_matrix* m, int go, int ge, orc_region* r) {
    size_t size = alen > blen ? alen : blen;
    int64_t* HH = (int64_t*)calloc(size + 1, sizeof(int64_t));
    int64_t* EE = (int64_t*)calloc(size + 1, sizeof(int64_t));
    int64_t score = 0;
    int found = 0;
    for (size_t j = 0; j < alen; j++) { HH[j] = 0; EE[j] = go; }
    for (size_t i = 0; i < blen; i++) {
        int64_t h = 0, p = 0, f = go;
        for (size_t j = 0; j < alen; j++) {
            f = MX(f, h + go) + ge;
            EE[j] = MX(EE[j], HH[j] + go) + ge;
            h = p + SC(m, a[j], b[i]);
            if (h < 0) h = 0;
            if (f > h) h = f;
            if (EE[j] > h) h = EE[j];
            p = HH[j];
            HH[j] = h;
            if (h > score) {
                score = h;
                r->a_end = (int64_t)j;
                r->b_end = (int64_t)i;
                found = 1;
            }
        }
    }
    if (!found) { free(HH); free(EE); return -1; }
    for (int64_t j = r->b_end; j >= 0; j--) { HH[j] = -1; EE[j] = -1; }
    int64_t cost = 0;
    for (int64_t i = r->b_end; i >= 0; i--) {
        int64_t h = -1, f = -1, p = (i == r->b_end) ? 0 : -1;
        for (int64_t j = r->a_end; j >= 0; j--) {
            f = MX(f, h + go) + ge;
            EE[j] = MX(EE[j], HH[j] + go) + ge;
            h = p + SC(m, a[j], b[i]);
            if (f > h) h = f;
            if (EE[j] > h) h = EE[j];
            p = HH[j];
            HH[j] = h;
            if (h > cost) {
                cost = h;
                r->a_begin = j;
                r->b_begin = i;
                if (cost >= score) { free(HH); free(EE); return 0; }
            }
        }
    }
    free(HH);
    free(EE);
    return -1;
}

/* src/algo/align.c:141-150 */
void orc_region_global(size_t alen, size_t blen, orc_region* r) {
    r->a_begin = 0;
    r->a_end = (int64_t)alen - 1;
    r->b_begin = 0;
    r->b_end = (int64_t)blen - 1;
}

#define GAP_UP 1
#define GAP_LEFT 2
#define EXT_UP 4
#define EXT_LEFT 8

/* src/algo/cigar.c:48-198: direction byte per cell, index a_len*j + i with
 * j over the DB sequence and i over the query; scores S(db, query). */
static uint8_t* directions(int type, const uint8_t* a, size_t alen, const uint8_t* b, size_t blen,
                           const orc_matrix* m, int go, int ge) {
    uint8_t* dir = (uint8_t*)calloc(alen * blen + 1, 1);
    int64_t* hh = (int64_t*)calloc(alen + 1, sizeof(int64_t));
    int64_t* ee = (int64_t*)calloc(alen + 1, sizeof(int64_t));
    if (type == ORC_NW) {
        for (size_t i = 0; i < alen; i++) {
            hh[i] = go + (int64_t)(i + 1) * ge;
            ee[i] = 2 * (int64_t)go + (int64_t)(i + 2) * ge;
        }
    }
    for (size_t j = 0; j < blen; j++) {
        int64_t f, h;
        if (type == ORC_NW) {
            f = 2 * (int64_t)go + (int64_t)(j + 2) * ge;
            h = (j == 0) ? 0 : go + (int64_t)j * ge;
        } else {
            f = 0;
            h = 0;
        }
        for (size_t i = 0; i < alen; i++) {
            size_t idx = alen * j + i;
            int64_t nn = hh[i];
            int64_t e = ee[i];
            h += SC(m, b[j], a[i]);
            if (f > h) { dir[idx] |= GAP_UP; h = f; }
            if (e > h) { h = e; dir[idx] |= GAP_LEFT; }
            if (type == ORC_SW && h < 0) h = 0;
            hh[i] = h;
            h += go + ge;
            e += ge;
            f += ge;
            if (f > h) dir[idx] |= EXT_UP; else f = h;
            if (e > h) dir[idx] |= EXT_LEFT; else e = h;
            ee[i] = e;
            h = nn;
        }
    }
    free(hh);
    free(ee);
    return dir;
}

/* add_to_cigar + reverse_cigar (src/algo/cigar.c:200-263): runs are written
 * back to front, count after op, and the whole string is reversed at the end,
 * giving "<count><op>" with the count omitted when it is 1. */
static void put_run(char* buf, size_t* len, char op, size_t count) {
    if (count == 0) return;
    buf[(*len)++] = op;
    if (count > 1) {
        char d[32];
        int k = snprintf(d, sizeof d, "%zu", count);
        for (int x = k - 1; x >= 0; x--) buf[(*len)++] = d[x];
    }
}

/* compute_cigar_string (src/algo/cigar.c:265-346) */
size_t orc_cigar(int type, const uint8_t* a, size_t alen, const uint8_t* b, size_t blen,
                 const orc_region* r, const orc_matrix* m, int go, int ge, char* out, size_t cap) {
    uint8_t* dir = directions(type, a, alen, b, blen, m, go, ge);
    char* rev = (char*)malloc(3 * (alen + blen) + 64);
    size_t rl = 0;
    size_t i = (size_t)r->a_end, j = (size_t)r->b_end;
    char prev = 0, op = 0;
    size_t cnt = 0;
    while ((i + 1 > 0) && (j + 1 > 0) && (i >= (size_t)r->a_begin) && (j >= (size_t)r->b_begin)) {
        uint8_t d = dir[alen * j + i];
        if ((d & GAP_LEFT) || (d & EXT_LEFT)) { j--; op = 'I'; }
        else if ((d & GAP_UP) || (d & EXT_UP)) { i--; op = 'D'; }
        else { i--; j--; op = 'M'; }
        if (op == prev) cnt++;
        else {
            put_run(rev, &rl, prev, cnt);
            prev = op;
            cnt = 1;
        }
    }
    put_run(rev, &rl, prev, cnt);
    size_t n = rl < cap - 1 ? rl : cap - 1;
    for (size_t k = 0; k < n; k++) out[k] = rev[rl - 1 - k];
    out[n] = 0;
    free(rev);
    free(dir);
    return n;
}
