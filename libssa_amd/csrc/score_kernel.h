// score_kernel.h -- the score kernels' device code, shared by the
// translation units that instantiate them (score_sw.hip, score_swf.hip,
// score_nw.hip, score_pair.hip).
//
// score_kernel<R, G, MODE>: SW (MODE 0) / NW (MODE 1) over packed int16 lanes;
// MODE 2 is NW without the overflow trackers, for 16-bit launches where the
// host has shown that no value of any pair can reach either end of the int16
// range (then no block can be flagged and the trackers are dead work).
// MODE 3 is the SW first tier in packed fp16 windows and MODE 4 NW without
// the trackers in packed fp16 (see cell arithmetic below).
// Restates the per-lane arithmetic of the reference's width kernels
// (src/algo/simd/search_simd_sw.c:172-277, search_simd_nw.c:180-312):
// saturating adds, the H/E/F recurrence and the overflow predicates, but not
// their schedule: the host replays the reference's lane schedule from the
// per-pair results.
//
// Work shape: a lane group of G lanes scores two streams of pairs (the lo and
// hi halves of the packed registers).  The query is cut into passes of G*R
// rows; lane g of the group owns R consecutive rows of the pass, keeps their H
// and E in VGPRs and walks the streams one 4-column block behind lane g-1,
// which hands it the H and F of its bottom row through a DPP wave shift.  The
// top lane of a group takes its top boundary from the matrix border (first
// pass) or from the bottom lane of the previous pass through HBM.  A block
// costs each lane 4*R cells per half, so the cross-lane traffic is 9 DPP moves
// per 8*R cells.
//
// The pass's score profile is staged in LDS as [code][row] int16; a lane reads
// the R scores of its rows for one DB code with one 16-byte (R = 8) read per
// half and interleaves the two halves with v_perm.  With the profile in VGPRs
// (VP: nucleotide DBs of A, C, G and T only) a lane holds its rows' scores
// for the four codes, two dwords per row, and one LDS read per column fetches
// the v_perm selector of the column's two codes from a table of 32 x 32
// selectors; padding columns (code 0) score 0 there, which changes no result
// of the modes that run it (SW: a padding cell never exceeds a real one; NW
// without trackers: the score cell is the last real column).  A lane runs one block per
// step; lane g is idle for the first g steps of a pass (and the last G-1-g),
// so a pass takes nblocks + G - 1 steps.
//
// Per-pair results combine over lanes and passes: SW takes the maximum cell
// (atomic max) and the first flagged block (atomic min); a block is flagged
// when any row of it is.  The NW score is the last real row's cell at the
// pair's last column, written by the lane holding that row.
//
// The host pads the query to whole passes.  Padding rows sit below the last
// query row, so no real cell depends on them, and their profile values keep
// them from changing any result: SW pads with the lane minimum (a padded cell
// never exceeds the real maximum), NW pads with 0 and resets padded E and H
// to 0 at each pair start (a padded cell stays between min(0, earlier real
// cells) and max(0, earlier real cells), so it raises no overflow flag the
// real cells do not).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "gpu.h"
#include "kernels.h"

namespace ssa {

typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 as_v(uint32_t u) { return __builtin_bit_cast(s16x2, u); }
__device__ __forceinline__ uint32_t as_u(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s16x2 adds(s16x2 a, s16x2 b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ s16x2 vmax(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s16x2 vmin(s16x2 a, s16x2 b) { return __builtin_elementwise_min(a, b); }
// per-half select: halves whose mask bits are set take a, the others b
__device__ __forceinline__ s16x2 sel(uint32_t m, s16x2 a, s16x2 b) {
    return as_v((m & as_u(a)) | (~m & as_u(b)));
}
// Cell arithmetic.  Modes 0-2 use packed saturating int16 (the 8-bit tier
// holds values shifted left by 8, so int16 saturation reproduces the
// reference's int8 saturation).  Mode 3 is the SW first tier (the 16-bit
// tier, which records the 8-bit overflow block too) in packed fp16, for the
// three-input packed maximum (v_pk_maximum3_f16: 7.5 packed operations per
// SW cell instead of 9).  fp16 holds integers exactly only below 2048, so
// each lane keeps its values relative to an offset of its own (per half, a
// multiple of 64): in a DP matrix a cell differs from its neighbours in the
// row or the column by at most max score + |gap open + extend|, so the R rows
// of a lane stay within a window of (R - 1) times that, and a block of four
// columns raises a lane's maximum by at most four times that.  A lane
// re-centres its window (maximum near 1024) when its running maximum since
// the last re-centring passes 1536 or every 128 columns while its offset is
// above 0; values handed down from the lane above (or the previous pass's
// bottom row, stored with its offset) are moved by the difference of the
// offsets.  The overflow predicates (a real maximum >= 255, >= 65535) compare
// the window against the thresholds less the offset, and the pair's running
// maximum is folded into int32 at each re-centring.  The host enables the
// mode where the matrix and gaps keep every value within +-2047 of the
// window (BLOSUM62 with gaps 10/1: the window spans at most 700).  Mode 4 is
// the NW kernel without overflow trackers (mode 2) in plain packed fp16: the
// host enables it where every value of every pair lies within +-2047 (an NW
// cell is at least the two-gap path to it, 2 gap open + (i + j) gap extend,
// and at most max score * min(i, j)), and the three-input maximum takes the
// NW cell from 8 packed operations to 7.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f16x2 as_h(s16x2 v) { return __builtin_bit_cast(f16x2, v); }
__device__ __forceinline__ s16x2 of_h(f16x2 v) { return __builtin_bit_cast(s16x2, v); }
template <bool F>
__device__ __forceinline__ s16x2 c_add(s16x2 a, s16x2 b) {
    if constexpr (F) return of_h(as_h(a) + as_h(b));
    else return adds(a, b);
}
template <bool F>
__device__ __forceinline__ s16x2 c_max(s16x2 a, s16x2 b) {
    if constexpr (F) return of_h(__builtin_elementwise_maximum(as_h(a), as_h(b)));
    else return vmax(a, b);
}
template <bool F>
__device__ __forceinline__ s16x2 c_min(s16x2 a, s16x2 b) {
    if constexpr (F) return of_h(__builtin_elementwise_minimum(as_h(a), as_h(b)));
    else return vmin(a, b);
}
template <bool F>
__device__ __forceinline__ s16x2 c_max3(s16x2 a, s16x2 b, s16x2 c) {
    if constexpr (F) return of_h(__builtin_elementwise_maximum(__builtin_elementwise_maximum(as_h(a), as_h(b)), as_h(c)));
    else return vmax(vmax(a, b), c);
}
// mode 3: an fp16 value (the half of a packed pair) as an integer (values
// below -65536, -inf included, read as -65536, values above 65536, +inf
// included, as 65536: a half with no pair keeps raising its running maximum
// by e per column and may reach +inf; its readings are never used, and the
// clamp keeps the integer arithmetic on them in range)
__device__ __forceinline__ int f16_int(uint32_t bits16) {
    return (int)fminf(fmaxf((float)__builtin_bit_cast(_Float16, (uint16_t)bits16), -65536.0f), 65536.0f);
}
// two integers as a packed fp16 pair (exact for the multiples of 64 and the
// window values mode 3 converts)
__device__ __forceinline__ uint32_t f16_pair(int lo, int hi) {
    const f16x2 v = {(_Float16)(float)lo, (_Float16)(float)hi};
    return __builtin_bit_cast(uint32_t, v);
}
constexpr uint32_t kF16Inf2 = 0x7c007c00u;  // +inf in both halves: a threshold that never fires
constexpr uint32_t kF16Trigger2 = 0x66006600u;  // kWinTrigger (1,536) in both halves
constexpr int kWinTrigger = 1536;           // re-centre once the window's maximum passes this
constexpr int kWinCentre = 1024;            // the window's maximum after re-centring

__host__ __device__ __forceinline__ uint32_t pack2d(int v) { return ((uint32_t)(uint16_t)v << 16) | (uint16_t)v; }
__device__ __forceinline__ int lo16(s16x2 v) { return (int)(short)(as_u(v) & 0xffffu); }
__device__ __forceinline__ int hi16(s16x2 v) { return (int)(short)(as_u(v) >> 16); }
// Moves each lane's value one lane down within its group of G lanes: lane g
// receives lane g-1's value and the group's top lane receives `top`.  G = 64
// uses the wave shift and G = 16 the row shift, whose vacated lane takes the
// old value; G = 8 and 32 shift across group boundaries and fix the top lanes.
template <int G>
__device__ __forceinline__ uint32_t shift_down(uint32_t top, uint32_t v, bool top_lane) {
    if constexpr (G == 64) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)top, (int)v, 0x138, 0xf, 0xf, false);  // wave_shr:1
    } else if constexpr (G == 16) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)top, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    } else if constexpr (G == 32) {
        const uint32_t s = (uint32_t)__builtin_amdgcn_update_dpp((int)top, (int)v, 0x138, 0xf, 0xf, false);
        return top_lane ? top : s;
    } else {
        const uint32_t s = (uint32_t)__builtin_amdgcn_update_dpp((int)top, (int)v, 0x111, 0xf, 0xf, false);
        return top_lane ? top : s;
    }
}

template <bool B>
struct Tag {
    static constexpr bool value = B;
};

template <int N>
struct Log2 {
    static constexpr int value = 1 + Log2<N / 2>::value;
};
template <>
struct Log2<1> {
    static constexpr int value = 0;
};

// Part q of the profile scores of one lane for one DB code: rows 8q .. 8q+7
// (R >= 8) or rows 0 .. 3 (R = 4), two rows per dword.  q is a constant once
// the callers' loops are unrolled.
template <int R>
__device__ __forceinline__ void read_part(const char* p, uint32_t* d, int q) {
    if constexpr (R == 4) {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        d[0] = v.x;
        d[1] = v.y;
    } else {
        const uint4 v = *reinterpret_cast<const uint4*>(p + 16 * q);
        d[4 * q + 0] = v.x;
        d[4 * q + 1] = v.y;
        d[4 * q + 2] = v.z;
        d[4 * q + 3] = v.w;
    }
}

template <int R>
__device__ __forceinline__ void read_rows(const char* p, uint32_t* d) {
#pragma unroll
    for (int q = 0; q < (R >= 8 ? R / 8 : 1); ++q) read_part<R>(p, d, q);
}

// Part q of a paired profile (lo, hi int16 per row, one dword per row): rows
// 8q .. 8q+7 (R >= 8) or rows 0 .. 3 (R = 4).
template <int R>
__device__ __forceinline__ void read_part_pair(const char* p, uint32_t* d, int q) {
    if constexpr (R == 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    } else {
        const uint4 a = *reinterpret_cast<const uint4*>(p + 32 * q);
        const uint4 b = *reinterpret_cast<const uint4*>(p + 32 * q + 16);
        d[8 * q + 0] = a.x;
        d[8 * q + 1] = a.y;
        d[8 * q + 2] = a.z;
        d[8 * q + 3] = a.w;
        d[8 * q + 4] = b.x;
        d[8 * q + 5] = b.y;
        d[8 * q + 6] = b.z;
        d[8 * q + 7] = b.w;
    }
}

// the packed (lo, hi) score pair of row r from the two halves' profile dwords
__device__ __forceinline__ s16x2 pair_row(const uint32_t* lo, const uint32_t* hi, int r) {
    return as_v(__builtin_amdgcn_perm(hi[r / 2], lo[r / 2], (r & 1) ? 0x07060302u : 0x05040100u));
}


// SW: the group's bottom lane adds the maximum over the group's lanes (handed
// down with the pair's last block); NW: the lane holding the last query row
// writes the score.
template <int MODE>
__device__ __forceinline__ void finish_pair(PairResult* results, int pair, bool score_lane, int score, int fa,
                                            int fb) {
    constexpr bool SW = MODE == 0 || MODE == 3;
    if (score_lane) {
        if (SW) atomicMax(&results[pair].score, score);
        else results[pair].score = score;
    }
    if (fa != kNoFlag) atomicMin(&results[pair].flag_a, fa);
    if (SW && fb != kNoFlag) atomicMin(&results[pair].flag_b, fb);
}

// All passes of a wave over its streams, as one pipeline: lane g handles
// block t - g of the passes laid end to end (nbp blocks per pass), so lane 0
// starts pass p + 1 while the lanes below still finish pass p, and a task
// takes npass * nbp + G - 1 steps instead of npass * (nbp + G - 1).  Lane 0
// of pass p + 1 reads the bottom row that lane G - 1 wrote for the same block
// in pass p, nbp - G + 1 steps earlier (nbp > G).  Each lane keeps its own
// pass: the staged profile and NW left border are double-buffered, and
// stage(p) loads pass p's into buffer p & 1 once every lane has left pass
// p - 2.  Steps run two per iteration (an even count; an extra step is idle)
// so that the one-step-ahead loads alternate between two register sets.
//
// PAIR (SW, packed queries): both halves of a lane group walk the same
// stream (the layout gives the lo and hi slots the same entries) for two
// packed queries; the staged profile holds the two queries' rows interleaved,
// one dword (lo, hi) per row, so no v_perm is needed, and each half has its
// own segment table.
//
// NB (the column tile): blocks per step, 1 or 2 (4 or 8 columns).  A step
// runs its NB blocks one after the other with everything that belongs to a
// block -- pair starts, the overflow predicates and first flagged block
// (still per 4-column block, the reference's unit), pass boundaries, pair
// ends -- and pays the per-step work (residue and boundary loads of the next
// step, pass bookkeeping) once; the lane above hands down the bottom row of
// all 4 * NB columns.  Pass lengths are whole numbers of steps.
template <int NB>
struct StepIn {
    uint2 w[NB];        // residue words of the step's blocks
    uint4 t[2 * NB];    // top lane: the previous pass's bottom row over them
    uint32_t toff[NB];  // mode 3: the offsets of that bottom row (offset / 64, per half)
};

template <int R, int G, int MODE, int NB, bool MULTI, bool SEG, bool PAIR, bool VP, class Stage>
__device__ __forceinline__ void run_task(const ScoreParams& P, uint32_t gbase0, const char* sp,
                                         const uint32_t* __restrict__ slut,
                                         const uint32_t* slbh0, const uint32_t* slbe0,
                                         const uint2* __restrict__ wres, uint4* __restrict__ wbnd,
                                         uint32_t* __restrict__ wboff,
                                         const int32_t* __restrict__ slot_entries, int pair_off,
                                         const SegLane* __restrict__ segs, const SegLane* __restrict__ segs_hi,
                                         PairResult* __restrict__ results, int sb_lo, int sb_hi, int k, int g,
                                         int nbw, int nbp, int npass, int last_g, int last_r, const Stage& stage) {
    constexpr bool SW = MODE == 0 || MODE == 3;  // SW semantics (mode 3: in fp16)
    constexpr bool F16 = MODE == 3;               // SW in fp16 windows
    constexpr bool FP = MODE == 3 || MODE == 4;   // fp16 cell arithmetic
    static_assert(!PAIR || (SEG && MULTI && SW), "paired queries are packed SW queries");
    static_assert(!VP || (!PAIR && NB == 1 && R <= 16 && MODE != 1), "the VGPR profile's shapes and modes");
    constexpr int NG = 64 / G;
    constexpr int PR = G * R;
    constexpr int LPR = Log2<PR>::value + (PAIR ? 1 : 0);  // log2 of a profile code row's int16 count
    constexpr uint32_t kProfBytes = 32 * PR * (PAIR ? 4 : 2);  // one staged profile buffer
    const s16x2 QR = as_v(P.qr), RR = as_v(P.r), Z = as_v(P.zero);
    // mode 3 holds the values of DB column j (the pair's j-th) less j g (g =
    // the gap extend, P.shift): E extends without an add, F as max(F, H +
    // open) + g, the floor (real 0) rises by -g = e per column; EP = e
    // packed, E4 = 4 e (a block), ei = e as an integer
    const s16x2 EP = as_v(P.r ^ 0x80008000u);
    const s16x2 E4 = as_v(P.e4);
    const int ei = -P.shift;
    // per-half thresholds of the overflow predicates; a half whose predicate
    // cannot fire (flag taken, or no pair running) holds the neutral value
    const uint32_t TA_on = SW ? pack2d(P.thr_a - 1) : pack2d(P.thr_a);
    const uint32_t TB_on = pack2d(P.thr_b - 1 > 32767 ? 32767 : P.thr_b - 1);
    const uint32_t TA_off = SW ? 0x7fff7fffu : 0x80008000u;
    const uint32_t TB_off = 0x7fff7fffu;

    const bool top_lane = g == 0;
    // the lane's pass and block within it (negative before the lane's first block)
    int pass = 0, b = -NB * g;
    // SEG: the lane's segment in its pass, per half (a plain query: one
    // segment per pass); seg_top holds the halves that start a query there
    int po_lo = pair_off, po_hi = pair_off;
    uint32_t seg_top = 0u;
    bool bot_lo = g == G - 1, bot_hi = g == G - 1, none_lo = false, none_hi = false;
    auto load_seg = [&](int p) {
        if constexpr (SEG) {
            if (p < npass) {
                const SegLane sl = segs[p * G + g];
                const SegLane sh = PAIR ? segs_hi[p * G + g] : sl;
                po_lo = sl.pair_off;
                po_hi = sh.pair_off;
                seg_top = ((sl.flags & kSegTop) ? 0x0000ffffu : 0u) | ((sh.flags & kSegTop) ? 0xffff0000u : 0u);
                bot_lo = (sl.flags & kSegBottom) != 0;
                bot_hi = (sh.flags & kSegBottom) != 0;
                none_lo = (sl.flags & kSegNone) != 0;
                none_hi = (sh.flags & kSegNone) != 0;
            }
        }
    };
    load_seg(0);
    uint32_t gbase = gbase0;  // the lane's rows in the profile buffer of its pass
    const uint32_t* slbh = slbh0;
    const uint32_t* slbe = slbe0;
    // byte offset of the lane's rows for column c's DB code in the staged profile
    // (the code field lies above the lane's rows and below the buffer bit of
    // gbase, so the two combine with an or)
    auto colofs_at = [](uint32_t word, int c, uint32_t base) { return (((word >> (8 * c)) & 0x3eu) << LPR) | base; };
    auto colofs = [&gbase, &colofs_at](uint32_t word, int c) { return colofs_at(word, c, gbase); };

    s16x2 Hl[R], El[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        Hl[r] = Z;
        El[r] = Z;
    }
    // VP: the lane's rows' scores for A, C (dword 2r) and G, T (dword 2r + 1)
    // in the pass it is in, from the staged rows of buffer pass & 1
    uint32_t vp[VP ? 2 * R : 1];
    auto load_vp = [&](int p) {
        if constexpr (VP) {
            const uint4* src = reinterpret_cast<const uint4*>(sp) + ((p & 1) * PR + g * R) / 2;
#pragma unroll
            for (int q = 0; q < R / 2; ++q) {
                const uint4 t = src[q];
                vp[4 * q] = t.x;
                vp[4 * q + 1] = t.y;
                vp[4 * q + 2] = t.z;
                vp[4 * q + 3] = t.w;
            }
        }
    };
    load_vp(0);
    s16x2 S = Z;     // SW: running maximum of the lane's cells of the current pair
    s16x2 htop = Z;  // H of the row above at the previous column
    s16x2 cap = Z;   // NW: the last row's H at the pair's last column
    s16x2 hb[4], fb[4];  // NW: the matrix's top border over the current block (top lane, first pass)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        hb[c] = as_v(P.hb_init[c]);
        fb[c] = as_v(P.fb_init[c]);
    }
    uint32_t obh[4 * NB], obf[4 * NB];  // bottom row H, F of the last step's blocks
#pragma unroll
    for (int c = 0; c < 4 * NB; ++c) {
        obh[c] = 0u;
        obf[c] = 0u;
    }
    uint32_t TA = TA_off, TB = TB_off;
    // SW: max over this lane and the lanes above for the pair it finished
    // last in block j of a step (the lane below reads entry j one step later);
    // mode 3: real maxima, per half (scout, scout_hi)
    uint32_t scout[NB], scout_hi[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) scout[j] = scout_hi[j] = F16 ? 0u : as_u(Z);
    // mode 3: the window's offset O (O / 64 per half, packed int16); the
    // floor flr (real 0, fp16): j e - O for the next column j (so that E,
    // which feeds column j, carries it), raised by e at each column's start;
    // the running maximum folded at re-centring (real, per half); the
    // threshold of the fire test TT = min(TW, TTt): TW the re-centring
    // trigger, TTt the overflow thresholds, raised by 4 e at each block's
    // start (they are tested against S at the block's end)
    uint32_t off = 0u, flr = 0u, TW = kF16Inf2, TTt = kF16Inf2;
    // mode 3: the offsets of the bottom rows in obh / obf (the lane's offset
    // at the end of block j of its last step: a pair starting in the next
    // block resets `off` before the lane below has read those rows)
    uint32_t ob_off[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) ob_off[j] = 0u;
    int s32_lo = 0, s32_hi = 0;
    const int thr_a = P.thr_a, thr_b = P.thr_b;
    int kp_lo = 0, kp_hi = 0, pair_lo = -1, pair_hi = -1, blk_lo = 0, blk_hi = 0;
    int fa_lo = kNoFlag, fa_hi = kNoFlag, fb_lo = kNoFlag, fb_hi = kNoFlag;
    // mode 3: the overflow thresholds of one half not yet passed, in the
    // frame of S at the end of the block (column 4 blk + 4), less the offset
    auto thr_half = [&](int pair, int fa, int fb, int o, int blk) {
        int t = 100000;
        if (fa == kNoFlag) t = min(t, thr_a - 1 - o + (4 * blk + 4) * ei);
        if (fb == kNoFlag) t = min(t, thr_b - 1 - o + (4 * blk + 4) * ei);
        return pair >= 0 ? t : 100000;
    };
    auto f16_thr = [](int lo, int hi) {
        const uint32_t v = f16_pair(min(lo, 4096), min(hi, 4096));
        return (lo > 4096 ? 0x7c00u : (v & 0xffffu)) | (hi > 4096 ? 0x7c000000u : (v & 0xffff0000u));
    };
    auto set_tt = [&]() {
        const int olo = (int)(short)(off & 0xffffu) * 64, ohi = (int)(short)(off >> 16) * 64;
        TTt = f16_thr(thr_half(pair_lo, fa_lo, fb_lo, olo, blk_lo), thr_half(pair_hi, fa_hi, fb_hi, ohi, blk_hi));
        TW = (pair_lo >= 0 ? (kF16Trigger2 & 0xffffu) : 0x7c00u) | (pair_hi >= 0 ? (kF16Trigger2 & 0xffff0000u) : 0x7c000000u);
    };
    // the same for the halves in `mv` that start a pair in this block (no
    // flag, offset 0, the block's end at column 4: one value per launch),
    // the halves in `mi` that have no pair (the neutral value); the others
    // keep theirs.  Pair starts and ends come nearly every block in a wave
    // (8 streams per wave, each seen by its G lanes one step apart), so
    // these stay selects.
    auto set_tt_masked = [&](uint32_t mv, uint32_t mi) {
        const uint32_t keep = ~(mv | mi);
        TTt = (mv & P.thr_start) | (mi & kF16Inf2) | (keep & TTt);
        TW = (mv & kF16Trigger2) | (mi & kF16Inf2) | (keep & TW);
    };

    // profile dwords of the current column (lo, hi halves), read while the
    // column before is scored; a step's first column at the step's start, or
    // (EARLY, 32-row lanes, whose steps are long enough to cover the global
    // load of the next step's residues) during the last column of the step
    // before, the first step's below
    // (PAIR: pl holds one dword per row, both halves' scores)
    constexpr bool EARLY = R >= 32;
    uint32_t pl[PAIR ? R : (VP ? 1 : R / 2)], ph[PAIR || VP ? 1 : R / 2];
    auto read_first = [&](uint2 w0) {
        if constexpr (PAIR) {
#pragma unroll
            for (int q = 0; q < (R >= 8 ? R / 8 : 1); ++q) read_part_pair<R>(sp + colofs(w0.x, 0), pl, q);
        } else if constexpr (!VP) {
            read_rows<R>(sp + colofs(w0.x, 0), pl);
            read_rows<R>(sp + colofs(w0.y, 0), ph);
        }
    };

    // FIRST: the top lane is in the first pass and takes the matrix border
    auto step = [&](auto first_tag, const StepIn<NB>& in, StepIn<NB>& nx) {
        constexpr bool FIRST = decltype(first_tag)::value;
        if constexpr (!EARLY) read_first(in.w[0]);
        const bool lane_last = pass == npass - 1;
        const bool bot_out = g == G - 1 && pass < npass - 1;  // bottom row to the next pass
        const bool score_lo = SW ? bot_lo : (lane_last && g == last_g);
        const bool score_hi = SW ? bot_hi : (lane_last && g == last_g);
        // the lane's next step, in this pass or (MULTI) at the start of the next
        const bool wrap = MULTI && b + NB == nbp;
        const int b_n = wrap ? 0 : b + NB, pass_n = wrap ? pass + 1 : pass;
        // the profile buffer of the lane's next step (a lane entering its next
        // pass reads that pass's buffer, staged before the pass's first lane
        // entered it)
        const uint32_t gbase_n = wrap ? gbase ^ kProfBytes : gbase;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (F16) nx.toff[j] = 0u;
            if ((!MULTI || pass_n < npass) && (unsigned)(b_n + j) < (unsigned)nbw) {
                const uint32_t o = (uint32_t)((b_n + j) * NG + k);
                nx.w[j] = wres[o];
                if (MULTI && top_lane && pass_n > 0) {
                    nx.t[2 * j] = wbnd[2 * o];
                    nx.t[2 * j + 1] = wbnd[2 * o + 1];
                    if (F16) nx.toff[j] = wboff[o];
                }
            } else {
                nx.w[j] = make_uint2(0, 0);
            }
        }

#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const uint2 w = in.w[j];
            const uint4 t0 = in.t[2 * j], t1 = in.t[2 * j + 1];
            const bool valid = pass < npass && (unsigned)(b + j) < (unsigned)nbw;

            if constexpr (F16) TTt = as_u(c_add<true>(as_v(TTt), E4));  // in the frame of this block's end
            // the wave-level tests read both halves' flags with one or; the
            // per-half flags are only taken inside the rare branches
            if (__any(((w.x | w.y) & kResStart) != 0)) {
                const bool st_lo = (w.x & kResStart) != 0, st_hi = (w.y & kResStart) != 0;
                const uint32_t m = (st_lo ? 0x0000ffffu : 0u) | (st_hi ? 0xffff0000u : 0u);
                if (st_lo) {
                    pair_lo = none_lo ? -1 : po_lo + slot_entries[sb_lo + kp_lo];
                    kp_lo++;
                    blk_lo = 0;
                    fa_lo = kNoFlag;
                    fb_lo = kNoFlag;
                }
                if (st_hi) {
                    pair_hi = none_hi ? -1 : po_hi + slot_entries[sb_hi + kp_hi];
                    kp_hi++;
                    blk_hi = 0;
                    fa_hi = kNoFlag;
                    fb_hi = kNoFlag;
                }
                TA = (m & TA_on) | (~m & TA);
                TB = (m & TB_on) | (~m & TB);
                if (SW) {
                    // mode 3: H of column 0 and S in the frame of column 0
                    // (real 0: 0), E of column 1 and the floor in that of
                    // column 1 (real 0: e)
                    const s16x2 e0 = F16 ? EP : Z;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        Hl[r] = sel(m, Z, Hl[r]);
                        El[r] = sel(m, e0, El[r]);
                    }
                    S = sel(m, Z, S);
                    htop = sel(m, Z, htop);
                    if constexpr (F16) {  // a new pair starts in a window at offset 0
                        off &= ~m;
                        flr = (m & as_u(EP)) | (~m & flr);
                        if (st_lo) s32_lo = 0;
                        if (st_hi) s32_hi = 0;
                        const uint32_t mv = (st_lo && pair_lo >= 0 ? 0x0000ffffu : 0u) |
                                            (st_hi && pair_hi >= 0 ? 0xffff0000u : 0u);
                        set_tt_masked(mv, m & ~mv);
                    }
                } else {
                    const int r0 = g * R;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        Hl[r] = sel(m, as_v(slbh[1 + r0 + r]), Hl[r]);
                        El[r] = sel(m, as_v(slbe[r0 + r]), El[r]);
                    }
                    htop = sel(m, as_v(slbh[r0]), htop);  // row -1 of the first pass: the corner, 0
                    if (FIRST) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            hb[c] = sel(m, as_v(P.hb_init[c]), hb[c]);
                            fb[c] = sel(m, as_v(P.fb_init[c]), fb[c]);
                        }
                    }
                }
            }

            // top boundary of this block: the lane above, one step earlier; the
            // group's top lane takes the previous pass's bottom row or the matrix
            // border.  tbh[3] also carries the next block's diagonal.
            // mode 3: the floor (real 0) of the block's four columns, the
            // border of the first pass's top lane and of a packed query's
            // first row (in this lane's window)
            uint32_t fz[4];
            if constexpr (F16 && (FIRST || SEG)) {
                fz[0] = flr;
#pragma unroll
                for (int c = 1; c < 4; ++c) fz[c] = as_u(c_add<true>(as_v(fz[c - 1]), EP));
            }
            uint32_t toph[4], topf[4];
            if (!FIRST) {
                toph[0] = t0.x;
                topf[0] = t0.y;
                toph[1] = t0.z;
                topf[1] = t0.w;
                toph[2] = t1.x;
                topf[2] = t1.y;
                toph[3] = t1.z;
                topf[3] = t1.w;
            } else if (SW) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    toph[c] = F16 ? fz[c] : as_u(Z);
                    topf[c] = F16 ? fz[c] : as_u(Z);
                }
            } else {
                toph[0] = as_u(hb[1]);
                toph[1] = as_u(hb[2]);
                toph[2] = as_u(hb[3]);
                toph[3] = as_u(c_add<FP>(hb[3], RR));
#pragma unroll
                for (int c = 0; c < 4; ++c) topf[c] = as_u(c_add<FP>(fb[c], QR));
            }
            uint32_t tbh[4], tbf[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                tbh[c] = shift_down<G>(toph[c], obh[4 * j + c], top_lane);
                tbf[c] = shift_down<G>(topf[c], obf[4 * j + c], top_lane);
            }
            if constexpr (F16) {
                // the values above are in their source's window (the matrix
                // border: already in this lane's): move them into this lane's
                const uint32_t osrc = shift_down<G>(FIRST ? off : in.toff[j], ob_off[j], top_lane);
                if (__any(osrc != off)) {
                    const int dlo = ((int)(short)(osrc & 0xffffu) - (int)(short)(off & 0xffffu)) * 64;
                    const int dhi = ((int)(short)(osrc >> 16) - (int)(short)(off >> 16)) * 64;
                    const s16x2 d = as_v(f16_pair(dlo, dhi));
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        tbh[c] = as_u(c_add<true>(as_v(tbh[c]), d));
                        tbf[c] = as_u(c_add<true>(as_v(tbf[c]), d));
                    }
                }
            }
            if (SEG) {  // a packed query's first row: the SW matrix border (real 0)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint32_t zb = F16 ? fz[c] : as_u(Z);
                    tbh[c] = (seg_top & zb) | (~seg_top & tbh[c]);
                    tbf[c] = (seg_top & zb) | (~seg_top & tbf[c]);
                }
            }

            s16x2 hmn = as_v(0u), hmx = as_v(0u);  // NW trackers, reset value of the reference
            const bool any_last = __any(((w.x | w.y) & kResLast) != 0);
            const int z_lo = (int)((w.x >> kResZShift) & 3u), z_hi = (int)((w.y >> kResZShift) & 3u);
            // VP: the selectors of the block's four columns (codes of the lo
            // and hi streams: bits 1-5 of each byte)
            uint32_t vsel[4];
            if constexpr (VP) {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    vsel[c] = slut[(__builtin_amdgcn_ubfe(w.x, 8 * c + 1, 5) << 5) | __builtin_amdgcn_ubfe(w.y, 8 * c + 1, 5)];
            }
            // the block's four columns; CAP also records the NW score cell (the
            // last query row at the pair's last column), a select per row on a
            // wave-uniform condition, so the loop stays straight-line code
            auto columns = [&](auto cap_tag) {
                constexpr bool CAP = decltype(cap_tag)::value;
                uint32_t cm_lo = 0u, cm_hi = 0u;
                if (CAP && g == last_g && lane_last) {
                    cm_lo = (w.x & kResLast) ? 0x0000ffffu : 0u;
                    cm_hi = (w.y & kResLast) ? 0xffff0000u : 0u;
                }
                // The cell statements are written so that every dependent packed
                // op has an independent one in front of it (gfx950 needs a wait
                // state between a packed op and a packed op that reads its
                // result).  The order is not pinned with scheduling barriers:
                // instruction selection already reorders pure ops, and left free
                // the post-RA scheduler removes most of the s_nop padding (what
                // it leaves is hidden by the other waves of the SIMD).
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if constexpr (F16) {  // into this column's frame; the floor of the next
                        S = c_add<true>(S, EP);
                        flr = as_u(c_add<true>(as_v(flr), EP));
                    }
                    s16x2 f = as_v(tbf[c]);
                    s16x2 fR = c_add<FP>(f, RR);
                    s16x2 v;
                    if constexpr (PAIR) v = as_v(pl[0]);
                    else if constexpr (VP) v = as_v(__builtin_amdgcn_perm(vp[1], vp[0], vsel[c]));
                    else v = pair_row(pl, ph, 0);
                    s16x2 h1 = c_add<FP>(htop, v);  // row 0: diagonal H from the row above
                    htop = as_v(tbh[c]);
                    s16x2 hq = Z;
                    s16x2 hp = Z;  // mode 3: the previous row's H (S takes two rows per maximum)
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const bool more = r + 1 < R;
                        if (more) {
                            if constexpr (PAIR) v = as_v(pl[r + 1 < R ? r + 1 : 0]);
                            else if constexpr (VP)
                                v = as_v(__builtin_amdgcn_perm(vp[2 * (r + 1 < R ? r + 1 : 0) + 1],
                                                               vp[2 * (r + 1 < R ? r + 1 : 0)], vsel[c]));
                            else v = pair_row(pl, ph, r + 1);
                        }
                        // the next column's profile rows (after column 3 the
                        // step's next block, or the next step's first, whose
                        // residues were loaded at this step's start), once
                        // this column's dwords are consumed
                        if (!VP && (c < 3 || j + 1 < NB || EARLY) && ((R == 4 && r == 3) || (R >= 8 && r % 8 == 7))) {
                            const bool nstep = c == 3 && j + 1 == NB;
                            const uint2 wc = c < 3 ? w : nstep ? nx.w[0] : in.w[j + 1 < NB ? j + 1 : j];
                            const int cn = c < 3 ? c + 1 : 0;
                            const uint32_t base = nstep ? gbase_n : gbase;
                            if constexpr (PAIR) {
                                read_part_pair<R>(sp + colofs_at(wc.x, cn, base), pl, r / 8);
                            } else {
                                read_part<R>(sp + colofs_at(wc.x, cn, base), pl, r / 8);
                                read_part<R>(sp + colofs_at(wc.y, cn, base), ph, r / 8);
                            }
                            // keep the reads here: left free, the scheduler
                            // sinks all of a column's reads to its end, and
                            // the next column's first rows wait for them
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        if constexpr (F16) {
                            // E carries the floor at real 0 (E >= 0 makes H >=
                            // 0; an F below 0 never wins over that E, as with
                            // saturation, and is not clamped).  Values of
                            // column j less j g: the next column's E is
                            // max(E, H + open) (QR is the gap open), the next
                            // row's F is max(F, H + open) + g, and the
                            // profile holds the score + e
                            const s16x2 h = c_max3<true>(h1, f, El[r]);
                            if (more) h1 = c_add<true>(Hl[r], v);  // next row's diagonal term
                            const s16x2 tt = c_add<true>(h, QR);
                            if (r & 1) S = c_max3<true>(S, hp, h);
                            else hp = h;
                            f = c_add<true>(c_max<true>(f, tt), RR);
                            El[r] = c_max3<true>(El[r], tt, as_v(flr));
                            Hl[r] = h;
                        } else if constexpr (MODE == 4) {
                            // NW in fp16, no trackers, every value of cell
                            // (i, j) held less (i + j) gap extend: E and F
                            // extend without an add, QR is the gap open and
                            // the profile holds the score - 2 gap extend
                            const s16x2 h = c_max3<true>(h1, f, El[r]);
                            if (more) h1 = c_add<true>(Hl[r], v);  // next row's diagonal term
                            const s16x2 tt = c_add<true>(h, QR);
                            f = c_max<true>(f, tt);
                            El[r] = c_max<true>(El[r], tt);
                            Hl[r] = h;
                            if (CAP && r == last_r) hq = h;
                        } else {
                            s16x2 h = vmax(h1, f);
                            const s16x2 eR = adds(El[r], RR);
                            h = vmax(h, El[r]);
                            if (more) h1 = adds(Hl[r], v);  // next row's diagonal term
                            const s16x2 tt = adds(h, QR);
                            if (MODE == 0) {
                                S = vmax(S, h);
                            } else if (MODE == 1) {
                                hmn = vmin(hmn, h);
                                hmx = vmax(hmx, h);
                            }
                            f = vmax(fR, tt);
                            El[r] = vmax(eR, tt);
                            if (more) fR = adds(f, RR);
                            Hl[r] = h;
                            if (CAP && r == last_r) hq = h;
                        }
                    }
                    obh[4 * j + c] = as_u(Hl[R - 1]);
                    obf[4 * j + c] = as_u(f);
                    if (CAP) {
                        const uint32_t cm = (z_lo == c ? cm_lo : 0u) | (z_hi == c ? cm_hi : 0u);
                        cap = sel(cm, hq, cap);
                    }
                }
            };
            // only the lane holding the last query row records a score cell
            if (!SW && __any(lane_last && g == last_g && ((w.x | w.y) & kResLast) != 0)) columns(Tag<true>{});
            else columns(Tag<false>{});
            // overflow predicates at the end of the block: a packed test here (in
            // the cell loop's block, so the running reductions stay in the loop),
            // the exact per-half bookkeeping below only where a predicate fired
            bool fire;
            if (F16) {
                const uint32_t TT = as_u(c_min<true>(as_v(TTt), as_v(TW)));
                fire = as_u(c_max<true>(S, as_v(TT))) != TT;
            } else if (SW) {
                fire = as_u(vmax(S, as_v(TA))) != TA || as_u(vmax(S, as_v(TB))) != TB;
            } else if (MODE == 1) {
                fire = as_u(vmin(hmn, as_v(TA))) != TA || as_u(vmax(hmx, as_v(TB))) != TB;
            } else {
                fire = false;
            }
            // mode 3: every 32 blocks of a pair (128 columns) a half whose
            // offset is above 0 re-centres its window (values fall by at most
            // |gap open + extend| + 127 |gap extend| over that many columns)
            const bool due = F16 && ((((blk_lo & 31) == 31) && (off & 0xffffu) != 0u) ||
                                     (((blk_hi & 31) == 31) && (off >> 16) != 0u));
            if (F16 && __any(fire || due)) {
                const int olo = (int)(short)(off & 0xffffu) * 64, ohi = (int)(short)(off >> 16) * 64;
                // S is in the frame of the block's last column, 4 blk + 4:
                // real = held + O - (4 blk + 4) e
                const int jlo = (4 * blk_lo + 4) * ei, jhi = (4 * blk_hi + 4) * ei;
                const int wlo = f16_int(as_u(S) & 0xffffu), whi = f16_int(as_u(S) >> 16);
                const int slo = wlo + olo - jlo, shi = whi + ohi - jhi;
                if (pair_lo >= 0) {
                    const int v = max(s32_lo, slo);
                    if (v >= thr_a && fa_lo == kNoFlag) fa_lo = blk_lo;
                    if (v >= thr_b && fb_lo == kNoFlag) fb_lo = blk_lo;
                }
                if (pair_hi >= 0) {
                    const int v = max(s32_hi, shi);
                    if (v >= thr_a && fa_hi == kNoFlag) fa_hi = blk_hi;
                    if (v >= thr_b && fb_hi == kNoFlag) fb_hi = blk_hi;
                }
                const bool re_lo = wlo >= kWinTrigger || ((blk_lo & 31) == 31 && olo > 0);
                const bool re_hi = whi >= kWinTrigger || ((blk_hi & 31) == 31 && ohi > 0);
                if (__any(re_lo || re_hi)) {
                    // fold the running maximum, centre the window on the
                    // lane's largest H and move every value the lane holds
                    // (the whole wave runs this: lanes that keep their
                    // window add 0, and no value needs a second register)
                    if (re_lo) s32_lo = max(s32_lo, slo);
                    if (re_hi) s32_hi = max(s32_hi, shi);
                    s16x2 mx = Hl[0];
#pragma unroll
                    for (int r = 1; r + 1 < R; r += 2) mx = c_max3<true>(mx, Hl[r], Hl[r + 1]);
                    mx = c_max<true>(mx, Hl[R - 1]);
                    const int nlo = re_lo ? (max(0, f16_int(as_u(mx) & 0xffffu) + olo - kWinCentre) & ~63) : olo;
                    const int nhi = re_hi ? (max(0, f16_int(as_u(mx) >> 16) + ohi - kWinCentre) & ~63) : ohi;
                    const s16x2 d = as_v(f16_pair(olo - nlo, ohi - nhi));
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        Hl[r] = c_add<true>(Hl[r], d);
                        El[r] = c_add<true>(El[r], d);
                    }
                    htop = c_add<true>(htop, d);
#pragma unroll
                    for (int c = 0; c < 4 * (j + 1); ++c) {  // this step's bottom rows so far
                        obh[c] = as_u(c_add<true>(as_v(obh[c]), d));
                        obf[c] = as_u(c_add<true>(as_v(obf[c]), d));
                    }
                    // real 0 in the new window, in the frame of S and in that
                    // of the next column (the floor)
                    const uint32_t nzero = f16_pair(jlo - nlo, jhi - nhi);
                    // the running maximum since this re-centring (real 0 at
                    // least; the folded maximum keeps the rest)
                    const uint32_t rm = (re_lo ? 0x0000ffffu : 0u) | (re_hi ? 0xffff0000u : 0u);
                    S = as_v((rm & nzero) | (~rm & as_u(c_add<true>(S, d))));
                    flr = f16_pair(jlo + ei - nlo, jhi + ei - nhi);
                    off = ((uint32_t)(uint16_t)(nlo / 64)) | ((uint32_t)(uint16_t)(nhi / 64) << 16);
                    // the step's earlier bottom rows moved with the halves that
                    // re-centred (a half that started a pair in this block keeps
                    // the offset of the pair those rows belong to)
#pragma unroll
                    for (int jj = 0; jj < j; ++jj) ob_off[jj] = (rm & off) | (~rm & ob_off[jj]);
                }
                set_tt();
            }
            if (!F16 && __any(fire)) {
                if (SW) {
                    if (pair_lo >= 0) {
                        const int v = lo16(S);
                        if (v >= P.thr_a && fa_lo == kNoFlag) fa_lo = blk_lo;
                        if (v >= P.thr_b && fb_lo == kNoFlag) fb_lo = blk_lo;
                    }
                    if (pair_hi >= 0) {
                        const int v = hi16(S);
                        if (v >= P.thr_a && fa_hi == kNoFlag) fa_hi = blk_hi;
                        if (v >= P.thr_b && fb_hi == kNoFlag) fb_hi = blk_hi;
                    }
                } else {
                    if (pair_lo >= 0 && fa_lo == kNoFlag && (lo16(hmn) < P.thr_a || lo16(hmx) >= P.thr_b))
                        fa_lo = blk_lo;
                    if (pair_hi >= 0 && fa_hi == kNoFlag && (hi16(hmn) < P.thr_a || hi16(hmx) >= P.thr_b))
                        fa_hi = blk_hi;
                }
                const uint32_t offa = (fa_lo != kNoFlag ? 0x0000ffffu : 0u) | (fa_hi != kNoFlag ? 0xffff0000u : 0u);
                const uint32_t offb =
                    SW ? ((fb_lo != kNoFlag ? 0x0000ffffu : 0u) | (fb_hi != kNoFlag ? 0xffff0000u : 0u)) : offa;
                TA = (offa & TA_off) | (~offa & TA);
                TB = (offb & TB_off) | (~offb & TB);
            }
            if (F16) ob_off[j] = off;
            if (MULTI && bot_out && valid) {
                const uint32_t o = (uint32_t)((b + j) * NG + k);
                wbnd[2 * o] = make_uint4(obh[4 * j], obf[4 * j], obh[4 * j + 1], obf[4 * j + 1]);
                wbnd[2 * o + 1] = make_uint4(obh[4 * j + 2], obf[4 * j + 2], obh[4 * j + 3], obf[4 * j + 3]);
                if (F16) wboff[o] = off;
            }

            if (!SW && FIRST) {  // the reference advances the top border by four columns per block
                fb[0] = c_add<FP>(fb[3], RR);
                fb[1] = c_add<FP>(fb[0], RR);
                fb[2] = c_add<FP>(fb[1], RR);
                fb[3] = c_add<FP>(fb[2], RR);
                hb[0] = c_add<FP>(hb[3], RR);
                hb[1] = c_add<FP>(hb[0], RR);
                hb[2] = c_add<FP>(hb[1], RR);
                hb[3] = c_add<FP>(hb[2], RR);
            }
            if (any_last) {
                const bool ls_lo = (w.x & kResLast) != 0, ls_hi = (w.y & kResLast) != 0;
                int sc_lo, sc_hi;  // the pair's score in the tier's int16 form
                if constexpr (F16) {
                    // real maxima (the folded one and the window's), handed
                    // down the group as int32 per half
                    const int olo = (int)(short)(off & 0xffffu) * 64, ohi = (int)(short)(off >> 16) * 64;
                    // (S in the frame of the block's last column)
                    const int rlo = max(s32_lo, f16_int(as_u(S) & 0xffffu) + olo - (4 * blk_lo + 4) * ei);
                    const int rhi = max(s32_hi, f16_int(as_u(S) >> 16) + ohi - (4 * blk_hi + 4) * ei);
                    int inlo = (int)shift_down<G>(0u, scout[j], top_lane);
                    int inhi = (int)shift_down<G>(0u, scout_hi[j], top_lane);
                    if (SEG) {
                        if (seg_top & 0x1u) inlo = 0;
                        if (seg_top & 0x10000u) inhi = 0;
                    }
                    scout[j] = (uint32_t)max(rlo, inlo);
                    scout_hi[j] = (uint32_t)max(rhi, inhi);
                    sc_lo = min((int)scout[j], 65535) - 32768;
                    sc_hi = min((int)scout_hi[j], 65535) - 32768;
                } else {
                    if (SW) {  // the lane above finished the same pair one step earlier
                        uint32_t scin = shift_down<G>(as_u(Z), scout[j], top_lane);
                        if (SEG) scin = (seg_top & as_u(Z)) | (~seg_top & scin);
                        scout[j] = as_u(vmax(S, as_v(scin)));
                    }
                    const s16x2 fin = SW ? as_v(scout[j]) : cap;
                    sc_lo = MODE == 4 ? f16_int(as_u(fin) & 0xffffu) : lo16(fin);
                    sc_hi = MODE == 4 ? f16_int(as_u(fin) >> 16) : hi16(fin);
                    if constexpr (MODE == 4) {
                        // back from the shifted values: the score cell is
                        // (qlen, column z + 1 of the pair's block blk)
                        const int qlen = (npass - 1) * PR + last_g * R + last_r + 1;
                        sc_lo += (qlen + 4 * blk_lo + z_lo + 1) * P.shift;
                        sc_hi += (qlen + 4 * blk_hi + z_hi + 1) * P.shift;
                    }
                }
                uint32_t m = 0u;
                if (ls_lo && pair_lo >= 0) {
                    finish_pair<MODE>(results, pair_lo, score_lo, sc_lo, fa_lo, fb_lo);
                    pair_lo = -1;
                    m |= 0x0000ffffu;
                }
                if (ls_hi && pair_hi >= 0) {
                    finish_pair<MODE>(results, pair_hi, score_hi, sc_hi, fa_hi, fb_hi);
                    pair_hi = -1;
                    m |= 0xffff0000u;
                }
                TA = (m & TA_off) | (~m & TA);
                TB = (m & TB_off) | (~m & TB);
                if (F16) set_tt_masked(0u, m);  // the finished halves run no pair
            }
            blk_lo++;
            blk_hi++;
        }
        b = b_n;
        if (MULTI && __any(wrap)) {  // lanes entering their next pass: its entry list and staged buffers
            if (wrap) {
                pass = pass_n;
                kp_lo = 0;
                kp_hi = 0;
                gbase ^= kProfBytes;
                load_seg(pass);
                const bool odd = pass & 1;
                slbh = odd ? slbh0 + (PR + 1) : slbh0;
                slbe = odd ? slbe0 + PR : slbe0;
            }
            // (the whole wave: lanes still in the earlier pass read its rows
            // again, which are still staged)
            load_vp(pass);
        }
    };

    // the blocks of step 0 (-NB g ..): only the top lanes start with real blocks
    StepIn<NB> sa, sb;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        sa.w[j] = sb.w[j] = make_uint2(0, 0);
        sa.t[2 * j] = sa.t[2 * j + 1] = sb.t[2 * j] = sb.t[2 * j + 1] = make_uint4(0, 0, 0, 0);
        sa.toff[j] = sb.toff[j] = 0u;
        if (g == 0 && j < nbw) sa.w[j] = wres[j * NG + k];
    }
    if constexpr (EARLY) read_first(sa.w[0]);
    // steps with the top lane in the first pass (an even number: nbp is a
    // multiple of 2 * NB)
    const int spp = nbp / NB;  // steps per pass
    for (int t = 0; t < spp; t += 2) {
        step(Tag<true>{}, sa, sb);
        step(Tag<true>{}, sb, sa);
    }
    // the rest: pass p + 2 is staged once lane G - 1 has left pass p
    const int nsteps = (npass * spp + G) & ~1;
    int next_stage = ((spp + G - 1) + 1) & ~1, staged = 2;
    for (int t = spp; t < nsteps; t += 2) {
        if (MULTI && t == next_stage && staged < npass) {  // the same for the whole workgroup
            stage(staged);
            ++staged;
            next_stage += spp;
        }
        step(Tag<false>{}, sa, sb);
        step(Tag<false>{}, sb, sa);
    }
}

// Wave priority from the steps the workgroup's task has left, in units of
// L.prio_steps (about half a typical task): 0 below one unit, then 1, 2 and 3
// from 1, 2 and 4 units.  The two workgroups of a CU share its SIMDs, and at
// equal priority a SIMD issues from its older wave first, so the CU's issue
// slots go to whichever task has more work left: a long task that a second
// workgroup takes late does not run at the leftover rate while its CU's
// first workgroup works through short ones (the end of the launch waited on
// such tasks).
__device__ __forceinline__ void set_task_priority(int64_t steps_left, int32_t unit) {
    if (unit <= 0) return;
    if (steps_left >= 4 * (int64_t)unit) __builtin_amdgcn_s_setprio(3);
    else if (steps_left >= 2 * (int64_t)unit) __builtin_amdgcn_s_setprio(2);
    else if (steps_left >= unit) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// Task trace (thread 0 of a workgroup, at each fetch): the end of the
// workgroup's previous task and the start of the one it fetched, on the GPU's
// constant-rate wall clock.
__device__ __forceinline__ void trace_fetch(const ScoreLaunch& L, int prev, int task) {
    const uint64_t now = wall_clock64();
    if (prev >= 0) L.trace[prev].end = now;
    if (task < L.ntasks) {
        L.trace[task].start = now;
        L.trace[task].wg = (int32_t)blockIdx.x;
    }
}

template <int R, int G, int MODE, int NB, bool VP>
__global__ __launch_bounds__(64 * kWavesPerWG) void score_kernel(const ScoreLaunch L, const ScoreParams P) {
    constexpr int PR = G * R;  // rows per pass
    static_assert((1 << Log2<PR>::value) == PR, "rows per pass must be a power of two");
    static_assert(PR <= 512, "two staged profiles of 32 codes x PR rows must fit the LDS budget");
    constexpr bool NW = MODE == 1 || MODE == 2 || MODE == 4;
    __shared__ __align__(16) int16_t sprof[VP ? 8 : 2 * 32 * PR];  // profiles of passes p (even, odd)
    __shared__ __align__(16) uint2 svp[VP ? 2 * PR : 1];  // VP: rows' scores for A, C, G, T of passes p
    __shared__ uint32_t slut[VP ? 32 * 32 : 1];           // VP: v_perm selector of a column's two codes
    __shared__ uint32_t slbh[NW ? 2 * (PR + 1) : 1];     // NW left boundary H of rows base-1 .. base+PR-1
    __shared__ uint32_t slbe[NW ? 2 * PR : 1];           // NW left boundary E of rows base .. base+PR-1
    __shared__ int s_task;

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = lane / G, g = lane % G;
    const uint32_t gbase = (uint32_t)(g * R * 2);  // byte offset of the lane's rows in a profile code row
    const char* sp = VP ? reinterpret_cast<const char*>(svp) : reinterpret_cast<const char*>(sprof);
    if constexpr (VP) {
        // selector of (lo code, hi code): the bytes of the code's score in
        // the row's dword pair (A 0-1, C 2-3, G 4-5, T 6-7); other codes
        // (the padding) select 0
        for (int i = threadIdx.x; i < 32 * 32; i += 64 * kWavesPerWG) {
            auto half = [](int code) -> uint32_t {
                const int slot = code == 1 ? 0 : code == 2 ? 1 : code == 4 ? 2 : code == 8 ? 3 : -1;
                return slot < 0 ? 0x0c0cu : (uint32_t)(2 * slot) | ((uint32_t)(2 * slot + 1) << 8);
            };
            slut[i] = half(i >> 5) | (half(i & 31) << 16);
        }
    }
    uint4* __restrict__ wbnd = L.bnd + ((int64_t)blockIdx.x * kWavesPerWG + wv) * L.bnd_per_wave;
    uint32_t* __restrict__ wboff = reinterpret_cast<uint32_t*>(wbnd + L.bnd_rows);

    // Each workgroup's first task is its own index, later ones come from the
    // counter (past the grid).  The hardware places workgroups 0 .. CUs - 1
    // first, one per CU, and at equal priority a SIMD issues from its older
    // wave first, so while both workgroups of a CU are busy the second runs at
    // about a third of the first one's rate.  The tasks come costliest first,
    // so this starts the longest ones on first workgroups; drawn from the
    // counter, some landed on second workgroups, where one outlasted the rest
    // of the launch (62 instead of 57 ms for the 5,000-residue SW query).
    bool first_task = true;
    int prev_task = -1;
    for (;;) {
        __syncthreads();  // the previous task is done with s_task and the staged profiles
        if (threadIdx.x == 0) {
            s_task = first_task ? (int)blockIdx.x : (int)gridDim.x + atomicAdd(L.counter, 1);
            if (L.trace) trace_fetch(L, prev_task, s_task);
        }
        first_task = false;
        __syncthreads();
        const int task = s_task;
        prev_task = task;
        if (task >= L.ntasks) break;  // the same for the whole workgroup
        const ScoreTask T = L.tasks[task];
        const ScoreWave Wv = L.waves[T.wave0 + wv];
        const QueryDesc Q = L.queries[T.query];
        const int npass = (Q.qlen + PR - 1) / PR;
        // the workgroup's waves share the pass length: the longest of them,
        // at least G + 2 steps (so that a pass's bottom rows are written
        // before the next pass reads them) and an even number of steps
        int nbmax = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerWG; ++i) nbmax = max(nbmax, L.waves[T.wave0 + i].nblocks);
        if (nbmax == 0) continue;
        const int nbp = ((npass > 1 ? max(nbmax, NB * (G + 2)) : nbmax) + 2 * NB - 1) / (2 * NB) * (2 * NB);
        if (L.trace && threadIdx.x == 0) L.trace[task].steps = npass * (nbp / NB);
        set_task_priority((int64_t)npass * (nbp / NB), L.prio_steps);
        const uint2* __restrict__ wres = L.res + Wv.res_off;
        const int sb_lo = L.slot_first[Wv.slot_base + 2 * k];
        const int sb_hi = L.slot_first[Wv.slot_base + 2 * k + 1];
        const int last_lr = Q.qlen - 1 - (npass - 1) * PR;  // last real row within the last pass
        const int last_g = last_lr / R, last_r = last_lr % R;

        // pass p's profile (and NW left border) into buffer p & 1
        auto load_pass = [&](int pass) {
            const int buf = pass & 1;
            if constexpr (VP) {  // each row's scores for codes 1, 2, 4 and 8 (A, C, G, T)
                const int16_t* src = L.qp + Q.qp_off + (int64_t)pass * 32 * PR;
                for (int i = threadIdx.x; i < PR; i += 64 * kWavesPerWG)
                    svp[buf * PR + i] = make_uint2((uint16_t)src[PR + i] | ((uint32_t)(uint16_t)src[2 * PR + i] << 16),
                                                   (uint16_t)src[4 * PR + i] | ((uint32_t)(uint16_t)src[8 * PR + i] << 16));
            } else {
                const uint4* src = reinterpret_cast<const uint4*>(L.qp + Q.qp_off + (int64_t)pass * 32 * PR);
                uint4* dst = reinterpret_cast<uint4*>(sprof + buf * 32 * PR);
                for (int i = threadIdx.x; i < 32 * PR / 8; i += 64 * kWavesPerWG) dst[i] = src[i];
            }
            if (NW) {
                const int base = pass * PR;
                uint32_t* bh = slbh + buf * (PR + 1);
                uint32_t* be = slbe + buf * PR;
                for (int i = threadIdx.x; i < PR; i += 64 * kWavesPerWG) {
                    bh[1 + i] = L.lb[Q.lb_off + 2 * (base + i)];
                    be[i] = L.lb[Q.lb_off + 2 * (base + i) + 1];
                }
                if (threadIdx.x == 0) bh[0] = base > 0 ? L.lb[Q.lb_off + 2 * (base - 1)] : 0u;
            }
        };
        load_pass(0);
        if (npass > 1) load_pass(1);
        __syncthreads();
        auto stage = [&](int pass) {
            __syncthreads();  // every lane of the workgroup has left pass - 2
            load_pass(pass);
            __syncthreads();
            // the lanes are in passes pass - 2 and pass - 1
            set_task_priority((int64_t)(npass - pass + 1) * (nbp / NB), L.prio_steps);
        };
        if (!NW && Q.seg_off >= 0)
            run_task<R, G, MODE, NB, true, !NW, false, VP>(P, gbase, sp, slut, slbh, slbe, wres, wbnd, wboff, L.slot_entries,
                                                         Q.pair_off, L.segs + Q.seg_off, nullptr, L.results, sb_lo,
                                                         sb_hi, k, g, Wv.nblocks, nbp, npass, last_g, last_r, stage);
        else if (npass > 1)
            run_task<R, G, MODE, NB, true, false, false, VP>(P, gbase, sp, slut, slbh, slbe, wres, wbnd, wboff, L.slot_entries,
                                                     Q.pair_off, nullptr, nullptr, L.results, sb_lo, sb_hi, k, g,
                                                     Wv.nblocks, nbp, npass, last_g, last_r, stage);
        else
            run_task<R, G, MODE, NB, false, false, false, VP>(P, gbase, sp, slut, slbh, slbe, wres, wbnd, wboff, L.slot_entries,
                                                      Q.pair_off, nullptr, nullptr, L.results, sb_lo, sb_hi, k, g,
                                                      Wv.nblocks, nbp, npass, last_g, last_r, stage);
    }
}

// Paired packed SW queries (run_task's PAIR): two packed queries share every
// stream, one per half.  The staged profile interleaves their rows (one
// dword per row), twice the bytes per pass of score_kernel, so a pass holds at
// most 256 rows.
template <int R, int G, bool F16>
__global__ __launch_bounds__(64 * kWavesPerWG) void score_kernel_pair(const ScoreLaunch L, const ScoreParams P) {
    constexpr int PR = G * R;
    static_assert((1 << Log2<PR>::value) == PR, "rows per pass must be a power of two");
    static_assert(PR <= 256, "two staged paired profiles of 32 codes x PR rows must fit the LDS budget");
    __shared__ __align__(16) int16_t sprof[2 * 2 * 32 * PR];  // paired profiles of passes p (even, odd)
    __shared__ uint32_t slb_unused[1];                         // (the NW left border of score_kernel)
    __shared__ int s_task;

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = lane / G, g = lane % G;
    const uint32_t gbase = (uint32_t)(g * R * 4);  // byte offset of the lane's rows in a paired profile code row
    const char* sp = reinterpret_cast<const char*>(sprof);
    uint4* __restrict__ wbnd = L.bnd + ((int64_t)blockIdx.x * kWavesPerWG + wv) * L.bnd_per_wave;
    uint32_t* __restrict__ wboff = reinterpret_cast<uint32_t*>(wbnd + L.bnd_rows);

    bool first_task = true;  // as in score_kernel: the first task is the workgroup's index
    int prev_task = -1;
    for (;;) {
        __syncthreads();  // the previous task is done with s_task and the staged profiles
        if (threadIdx.x == 0) {
            s_task = first_task ? (int)blockIdx.x : (int)gridDim.x + atomicAdd(L.counter, 1);
            if (L.trace) trace_fetch(L, prev_task, s_task);
        }
        first_task = false;
        __syncthreads();
        const int task = s_task;
        prev_task = task;
        if (task >= L.ntasks) break;  // the same for the whole workgroup
        const ScoreTask T = L.tasks[task];
        const ScoreWave Wv = L.waves[T.wave0 + wv];
        const QueryDesc Q = L.queries[T.query];
        const int npass = (Q.qlen + PR - 1) / PR;
        int nbmax = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerWG; ++i) nbmax = max(nbmax, L.waves[T.wave0 + i].nblocks);
        if (nbmax == 0) continue;
        const int nbp = npass > 1 ? (max(nbmax, G + 2) + 1) & ~1 : (nbmax + 1) & ~1;
        if (L.trace && threadIdx.x == 0) L.trace[task].steps = npass * nbp;
        set_task_priority((int64_t)npass * nbp, L.prio_steps);
        const uint2* __restrict__ wres = L.res + Wv.res_off;
        const int sb_lo = L.slot_first[Wv.slot_base + 2 * k];
        const int sb_hi = L.slot_first[Wv.slot_base + 2 * k + 1];
        auto load_pass = [&](int pass) {
            const int buf = pass & 1;
            const uint4* src = reinterpret_cast<const uint4*>(L.qp + Q.qp_off + (int64_t)pass * 2 * 32 * PR);
            uint4* dst = reinterpret_cast<uint4*>(sprof + buf * 2 * 32 * PR);
            for (int i = threadIdx.x; i < 2 * 32 * PR / 8; i += 64 * kWavesPerWG) dst[i] = src[i];
        };
        load_pass(0);
        if (npass > 1) load_pass(1);
        __syncthreads();
        auto stage = [&](int pass) {
            __syncthreads();  // every lane of the workgroup has left pass - 2
            load_pass(pass);
            __syncthreads();
            set_task_priority((int64_t)(npass - pass + 1) * nbp, L.prio_steps);
        };
        run_task<R, G, F16 ? 3 : 0, 1, true, true, true, false>(P, gbase, sp, nullptr, slb_unused, slb_unused, wres, wbnd, wboff,
                                                         L.slot_entries,
                                            Q.pair_off, L.segs + Q.seg_off, L.segs + Q.seg_off_hi, L.results, sb_lo,
                                            sb_hi, k, g, Wv.nblocks, nbp, npass, 0, 0, stage);
    }
}

// ---------------------------------------------------------------------------
// Instantiation: one translation unit per mode (score_sw.hip: 0, score_nw1.hip:
// 1, score_nw2.hip: 2, score_swf.hip: 3, score_nwf.hip: 4; with the profile in
// VGPRs score_swv.hip: 0 and 3, score_nwv.hip: 2 and 4; score_pair.hip: the
// paired kernel) defines SSA_SCORE_INSTANTIATE and instantiates launch_mode /
// resident_mode for its modes; the dispatch (score_kernels.hip) sees the
// declarations only.
// ---------------------------------------------------------------------------
hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, int64_t n, uint32_t v);

// score_kernel<R, G, MODE, NB, VP> for lanes = G, rows = R, cols = 4 * NB:
// launch on a persistent grid of at most max_wg workgroups / resident
// workgroups of the chip (0 for a shape the mode does not build)
template <int MODE, bool VP>
hipError_t launch_mode(int lanes, int rows, int cols, hipStream_t st, const ScoreLaunch& L, const ScoreParams& P,
                       int max_wg);
template <int MODE, bool VP>
int resident_mode(int lanes, int rows, int cols);
// score_kernel_pair<R, G, F16>
template <bool F16>
hipError_t launch_pair_mode(int lanes, int rows, hipStream_t st, const ScoreLaunch& L, const ScoreParams& P,
                            int max_wg);
template <bool F16>
int resident_pair_mode(int lanes, int rows);

// NW without trackers (MODE 2) exists for the common NW shapes only
template <int R, int G>
constexpr bool nw_untracked_shape() {
    return (G == 8 && (R == 8 || R == 16)) || (G == 16 && (R == 8 || R == 16)) || (G == 32 && R == 8);
}

// 8-column tiles (two blocks per step) for lanes of at most 16 rows in groups
// of 8 or 16 lanes: the step's second block holds 4 more bottom-row H/F pairs
// and residue/boundary loads, which 32-row lanes have no registers for
template <int R, int G>
constexpr bool tile8_shape() {
    return R <= 16 && (G == 8 || G == 16);
}

// 32 rows per lane for SW only (NW holds more state per row and would run one
// wave per SIMD); the profile in VGPRs for 4-column tiles of at most 16 rows
// per lane, SW and NW without trackers (2R more registers per lane)
template <int R, int G, int MODE, int NB, bool VP>
constexpr bool shape_built() {
    return !(R == 32 && (MODE == 1 || MODE == 2 || MODE == 4)) &&
           !((MODE == 2 || MODE == 4) && !nw_untracked_shape<R, G>()) &&
           (NB == 1 || tile8_shape<R, G>()) && !(VP && (NB != 1 || R > 16 || MODE == 1));
}

#define SSA_SHAPES(X)                                                                                   \
    X(8, 4) X(8, 8) X(8, 16) X(8, 32) X(16, 4) X(16, 8) X(16, 16) X(16, 32) X(32, 4) X(32, 8) X(32, 16) \
    X(64, 4) X(64, 8)

// shapes of the paired kernel (lanes per stream pair, rows per lane)
#define SSA_PAIR_SHAPES(X) X(8, 16) X(8, 32) X(16, 8) X(16, 16) X(32, 8)

#ifdef SSA_SCORE_INSTANTIATE
template <int R, int G, int MODE, int NB, bool VP>
static int resident_wg_t() {
    if constexpr (!shape_built<R, G, MODE, NB, VP>()) {
        return 0;
    } else {
        static int cap = 0;
        if (!cap) {
            int per_cu = 0, dev = 0;
            hipDeviceProp_t prop;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, score_kernel<R, G, MODE, NB, VP>,
                                                             64 * kWavesPerWG, 0) != hipSuccess ||
                hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
                return 0;
            cap = std::max(1, per_cu) * prop.multiProcessorCount;
        }
        return cap;
    }
}

template <int R, int G, int MODE, int NB, bool VP>
static hipError_t launch_score_t(hipStream_t st, const ScoreLaunch& L, const ScoreParams& P, int max_wg) {
    if constexpr (!shape_built<R, G, MODE, NB, VP>()) {
        return hipErrorInvalidValue;
    } else {
        const int nwg = std::max(1, std::min(L.ntasks, max_wg));
        hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(L.counter), 1, 0u);
        if (e == hipSuccess && L.trace)
            e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(L.trace), (int64_t)L.ntasks * (sizeof(TaskTrace) / 4), 0u);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((score_kernel<R, G, MODE, NB, VP>), dim3(nwg), dim3(64 * kWavesPerWG), 0, st, L, P);
        return hipGetLastError();
    }
}

template <int MODE, bool VP>
hipError_t launch_mode(int lanes, int rows, int cols, hipStream_t st, const ScoreLaunch& L, const ScoreParams& P,
                       int max_wg) {
#define SSA_CASE(GG, RR)                                                              \
    if (lanes == GG && rows == RR)                                                  \
        return cols == 8 ? launch_score_t<RR, GG, MODE, 2, VP>(st, L, P, max_wg)    \
                         : (cols == 4 ? launch_score_t<RR, GG, MODE, 1, VP>(st, L, P, max_wg) : hipErrorInvalidValue);
    SSA_SHAPES(SSA_CASE)
#undef SSA_CASE
    return hipErrorInvalidValue;
}

template <int MODE, bool VP>
int resident_mode(int lanes, int rows, int cols) {
#define SSA_CASE(GG, RR)                                                                     \
    if (lanes == GG && rows == RR)                                                         \
        return cols == 8 ? resident_wg_t<RR, GG, MODE, 2, VP>() : resident_wg_t<RR, GG, MODE, 1, VP>();
    SSA_SHAPES(SSA_CASE)
#undef SSA_CASE
    return 0;
}

template <int R, int G, bool F16>
static hipError_t launch_pair_t(hipStream_t st, const ScoreLaunch& L, const ScoreParams& P, int max_wg) {
    const int nwg = std::max(1, std::min(L.ntasks, max_wg));
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(L.counter), 1, 0u);
    if (e == hipSuccess && L.trace)
        e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(L.trace), (int64_t)L.ntasks * (sizeof(TaskTrace) / 4), 0u);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((score_kernel_pair<R, G, F16>), dim3(nwg), dim3(64 * kWavesPerWG), 0, st, L, P);
    return hipGetLastError();
}

template <int R, int G, bool F16>
static int resident_pair_t() {
    static int cap = 0;
    if (!cap) {
        int per_cu = 0, dev = 0;
        hipDeviceProp_t prop;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, score_kernel_pair<R, G, F16>, 64 * kWavesPerWG,
                                                         0) != hipSuccess ||
            hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
            return 0;
        cap = std::max(1, per_cu) * prop.multiProcessorCount;
    }
    return cap;
}

template <bool F16>
hipError_t launch_pair_mode(int lanes, int rows, hipStream_t st, const ScoreLaunch& L, const ScoreParams& P,
                            int max_wg) {
#define SSA_CASE(GG, RR) \
    if (lanes == GG && rows == RR) return launch_pair_t<RR, GG, F16>(st, L, P, max_wg);
    SSA_PAIR_SHAPES(SSA_CASE)
#undef SSA_CASE
    return hipErrorInvalidValue;
}

template <bool F16>
int resident_pair_mode(int lanes, int rows) {
#define SSA_CASE(GG, RR) \
    if (lanes == GG && rows == RR) return resident_pair_t<RR, GG, F16>();
    SSA_PAIR_SHAPES(SSA_CASE)
#undef SSA_CASE
    return 0;
}
#endif  // SSA_SCORE_INSTANTIATE

}  // namespace ssa
