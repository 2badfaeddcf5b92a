// score_kernels.hip -- gfx950 kernels of the database search.
//
// score_kernel<R, MODE>: inter-sequence SW (MODE 0) / NW (MODE 1) over packed
// int16 lanes.  Restates the per-lane arithmetic of the reference's width
// kernels (src/algo/simd/search_simd_sw.c:172-277, search_simd_nw.c:180-312):
// saturating adds, the H/E/F recurrence and the overflow predicates, but not
// their schedule: each lane half owns a fixed stream of pairs and the host
// replays the reference's lane schedule from the per-pair results.
//
// Work shape: one wave = 64 lanes x 2 halves = 128 pair streams sharing one
// query.  The query is cut into strips of R rows; a strip's H and E columns
// live in VGPRs (2 x R packed registers), DB columns stream through in blocks
// of four.  The strip's score profile is staged in LDS and read with the DB
// code as a per-lane byte offset and the row as an immediate offset.  Strip
// boundaries (H and F of the strip's last row, per column) go through HBM.
//
// The host pads the query to whole strips.  Padding rows sit below the last
// query row, so no real cell depends on them, and their profile values keep
// them from changing any result: SW pads with the lane minimum (a padded cell
// never exceeds the real maximum), NW pads with 0 and resets padded E and H
// to 0 at each pair start (a padded cell stays between min(0, earlier real
// cells) and max(0, earlier real cells), so it raises no overflow flag the
// real cells do not).  The NW score is read from the last real row.
#include <hip/hip_runtime.h>

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
__device__ __forceinline__ int lo16(s16x2 v) { return (int)(short)(as_u(v) & 0xffffu); }
__device__ __forceinline__ int hi16(s16x2 v) { return (int)(short)(as_u(v) >> 16); }

template <int MODE>
__device__ __forceinline__ void finish_pair(PairResult* results, int pair, bool first, bool last, int score,
                                            int fa, int fb) {
    PairResult o;
    if (first) {
        o.score = score;
        o.flag_a = fa;
        o.flag_b = fb;
    } else {
        const PairResult old = results[pair];
        o.score = MODE == 0 ? max(old.score, score) : (last ? score : old.score);
        o.flag_a = min(old.flag_a, fa);
        o.flag_b = min(old.flag_b, fb);
    }
    o.pad = 0;
    results[pair] = o;
}

template <int R, int MODE>
__global__ __launch_bounds__(64) void score_kernel(const ScoreGroup* __restrict__ groups,
                                                   const QueryDesc* __restrict__ queries,
                                                   const int16_t* __restrict__ qp, const uint32_t* __restrict__ lb,
                                                   const uint2* __restrict__ res, uint4* __restrict__ bnd,
                                                   const int32_t* __restrict__ slot_begin,
                                                   const int32_t* __restrict__ pair_list,
                                                   PairResult* __restrict__ results, const ScoreParams P) {
    __shared__ int16_t sqp[R * 32];
    __shared__ uint32_t slb[2 * R + 1];

    const int lane = threadIdx.x;
    const ScoreGroup G = groups[blockIdx.x];
    const QueryDesc Q = queries[G.query];
    const int qlen = Q.qlen;                   // real rows
    const int nstrips = (qlen + R - 1) / R;    // profile rows are padded to nstrips * R
    const int last_row = qlen - 1 - (nstrips - 1) * R;  // last real row within the last strip
    const uint2* __restrict__ gres = res + G.res_off;
    uint4* __restrict__ gbnd = bnd + G.bnd_off;
    const int sb_lo = slot_begin[G.slot_base + 2 * lane];
    const int sb_hi = slot_begin[G.slot_base + 2 * lane + 1];

    const s16x2 QR = as_v(P.qr), RR = as_v(P.r), Z = as_v(P.zero);

    for (int s = 0; s < nstrips; ++s) {
        const int i0 = s * R;
        const bool first = (s == 0), last = (s == nstrips - 1);

        __syncthreads();
        {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(qp + Q.qp_off) + (int64_t)i0 * 16;
            uint32_t* dst = reinterpret_cast<uint32_t*>(sqp);
            for (int k = lane; k < R * 16; k += 64) dst[k] = src[k];
            if (MODE == 1) {
                for (int k = lane; k < 2 * R; k += 64) slb[k] = lb[Q.lb_off + 2 * i0 + k];
                if (lane == 0) slb[2 * R] = (i0 > 0) ? lb[Q.lb_off + 2 * (i0 - 1)] : 0u;
            }
        }
        __syncthreads();

        s16x2 Hl[R], El[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Hl[r] = Z;
            El[r] = Z;
        }
        s16x2 S = Z;     // SW: running maximum of the strip's cells
        s16x2 htop = Z;  // strips after the first: H(i0-1, j-1) for the next column
        s16x2 cap = Z;   // NW: last-row H at the pair's last column
        s16x2 hb[4], fb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            hb[k] = as_v(P.hb_init[k]);
            fb[k] = as_v(P.fb_init[k]);
        }
        int kp_lo = 0, kp_hi = 0, pair_lo = -1, pair_hi = -1, blk_lo = 0, blk_hi = 0;
        int fa_lo = kNoFlag, fa_hi = kNoFlag, fb_lo = kNoFlag, fb_hi = kNoFlag;

        for (int b = 0; b < G.nblocks; ++b) {
            const uint2 w = gres[(int64_t)b * 64 + lane];
            const bool st_lo = (w.x & kResStart) != 0, st_hi = (w.y & kResStart) != 0;
            uint4 tb0 = make_uint4(0, 0, 0, 0), tb1 = make_uint4(0, 0, 0, 0);
            if (!first) {
                tb0 = gbnd[((int64_t)b * 64 + lane) * 2];
                tb1 = gbnd[((int64_t)b * 64 + lane) * 2 + 1];
            }
            if (__any(st_lo || st_hi)) {
                const uint32_t m = (st_lo ? 0x0000ffffu : 0u) | (st_hi ? 0xffff0000u : 0u);
                if (st_lo) {
                    pair_lo = pair_list[sb_lo + kp_lo];
                    kp_lo++;
                    blk_lo = 0;
                    fa_lo = kNoFlag;
                    fb_lo = kNoFlag;
                }
                if (st_hi) {
                    pair_hi = pair_list[sb_hi + kp_hi];
                    kp_hi++;
                    blk_hi = 0;
                    fa_hi = kNoFlag;
                    fb_hi = kNoFlag;
                }
                if (MODE == 0) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        Hl[r] = sel(m, Z, Hl[r]);
                        El[r] = sel(m, Z, El[r]);
                    }
                    S = sel(m, Z, S);
                    htop = sel(m, Z, htop);
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        Hl[r] = sel(m, as_v(slb[2 * r]), Hl[r]);
                        El[r] = sel(m, as_v(slb[2 * r + 1]), El[r]);
                    }
                    htop = sel(m, as_v(slb[2 * R]), htop);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        hb[k] = sel(m, as_v(P.hb_init[k]), hb[k]);
                        fb[k] = sel(m, as_v(P.fb_init[k]), fb[k]);
                    }
                }
            }
            const uint32_t tbh[4] = {tb0.x, tb0.z, tb1.x, tb1.z};
            const uint32_t tbf[4] = {tb0.y, tb0.w, tb1.y, tb1.w};
            uint32_t obh[4], obf[4];
            s16x2 hmn = as_v(0u), hmx = as_v(0u);  // NW trackers, reset value of the reference
            const bool ls_lo = (w.x & kResLast) != 0, ls_hi = (w.y & kResLast) != 0;
            const int z_lo = (int)((w.x >> kResZShift) & 3u), z_hi = (int)((w.y >> kResZShift) & 3u);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t alo = (w.x >> (8 * c)) & 0x3eu;
                const uint32_t ahi = (w.y >> (8 * c)) & 0x3eu;
                s16x2 hd, f;
                if (first) {
                    if (MODE == 0) {
                        hd = Z;
                        f = Z;
                    } else {
                        hd = hb[c];
                        f = adds(fb[c], QR);
                    }
                } else {
                    hd = htop;
                    f = as_v(tbf[c]);
                    htop = as_v(tbh[c]);
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const char* rowp = reinterpret_cast<const char*>(sqp) + r * 64;
                    s16x2 v;
                    v.x = *reinterpret_cast<const int16_t*>(rowp + alo);
                    v.y = *reinterpret_cast<const int16_t*>(rowp + ahi);
                    s16x2 h = adds(hd, v);
                    h = vmax(h, f);
                    h = vmax(h, El[r]);
                    if (MODE == 0) {
                        S = vmax(S, h);
                    } else {
                        hmn = vmin(hmn, h);
                        hmx = vmax(hmx, h);
                    }
                    hd = Hl[r];
                    Hl[r] = h;
                    const s16x2 t = adds(h, QR);
                    El[r] = vmax(adds(El[r], RR), t);
                    f = vmax(adds(f, RR), t);
                }
                obh[c] = as_u(Hl[R - 1]);
                obf[c] = as_u(f);
                if (MODE == 1 && last) {
                    const uint32_t cm = ((ls_lo && z_lo == c) ? 0x0000ffffu : 0u) |
                                        ((ls_hi && z_hi == c) ? 0xffff0000u : 0u);
                    if (__any(cm != 0u)) {
                        s16x2 hq = Hl[0];
#pragma unroll
                        for (int r = 1; r < R; ++r)
                            if (r == last_row) hq = Hl[r];
                        cap = sel(cm, hq, cap);
                    }
                }
            }
            if (!last) {
                gbnd[((int64_t)b * 64 + lane) * 2] = make_uint4(obh[0], obf[0], obh[1], obf[1]);
                gbnd[((int64_t)b * 64 + lane) * 2 + 1] = make_uint4(obh[2], obf[2], obh[3], obf[3]);
            }

            // overflow predicates at the end of the block
            if (MODE == 0) {
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
                if (pair_lo >= 0 && fa_lo == kNoFlag && (lo16(hmn) < P.thr_a || lo16(hmx) >= P.thr_b)) fa_lo = blk_lo;
                if (pair_hi >= 0 && fa_hi == kNoFlag && (hi16(hmn) < P.thr_a || hi16(hmx) >= P.thr_b)) fa_hi = blk_hi;
                if (first) {  // the reference advances the top boundary by four columns per block
                    fb[0] = adds(fb[3], RR);
                    fb[1] = adds(fb[0], RR);
                    fb[2] = adds(fb[1], RR);
                    fb[3] = adds(fb[2], RR);
                    hb[0] = adds(hb[3], RR);
                    hb[1] = adds(hb[0], RR);
                    hb[2] = adds(hb[1], RR);
                    hb[3] = adds(hb[2], RR);
                }
            }
            if (ls_lo && pair_lo >= 0) {
                finish_pair<MODE>(results, pair_lo, first, last, MODE == 0 ? lo16(S) : lo16(cap), fa_lo, fb_lo);
                pair_lo = -1;
            }
            if (ls_hi && pair_hi >= 0) {
                finish_pair<MODE>(results, pair_hi, first, last, MODE == 0 ? hi16(S) : hi16(cap), fa_hi, fb_hi);
                pair_hi = -1;
            }
            blk_lo++;
            blk_hi++;
        }
    }
}

// ---------------------------------------------------------------------------
// Exact tier: int32 Gotoh, one thread per pair.  Restates full_sw / full_nw
// (src/algo/64/smith_waterman_63.c:32-98, needleman_wunsch_64.c:32-98); the
// host checks that every value of the pair fits in 32 bits.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void exact_kernel(const ExactPair* __restrict__ pairs, int npairs,
                                                   const QueryDesc* __restrict__ queries,
                                                   const int32_t* __restrict__ qp32,
                                                   const uint8_t* __restrict__ codes, int mode, int go, int ge,
                                                   int32_t* __restrict__ scratch, int64_t scratch_stride,
                                                   int64_t* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const ExactPair p = pairs[t];
    const QueryDesc Q = queries[p.query];
    const int32_t* prof = qp32 + Q.qp_off;  // [qlen][32]
    const uint8_t* d = codes + p.res_off;
    int32_t* hh = scratch + (int64_t)t * scratch_stride;
    int32_t* ee = hh + Q.qlen;
    const int qlen = Q.qlen;
    if (mode == 0) {
        for (int i = 0; i < qlen; ++i) {
            hh[i] = 0;
            ee[i] = 0;
        }
        int32_t best = 0;
        for (int j = 0; j < p.dlen; ++j) {
            const int dc = d[j];
            int32_t f = 0, diag = 0;
            for (int i = 0; i < qlen; ++i) {
                const int32_t up_left = hh[i];
                int32_t e = ee[i];
                int32_t h = diag + prof[i * 32 + dc];
                h = max(h, e);
                h = max(h, f);
                h = max(h, 0);
                best = max(best, h);
                hh[i] = h;
                const int32_t open = h + go + ge;
                ee[i] = max(e + ge, open);
                f = max(f + ge, open);
                diag = up_left;
            }
        }
        out[p.out] = best;
    } else {
        for (int i = 0; i < qlen; ++i) {
            hh[i] = go + (i + 1) * ge;
            ee[i] = 2 * go + (i + 2) * ge;
        }
        for (int j = 0; j < p.dlen; ++j) {
            const int dc = d[j];
            int32_t f = 2 * go + (j + 2) * ge;
            int32_t diag = (j == 0) ? 0 : go + j * ge;
            for (int i = 0; i < qlen; ++i) {
                const int32_t up_left = hh[i];
                int32_t e = ee[i];
                int32_t h = diag + prof[i * 32 + dc];
                h = max(h, f);
                h = max(h, e);
                hh[i] = h;
                const int32_t open = h + go + ge;
                ee[i] = max(e + ge, open);
                f = max(f + ge, open);
                diag = up_left;
            }
        }
        out[p.out] = qlen ? hh[qlen - 1] : 0;
    }
}

// ---------------------------------------------------------------------------
// Interleaves the slots of each group into the score kernel's residue layout.
// One workgroup per group, one lane per lane: the writes of a block are one
// coalesced 512-byte row.  Codes are doubled into LDS byte offsets (codes are
// below 32, so doubling a packed word never carries across bytes).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pack_word(const uint8_t* codes, const PackSlot& s, int b) {
    if (b >= s.nblocks) return 0u;
    uint32_t w = *reinterpret_cast<const uint32_t*>(codes + s.code_off + 4 * (int64_t)b) << 1;
    if (b == 0) w |= kResStart;
    if (b == s.nblocks - 1) w |= kResLast | ((uint32_t)s.last_col << kResZShift);
    return w;
}

__global__ __launch_bounds__(64) void pack_kernel(const ScoreGroup* __restrict__ groups,
                                                  const PackSlot* __restrict__ slots,
                                                  const uint8_t* __restrict__ codes, uint2* __restrict__ res) {
    const ScoreGroup G = groups[blockIdx.x];
    const int lane = threadIdx.x;
    const PackSlot lo = slots[G.slot_base + 2 * lane];
    const PackSlot hi = slots[G.slot_base + 2 * lane + 1];
    uint2* out = res + G.res_off;
    for (int b = 0; b < G.nblocks; ++b) out[(int64_t)b * 64 + lane] = make_uint2(pack_word(codes, lo, b), pack_word(codes, hi, b));
}

// ---------------------------------------------------------------------------
// launch wrappers (host side)
// ---------------------------------------------------------------------------
template <int R, int MODE>
static hipError_t launch_score_r(int ngroups, hipStream_t st, const ScoreGroup* groups, const QueryDesc* queries,
                                 const int16_t* qp, const uint32_t* lb, const uint2* res, uint4* bnd,
                                 const int32_t* slot_begin, const int32_t* pair_list, PairResult* results,
                                 const ScoreParams& P) {
    hipLaunchKernelGGL((score_kernel<R, MODE>), dim3(ngroups), dim3(64), 0, st, groups, queries, qp, lb, res, bnd,
                       slot_begin, pair_list, results, P);
    return hipGetLastError();
}

hipError_t launch_score(int rows, int mode, int ngroups, hipStream_t st, const ScoreGroup* groups,
                        const QueryDesc* queries, const int16_t* qp, const uint32_t* lb, const uint2* res, uint4* bnd,
                        const int32_t* slot_begin, const int32_t* pair_list, PairResult* results,
                        const ScoreParams& P) {
    if (ngroups <= 0) return hipSuccess;
#define SSA_CASE(RR)                                                                                        \
    case RR:                                                                                                \
        return mode == 0 ? launch_score_r<RR, 0>(ngroups, st, groups, queries, qp, lb, res, bnd, slot_begin, \
                                                 pair_list, results, P)                                     \
                         : launch_score_r<RR, 1>(ngroups, st, groups, queries, qp, lb, res, bnd, slot_begin, \
                                                 pair_list, results, P);
    switch (rows) {
        SSA_CASE(8)
        SSA_CASE(16)
        SSA_CASE(24)
        SSA_CASE(32)
        SSA_CASE(48)
        default:
            return hipErrorInvalidValue;
    }
#undef SSA_CASE
}

hipError_t launch_pack(int ngroups, hipStream_t st, const ScoreGroup* groups, const PackSlot* slots,
                       const uint8_t* codes, uint2* res) {
    if (ngroups <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, dim3(ngroups), dim3(64), 0, st, groups, slots, codes, res);
    return hipGetLastError();
}

hipError_t launch_exact(int npairs, hipStream_t st, const ExactPair* pairs, const QueryDesc* queries,
                        const int32_t* qp32, const uint8_t* codes, int mode, int go, int ge, int32_t* scratch,
                        int64_t scratch_stride, int64_t* out) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(exact_kernel, dim3((npairs + 63) / 64), dim3(64), 0, st, pairs, npairs, queries, qp32, codes,
                       mode, go, ge, scratch, scratch_stride, out);
    return hipGetLastError();
}

}  // namespace ssa
