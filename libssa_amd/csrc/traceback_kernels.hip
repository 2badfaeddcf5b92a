// traceback_kernels.hip -- recomputes each hit and emits its CIGAR runs.
//
// traceback_wg_kernel: one workgroup per hit, the query's passes spread over
// its waves, int32 (the host checks that every value of the hit fits).
// traceback_kernel: one thread per hit in int64, for hits whose values might
// not fit int32.
//
// Both restate the reference's region
// search (src/algo/align.c:39-150), direction matrix (src/algo/cigar.c:48-198)
// and back-trace (src/algo/cigar.c:265-346), including their quirks: the local
// end cell is the first strict maximum in DB-row-major order, the reverse
// pass re-initialises only indices up to b_end, and the walk follows a gap
// flag of the current cell before a diagonal step.  Only the rectangle up to
// the end cell is computed; cells past it cannot influence the walk.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ssa {

constexpr uint8_t kGapUp = 1, kGapLeft = 2, kExtUp = 4, kExtLeft = 8;

__device__ __forceinline__ int64_t mx64(int64_t a, int64_t b) { return a > b ? a : b; }

__global__ __launch_bounds__(64) void traceback_kernel(const TraceTask* __restrict__ tasks, int ntasks,
                                                       const uint8_t* __restrict__ codes,
                                                       const int64_t* __restrict__ mat, int go, int ge,
                                                       uint8_t* __restrict__ dirs, int64_t* __restrict__ scratch,
                                                       int32_t* __restrict__ ops, TraceResult* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntasks) return;
    const TraceTask T = tasks[t];
    const uint8_t* a = codes + T.q_off;  // query
    const uint8_t* b = codes + T.d_off;  // DB sequence
    const int64_t alen = T.qlen, blen = T.dlen;
    int64_t* HH = scratch + T.hh_off;
    const int64_t size = alen > blen ? alen : blen;
    int64_t* EE = HH + size;
    TraceResult R;
    R.nops = 0;
    R.status = 0;

    if (T.type == 0) {
        // forward pass: rows over the DB sequence, columns over the query, S(query, db)
        int64_t score = 0;
        bool found = false;
        for (int64_t j = 0; j < alen; ++j) {
            HH[j] = 0;
            EE[j] = go;
        }
        for (int64_t i = 0; i < blen; ++i) {
            int64_t h = 0, p = 0, f = go;
            for (int64_t j = 0; j < alen; ++j) {
                f = mx64(f, h + go) + ge;
                EE[j] = mx64(EE[j], HH[j] + go) + ge;
                h = p + mat[a[j] * 32 + b[i]];
                if (h < 0) h = 0;
                if (f > h) h = f;
                if (EE[j] > h) h = EE[j];
                p = HH[j];
                HH[j] = h;
                if (h > score) {
                    score = h;
                    R.a_end = j;
                    R.b_end = i;
                    found = true;
                }
            }
        }
        if (!found) {
            R.status = 1;
            out[t] = R;
            return;
        }
        // reverse pass from the end cell
        for (int64_t j = R.b_end; j >= 0; --j) {
            HH[j] = -1;
            EE[j] = -1;
        }
        int64_t cost = 0;
        bool done = false;
        for (int64_t i = R.b_end; i >= 0 && !done; --i) {
            int64_t h = -1, f = -1, p = (i == R.b_end) ? 0 : -1;
            for (int64_t j = R.a_end; j >= 0; --j) {
                f = mx64(f, h + go) + ge;
                EE[j] = mx64(EE[j], HH[j] + go) + ge;
                h = p + mat[a[j] * 32 + b[i]];
                if (f > h) h = f;
                if (EE[j] > h) h = EE[j];
                p = HH[j];
                HH[j] = h;
                if (h > cost) {
                    cost = h;
                    R.a_begin = j;
                    R.b_begin = i;
                    if (cost >= score) {
                        done = true;
                        break;
                    }
                }
            }
        }
        if (!done) {
            R.status = 1;
            out[t] = R;
            return;
        }
    } else {
        R.a_begin = 0;
        R.a_end = alen - 1;
        R.b_begin = 0;
        R.b_end = blen - 1;
    }

    // direction bits for columns j <= b_end (DB) and rows i <= a_end (query), S(db, query)
    const int64_t rows = R.a_end + 1, cols = R.b_end + 1;
    uint8_t* D = dirs + T.dir_off;
    int64_t* hh = HH;
    int64_t* ee = HH + size;
    for (int64_t i = 0; i < rows; ++i) {
        if (T.type == 1) {
            hh[i] = go + (i + 1) * ge;
            ee[i] = 2 * (int64_t)go + (i + 2) * ge;
        } else {
            hh[i] = 0;
            ee[i] = 0;
        }
    }
    for (int64_t j = 0; j < cols; ++j) {
        int64_t f, h;
        if (T.type == 1) {
            f = 2 * (int64_t)go + (j + 2) * ge;
            h = (j == 0) ? 0 : go + j * ge;
        } else {
            f = 0;
            h = 0;
        }
        for (int64_t i = 0; i < rows; ++i) {
            uint8_t d = 0;
            const int64_t nn = hh[i];
            int64_t e = ee[i];
            h += mat[b[j] * 32 + a[i]];
            if (f > h) {
                d |= kGapUp;
                h = f;
            }
            if (e > h) {
                h = e;
                d |= kGapLeft;
            }
            if (T.type == 0 && h < 0) h = 0;
            hh[i] = h;
            h += go + ge;
            e += ge;
            f += ge;
            if (f > h) d |= kExtUp;
            else f = h;
            if (e > h) d |= kExtLeft;
            else e = h;
            ee[i] = e;
            h = nn;
            D[rows * j + i] = d;
        }
    }

    // back-trace: runs are written from the end of the alignment to its start
    int32_t* O = ops + T.ops_off * 2;
    int64_t i = R.a_end, j = R.b_end;
    char prev = 0;
    int32_t cnt = 0, nops = 0;
    while (i >= 0 && j >= 0 && i >= R.a_begin && j >= R.b_begin) {
        const uint8_t d = D[rows * j + i];
        char op;
        if ((d & kGapLeft) || (d & kExtLeft)) {
            --j;
            op = 'I';
        } else if ((d & kGapUp) || (d & kExtUp)) {
            --i;
            op = 'D';
        } else {
            --i;
            --j;
            op = 'M';
        }
        if (op == prev) {
            ++cnt;
        } else {
            if (cnt > 0) {
                if (nops >= T.ops_cap) {
                    R.status = 2;
                    break;
                }
                O[2 * nops] = prev;
                O[2 * nops + 1] = cnt;
                ++nops;
            }
            prev = op;
            cnt = 1;
        }
    }
    if (R.status == 0 && cnt > 0) {
        if (nops >= T.ops_cap) R.status = 2;
        else {
            O[2 * nops] = prev;
            O[2 * nops + 1] = cnt;
            ++nops;
        }
    }
    R.nops = nops;
    out[t] = R;
}

// ---------------------------------------------------------------------------
// Helpers of the workgroup traceback below.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t tb_shr(int32_t v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);  // wave_shr:1
}
__device__ __forceinline__ int32_t mx(int32_t a, int32_t b) { return a > b ? a : b; }

// (value, column, row) candidates ordered by value desc, then column, then row
__device__ __forceinline__ bool tb_better(int32_t v, int32_t c, int32_t r, int32_t bv, int32_t bc, int32_t br) {
    if (v != bv) return v > bv;
    if (c != bc) return c < bc;
    return r < br;
}

// ---------------------------------------------------------------------------
// Workgroup traceback.  The three dynamic programmes of a hit (the local
// forward and reverse region searches, the direction matrix) each run as a
// systolic pipeline: lane g owns R consecutive query rows of a pass and walks
// the DB columns one column behind lane g-1, which hands it the carries of its
// last row (DPP wave shift); the bottom lane stores its carries for the next
// pass.  Each programme keeps the reference's cell order for the quantities
// that depend on it: the first strict maximum in (DB column, query row) order
// for the end cell, the first cell reaching the score in reverse order for the
// start cell.  The query's passes are spread over the W waves of the
// workgroup.  A pass is
// 64 * R query rows (lane g owns R consecutive rows) and walks the columns in
// blocks of kTgS; pass p runs on wave p mod W, starting kTgL blocks after
// pass p - 1 (so that the bottom lane of pass p - 1 has written a column's
// carries before the top lane of pass p reads them: kTgL - 1 blocks cover the
// 63-step skew between them) and not before its wave has finished pass p - W.
// All waves step through the same phases, one column block each, separated by
// workgroup barriers; the carries between passes go through the hit's scratch
// in global memory.
// ---------------------------------------------------------------------------
constexpr int kTgS = 16;        // columns per phase
constexpr int kTgL = 5;         // phases between the starts of consecutive passes
constexpr int kTgMaxPasses = 256;
static_assert((kTgL - 1) * kTgS >= 63, "a pass's bottom lane must lead the next pass's top lane");

// start phase of every pass into s_start (thread 0), then a barrier
__device__ void tg_schedule(int P, int nblk, int W, int* s_start) {
    if (threadIdx.x == 0) {
        for (int p = 0; p < P; ++p) {
            int st = p == 0 ? 0 : s_start[p - 1] + kTgL;
            if (p >= W) st = max(st, s_start[p - W] + nblk);
            s_start[p] = st;
        }
    }
    __syncthreads();
}

// (value, column, row) best over the workgroup: each wave's lane 0 holds its
// wave's best after the shuffle reduction
__device__ void tg_reduce(int32_t& bv, int32_t& bcol, int32_t& brow, int32_t* s_red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, W = blockDim.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        const int32_t v2 = __shfl_xor(bv, off), c2 = __shfl_xor(bcol, off), r2 = __shfl_xor(brow, off);
        if (tb_better(v2, c2, r2, bv, bcol, brow)) {
            bv = v2;
            bcol = c2;
            brow = r2;
        }
    }
    if (lane == 0) {
        s_red[3 * w] = bv;
        s_red[3 * w + 1] = bcol;
        s_red[3 * w + 2] = brow;
    }
    __syncthreads();
    bv = s_red[0];
    bcol = s_red[1];
    brow = s_red[2];
    for (int i = 1; i < W; ++i)
        if (tb_better(s_red[3 * i], s_red[3 * i + 1], s_red[3 * i + 2], bv, bcol, brow)) {
            bv = s_red[3 * i];
            bcol = s_red[3 * i + 1];
            brow = s_red[3 * i + 2];
        }
    __syncthreads();  // s_red is free again
}

// Region search (src/algo/align.c:55-83 forward, :85-129 reverse) over Q rows
// and C columns, in the pass's own orientation: rows run along the query
// (index j of the reference), columns along the DB sequence (index i).
// FWD: rows 0..Q-1 = query 0..alen-1, columns = DB 0..blen-1, zero floor,
//      finds the first strict maximum (value, column, row).
// !FWD: rows = query a_end..0, columns = DB b_end..0, no floor, finds the
//      first cell in (column, row) order whose value reaches `target`.
template <int R, bool FWD>
__device__ void tg_region(const int32_t* smat, const uint8_t* qc, const uint8_t* dc, int Q, int C, int32_t go,
                          int32_t ge, int32_t target, int32_t* bnd, int32_t* stale_h, int32_t* stale_e,
                          int stale_from, int* s_start, int32_t* s_red, int32_t* res_v, int32_t* res_c,
                          int32_t* res_r) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, W = blockDim.x >> 6;
    constexpr int kRows = 64 * R;
    const int P = (Q + kRows - 1) / kRows;
    const int nblk = (C + 63 + kTgS - 1) / kTgS;
    tg_schedule(P, nblk, W, s_start);
    int32_t bv = 0, bcol = INT32_MAX, brow = INT32_MAX;
    bool found = false;
    int p = w, r0 = 0, nrow = 0;
    int32_t H[R], E[R];
    int qcode[R];
    int32_t oh = 0, of = 0, op = 0;
    const int total = P > 0 ? s_start[P - 1] + nblk : 0;
    for (int ph = 0; ph < total; ++ph) {
        if (p < P && ph >= s_start[p]) {
            const int blk = ph - s_start[p];
            if (blk == 0) {
                r0 = p * kRows + lane * R;
                nrow = max(0, min(R, Q - r0));
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int r = r0 + k;
                    qcode[k] = k < nrow ? (int)qc[FWD ? r : -r] : 0;
                    if (FWD) {
                        H[k] = 0;
                        E[k] = go;
                    } else {
                        const bool fresh = k >= nrow || r >= stale_from;
                        H[k] = fresh ? -1 : stale_h[r];
                        E[k] = fresh ? -1 : stale_e[r];
                    }
                }
                oh = of = op = 0;
            }
            for (int t = blk * kTgS; t < blk * kTgS + kTgS; ++t) {
                const int c = t - lane;
                int32_t ih = tb_shr(oh), ifr = tb_shr(of), ip = tb_shr(op);
                const bool valid = c >= 0 && c < C;
                if (lane == 0 && valid) {
                    if (p == 0) {
                        ih = FWD ? 0 : -1;
                        ifr = FWD ? go : -1;
                        ip = FWD ? 0 : (c == 0 ? 0 : -1);
                    } else {
                        ih = bnd[3 * c];
                        ifr = bnd[3 * c + 1];
                        ip = bnd[3 * c + 2];
                    }
                }
                if (valid) {
                    const int dcode = dc[FWD ? c : -c];
                    int32_t h = ih, f = ifr, pp = ip;
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        if (k < nrow) {
                            f = mx(f, h + go) + ge;
                            E[k] = mx(E[k], H[k] + go) + ge;
                            int32_t hn = pp + smat[qcode[k] * 32 + dcode];
                            if (FWD) hn = mx(hn, 0);
                            hn = mx(hn, f);
                            hn = mx(hn, E[k]);
                            pp = H[k];
                            H[k] = hn;
                            h = hn;
                            if (FWD) {
                                if (hn > bv || (hn == bv && hn > 0 && c < bcol)) {
                                    bv = hn;
                                    bcol = c;
                                    brow = r0 + k;
                                }
                            } else if (hn >= target && (!found || c < bcol)) {
                                found = true;
                                bcol = c;
                                brow = r0 + k;
                            }
                        }
                    }
                    oh = h;
                    of = f;
                    op = pp;
                    if (lane == 63 && p + 1 < P) {
                        bnd[3 * c] = h;
                        bnd[3 * c + 1] = f;
                        bnd[3 * c + 2] = pp;
                    }
                }
            }
            if (blk == nblk - 1) {
                if (FWD) {
#pragma unroll
                    for (int k = 0; k < R; ++k)
                        if (k < nrow) {
                            stale_h[r0 + k] = H[k];
                            stale_e[r0 + k] = E[k];
                        }
                }
                p += W;
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    if (!FWD && !found) {
        bv = INT32_MIN;
        bcol = INT32_MAX;
        brow = INT32_MAX;
    } else if (!FWD) {
        bv = 1;
    }
    tg_reduce(bv, bcol, brow, s_red);
    *res_v = bv;
    *res_c = bcol;
    *res_r = brow;
}

// Direction matrix (src/algo/cigar.c:48-198) over rows 0..Q-1 (query) and
// columns 0..C-1 (DB): D[rows_pad * column + row].
template <int R>
__device__ void tg_directions(const int32_t* smat, const uint8_t* qc, const uint8_t* dc, int Q, int C, int rows_pad,
                              int32_t go, int32_t ge, bool nw, int32_t* bnd, uint8_t* D, int* s_start) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, W = blockDim.x >> 6;
    constexpr int kRows = 64 * R;
    const int P = (Q + kRows - 1) / kRows;
    const int nblk = (C + 63 + kTgS - 1) / kTgS;
    tg_schedule(P, nblk, W, s_start);
    int p = w, r0 = 0, nrow = 0;
    int32_t hh[R], ee[R];
    int qcode[R];
    int32_t oh = 0, of = 0;
    const int total = P > 0 ? s_start[P - 1] + nblk : 0;
    for (int ph = 0; ph < total; ++ph) {
        if (p < P && ph >= s_start[p]) {
            const int blk = ph - s_start[p];
            if (blk == 0) {
                r0 = p * kRows + lane * R;
                nrow = max(0, min(R, Q - r0));
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int i = r0 + k;
                    qcode[k] = k < nrow ? (int)qc[i] : 0;
                    hh[k] = nw ? go + (i + 1) * ge : 0;
                    ee[k] = nw ? 2 * go + (i + 2) * ge : 0;
                }
                oh = of = 0;
            }
            for (int t = blk * kTgS; t < blk * kTgS + kTgS; ++t) {
                const int j = t - lane;
                int32_t ih = tb_shr(oh), ifr = tb_shr(of);
                const bool valid = j >= 0 && j < C;
                if (lane == 0 && valid) {
                    if (p == 0) {
                        ifr = nw ? 2 * go + (j + 2) * ge : 0;
                        ih = nw ? (j == 0 ? 0 : go + j * ge) : 0;
                    } else {
                        ih = bnd[2 * j];
                        ifr = bnd[2 * j + 1];
                    }
                }
                if (valid) {
                    const int dcode = dc[j];
                    int32_t h = ih, f = ifr;
                    uint32_t lo = 0, hi = 0;
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        if (k < nrow) {
                            uint32_t d = 0;
                            const int32_t nn = hh[k];
                            int32_t e = ee[k];
                            h += smat[dcode * 32 + qcode[k]];
                            if (f > h) {
                                d |= kGapUp;
                                h = f;
                            }
                            if (e > h) {
                                h = e;
                                d |= kGapLeft;
                            }
                            if (!nw && h < 0) h = 0;
                            hh[k] = h;
                            h += go + ge;
                            e += ge;
                            f += ge;
                            if (f > h) d |= kExtUp;
                            else f = h;
                            if (e > h) d |= kExtLeft;
                            else e = h;
                            ee[k] = e;
                            h = nn;
                            if (k < 4) lo |= d << (8 * k);
                            else hi |= d << (8 * (k - 4));
                        }
                    }
                    of = f;
                    oh = h;
                    if (nrow > 0) {
                        uint8_t* dst = D + (int64_t)rows_pad * j + r0;
                        if constexpr (R == 8) *reinterpret_cast<uint2*>(dst) = make_uint2(lo, hi);
                        else if constexpr (R == 4) *reinterpret_cast<uint32_t*>(dst) = lo;
                        else if constexpr (R == 2) *reinterpret_cast<uint16_t*>(dst) = (uint16_t)lo;
                        else *dst = (uint8_t)lo;
                    }
                    if (lane == 63 && p + 1 < P) {
                        bnd[2 * j] = h;
                        bnd[2 * j + 1] = f;
                    }
                }
            }
            if (blk == nblk - 1) p += W;
        }
        __threadfence_block();
        __syncthreads();
    }
}

template <int R>
__global__ __launch_bounds__(1024) void traceback_wg_kernel(const TraceTask* __restrict__ tasks,
                                                           const uint8_t* __restrict__ codes,
                                                           const int64_t* __restrict__ mat, int go, int ge,
                                                           uint8_t* __restrict__ dirs, int64_t* __restrict__ scratch,
                                                           int32_t* __restrict__ ops,
                                                           TraceResult* __restrict__ out) {
    __shared__ int32_t smat[32 * 32];
    __shared__ int s_start[kTgMaxPasses];
    __shared__ int32_t s_red[3 * 16];
    const TraceTask T = tasks[blockIdx.x];
    for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x) smat[i] = (int32_t)mat[i];
    __syncthreads();
    const uint8_t* a = codes + T.q_off;  // query
    const uint8_t* b = codes + T.d_off;  // DB sequence
    const int alen = T.qlen, blen = T.dlen;
    int32_t* sc = reinterpret_cast<int32_t*>(scratch + T.hh_off);
    int32_t* bnd = sc;  // 3 * blen carries between passes
    int32_t* stale_h = sc + 3 * (int64_t)blen;
    int32_t* stale_e = stale_h + alen;
    TraceResult Rs;
    Rs.nops = 0;
    Rs.status = 0;
    if (T.type == 0) {
        int32_t v, c, r;
        tg_region<R, true>(smat, a, b, alen, blen, go, ge, 0, bnd, stale_h, stale_e, 0, s_start, s_red, &v, &c, &r);
        if (v <= 0) {  // the same for the whole workgroup
            if (threadIdx.x == 0) {
                Rs.status = 1;
                out[blockIdx.x] = Rs;
            }
            return;
        }
        Rs.a_end = r;
        Rs.b_end = c;
        const int32_t score = v;
        int32_t v2, c2, r2;
        const int stale_rows = (int)Rs.a_end - (int)Rs.b_end;  // rows k < stale_rows keep stale values
        int32_t* rs_h = stale_e + alen;
        int32_t* rs_e = rs_h + alen;
        for (int k = threadIdx.x; k < (int)Rs.a_end + 1; k += blockDim.x) {
            rs_h[k] = stale_h[Rs.a_end - k];
            rs_e[k] = stale_e[Rs.a_end - k];
        }
        __threadfence_block();
        __syncthreads();
        tg_region<R, false>(smat, a + Rs.a_end, b + Rs.b_end, (int)Rs.a_end + 1, (int)Rs.b_end + 1, go, ge, score,
                            bnd, rs_h, rs_e, stale_rows > 0 ? stale_rows : 0, s_start, s_red, &v2, &c2, &r2);
        if (v2 == INT32_MIN) {
            if (threadIdx.x == 0) {
                Rs.status = 1;
                out[blockIdx.x] = Rs;
            }
            return;
        }
        Rs.a_begin = Rs.a_end - r2;
        Rs.b_begin = Rs.b_end - c2;
    } else {
        Rs.a_begin = 0;
        Rs.a_end = alen - 1;
        Rs.b_begin = 0;
        Rs.b_end = blen - 1;
    }
    const int rows = (int)Rs.a_end + 1, cols = (int)Rs.b_end + 1;
    const int rows_pad = (rows + 7) & ~7;
    uint8_t* D = dirs + T.dir_off;
    tg_directions<R>(smat, a, b, rows, cols, rows_pad, go, ge, T.type == 1, bnd, D, s_start);

    // Back-trace (src/algo/cigar.c:265-346) by thread 0, from LDS: the
    // workgroup copies the 64 x 64 block of direction bytes up and left of the
    // current cell, thread 0 walks it until the path leaves it, and so on, so
    // each step reads LDS instead of waiting on a global load.
    __shared__ uint8_t tile[64 * 64];  // [column - j0][row - i0]
    __shared__ int64_t s_walk[5];      // i, j, i0, j0, done
    int32_t* O = ops + T.ops_off * 2;
    int64_t i = Rs.a_end, j = Rs.b_end;
    char prev = 0;
    int32_t cnt = 0, nops = 0;
    bool done = false;
    for (;;) {
        if (threadIdx.x == 0) {
            done = done || !(i >= 0 && j >= 0 && i >= Rs.a_begin && j >= Rs.b_begin);
            s_walk[0] = i;
            s_walk[1] = j;
            s_walk[2] = max((int64_t)Rs.a_begin, i - 63);
            s_walk[3] = max((int64_t)Rs.b_begin, j - 63);
            s_walk[4] = done ? 1 : 0;
        }
        __syncthreads();
        if (s_walk[4]) break;  // the same for the whole workgroup
        const int64_t ti = s_walk[0], tj = s_walk[1], i0 = s_walk[2], j0 = s_walk[3];
        for (int idx = threadIdx.x; idx < 64 * 64; idx += blockDim.x) {
            const int64_t jj = j0 + idx / 64, ii = i0 + idx % 64;
            if (jj <= tj && ii <= ti) tile[idx] = D[(int64_t)rows_pad * jj + ii];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            while (i >= i0 && j >= j0) {  // i0, j0 >= the region's begin and >= 0
                const uint8_t d = tile[(j - j0) * 64 + (i - i0)];
                char op;
                if ((d & kGapLeft) || (d & kExtLeft)) {
                    --j;
                    op = 'I';
                } else if ((d & kGapUp) || (d & kExtUp)) {
                    --i;
                    op = 'D';
                } else {
                    --i;
                    --j;
                    op = 'M';
                }
                if (op == prev) {
                    ++cnt;
                } else {
                    if (cnt > 0) {
                        if (nops >= T.ops_cap) {
                            Rs.status = 2;
                            done = true;
                            break;
                        }
                        O[2 * nops] = prev;
                        O[2 * nops + 1] = cnt;
                        ++nops;
                    }
                    prev = op;
                    cnt = 1;
                }
            }
        }
        __syncthreads();  // the tile is read before the next one is copied
    }
    if (threadIdx.x != 0) return;
    if (Rs.status == 0 && cnt > 0) {
        if (nops >= T.ops_cap) Rs.status = 2;
        else {
            O[2 * nops] = prev;
            O[2 * nops + 1] = cnt;
            ++nops;
        }
    }
    Rs.nops = nops;
    out[blockIdx.x] = Rs;
}

int traceback_wg_max_passes() { return kTgMaxPasses; }

hipError_t launch_traceback_wg(int rows_per_lane, int waves, int ntasks, hipStream_t st, const TraceTask* tasks,
                               const uint8_t* codes, const int64_t* mat, int go, int ge, uint8_t* dirs,
                               int64_t* scratch, int32_t* ops, TraceResult* out) {
    if (ntasks <= 0) return hipSuccess;
    if (waves < 1 || waves > 16) return hipErrorInvalidValue;
    const dim3 grid(ntasks), block(64 * waves);
    switch (rows_per_lane) {
        case 1: hipLaunchKernelGGL(traceback_wg_kernel<1>, grid, block, 0, st, tasks, codes, mat, go, ge, dirs, scratch, ops, out); break;
        case 2: hipLaunchKernelGGL(traceback_wg_kernel<2>, grid, block, 0, st, tasks, codes, mat, go, ge, dirs, scratch, ops, out); break;
        case 4: hipLaunchKernelGGL(traceback_wg_kernel<4>, grid, block, 0, st, tasks, codes, mat, go, ge, dirs, scratch, ops, out); break;
        case 8: hipLaunchKernelGGL(traceback_wg_kernel<8>, grid, block, 0, st, tasks, codes, mat, go, ge, dirs, scratch, ops, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_traceback(int ntasks, hipStream_t st, const TraceTask* tasks, const uint8_t* codes,
                            const int64_t* mat, int go, int ge, uint8_t* dirs, int64_t* scratch, int32_t* ops,
                            TraceResult* out) {
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(traceback_kernel, dim3((ntasks + 63) / 64), dim3(64), 0, st, tasks, ntasks, codes, mat, go, ge,
                       dirs, scratch, ops, out);
    return hipGetLastError();
}

}  // namespace ssa
