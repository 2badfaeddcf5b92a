// traceback_kernels.hip -- recomputes each hit and emits its CIGAR runs.
//
// One thread per hit restates, in int64 arithmetic, the reference's region
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

hipError_t launch_traceback(int ntasks, hipStream_t st, const TraceTask* tasks, const uint8_t* codes,
                            const int64_t* mat, int go, int ge, uint8_t* dirs, int64_t* scratch, int32_t* ops,
                            TraceResult* out) {
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(traceback_kernel, dim3((ntasks + 63) / 64), dim3(64), 0, st, tasks, ntasks, codes, mat, go, ge,
                       dirs, scratch, ops, out);
    return hipGetLastError();
}

}  // namespace ssa
