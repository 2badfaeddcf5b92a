// score_kernels.hip -- gfx950 kernels of the database search other than the
// score kernel's instantiations (score_kernel.h holds its device code):
// the launchers that dispatch a shape to them, fills, the exact tier, residue
// packing and the no-allocation scratch array.
#include "score_kernel.h"

namespace ssa {


__global__ void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Sets every pair's result to the neutral value of the atomic combination.
__global__ void init_results_kernel(PairResult* __restrict__ r, int64_t n, int32_t score) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = PairResult{score, kNoFlag, kNoFlag};
}

// ---------------------------------------------------------------------------
// Exact tier: the reference's full_sw / full_nw (src/algo/64/smith_waterman_63.c:32-98,
// needleman_wunsch_64.c:32-98): textbook Gotoh with the same recurrence, border
// and result, in T = int32 where the host has shown that every value of the
// pair fits, else int64 as the reference computes.  One wave per pair, as a
// systolic pipeline: lane g owns kExactR consecutive query rows of a pass of
// kExactRows rows and walks the DB columns one column behind lane g - 1,
// which hands it, per column, the vertical gap value F after its last row
// and that row's H of the previous column (the next row's diagonal); the
// bottom lane stores both per column for the next pass.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wave_shr1(T v) {
    if constexpr (sizeof(T) == 4) {
        return (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);  // wave_shr:1
    } else {
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(uint64_t)v, 0x138, 0xf, 0xf, false);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)((uint64_t)v >> 32), 0x138, 0xf, 0xf, false);
        return (T)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
    }
}

template <class T>
__global__ __launch_bounds__(64) void exact_wave_kernel(const ExactPair* __restrict__ pairs,
                                                        const QueryDesc* __restrict__ queries,
                                                        const uint8_t* __restrict__ qcodes,
                                                        const int64_t* __restrict__ mat,
                                                        const uint8_t* __restrict__ codes, int mode, int go_, int ge_,
                                                        T* __restrict__ bnd_all, int64_t* __restrict__ out) {
    __shared__ T smat[32 * 32];  // [db code * 32 + query code]
    for (int i = threadIdx.x; i < 32 * 32; i += 64) smat[i] = (T)mat[i];
    __syncthreads();
    const ExactPair P = pairs[blockIdx.x];
    const QueryDesc Q = queries[P.query];
    const uint8_t* qc = qcodes + Q.qp_off;
    const uint8_t* dc = codes + P.res_off;
    T* bnd = bnd_all + P.bnd_off;
    const int lane = threadIdx.x;
    const int qlen = Q.qlen, dlen = P.dlen;
    const T go = (T)go_, ge = (T)ge_;
    const int npass = (qlen + kExactRows - 1) / kExactRows;
    T best = 0, last = 0;
    bool has_last = false;
    for (int p = 0; p < npass; ++p) {
        const int r0 = p * kExactRows + lane * kExactR;
        const int nrow = max(0, min(kExactR, qlen - r0));
        T hh[kExactR], ee[kExactR];
        int qrow[kExactR];
#pragma unroll
        for (int k = 0; k < kExactR; ++k) {
            const T i = (T)(r0 + k);
            qrow[k] = k < nrow ? (int)qc[r0 + k] : 0;
            hh[k] = mode == 0 ? (T)0 : go + (i + 1) * ge;
            ee[k] = mode == 0 ? (T)0 : 2 * go + (i + 2) * ge;
        }
        T of = 0, od = 0;  // out of this lane's last row at its last column: F, and H of the column before
        for (int t = 0; t < dlen + 63; ++t) {
            const int j = t - lane;
            T f = wave_shr1(of), diag = wave_shr1(od);
            const bool valid = j >= 0 && j < dlen;
            if (lane == 0 && valid) {
                if (p == 0) {
                    f = mode == 0 ? (T)0 : 2 * go + (T)(j + 2) * ge;
                    diag = mode == 0 ? (T)0 : (j == 0 ? (T)0 : go + (T)j * ge);
                } else {
                    f = bnd[2 * j];
                    diag = bnd[2 * j + 1];
                }
            }
            if (valid) {
                const T* srow = smat + 32 * (int)dc[j];
#pragma unroll
                for (int k = 0; k < kExactR; ++k) {
                    if (k < nrow) {
                        const T up_left = hh[k];
                        const T e = ee[k];
                        T h = diag + srow[qrow[k]];
                        if (mode == 0) {
                            h = max(h, e);
                            h = max(h, f);
                            h = max(h, (T)0);
                            best = max(best, h);
                        } else {
                            h = max(h, f);
                            h = max(h, e);
                        }
                        hh[k] = h;
                        const T open = h + go + ge;
                        ee[k] = max(e + ge, open);
                        f = max(f + ge, open);
                        diag = up_left;
                    }
                }
                of = f;
                od = diag;
                if (lane == 63 && p + 1 < npass) {
                    bnd[2 * j] = f;
                    bnd[2 * j + 1] = od;
                }
            }
        }
        // NW: the last query row's H after the last column
        if (mode != 0 && p == npass - 1) {
#pragma unroll
            for (int k = 0; k < kExactR; ++k)
                if (r0 + k == qlen - 1) {
                    last = hh[k];
                    has_last = true;
                }
        }
        __threadfence_block();
        __builtin_amdgcn_wave_barrier();
    }
    if (mode == 0) {
        for (int o = 32; o > 0; o >>= 1) {
            if constexpr (sizeof(T) == 4) {
                best = max(best, (T)__shfl_xor((int)best, o, 64));
            } else {
                const int lo = __shfl_xor((int)(uint32_t)(uint64_t)best, o, 64);
                const int hi = __shfl_xor((int)((uint64_t)best >> 32), o, 64);
                best = max(best, (T)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo));
            }
        }
        if (lane == 0) out[P.out] = (int64_t)best;
    } else if (has_last) {
        out[P.out] = (int64_t)last;
    } else if (qlen == 0 && lane == 0) {
        out[P.out] = 0;
    }
}

// ---------------------------------------------------------------------------
// Residue interleaving: one wave per (stream, entry), lanes over the entry's
// blocks.  Codes are doubled into LDS byte offsets (codes are below 32, so
// doubling a packed word never carries across bytes).  Words past the end of
// a stream stay zero (cleared by fill_u32_kernel).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void pack_kernel(const PackEntry* __restrict__ entries,
                                                  const uint8_t* __restrict__ codes, uint32_t* __restrict__ res) {
    const PackEntry E = entries[blockIdx.x];
    for (int b = threadIdx.x; b < E.nblocks; b += 64) {
        uint32_t w = *reinterpret_cast<const uint32_t*>(codes + E.code_off + 4 * (int64_t)b) << 1;
        if (b == 0) w |= kResStart;
        if (b == E.nblocks - 1) w |= kResLast | ((uint32_t)E.last_col << kResZShift);
        res[E.dst + (int64_t)b * E.stride] = w;
    }
}


// ---------------------------------------------------------------------------
// launch wrappers (host side): dispatch to the per-mode instantiations
// ---------------------------------------------------------------------------
hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, int64_t n, uint32_t v) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

// Passes of at most 512 rows (two staged profiles of 32 KiB); 32 rows per
// lane for SW only (NW holds more state per row and would run one wave per SIMD)
bool score_shape_supported(int lanes, int rows, int mode, int cols) {
    const bool sw = mode == 0 || mode == 3;
    return (lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64) &&
           (rows == 4 || rows == 8 || rows == 16 || (rows == 32 && sw)) && lanes * rows <= 512 &&
           (cols == 4 || (cols == 8 && rows <= 16 && (lanes == 8 || lanes == 16)));
}

bool score_nw_untracked_supported(int lanes, int rows) {
    return (lanes == 8 && (rows == 8 || rows == 16)) || (lanes == 16 && (rows == 8 || rows == 16)) ||
           (lanes == 32 && rows == 8);
}

// the profile in VGPRs: 4-column tiles of at most 16 rows per lane, SW and
// NW without overflow trackers
bool score_vp_supported(int lanes, int rows, int mode, int cols) {
    if (cols != 4 || rows > 16 || !(mode == 0 || mode == 2 || mode == 3 || mode == 4)) return false;
    if (!score_shape_supported(lanes, rows, mode, cols)) return false;
    return (mode == 0 || mode == 3) || score_nw_untracked_supported(lanes, rows);
}

int score_resident_workgroups(int lanes, int rows, int mode, int cols, bool vp) {
    if (vp) {
        switch (mode) {
        case 0: return resident_mode<0, true>(lanes, rows, cols);
        case 2: return resident_mode<2, true>(lanes, rows, cols);
        case 3: return resident_mode<3, true>(lanes, rows, cols);
        case 4: return resident_mode<4, true>(lanes, rows, cols);
        default: return 0;
        }
    }
    switch (mode) {
    case 0: return resident_mode<0, false>(lanes, rows, cols);
    case 1: return resident_mode<1, false>(lanes, rows, cols);
    case 2: return resident_mode<2, false>(lanes, rows, cols);
    case 3: return resident_mode<3, false>(lanes, rows, cols);
    case 4: return resident_mode<4, false>(lanes, rows, cols);
    default: return 0;
    }
}

hipError_t launch_score(int lanes, int rows, int mode, int cols, bool vp, hipStream_t st, const ScoreLaunch& L,
                        const ScoreParams& P, int max_wg) {
    if (L.ntasks <= 0) return hipSuccess;
    if (vp) {
        switch (mode) {
        case 0: return launch_mode<0, true>(lanes, rows, cols, st, L, P, max_wg);
        case 2: return launch_mode<2, true>(lanes, rows, cols, st, L, P, max_wg);
        case 3: return launch_mode<3, true>(lanes, rows, cols, st, L, P, max_wg);
        case 4: return launch_mode<4, true>(lanes, rows, cols, st, L, P, max_wg);
        default: return hipErrorInvalidValue;
        }
    }
    switch (mode) {
    case 0: return launch_mode<0, false>(lanes, rows, cols, st, L, P, max_wg);
    case 1: return launch_mode<1, false>(lanes, rows, cols, st, L, P, max_wg);
    case 2: return launch_mode<2, false>(lanes, rows, cols, st, L, P, max_wg);
    case 3: return launch_mode<3, false>(lanes, rows, cols, st, L, P, max_wg);
    case 4: return launch_mode<4, false>(lanes, rows, cols, st, L, P, max_wg);
    default: return hipErrorInvalidValue;
    }
}

bool score_pair_supported(int lanes, int rows) {
#define SSA_CASE(GG, RR) \
    if (lanes == GG && rows == RR) return true;
    SSA_PAIR_SHAPES(SSA_CASE)
#undef SSA_CASE
    return false;
}

int score_pair_resident_workgroups(int lanes, int rows, bool f16) {
    return f16 ? resident_pair_mode<true>(lanes, rows) : resident_pair_mode<false>(lanes, rows);
}

hipError_t launch_score_pair(int lanes, int rows, bool f16, hipStream_t st, const ScoreLaunch& L, const ScoreParams& P,
                             int max_wg) {
    if (L.ntasks <= 0) return hipSuccess;
    return f16 ? launch_pair_mode<true>(lanes, rows, st, L, P, max_wg)
               : launch_pair_mode<false>(lanes, rows, st, L, P, max_wg);
}

hipError_t launch_init_results(hipStream_t st, PairResult* r, int64_t n, int32_t score) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(init_results_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, r, n, score);
    return hipGetLastError();
}

hipError_t launch_pack(int nentries, hipStream_t st, const PackEntry* entries, const uint8_t* codes,
                       uint32_t* res, int64_t res_words) {
    const hipError_t e = launch_fill_u32(st, res, res_words, 0u);
    if (e != hipSuccess) return e;
    if (nentries <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, dim3(nentries), dim3(64), 0, st, entries, codes, res);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point: every per-call device buffer of a
// search run through it (pair results, task list, profile, pass boundaries,
// counters) is a piece of this array.  Each device holds its own copy from
// the time the library's code object is loaded until the process ends.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

void* scratch_setup(size_t* bytes) {
    // the copy of the device that is current (the caller's set-up runs once
    // per device with it current); host code passes this address to kernels
    // and copies, never the array's name, which is a host-side stand-in
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) fatal("HIP error %d (%s) resolving the device scratch array", (int)e, hipGetErrorString(e));
    *bytes = kScratchBytes;
    return ptr;
}

hipError_t launch_exact(bool wide, int npairs, hipStream_t st, const ExactPair* pairs, const QueryDesc* queries,
                        const uint8_t* qcodes, const int64_t* mat, const uint8_t* codes, int mode, int go, int ge,
                        void* bnd, int64_t* out) {
    if (npairs <= 0) return hipSuccess;
    if (wide)
        hipLaunchKernelGGL(exact_wave_kernel<int64_t>, dim3(npairs), dim3(64), 0, st, pairs, queries, qcodes, mat,
                           codes, mode, go, ge, (int64_t*)bnd, out);
    else
        hipLaunchKernelGGL(exact_wave_kernel<int32_t>, dim3(npairs), dim3(64), 0, st, pairs, queries, qcodes, mat,
                           codes, mode, go, ge, (int32_t*)bnd, out);
    return hipGetLastError();
}

}  // namespace ssa
