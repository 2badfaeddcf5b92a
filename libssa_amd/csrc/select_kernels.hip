// select_kernels.hip -- the first pass of the host's replay of the
// reference's heap, on the GPU: per chunk of DB entries, the largest inserts
// and the overflow counts, reduced over a wavefront.
//
// The reference inserts every accepted score of a chunk into a bounded
// min-heap (src/util/minheap.c:60-91, src/algo/manager.c:141-153) in an order
// set by its lane schedule.  The heap always holds the k largest scores
// inserted so far, so a chunk's k + 1 largest inserts (in any order) bound
// what it can change; the host replays the lane schedule only for chunks
// whose largest inserts exceed the running k-th score.  This kernel computes
// those per-chunk inserts where the pair results already are, so only a few
// bytes per chunk cross to the host instead of 12 per pair.
#include <hip/hip_runtime.h>

#include <climits>

#include "kernels.h"

namespace ssa {

namespace {

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ long long wave_sum(long long v) {
    int lo = (int)(unsigned)(v & 0xffffffffll), hi = (int)(v >> 32);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned a = (unsigned)__shfl_xor(lo, o, 64);
        const int b = __shfl_xor(hi, o, 64);
        const unsigned long long s = (unsigned long long)(unsigned)lo + a;
        hi = hi + b + (int)(s >> 32);
        lo = (int)(unsigned)s;
    }
    return ((long long)hi << 32) | (long long)(unsigned)lo;
}

// a tier result decoded as the host decodes it (the accept rule and the score
// of search.cpp:decode for the 16-bit representation)
struct Dec {
    bool accept;
    int score;
};

__device__ __forceinline__ Dec dec16(const PairResult& r, int type) {
    if (type == 0) return Dec{r.flag_b == kNoFlag, r.score + 32768};
    return Dec{r.flag_a == kNoFlag && r.score > -32768 && r.score < 32767, r.score};
}

}  // namespace

// One wave per chunk.  Each lane takes every 64th entry of the chunk and keeps
// its own m largest inserts, sorted, in LDS (element i of lane l at
// [i * 64 + l]); the wave then takes the chunk's m largest from the 64 lists
// by m rounds of a wave maximum over the lists' heads, the lowest lane holding
// it giving up its head.  The overflow counts are wave sums.
__global__ __launch_bounds__(64) void chunk_top_kernel(const TopSource S, const uint2* __restrict__ chunks,
                                                       ChunkTop* __restrict__ out) {
    __shared__ int lists[kChunkTopMax * 64];
    const int lane = threadIdx.x;
    const uint2 ch = chunks[blockIdx.x];
    const int m = S.m;
    int n = 0;
    auto L = [&](int i) -> int& { return lists[i * 64 + lane]; };
    auto insert = [&](int v, int copies) {
        for (int r = 0; r < copies; ++r) {
            if (n == m && !(v > L(m - 1))) break;
            int i = n < m ? n++ : m - 1;
            while (i > 0 && L(i - 1) < v) {
                L(i) = L(i - 1);
                --i;
            }
            L(i) = v;
        }
    };
    long long ov8 = 0, ov16 = 0;
    int need = 0;
    for (uint32_t e = ch.x + lane; e < ch.y; e += 64) {
        // the inserts of one entry, as the host's replay enumerates them
        int m8 = 1;
        if (S.mode != kTopMode16) {
            m8 = 0;
            for (int q = 0; q < S.nq; ++q) {
                if (S.mode == kTopModeSwSingle) {
                    const PairResult r = S.res16[q][e ^ ((S.flip >> q) & 1u)];
                    if (r.flag_a == kNoFlag) insert(r.score + 32768, 1);
                    else ++m8;
                } else {
                    ++m8;
                }
            }
            ov8 += m8;
            if (m8 == 0) continue;
        }
        int m16 = 0;
        for (int q = 0; q < S.nq; ++q) {
            const Dec d = dec16(S.res16[q][e ^ ((S.flip >> q) & 1u)], S.type);
            if (d.accept) insert(d.score, m8);
            else ++m16;
        }
        ov16 += (long long)m8 * m16;
        need |= m16 != 0;
    }
    // the chunk's m largest over the lanes' lists
    __shared__ int s_top[kChunkTopMax];
    int hd = 0, cnt = 0;
    for (int o = 0; o < m; ++o) {
        const bool valid = hd < n;
        if (!__any(valid)) break;
        const int v = valid ? L(hd) : INT_MIN;
        const int mx = wave_max(v);
        const unsigned long long b = __ballot(valid && v == mx);
        if (lane == __ffsll((long long)b) - 1) ++hd;
        if (lane == 0) s_top[o] = mx;
        cnt = o + 1;
    }
    ov8 = wave_sum(ov8);
    ov16 = wave_sum(ov16);
    need = __any(need != 0) ? 1 : 0;
    __syncthreads();
    ChunkTop* o = out + blockIdx.x;
    if (lane < cnt) o->top[lane] = s_top[lane];
    if (lane == 0) {
        o->ov8 = ov8;
        o->ov16 = ov16;
        o->ntop = cnt;
        o->need_exact = need;
    }
}

// One workgroup per range: copies segs[i].n pair results to out[segs[i].dst ..].
__global__ __launch_bounds__(256) void gather_kernel(const GatherSeg* __restrict__ segs, PairResult* __restrict__ out) {
    const GatherSeg g = segs[blockIdx.x];
    for (int i = threadIdx.x; i < g.n; i += 256) out[g.dst + i] = g.src[i];
}

hipError_t launch_chunk_top(int nchunks, hipStream_t st, const TopSource& src, const uint2* chunks, ChunkTop* out) {
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chunk_top_kernel, dim3(nchunks), dim3(64), 0, st, src, chunks, out);
    return hipGetLastError();
}

hipError_t launch_gather(int nsegs, hipStream_t st, const GatherSeg* segs, PairResult* out) {
    if (nsegs <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(nsegs), dim3(256), 0, st, segs, out);
    return hipGetLastError();
}

}  // namespace ssa
