// gpu.h -- device context of the host runtime: one HIP stream, growable
// device arenas, pinned staging, and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"

namespace ssa {

void hip_check(hipError_t e, const char* what);
#define SSA_HIP(x) ::ssa::hip_check((x), #x)

// A pinned host allocation that only grows; contents are not preserved on growth.
struct PinnedArena {
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t need);
    void release();
};

// A device allocation that only grows; contents are not preserved on growth.
struct DeviceArena {
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t need);
    void release();
};

// Thrown by a fixed-region context (Exec::use_region) when a call needs more
// device memory than the region holds, or device state it may not build.
struct FixedRegionError {
    int code;            // -3: device state missing, -4: region too small
    const char* what;
};

// A device work context: one stream and the growable buffers of the work
// issued on it.  Work on different contexts may run concurrently; one context
// is used by one host thread at a time.
//
// A context given a fixed region (use_region) allocates nothing: arena()
// hands out consecutive 256-byte aligned pieces of the region, the same piece
// for a name asked again within one call, until begin_call() starts over;
// pinned() hands out pageable host blocks.  Code that would otherwise build
// device state asks fixed() and takes another route or throws.
class Exec {
public:
    Exec();
    ~Exec();
    Exec(const Exec&) = delete;
    Exec& operator=(const Exec&) = delete;
    hipStream_t stream() const { return stream_; }
    hipStream_t aux(size_t i);  // extra streams for graph branches (4, created on first use)
    void* arena(const std::string& name, size_t bytes);
    void* pinned(const std::string& name, size_t bytes);
    void upload(void* dst, const void* src, size_t bytes);
    void download(void* dst, const void* src, size_t bytes);
    void sync();
    void release_all();

    // the instantiated graph of this context's last batch (reused for the
    // next batch through hipGraphExecUpdate when its topology is the same)
    hipGraphExec_t graph_exec() const { return graph_exec_; }
    void set_graph_exec(hipGraphExec_t g) { graph_exec_ = g; }

    void use_region(void* base, size_t bytes);  // from now on, arena() carves this region
    bool fixed() const { return region_ != nullptr; }
    void begin_call();                          // fixed: the region is free again
    // fixed: every piece is free again (the caller has synchronised the
    // stream and keeps no device pointer of the pieces); otherwise nothing
    void recycle() {
        if (region_) begin_call();
    }
    size_t region_left() const { return region_bytes_ - region_used_; }

private:
    hipStream_t stream_ = nullptr;
    hipStream_t aux_[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t graph_exec_ = nullptr;
    std::unordered_map<std::string, DeviceArena> arenas_;
    std::unordered_map<std::string, PinnedArena> pinned_;
    // fixed region: base, size, bytes handed out this call, pieces by name
    uint8_t* region_ = nullptr;
    size_t region_bytes_ = 0, region_used_ = 0;
    std::unordered_map<std::string, std::pair<size_t, size_t>> pieces_;  // name -> (offset, bytes)
    struct HostBlock {
        void* p = nullptr;
        size_t cap = 0;
    };
    std::unordered_map<std::string, HostBlock> host_;
};

class Gpu {
public:
    static Gpu& get();  // initialises the device on first use
    Exec& exec() { return *exec_; }  // the context of the libssa.h entry points
    // the context of the no-allocation entry point: its arenas are the
    // device's copy of the scratch array (score_kernels.hip)
    Exec& fixed_exec() { return *fixed_; }
    int device() const { return device_; }
    std::string arch() const { return arch_; }
    int cu_count() const { return cus_; }

private:
    Gpu();
    int device_ = 0;
    int cus_ = 0;
    std::string arch_;
    Exec* exec_ = nullptr;
    Exec* fixed_ = nullptr;
};

// event pair timing a span of work on one stream
class GpuTimer {
public:
    explicit GpuTimer(hipStream_t st);
    ~GpuTimer();
    void start();
    void stop();
    float ms();  // synchronises on the stop event

private:
    hipStream_t st_ = nullptr;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

// ---- kernel launchers (score_kernels.hip, traceback_kernels.hip, pack_kernels.hip) ----
// modes: 0 SW (int16), 1 NW, 2 NW without overflow trackers, 3 SW in fp16
// windows, 4 NW without overflow trackers in fp16
bool score_shape_supported(int lanes, int rows, int mode, int cols);
// the query profile in VGPRs (nucleotide DBs of A, C, G and T only)
bool score_vp_supported(int lanes, int rows, int mode, int cols);
int score_resident_workgroups(int lanes, int rows, int mode, int cols, bool vp);  // persistent grid size of a shape
// NW shapes with an untracked variant (mode 2: no value of any pair can reach
// either end of the int16 range)
bool score_nw_untracked_supported(int lanes, int rows);
// paired packed SW queries (score_kernel_pair): shapes, grid size, launch
bool score_pair_supported(int lanes, int rows);
int score_pair_resident_workgroups(int lanes, int rows, bool f16);
hipError_t launch_score_pair(int lanes, int rows, bool f16, hipStream_t st, const ScoreLaunch& L, const ScoreParams& P,
                             int max_wg);
hipError_t launch_score(int lanes, int rows, int mode, int cols, bool vp, hipStream_t st, const ScoreLaunch& L,
                        const ScoreParams& P, int max_wg);
hipError_t launch_init_results(hipStream_t st, PairResult* r, int64_t n, int32_t score);
// exact tier: one wave per pair, int64 (wide) or int32; bnd holds each pair's
// pass boundaries (elements of that width, at ExactPair::bnd_off)
hipError_t launch_exact(bool wide, int npairs, hipStream_t st, const ExactPair* pairs, const QueryDesc* queries,
                        const uint8_t* qcodes, const int64_t* mat, const uint8_t* codes, int mode, int go, int ge,
                        void* bnd, int64_t* out);
// workgroup traceback: a hit's passes of 64 * rows_per_lane query rows (1, 2,
// 4 or 8) spread over `waves` waves (1-16); at most traceback_wg_max_passes()
// passes per dynamic programme
hipError_t launch_traceback_wg(int rows_per_lane, int waves, int ntasks, hipStream_t st, const TraceTask* tasks,
                               const uint8_t* codes, const int64_t* mat, int go, int ge, uint8_t* dirs,
                               int64_t* scratch, int32_t* ops, TraceResult* out);
int traceback_wg_max_passes();
hipError_t launch_traceback(int ntasks, hipStream_t st, const TraceTask* tasks, const uint8_t* codes,
                            const int64_t* mat, int go, int ge, uint8_t* dirs, int64_t* scratch, int32_t* ops,
                            TraceResult* out);

hipError_t launch_pack(int nentries, hipStream_t st, const PackEntry* entries, const uint8_t* codes,
                       uint32_t* res, int64_t res_words);

// select_kernels.hip: per-chunk top inserts (chunks: entry ranges [x, y)),
// and the copy of pair-result ranges into a compact buffer
hipError_t launch_chunk_top(int nchunks, hipStream_t st, const TopSource& src, const uint2* chunks, ChunkTop* out);
hipError_t launch_gather(int nsegs, hipStream_t st, const GatherSeg* segs, PairResult* out);

// The device scratch array of the no-allocation entry point (a file-scope
// __device__ array in score_kernels.hip, one copy per device): resolves the
// current device's copy once (hipGetSymbolAddress) and returns its address
// and size.  Host code uses only this address, never the array's name.
void* scratch_setup(size_t* bytes);

}  // namespace ssa
