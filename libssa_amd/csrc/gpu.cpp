// gpu.cpp -- device context: stream, growable arenas, timing.
#include "gpu.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "common.h"

namespace ssa {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fatal("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
}

void* DeviceArena::get(size_t need) {
    if (need == 0) need = 256;
    if (need > bytes) {
        if (ptr) SSA_HIP(hipFree(ptr));
        size_t grow = bytes + bytes / 2;
        bytes = need > grow ? need : grow;
        bytes = (bytes + 4095) & ~(size_t)4095;
        SSA_HIP(hipMalloc(&ptr, bytes));
    }
    return ptr;
}

void* PinnedArena::get(size_t need) {
    if (need == 0) need = 256;
    if (need > bytes) {
        if (ptr) SSA_HIP(hipHostFree(ptr));
        size_t grow = bytes + bytes / 2;
        bytes = need > grow ? need : grow;
        bytes = (bytes + 4095) & ~(size_t)4095;
        SSA_HIP(hipHostMalloc(&ptr, bytes, hipHostMallocDefault));
    }
    return ptr;
}

void PinnedArena::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void DeviceArena::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

Gpu::Gpu() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fatal("libssa_amd: no HIP device available (the search runs on the GPU only)");
    SSA_HIP(hipGetDevice(&device_));
    hipDeviceProp_t prop;
    SSA_HIP(hipGetDeviceProperties(&prop, device_));
    cus_ = prop.multiProcessorCount;
    arch_ = prop.gcnArchName;
    exec_ = new Exec();
    // the no-allocation entry point's context: its own stream, arenas in the
    // device's copy of the scratch array, resolved here once with the device
    // current
    size_t scratch_bytes = 0;
    void* scratch = scratch_setup(&scratch_bytes);
    fixed_ = new Exec();
    fixed_->use_region(scratch, scratch_bytes);
}

void* Exec::arena(const std::string& name, size_t bytes) {
    if (!region_) return arenas_[name].get(bytes);
    auto it = pieces_.find(name);
    if (it != pieces_.end() && it->second.second >= bytes) return region_ + it->second.first;
    const size_t need = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    if (need > region_bytes_ - region_used_)
        throw FixedRegionError{-4, "the search needs more device memory than the no-allocation scratch holds"};
    const size_t off = region_used_;
    region_used_ += need;
    pieces_[name] = std::make_pair(off, need);
    return region_ + off;
}

void* Exec::pinned(const std::string& name, size_t bytes) {
    if (!region_) return pinned_[name].get(bytes);
    HostBlock& b = host_[name];
    if (bytes > b.cap) {
        if (b.p) host_block_release(b.p, b.cap);
        b.p = host_block_acquire(bytes, &b.cap);
    }
    return b.p;
}

void Exec::use_region(void* base, size_t bytes) {
    region_ = (uint8_t*)base;
    region_bytes_ = bytes;
    begin_call();
}

void Exec::begin_call() {
    region_used_ = 0;
    pieces_.clear();
}

Gpu& Gpu::get() {
    static Gpu g;
    return g;
}

Exec::Exec() { SSA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)); }

Exec::~Exec() {
    release_all();
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    for (hipStream_t& a : aux_)
        if (a) (void)hipStreamDestroy(a);
    if (stream_) (void)hipStreamDestroy(stream_);
}

hipStream_t Exec::aux(size_t i) {
    hipStream_t& a = aux_[i % 4];
    if (!a) SSA_HIP(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
    return a;
}

void Exec::upload(void* dst, const void* src, size_t bytes) {
    if (bytes) SSA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
}

void Exec::download(void* dst, const void* src, size_t bytes) {
    if (bytes) SSA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
}

void Exec::sync() { SSA_HIP(hipStreamSynchronize(stream_)); }

void Exec::release_all() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto& kv : arenas_) kv.second.release();
    arenas_.clear();
    for (auto& kv : pinned_) kv.second.release();
    pinned_.clear();
    for (auto& kv : host_) host_block_release(kv.second.p, kv.second.cap);
    host_.clear();
    pieces_.clear();
    region_used_ = 0;
}

GpuTimer::GpuTimer(hipStream_t st) : st_(st) {
    SSA_HIP(hipEventCreate(&a_));
    SSA_HIP(hipEventCreate(&b_));
}

GpuTimer::~GpuTimer() {
    if (a_) (void)hipEventDestroy(a_);
    if (b_) (void)hipEventDestroy(b_);
}

void GpuTimer::start() { SSA_HIP(hipEventRecord(a_, st_)); }
void GpuTimer::stop() { SSA_HIP(hipEventRecord(b_, st_)); }

float GpuTimer::ms() {
    SSA_HIP(hipEventSynchronize(b_));
    float t = 0.f;
    SSA_HIP(hipEventElapsedTime(&t, a_, b_));
    return t;
}

}  // namespace ssa
