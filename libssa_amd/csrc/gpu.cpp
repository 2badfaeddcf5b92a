// gpu.cpp -- device context: stream, growable arenas, timing.
#include "gpu.h"

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
}

Gpu& Gpu::get() {
    static Gpu g;
    return g;
}

Exec::Exec() { SSA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)); }

Exec::~Exec() {
    release_all();
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
