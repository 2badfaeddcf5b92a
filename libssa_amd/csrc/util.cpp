// util.cpp -- messages, fatal errors, process-wide configuration and the
// host worker pool.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "common.h"

namespace ssa {

namespace {

// One parallel_for call: items are claimed by index; the caller waits for done.
struct WorkBatch {
    const std::function<void(size_t)>* fn;
    size_t n;
    std::atomic<size_t> next{0};
    std::atomic<size_t> done{0};
    std::mutex mu;
    std::condition_variable cv;
};

class WorkerPool {
public:
    static WorkerPool& get() {
        static WorkerPool* p = new WorkerPool();  // never destroyed: workers may outlive static teardown
        return *p;
    }
    size_t width() const { return threads_.size() + 1; }

    void run(size_t n, const std::function<void(size_t)>& fn) {
        if (n == 0) return;
        if (n == 1 || threads_.empty()) {
            for (size_t i = 0; i < n; ++i) fn(i);
            return;
        }
        auto b = std::make_shared<WorkBatch>();
        b->fn = &fn;
        b->n = n;
        {
            std::lock_guard<std::mutex> lk(mu_);
            queue_.push_back(b);
        }
        cv_.notify_all();
        work(*b);  // the caller takes items too
        std::unique_lock<std::mutex> lk(b->mu);
        b->cv.wait(lk, [&] { return b->done.load() == b->n; });
    }

private:
    WorkerPool() {
        const size_t hw = std::thread::hardware_concurrency();
        const size_t n = std::min<size_t>(16, hw ? hw : 1);
        for (size_t i = 0; i + 1 < n; ++i) threads_.emplace_back([this] { loop(); });
        for (std::thread& t : threads_) t.detach();
    }

    // claims and runs items of b until none are left
    static void work(WorkBatch& b) {
        for (;;) {
            const size_t i = b.next.fetch_add(1);
            if (i >= b.n) return;
            (*b.fn)(i);
            if (b.done.fetch_add(1) + 1 == b.n) {
                std::lock_guard<std::mutex> lk(b.mu);
                b.cv.notify_all();
            }
        }
    }

    void loop() {
        for (;;) {
            std::shared_ptr<WorkBatch> b;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return !queue_.empty(); });
                b = queue_.front();
                // a batch whose items are all claimed leaves the queue
                if (b->next.load() >= b->n) {
                    queue_.pop_front();
                    continue;
                }
            }
            work(*b);
        }
    }

    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<WorkBatch>> queue_;
};

}  // namespace

void parallel_for(size_t n, const std::function<void(size_t)>& fn) { WorkerPool::get().run(n, fn); }

size_t parallel_width() { return WorkerPool::get().width(); }

namespace {

struct BlockCache {
    static constexpr size_t kMaxCached = size_t(4) << 30;
    std::mutex mu;
    std::multimap<size_t, void*> free;  // capacity -> block
    size_t cached = 0;
    static BlockCache& get() {
        static BlockCache* c = new BlockCache();  // never destroyed: blocks may be released during teardown
        return *c;
    }
};

}  // namespace

void* host_block_acquire(size_t bytes, size_t* capacity) {
    bytes = std::max<size_t>(bytes, 64);
    BlockCache& bc = BlockCache::get();
    {
        std::lock_guard<std::mutex> lk(bc.mu);
        // the smallest cached block that fits, if it is at most twice the request
        auto it = bc.free.lower_bound(bytes);
        if (it != bc.free.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            *capacity = it->first;
            bc.cached -= it->first;
            bc.free.erase(it);
            return p;
        }
    }
    void* p = malloc(bytes);
    if (!p) fatal("Unable to allocate enough memory: %zu.", bytes);
    *capacity = bytes;
    return p;
}

void host_block_release(void* p, size_t capacity) {
    if (!p) return;
    BlockCache& bc = BlockCache::get();
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> lk(bc.mu);
        bc.free.emplace(capacity, p);
        bc.cached += capacity;
        // over the limit: give back the largest blocks
        while (bc.cached > BlockCache::kMaxCached && !bc.free.empty()) {
            auto it = std::prev(bc.free.end());
            bc.cached -= it->first;
            drop.push_back(it->second);
            bc.free.erase(it);
        }
    }
    for (void* d : drop) ::free(d);
}

Config& config() {
    static Config c;
    return c;
}

// src/util/util.c:36-45: message to stderr, then exit(1)
void fatal(const char* fmt, ...) {
    if (fmt) {
        va_list ap;
        va_start(ap, fmt);
        vfprintf(stderr, fmt, ap);
        va_end(ap);
        fprintf(stderr, "\n");
    }
    exit(1);
}

static bool shown(int level) { return level <= config().output_mode; }

// src/util/util.c:54-64 (no newline added, as in the reference)
void print_info(const char* fmt, ...) {
    if (!shown(OUTPUT_INFO)) return;
    printf("libssa INFO: ");
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
}

// src/util/util.c:66-77
void print_warning(const char* fmt, ...) {
    if (!shown(OUTPUT_WARNING)) return;
    printf("libssa WARNING: ");
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    printf("\n");
}

// src/util/util.c:79-90
void print_error(const char* fmt, ...) {
    if (!shown(OUTPUT_ERROR)) return;
    printf("libssa ERROR: ");
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    printf("\n");
}

// The reference's 8-bit kernels exist for AVX2 (32 lanes) and SSE4.1 (16
// lanes), its 16-bit kernels for AVX2 (16) and SSE2 (8); the selection follows
// set_max_compute_capability (src/cpu_config.c:29-33, src/algo/8/search_8.c:36-60).
int lanes8() { return config().simd_mode >= COMPUTE_ON_AVX2 ? 32 : 16; }
int lanes16() { return config().simd_mode >= COMPUTE_ON_AVX2 ? 16 : 8; }

}  // namespace ssa
