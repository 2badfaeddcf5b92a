// server.cpp -- continuous-batching request loop over the resident DB.
//
// Requests (sw_align / nw_align arguments) are queued as they arrive.  Each of
// `slots` workers owns a device context (stream and buffers); a free worker
// takes waiting requests as one batch (run_batch: one captured graph for the
// batch's first tiers): first the oldest request if it has waited longer than
// the aging limit, then the cheapest ones (estimated cells, shortest first)
// while the batch stays within the cell budget, the batch size and the
// worker's share of the device memory budget.  Batches of different workers
// run on the GPU at the same time, so requests that arrive while others are
// on the GPU start as soon as a worker and memory free up.  A worker joins
// another one on the GPU only once the oldest waiting request has waited past
// the assist limit: below saturation a request that arrives during a short
// search then starts alone a little later instead of sharing the device (and
// its time) with it, and under load every worker stays busy.  Every result is
// what the search returns alone (run_batch's contract).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "gpu.h"
#include "search.h"

extern "C" {
#include "libssa_amd.h"
}

namespace ssa {
namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Ticket {
    SearchRequest rq;
    SearchResult res;
    RequestPlan plan;  // cells, device bytes, occupancy of its first tier
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    bool claimed = false;  // a thread waits for the result
};

}  // namespace

// The admission policy (ssa_amd_admission_plan): which waiting requests a free
// worker takes as its next batch.  Candidates in order: the oldest request
// first once it has waited past the aging limit, then the cheapest (cells),
// shortest first.  A candidate joins while the batch stays within the batch
// size and cell budget and its device bytes fit what the budget leaves beside
// the batches on the GPU (a request larger than the whole budget runs alone
// on an idle device).  Beside running batches, a batch starts only if its
// workgroups fit into the resident slots theirs leave free (occupancy, set
// by the kernel shapes' LDS and VGPR use), or once the oldest request has
// waited past the assist limit; otherwise the worker waits.
std::vector<size_t> plan_batch(const ssa_amd_admission_request_t* q, size_t n, const ssa_amd_admission_state_t& st) {
    std::vector<size_t> out;
    if (n == 0) return out;
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    const bool aged = q[0].wait_ms > st.aging_ms;
    std::stable_sort(order.begin() + (aged ? 1 : 0), order.end(),
                     [&](size_t a, size_t b) { return q[a].cells < q[b].cells; });
    const size_t free_bytes = st.device_budget > st.device_in_use ? st.device_budget - st.device_in_use : 0;
    size_t bytes = 0;
    double cells = 0, occ = 0;
    for (size_t i : order) {
        if (out.size() >= std::max<size_t>(st.max_batch, 1)) break;
        const ssa_amd_admission_request_t& r = q[i];
        if (!out.empty() && cells + r.cells > st.batch_cells) break;
        if (bytes + r.device_bytes > free_bytes) {
            if (out.empty() && st.running == 0) out.push_back(i);  // larger than the budget: alone, on an idle device
            break;
        }
        bytes += r.device_bytes;
        cells += r.cells;
        occ += r.occupancy;
        out.push_back(i);
    }
    if (out.empty()) return out;
    if (st.running > 0 && st.occupancy_in_use + std::min(occ, 1.0) > 1.0 + 1e-9 && q[0].wait_ms < st.assist_ms)
        out.clear();  // its workgroups would queue behind the running ones': wait
    return out;
}

class Server {
public:
    Server(int slots, size_t max_batch, size_t device_bytes) : max_batch_(std::max<size_t>(1, max_batch)) {
        size_t free_b = 0, total_b = 0;
        SSA_HIP(hipMemGetInfo(&free_b, &total_b));
        budget_ = device_bytes ? device_bytes : free_b / 2;
        for (int i = 0; i < slots; ++i) workers_.emplace_back([this] { work(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stopping_ = true;
        }
        cv_work_.notify_all();
        for (auto& t : workers_) t.join();
        for (auto& kv : tickets_)
            if (kv.second.res.list) free_alignment(kv.second.res.list);
    }

    void tune(double batch_cells, double aging_ms) {
        std::lock_guard<std::mutex> lk(mu_);
        if (batch_cells > 0) batch_cells_ = batch_cells;
        if (aging_ms > 0) aging_ms_ = aging_ms;
    }

    int64_t submit(const SearchRequest& rq, const RequestPlan& plan) {
        std::lock_guard<std::mutex> lk(mu_);
        const int64_t id = next_id_++;
        Ticket& t = tickets_[id];
        t.rq = rq;
        t.plan = plan;
        t.t_submit = now_ms();
        queue_.push_back(id);
        cv_work_.notify_one();
        return id;
    }

    // One waiter takes a ticket's result: the first to wait claims it, and
    // any other waiter on the same ticket returns false at once.  The ticket
    // is looked up again after every wake-up (the map may rehash meanwhile).
    bool wait(int64_t id, SearchResult* out, double* latency_ms, double* queue_ms) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = tickets_.find(id);
        if (it == tickets_.end() || it->second.claimed) return false;
        it->second.claimed = true;
        cv_done_.wait(lk, [&] { return tickets_.at(id).done; });
        it = tickets_.find(id);
        *out = it->second.res;
        if (latency_ms) *latency_ms = it->second.t_done - it->second.t_submit;
        if (queue_ms) *queue_ms = it->second.t_start - it->second.t_submit;
        tickets_.erase(it);
        return true;
    }

    void counters(ssa_amd_server_counters_t* c) {
        std::lock_guard<std::mutex> lk(mu_);
        c->batches = batches_;
        c->requests = requests_;
        c->max_batch_seen = max_batch_seen_;
        c->queued = queue_.size();
        c->device_budget = budget_;
        c->deferrals = deferrals_;
        c->device_in_use = in_use_;
    }

private:
    void work() {
        Exec ex;
        for (;;) {
            std::vector<int64_t> batch;
            size_t batch_bytes = 0;
            double batch_occ = 0;
            {
                std::unique_lock<std::mutex> lk(mu_);
                for (;;) {
                    cv_work_.wait(lk, [&] { return stopping_ || !queue_.empty(); });
                    if (queue_.empty()) return;  // stopping with nothing left
                    const double now = now_ms();
                    std::vector<ssa_amd_admission_request_t> q;
                    q.reserve(queue_.size());
                    for (int64_t id : queue_) {
                        const Ticket& t = tickets_[id];
                        q.push_back(ssa_amd_admission_request_t{now - t.t_submit, t.plan.cells, t.plan.hbm_bytes,
                                                                t.plan.occupancy});
                    }
                    ssa_amd_admission_state_t st{max_batch_, batch_cells_, aging_ms_, assist_ms_, budget_,
                                                 in_use_,    occ_in_use_, busy_};
                    const std::vector<size_t> pick = plan_batch(q.data(), q.size(), st);
                    if (!pick.empty()) {
                        for (size_t i : pick) batch.push_back(queue_[i]);
                        break;
                    }
                    // deferred: wait for a running batch to finish, or until the
                    // oldest request reaches the assist limit
                    ++deferrals_;
                    const double left = std::max(1.0, assist_ms_ - q[0].wait_ms);
                    cv_work_.wait_for(lk, std::chrono::duration<double, std::milli>(left));
                }
                for (int64_t id : batch) {
                    queue_.erase(std::find(queue_.begin(), queue_.end(), id));
                    batch_bytes += tickets_[id].plan.hbm_bytes;
                    batch_occ += tickets_[id].plan.occupancy;
                }
                batch_occ = std::min(batch_occ, 1.0);
                const double t0 = now_ms();
                for (int64_t id : batch) tickets_[id].t_start = t0;
                ++batches_;
                ++busy_;
                in_use_ += batch_bytes;
                occ_in_use_ += batch_occ;
                requests_ += batch.size();
                max_batch_seen_ = std::max<size_t>(max_batch_seen_, batch.size());
                if (!queue_.empty()) cv_work_.notify_one();
            }
            std::vector<SearchRequest> rq;
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (int64_t id : batch) rq.push_back(tickets_[id].rq);
            }
            std::vector<SearchResult> res;
            run_batch(ex, rq, res, true);
            {
                std::lock_guard<std::mutex> lk(mu_);
                const double t1 = now_ms();
                for (size_t i = 0; i < batch.size(); ++i) {
                    Ticket& t = tickets_[batch[i]];
                    t.res = res[i];
                    t.t_done = t1;
                    t.done = true;
                }
                --busy_;
                in_use_ -= batch_bytes;
                occ_in_use_ = std::max(0.0, occ_in_use_ - batch_occ);
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
        }
    }

    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<int64_t> queue_;
    std::unordered_map<int64_t, Ticket> tickets_;
    std::vector<std::thread> workers_;
    size_t max_batch_;
    size_t budget_ = 0;       // device bytes the batches may hold together
    size_t in_use_ = 0;       // device bytes of the batches on the GPU
    double occ_in_use_ = 0;   // their share of the resident workgroup slots
    size_t deferrals_ = 0;    // times a free worker waited instead of starting a batch
    double batch_cells_ = 2.5e11;  // GPU work per batch (about one average request, ~45 ms)
    double aging_ms_ = 750.0;    // a request waiting longer goes first
    double assist_ms_ = 50.0;    // a request waiting longer brings another worker onto the GPU
    int busy_ = 0;               // workers with a batch on the GPU
    int64_t next_id_ = 0;
    bool stopping_ = false;
    size_t batches_ = 0, requests_ = 0, max_batch_seen_ = 0;
};

// lengths of the query buffers a search makes of the query
// (src/algo/searcher.c:42-90): strands, six or three frames, or one
std::vector<size_t> query_lengths(p_query q) {
    const Config& cf = config();
    std::vector<size_t> v;
    if (cf.symtype == NUCLEOTIDE) {
        for (int s = 0; s < 2; ++s)
            if (((s + 1) & cf.strands) && q->has_nt[s]) v.push_back(q->nt[s].size());
    } else if (query_is_translated(cf.symtype)) {
        for (int s = 0; s < 2; ++s)
            if ((s + 1) & cf.strands)
                for (int f = 0; f < 3; ++f) v.push_back(q->aa[3 * s + f].size());
    } else {
        v.push_back(q->aa[0].size());
    }
    return v;
}

}  // namespace ssa

using namespace ssa;

extern "C" {

struct ssa_amd_server {
    Server* s;
};

ssa_amd_server* ssa_amd_server_start(int slots, size_t max_batch, size_t device_bytes) {
    if (slots < 1 || slots > 8) return nullptr;
    if (!matrix().valid) return nullptr;
    const Config& c = config();
    if (!matrix().constant && c.symtype == NUCLEOTIDE) return nullptr;
    ensure_device_db();  // resident before any request is planned against it
    ssa_amd_server* h = new ssa_amd_server;
    h->s = new Server(slots, max_batch, device_bytes);
    return h;
}

int64_t ssa_amd_server_submit(ssa_amd_server* h, const ssa_amd_request_t* r) {
    if (!h || !r) return -1;
    if (r->search_type != 0 && r->search_type != 1) return -1;
    if (!r->query) return -1;
    if (r->bit_width != BIT_WIDTH_8 && r->bit_width != BIT_WIDTH_16 && r->bit_width != BIT_WIDTH_64) return -1;
    if (r->align_type != COMPUTE_SCORE && r->align_type != COMPUTE_ALIGNMENT) return -1;
    if (r->bit_width == BIT_WIDTH_8 && config().simd_mode < COMPUTE_ON_SSE41) return -1;
    const SearchRequest rq{r->search_type == 0 ? kSW : kNW, r->query, r->hitcount, r->bit_width, r->align_type};
    return h->s->submit(rq, plan_request(rq, query_lengths(rq.query)));
}

p_alignment_list ssa_amd_server_wait(ssa_amd_server* h, int64_t ticket, ssa_amd_stats_t* stats, double* latency_ms,
                                     double* queue_ms) {
    if (!h) return nullptr;
    SearchResult res;
    if (!h->s->wait(ticket, &res, latency_ms, queue_ms)) return nullptr;
    if (stats) *stats = res.stats;
    return res.list;
}

void ssa_amd_server_tune(ssa_amd_server* h, double batch_cells, double aging_ms) {
    if (h) h->s->tune(batch_cells, aging_ms);
}

void ssa_amd_server_counters(ssa_amd_server* h, ssa_amd_server_counters_t* out) {
    if (h && out) h->s->counters(out);
}

size_t ssa_amd_admission_plan(const ssa_amd_admission_request_t* queue, size_t n,
                              const ssa_amd_admission_state_t* state, size_t* chosen) {
    if (!queue || !state || !chosen || n == 0) return 0;
    const std::vector<size_t> pick = plan_batch(queue, n, *state);
    std::copy(pick.begin(), pick.end(), chosen);
    return pick.size();
}

void ssa_amd_server_stop(ssa_amd_server* h) {
    if (!h) return;
    delete h->s;
    delete h;
}

}  // extern "C"
