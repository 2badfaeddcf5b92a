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
    size_t footprint = 0;
    double cells = 0;  // estimated GPU work
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    bool claimed = false;  // a thread waits for the result
};

}  // namespace

class Server {
public:
    Server(int slots, size_t max_batch, size_t device_bytes) : max_batch_(std::max<size_t>(1, max_batch)) {
        size_t free_b = 0, total_b = 0;
        SSA_HIP(hipMemGetInfo(&free_b, &total_b));
        budget_ = device_bytes ? device_bytes : free_b / 2;
        slot_budget_ = budget_ / (size_t)slots;
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

    int64_t submit(const SearchRequest& rq, size_t footprint, double cells) {
        std::lock_guard<std::mutex> lk(mu_);
        const int64_t id = next_id_++;
        Ticket& t = tickets_[id];
        t.rq = rq;
        t.footprint = footprint;
        t.cells = cells;
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
    }

private:
    void work() {
        Exec ex;
        for (;;) {
            std::vector<int64_t> batch;
            {
                std::unique_lock<std::mutex> lk(mu_);
                for (;;) {
                    cv_work_.wait(lk, [&] { return stopping_ || !queue_.empty(); });
                    if (queue_.empty()) return;  // stopping with nothing left
                    if (busy_ == 0 || stopping_) break;
                    // another worker is on the GPU: join it once the oldest
                    // request has waited past the assist limit (a worker
                    // that finishes wakes the others)
                    const double waited = now_ms() - tickets_[queue_.front()].t_submit;
                    if (waited >= assist_ms_) break;
                    cv_work_.wait_for(lk, std::chrono::duration<double, std::milli>(assist_ms_ - waited));
                }
                // the oldest request first once it has waited past the aging
                // limit, then the cheapest ones
                size_t bytes = 0;
                double cells = 0;
                const double now = now_ms();
                std::vector<int64_t> order(queue_.begin(), queue_.end());
                const Ticket& oldest = tickets_[order.front()];
                const bool aged = now - oldest.t_submit > aging_ms_;
                std::stable_sort(order.begin() + (aged ? 1 : 0), order.end(),
                                 [&](int64_t a, int64_t b) { return tickets_[a].cells < tickets_[b].cells; });
                for (int64_t id : order) {
                    if (batch.size() >= max_batch_) break;
                    const Ticket& t = tickets_[id];
                    // a request larger than the budgets runs alone
                    if (!batch.empty() && (bytes + t.footprint > slot_budget_ || cells + t.cells > batch_cells_)) break;
                    bytes += t.footprint;
                    cells += t.cells;
                    batch.push_back(id);
                }
                for (int64_t id : batch) queue_.erase(std::find(queue_.begin(), queue_.end(), id));
                const double t0 = now_ms();
                for (int64_t id : batch) tickets_[id].t_start = t0;
                ++batches_;
                ++busy_;
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
    size_t budget_ = 0, slot_budget_ = 0;
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

// search entries per DB sequence, and their residues per DB residue
// (src/db_adapter.c:124-138): 2 and 2 for nucleotides of both strands, 6 or 3
// frames of a third of the length for translated DBs, else 1 and 1
std::pair<double, double> entry_factors() {
    const Config& cf = config();
    if (cf.symtype == NUCLEOTIDE && (cf.strands & 2)) return {2.0, 2.0};
    if (db_is_translated(cf.symtype)) return cf.strands == BOTH_STRANDS ? std::make_pair(6.0, 2.0) : std::make_pair(3.0, 1.0);
    return {1.0, 1.0};
}

// GPU work of one request: cells of its first tier, twice for NW starting at
// 8 bit (its 16-bit pass re-runs nearly every entry); residues = residues of
// the DB's search entries
double request_cells(const SearchRequest& rq, double residues) {
    double qlen = 0;
    for (size_t l : query_lengths(rq.query)) qlen += (double)l;
    return qlen * residues * ((rq.type == kNW && rq.bit_width == BIT_WIDTH_8) ? 2.0 : 1.0);
}

// residues of the DB's search entries, and the longest DB sequence
struct DbSize {
    double residues = 0;
    size_t max_len = 0;
};

DbSize db_size() {
    DbSize d;
    const size_t n = ssa_db_get_sequence_count();
    for (size_t i = 0; i < n; ++i) {
        p_seqinfo s = ssa_db_get_sequence(i);
        if (!s) break;
        d.residues += (double)s->seqlen;
        d.max_len = std::max<size_t>(d.max_len, s->seqlen);
    }
    d.residues *= entry_factors().second;
    return d;
}

// device bytes one request needs in a batch: pair results and pass boundaries
// of its first tier, and the direction matrices of its hits' traceback (hits
// against the DB's longest sequence, at most one traceback launch's budget:
// the hits of a batch run in launches that fit it)
size_t request_footprint(const SearchRequest& rq, size_t db_max_len) {
    const size_t n = ssa_db_get_sequence_count();
    const size_t entries = (size_t)((double)n * entry_factors().first);
    const std::vector<size_t> ql = query_lengths(rq.query);
    const size_t nq = ql.size();
    size_t qlen = 0;
    for (size_t l : ql) qlen = std::max(qlen, l);
    const size_t mean_len = 400;  // a planning figure, not a limit
    size_t bytes = nq * entries * sizeof(PairResult) + nq * entries * (mean_len / 4) * 16;
    if (rq.align_type != COMPUTE_SCORE)
        bytes += std::min<size_t>(rq.hitcount * ((qlen + 7) & ~(size_t)7) * std::max<size_t>(db_max_len, 1),
                                  (size_t)1 << 30);
    return bytes;
}

}  // namespace ssa

using namespace ssa;

extern "C" {

struct ssa_amd_server {
    Server* s;
    DbSize db;  // of the DB at start
};

ssa_amd_server* ssa_amd_server_start(int slots, size_t max_batch, size_t device_bytes) {
    if (slots < 1 || slots > 8) return nullptr;
    if (!matrix().valid) return nullptr;
    const Config& c = config();
    if (!matrix().constant && c.symtype == NUCLEOTIDE) return nullptr;
    ssa_amd_server* h = new ssa_amd_server;
    h->db = db_size();
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
    return h->s->submit(rq, request_footprint(rq, h->db.max_len), request_cells(rq, h->db.residues));
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

void ssa_amd_server_stop(ssa_amd_server* h) {
    if (!h) return;
    delete h->s;
    delete h;
}

}  // extern "C"
