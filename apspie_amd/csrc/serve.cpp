// Continuous-batching query server over HBM-resident graphs.
//
// No reference counterpart (the reference answers one query per process).
// Graphs are admitted into a resident pool against an HBM budget.  Queries
// (BFS, SSSP, TC) enter one FIFO.  `slots` worker threads each own a context
// (its own HIP stream and scratch); a free worker takes the oldest query and
// coalesces every queued query of the same graph and kind into one launch:
// up to max_batch sources go through the batched multi-source traversal, and
// all queued TC queries on a graph share one count.  Queries that arrive while
// other batches run are picked up as soon as a worker frees, so their work
// overlaps on the device with the batches in flight.  Each worker's scratch
// budget is what the resident graphs leave of the server's budget, split
// evenly; a batch is sized to fit it (its result rows are held about three
// times over), and one that still does not fit is split in halves; a single
// query that does not fit fails with APSPIE_ERR_BUDGET.  Results land in host
// memory; latency is measured from submit to completion on the host clock.
// Result buffers are pinned host memory from a pool owned by the server: one
// direct D2H copy per query (a batch writes its rows to device memory and each
// row goes straight to its query's buffer), and the buffer returns to the
// pool when the caller collects the result.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace {

using Clock = std::chrono::steady_clock;

// Pinned result buffers, reused by size (a buffer serves any request up to
// its capacity and at most twice smaller, so big buffers are not tied up by
// small results).
class PinnedPool {
  public:
    ~PinnedPool() {
        for (auto &kv : cap_) (void)hipHostFree(kv.first);
    }
    void *get(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= 2 * bytes) {
                void *p = it->second;
                free_.erase(it);
                return p;
            }
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        std::lock_guard<std::mutex> lk(mu_);
        cap_[p] = bytes;
        return p;
    }
    void put(void *p) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(cap_[p], p);
    }

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> cap_;
};

struct Request {
    int64_t ticket = 0;
    int32_t graph = 0;
    int kind = 0;
    int32_t source = 0;
    int64_t bytes = 0;  // result size, fixed at submit
    const apspie_graph *g = nullptr;  // resident graph and its vertex count, fixed at submit
    int32_t m = 0;
    Clock::time_point t_submit, t_done;
    void *buf = nullptr;    // pinned result (BFS depths / SSSP distances)
    void *out = nullptr;    // caller-owned result memory (submit_into), else null
    int out_mem = APSPIE_MEM_HOST;
    uint64_t triangles = 0;  // TC result
    int status = APSPIE_OK;
    bool done = false;
};

// Work done per (graph, kind): batches launched, queries answered and the
// wall time workers spent on them (launch to results in host memory).
struct Busy {
    int64_t batches = 0, queries = 0, retries = 0;
    double busy_ms = 0, traverse_ms = 0, copy_ms = 0, alloc_ms = 0, retry_ms = 0;
};

// wall time of the phases of one batch
struct Phases {
    double traverse_ms = 0, copy_ms = 0, alloc_ms = 0;
};

double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct Resident {
    apspie_graph *g = nullptr;
    int32_t m = 0;
    int64_t bytes = 0;
    Busy busy[3];  // by query kind
};

}  // namespace

struct apspie_server {
    int device = 0;
    apspie_server_config cfg{};
    apspie_ctx *admin = nullptr;
    // a deque: admitting a graph never moves the residents that queued
    // requests and running batches refer to
    std::deque<Resident> graphs;
    int64_t graph_bytes = 0;
    std::mutex admin_mu;  // admissions share the admin context, one at a time
    // each worker's scratch budget: what the resident graphs leave of the
    // server's budget, split evenly (0 = unlimited); a worker applies it to
    // its own context before each batch
    int64_t slot_budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Request>> live;
    // results handed out by apspie_server_wait_view, until released
    std::unordered_map<int64_t, std::shared_ptr<Request>> viewed;
    std::vector<std::thread> workers;
    std::vector<apspie_ctx *> ctxs;
    PinnedPool pinned;
    int64_t next_ticket = 1;
    int busy = 0;
    bool stop = false;
};

namespace {

int64_t result_bytes(int kind, int32_t m) {
    return kind == APSPIE_Q_TC ? (int64_t)sizeof(uint64_t) : 4 * (int64_t)m;
}

// The batch's graph and size come from its requests (fixed at submit under
// the server lock), so a batch reads nothing of the resident pool while
// graphs are being admitted.
int run_batch(apspie_server *s, apspie_ctx *ctx, std::vector<std::shared_ptr<Request>> &batch,
              Phases &ph) {
    struct {
        const apspie_graph *g;
        int32_t m;
    } R{batch[0]->g, batch[0]->m};
    const int kind = batch[0]->kind;
    const size_t n = batch.size();
    const size_t row = 4 * (size_t)R.m;
    auto t0 = Clock::now();
    if (kind == APSPIE_Q_TC) {
        uint64_t total = 0;
        const int st = apspie_tc(ctx, R.g, &total, nullptr, APSPIE_MEM_HOST);
        for (auto &r : batch) r->triangles = total;
        ph.traverse_ms += ms_since(t0);
        return st;
    }
    if (n == 1) {
        auto &r = batch[0];
        void *dst = r->out;
        int mem = r->out_mem;
        if (!dst) {
            if (!r->buf) r->buf = s->pinned.get(row);
            ph.alloc_ms += ms_since(t0);
            if (!r->buf) return APSPIE_ERR_OOM;
            dst = r->buf;
            mem = APSPIE_MEM_HOST;
        }
        t0 = Clock::now();
        const int st = kind == APSPIE_Q_BFS
                           ? apspie_bfs(ctx, R.g, r->source, (int32_t *)dst, mem)
                           : apspie_sssp(ctx, R.g, r->source, (float *)dst, mem);
        ph.traverse_ms += ms_since(t0);  // a host result's copy included
        return st;
    }
    std::vector<int32_t> src(n);
    for (size_t i = 0; i < n; ++i) src[i] = batch[i]->source;
    // the batch's rows land in a device buffer the worker's context keeps
    // between batches: a hipMalloc / hipFree per batch would wait for the
    // whole device and stall the other workers' streams
    void *dout = nullptr;
    int st = apspie::guard(ctx, [&] {
        apspie::set_device(ctx);
        dout = apspie::scratch(ctx, "serve_rows", row * n);
    });
    if (st == APSPIE_OK)
        st = kind == APSPIE_Q_BFS
                 ? apspie_bfs_batch(ctx, R.g, src.data(), (int32_t)n, (int32_t *)dout,
                                    APSPIE_MEM_DEVICE)
                 : apspie_sssp_batch(ctx, R.g, src.data(), (int32_t)n, (float *)dout,
                                     APSPIE_MEM_DEVICE);
    ph.traverse_ms += ms_since(t0);
    // each row straight to its query's pinned buffer: the copies queue on the
    // worker's stream and one wait covers them all
    // (or to the caller's own memory, device or host)
    t0 = Clock::now();
    for (size_t i = 0; i < n && st == APSPIE_OK; ++i) {
        if (batch[i]->out || batch[i]->buf) continue;
        batch[i]->buf = s->pinned.get(row);
        if (!batch[i]->buf) st = APSPIE_ERR_OOM;
    }
    ph.alloc_ms += ms_since(t0);
    t0 = Clock::now();
    if (st == APSPIE_OK)
        st = apspie::guard(ctx, [&] {
            for (size_t i = 0; i < n; ++i) {
                const auto &r = batch[i];
                const void *rowp = (uint8_t *)dout + row * i;
                if (r->out && r->out_mem == APSPIE_MEM_DEVICE)
                    APSPIE_HIP(hipMemcpyAsync(r->out, rowp, row, hipMemcpyDeviceToDevice,
                                              ctx->stream));
                else
                    apspie::copy_from_device(ctx, r->out ? r->out : r->buf, rowp, row,
                                             APSPIE_MEM_HOST);
            }
            apspie::sync(ctx);
        });
    ph.copy_ms += ms_since(t0);
    return st;
}

// run one coalesced batch.  Out of budget or memory: the worker's context
// drops its cached scratch (buffers sized for earlier, larger batches) and
// tries again; a batch that still does not fit is split in halves.
void execute(apspie_server *s, apspie_ctx *ctx, std::vector<std::shared_ptr<Request>> &batch) {
    const auto t0 = Clock::now();
    Phases ph;
    int st = run_batch(s, ctx, batch, ph);
    int64_t retries = 0;
    double retry_ms = 0;  // the failed attempt and the trim after it
    if (st == APSPIE_ERR_BUDGET || st == APSPIE_ERR_OOM) {
        apspie_ctx_trim(ctx);
        retries = 1;
        retry_ms = ms_since(t0);
        ph = Phases{};
        st = run_batch(s, ctx, batch, ph);
        if ((st == APSPIE_ERR_BUDGET || st == APSPIE_ERR_OOM) && batch.size() > 1) {
            {
                std::lock_guard<std::mutex> lk(s->mu);
                Busy &b = s->graphs[batch[0]->graph].busy[batch[0]->kind];
                b.retries += 2;
                b.retry_ms += ms_since(t0);
            }
            const size_t n = batch.size();
            std::vector<std::shared_ptr<Request>> lo(batch.begin(), batch.begin() + n / 2),
                hi(batch.begin() + n / 2, batch.end());
            execute(s, ctx, lo);
            execute(s, ctx, hi);
            return;
        }
    }
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &r : batch) {
        r->status = st;
        r->t_done = now;
        r->done = true;
    }
    Busy &b = s->graphs[batch[0]->graph].busy[batch[0]->kind];
    b.batches += 1;
    b.queries += (int64_t)batch.size();
    b.busy_ms += std::chrono::duration<double, std::milli>(now - t0).count();
    b.traverse_ms += ph.traverse_ms;
    b.copy_ms += ph.copy_ms;
    b.alloc_ms += ph.alloc_ms;
    b.retries += retries;
    b.retry_ms += retry_ms;
    s->cv_done.notify_all();
}

void worker(apspie_server *s, int slot) {
    apspie_ctx *ctx = s->ctxs[slot];
    int64_t applied = -1;
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        int64_t budget = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            if (s->stop && s->queue.empty()) return;
            auto head = s->queue.front();
            s->queue.pop_front();
            batch.push_back(head);
            budget = s->slot_budget;
            // TC queries on one graph share one count; max_batch == 1 is the
            // plain one-query-per-launch form with no coalescing at all
            int cap = s->cfg.max_batch <= 1     ? 1
                      : head->kind == APSPIE_Q_TC ? INT32_MAX
                                                  : s->cfg.max_batch;
            // a batch holds its result rows twice (the traversal's matrix and
            // the rows copied out) plus frontier bits: keep it within the
            // worker's budget rather than find out by failing
            if (budget > 0 && head->kind != APSPIE_Q_TC) {
                const int64_t fit = budget / (3 * head->bytes);
                if (fit < cap) {
                    // a pass runs its sources in bit rows of 32, 64, 128, 256 or
                    // 512: a batch the budget cuts short is cut to a full row
                    // (80 sources would run in the 128-wide row at the cost of
                    // 128; 64 fill the 64-wide one)
                    int64_t full = fit >= 32 ? 32 : fit;
                    while (full * 2 <= fit) full *= 2;
                    cap = (int)std::max<int64_t>(1, full);
                }
            }
            for (auto it = s->queue.begin(); it != s->queue.end() && (int)batch.size() < cap;) {
                if ((*it)->graph == head->graph && (*it)->kind == head->kind) {
                    batch.push_back(*it);
                    it = s->queue.erase(it);
                } else {
                    ++it;
                }
            }
            ++s->busy;
        }
        if (budget != applied) {
            // the worker's scratch comes from one arena of its whole budget:
            // a batch that needs other buffers than the last then reuses the
            // arena's ranges instead of hipFree (which would wait for the
            // other workers' streams); without the memory for it the worker
            // keeps per-buffer allocations
            apspie_ctx_reserve_arena(ctx, 0);
            apspie_ctx_set_hbm_budget(ctx, budget);
            if (budget > 0 && apspie_ctx_reserve_arena(ctx, budget) != APSPIE_OK)
                apspie_ctx_reserve_arena(ctx, 0);
            applied = budget;
        }
        execute(s, ctx, batch);
        std::lock_guard<std::mutex> lk(s->mu);
        --s->busy;
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int apspie_server_create(int device, const apspie_server_config *cfg, apspie_server **out) {
    if (!cfg || !out || cfg->slots < 1 || cfg->slots > 16 || cfg->max_batch < 1 ||
        cfg->hbm_budget < 0)
        return APSPIE_ERR_INVALID;
    *out = nullptr;
    auto *s = new apspie_server();
    s->device = device;
    s->cfg = *cfg;
    int st = apspie_ctx_create(device, &s->admin);
    if (st != APSPIE_OK) {
        delete s;
        return st;
    }
    s->slot_budget = cfg->hbm_budget / cfg->slots;
    for (int i = 0; i < cfg->slots; ++i) {
        apspie_ctx *c = nullptr;
        st = apspie_ctx_create(device, &c);
        if (st != APSPIE_OK) break;
        apspie_ctx_set_hbm_budget(c, s->slot_budget);
        s->ctxs.push_back(c);
    }
    if (st != APSPIE_OK) {
        for (auto *c : s->ctxs) apspie_ctx_destroy(c);
        apspie_ctx_destroy(s->admin);
        delete s;
        return st;
    }
    for (int i = 0; i < cfg->slots; ++i) s->workers.emplace_back(worker, s, i);
    *out = s;
    return APSPIE_OK;
}

int apspie_server_destroy(apspie_server *s) {
    if (!s) return APSPIE_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &r : s->graphs) apspie_graph_destroy(r.g);
    for (auto *c : s->ctxs) apspie_ctx_destroy(c);
    apspie_ctx_destroy(s->admin);
    delete s;
    return APSPIE_OK;
}

}  // extern "C"

namespace {

// Resident graphs may take 60% of the budget; the workers split what the
// graphs leave.
bool fits(apspie_server *s, int64_t bytes) {
    std::lock_guard<std::mutex> lk(s->mu);
    return s->cfg.hbm_budget <= 0 || s->graph_bytes + bytes <= (int64_t)(0.6 * s->cfg.hbm_budget);
}

// Admit a graph the admin context has just built: check the budget with its
// actual size, build what the traversals and TC derive from a graph (kept
// with it: longest row, vertices without entries, symmetry and components,
// value range, rows below the diagonal) so that the workers, whose scratch
// fills their budget, never build it mid-batch, and publish it to the pool.
int admit(apspie_server *s, apspie_graph *g, int32_t *graph_id) {
    int32_t m = 0;
    int64_t bytes = 0;
    apspie_graph_info(g, &m, nullptr, nullptr, &bytes);
    if (!fits(s, bytes)) {
        apspie_graph_destroy(g);
        return APSPIE_ERR_BUDGET;
    }
    int st = apspie::guard(s->admin, [&] {
        apspie::graph_max_degree(s->admin, g);
        apspie::graph_iso_bits(s->admin, g);
        if (apspie::graph_is_symmetric(s->admin, g)) apspie::graph_components(s->admin, g);
        float lo = 0, hi = 0;
        apspie::graph_weight_range(s->admin, g, &lo, &hi);
        apspie_ctx_trim(s->admin);  // the admin's temporaries
    });
    if (st != APSPIE_OK) {
        apspie_graph_destroy(g);
        return st;
    }
    st = apspie::guard(s->admin, [&] { apspie::graph_prepare_tc(s->admin, g); });
    if (st != APSPIE_OK && st != APSPIE_ERR_INVALID) {  // unsorted rows: TC fails later
        apspie_graph_destroy(g);
        return st;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->cfg.hbm_budget > 0 && s->graph_bytes + bytes > (int64_t)(0.6 * s->cfg.hbm_budget)) {
        apspie_graph_destroy(g);  // another admission took the room meanwhile
        return APSPIE_ERR_BUDGET;
    }
    s->graphs.push_back({g, m, bytes});
    s->graph_bytes += bytes;
    if (s->cfg.hbm_budget > 0)
        s->slot_budget = std::max<int64_t>(1, (s->cfg.hbm_budget - s->graph_bytes) /
                                                  (int64_t)s->ctxs.size());
    *graph_id = (int32_t)s->graphs.size() - 1;
    return APSPIE_OK;
}

}  // namespace

extern "C" {

int apspie_server_add_graph(apspie_server *s, const apspie_csr *host, int32_t *graph_id) {
    if (!s || !host || !graph_id) return APSPIE_ERR_INVALID;
    // refuse before uploading when the arrays alone do not fit
    if (!fits(s, 4 * ((int64_t)host->m + 1) + (host->val ? 8 : 4) * host->nnz))
        return APSPIE_ERR_BUDGET;
    std::lock_guard<std::mutex> ak(s->admin_mu);
    apspie_graph *g = nullptr;
    int st = apspie_graph_upload(s->admin, host, &g);
    if (st != APSPIE_OK) return st;
    return admit(s, g, graph_id);
}

int apspie_server_add_graph_coo(apspie_server *s, const apspie_coo *coo, int mem, int symmetrize,
                                int32_t *graph_id) {
    if (!s || !coo || !graph_id) return APSPIE_ERR_INVALID;
    std::lock_guard<std::mutex> ak(s->admin_mu);
    apspie_graph *g = nullptr;
    int st = apspie_graph_from_coo(s->admin, coo, mem, symmetrize, &g);
    if (st != APSPIE_OK) return st;
    // the build's COO staging and sort buffers are not the graph's: release them
    apspie_ctx_trim(s->admin);
    return admit(s, g, graph_id);
}

int apspie_server_add_graph_file(apspie_server *s, const char *path, int32_t *graph_id) {
    if (!s || !path || !graph_id) return APSPIE_ERR_INVALID;
    std::lock_guard<std::mutex> ak(s->admin_mu);
    apspie_graph *g = nullptr;
    int st = apspie_graph_load(s->admin, path, &g);
    if (st != APSPIE_OK) return st;
    return admit(s, g, graph_id);
}

const char *apspie_server_last_error(apspie_server *s) {
    return s ? apspie_last_error(s->admin) : "no server";
}

static int submit(apspie_server *s, int32_t graph_id, int kind, int32_t source, void *out,
                  int out_mem, int64_t *ticket) {
    if (!s || !ticket || kind < APSPIE_Q_BFS || kind > APSPIE_Q_TC) return APSPIE_ERR_INVALID;
    if (out_mem != APSPIE_MEM_HOST && out_mem != APSPIE_MEM_DEVICE) return APSPIE_ERR_INVALID;
    auto r = std::make_shared<Request>();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return APSPIE_ERR_INVALID;
        const Resident &R = s->graphs[graph_id];
        if (kind != APSPIE_Q_TC && (source < 0 || source >= R.m)) return APSPIE_ERR_INVALID;
        if (s->cfg.queue_capacity > 0 && (int)s->queue.size() >= s->cfg.queue_capacity)
            return APSPIE_ERR_BUSY;
        r->ticket = s->next_ticket++;
        r->graph = graph_id;
        r->kind = kind;
        r->source = source;
        r->g = R.g;
        r->m = R.m;
        r->bytes = result_bytes(kind, R.m);
        r->out = kind == APSPIE_Q_TC ? nullptr : out;  // a count is returned by the wait
        r->out_mem = out_mem;
        r->t_submit = Clock::now();
        s->queue.push_back(r);
        s->live[r->ticket] = r;
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return APSPIE_OK;
}

int apspie_server_submit(apspie_server *s, int32_t graph_id, int kind, int32_t source,
                         int64_t *ticket) {
    return submit(s, graph_id, kind, source, nullptr, APSPIE_MEM_HOST, ticket);
}

int apspie_server_submit_into(apspie_server *s, int32_t graph_id, int kind, int32_t source,
                              void *out, int mem, int64_t *ticket) {
    if (kind != APSPIE_Q_TC && !out) return APSPIE_ERR_INVALID;
    return submit(s, graph_id, kind, source, out, mem, ticket);
}

int apspie_server_wait(apspie_server *s, int64_t ticket, void *out, int64_t out_bytes,
                       double *latency_ms) {
    if (!s) return APSPIE_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return APSPIE_ERR_INVALID;
    auto r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->live.erase(ticket);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    lk.unlock();
    int st = r->status;
    const int64_t need = r->bytes;
    if (st == APSPIE_OK && out) {
        if (out_bytes < need) st = APSPIE_ERR_INVALID;
        else if (r->kind == APSPIE_Q_TC) std::memcpy(out, &r->triangles, sizeof r->triangles);
        else if (r->out) st = APSPIE_ERR_INVALID;  // the result is in the submitter's memory
        else std::memcpy(out, r->buf, (size_t)need);
    }
    s->pinned.put(r->buf);
    r->buf = nullptr;
    return st;
}

int apspie_server_wait_view(apspie_server *s, int64_t ticket, const void **data, int64_t *bytes,
                            double *latency_ms) {
    if (!s || !data || !bytes) return APSPIE_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return APSPIE_ERR_INVALID;
    auto r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->live.erase(ticket);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    *data = nullptr;
    *bytes = 0;
    if (r->status != APSPIE_OK) {  // nothing to hold: the buffer goes back now
        lk.unlock();
        s->pinned.put(r->buf);
        r->buf = nullptr;
        return r->status;
    }
    *bytes = r->bytes;
    *data = r->kind == APSPIE_Q_TC ? (const void *)&r->triangles : r->out ? r->out : r->buf;
    s->viewed[ticket] = r;
    return APSPIE_OK;
}

int apspie_server_release(apspie_server *s, int64_t ticket) {
    if (!s) return APSPIE_ERR_INVALID;
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->viewed.find(ticket);
        if (it == s->viewed.end()) return APSPIE_ERR_INVALID;
        r = it->second;
        s->viewed.erase(it);
    }
    s->pinned.put(r->buf);
    r->buf = nullptr;
    return APSPIE_OK;
}

int apspie_server_busy(apspie_server *s, int32_t graph_id, int kind, int64_t *batches,
                       int64_t *queries, double *busy_ms, int reset) {
    if (!s || kind < APSPIE_Q_BFS || kind > APSPIE_Q_TC) return APSPIE_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return APSPIE_ERR_INVALID;
    Busy &b = s->graphs[graph_id].busy[kind];
    if (batches) *batches = b.batches;
    if (queries) *queries = b.queries;
    if (busy_ms) *busy_ms = b.busy_ms;
    if (reset) b = Busy{};
    return APSPIE_OK;
}

int apspie_server_busy_stats(apspie_server *s, int32_t graph_id, int kind,
                             apspie_server_stats *out, int reset) {
    if (!s || !out || kind < APSPIE_Q_BFS || kind > APSPIE_Q_TC) return APSPIE_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return APSPIE_ERR_INVALID;
    Busy &b = s->graphs[graph_id].busy[kind];
    *out = apspie_server_stats{b.batches, b.queries,   b.busy_ms, b.traverse_ms,
                               b.copy_ms, b.alloc_ms, b.retries, b.retry_ms};
    if (reset) b = Busy{};
    return APSPIE_OK;
}

int apspie_server_drain(apspie_server *s) {
    if (!s) return APSPIE_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->busy == 0; });
    return APSPIE_OK;
}

}  // extern "C"
