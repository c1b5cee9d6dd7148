// Served form: a request loop with continuous batching over the C ABI.
//
// Requests (one edge map each) are queued as they arrive.  Each worker thread
// owns one context (its own HIP streams and device pools, kept resident
// across batches); whenever a worker is free it admits the pending requests,
// oldest first, while their estimated HBM footprint (edges x bytes per edge
// at the expected neighbour density) fits its budget, and runs them as one
// batch (cvg_run_batch).  Requests arriving while other batches are on the
// GPU are admitted by the next free worker, so the work of consecutive batches
// overlaps on the device.  A map's results are those of cvg_run on it alone:
// the batch shares launches, not arithmetic.
//
// Host memory: request edges are copied at submit into page-locked blocks,
// and each batch's chains and info come back in one transfer into a
// page-locked result block shared by the batch's requests (released when the
// last of them is), both from a pool, so transfers run by DMA at link speed
// and a request's rows are read in place (cvg_server_view) or copied once.
//
// The reference has no serving layer; this is the loop around
// run_curvelet<double> (cpu_curvelet/main.cpp:20-96) a service would run.
#include <chrono>
#include <condition_variable>
#include <map>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/curvelet_gpu.h"

namespace {

using Clock = std::chrono::steady_clock;

// Page-locked host blocks, reused by size (smallest free block that fits).
// Sizes come in classes 1.25x apart, in whole MiB: maps and batches of
// varying size then reuse blocks instead of pinning new memory, which costs
// far more than the copies the blocks serve, and a block wastes at most a
// quarter of its size.
class PinnedPool {
  public:
    ~PinnedPool()
    {
        for (auto &f : free_) cvg_host_free(f.second);
    }
    void *get(size_t bytes, size_t *got)
    {
        const size_t mib = static_cast<size_t>(1) << 20;
        size_t cls = mib;
        while (cls < bytes) cls = (cls + cls / 4 + mib - 1) & ~(mib - 1);
        bytes = cls;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= 4 * bytes) {
                void *p = it->second;
                *got = it->first;
                free_.erase(it);
                return p;
            }
        }
        void *p = nullptr;
        if (cvg_host_alloc(bytes, &p) != CVG_OK) return nullptr;
        *got = bytes;
        return p;
    }
    void put(void *p, size_t bytes)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(bytes, p);
    }

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
};

struct Block {
    PinnedPool *pool = nullptr;
    void *mem = nullptr;
    size_t bytes = 0;
    ~Block()
    {
        if (pool) pool->put(mem, bytes);
    }
};

struct Request {
    unsigned long long id = 0;
    std::unique_ptr<Block> edges;  // page-locked copy of the map
    int n = 0;
    int sz = 4;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = CVG_OK;
    unsigned rows = 0, width = 0;
    std::shared_ptr<Block> result;  // the batch's chains and info
    const unsigned *chains = nullptr;
    const double *info = nullptr;
    int batch_maps = 0;
};

}  // namespace

struct cvg_server {
    PinnedPool pool;
    cvg_params params{};
    cvg_server_config cfg{};
    std::vector<cvg_context *> ctxs;
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<unsigned long long, std::shared_ptr<Request>> table;
    unsigned long long next_id = 1;
    bool stop = false;
    std::string err;
    // counters
    unsigned long long batches = 0, maps_done = 0, edges_done = 0;
};

namespace {

double footprint(const cvg_server_config &c, int n) { return static_cast<double>(n) * c.bytes_per_edge; }

void worker(cvg_server *s, int w)
{
    cvg_context *ctx = s->ctxs[w];
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->stop && s->pending.empty()) return;
            double fp = 0;
            while (!s->pending.empty() && static_cast<int>(batch.size()) < s->cfg.max_batch_maps) {
                const auto &r = s->pending.front();
                const double f = footprint(s->cfg, r->n);
                if (!batch.empty() && (fp + f > s->cfg.hbm_budget_bytes || r->sz != batch[0]->sz)) break;
                fp += f;
                batch.push_back(r);
                s->pending.pop_front();
            }
        }
        const auto t0 = Clock::now();
        std::vector<int> counts(batch.size());
        std::vector<const double *> ptrs(batch.size());
        for (size_t i = 0; i < batch.size(); ++i) {
            counts[i] = batch[i]->n;
            ptrs[i] = static_cast<const double *>(batch[i]->edges->mem);
            batch[i]->t_start = t0;
        }
        int rc = cvg_run_batch(ctx, static_cast<int>(batch.size()), counts.data(), batch[0]->sz, ptrs.data(),
                               &s->params, nullptr);
        // all of the batch's rows in one transfer each for chains and info
        unsigned total = 0, width = 0;
        std::shared_ptr<Block> blk;
        if (rc == CVG_OK) rc = cvg_curvelet_shape(ctx, &total, &width);
        if (rc == CVG_OK) {
            const size_t cbytes = (static_cast<size_t>(total) * width * sizeof(unsigned) + 63) & ~static_cast<size_t>(63);
            const size_t bytes = cbytes + static_cast<size_t>(total) * 10 * sizeof(double) + 64;
            blk = std::make_shared<Block>();
            blk->mem = s->pool.get(bytes, &blk->bytes);
            blk->pool = &s->pool;
            if (!blk->mem) {
                rc = CVG_ERR_OUT_OF_MEMORY;
            } else {
                unsigned *chains = static_cast<unsigned *>(blk->mem);
                double *info = reinterpret_cast<double *>(static_cast<char *>(blk->mem) + cbytes);
                rc = cvg_get_curvelets(ctx, chains, info);
                unsigned row = 0;
                for (size_t i = 0; i < batch.size() && rc == CVG_OK; ++i) {
                    Request &r = *batch[i];
                    unsigned rows = 0, w = 0;
                    rc = cvg_batch_map_shape(ctx, static_cast<int>(i), &rows, &w);
                    r.rows = rows;
                    r.width = w;
                    r.chains = chains + static_cast<size_t>(row) * width;
                    r.info = info + static_cast<size_t>(row) * 10;
                    r.result = blk;
                    row += rows;
                }
            }
        }
        for (auto &r : batch) {
            r->status = rc;
            r->batch_maps = static_cast<int>(batch.size());
            r->edges.reset();
            if (rc != CVG_OK) r->result.reset();
        }
        blk.reset();
        const auto t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto &r : batch) {
                r->t_done = t1;
                r->done = true;
                s->maps_done += 1;
                s->edges_done += static_cast<unsigned long long>(r->n);
            }
            s->batches += 1;
            if (rc != CVG_OK) s->err = cvg_last_error(ctx);
        }
        s->cv_done.notify_all();
    }
}

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

extern "C" {

void cvg_server_config_default(cvg_server_config *cfg)
{
    if (!cfg) return;
    cfg->workers = 2;
    cfg->max_batch_maps = 64;
    cfg->bytes_per_edge = 4096.0;           // ~16 neighbours per edge x ~250 B per pair, plus per-edge arrays
    cfg->hbm_budget_bytes = 24.0 * (1ull << 30);
}

int cvg_server_create(int device, const cvg_params *params, const cvg_server_config *cfg, cvg_server **out)
{
    if (!params || !out) return CVG_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto *s = new (std::nothrow) cvg_server;
    if (!s) return CVG_ERR_OUT_OF_MEMORY;
    s->params = *params;
    if (cfg)
        s->cfg = *cfg;
    else
        cvg_server_config_default(&s->cfg);
    if (s->cfg.workers < 1 || s->cfg.workers > 8 || s->cfg.max_batch_maps < 1 || !(s->cfg.bytes_per_edge > 0) ||
        !(s->cfg.hbm_budget_bytes > 0)) {
        delete s;
        return CVG_ERR_INVALID_ARGUMENT;
    }
    for (int w = 0; w < s->cfg.workers; ++w) {
        cvg_context *c = nullptr;
        const int rc = cvg_create(device, &c);
        if (rc != CVG_OK) {
            for (auto *x : s->ctxs) cvg_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(c);
    }
    for (int w = 0; w < s->cfg.workers; ++w) s->threads.emplace_back(worker, s, w);
    *out = s;
    return CVG_OK;
}

int cvg_server_submit(cvg_server *s, int num_edges, int sz_edge_data, const double *edges,
                      unsigned long long *request_id)
{
    if (!s || !edges || !request_id || num_edges <= 0 || sz_edge_data < 4) return CVG_ERR_INVALID_ARGUMENT;
    auto r = std::make_shared<Request>();
    r->t_submit = Clock::now();
    r->n = num_edges;
    r->sz = sz_edge_data;
    const size_t bytes = static_cast<size_t>(num_edges) * sz_edge_data * sizeof(double);
    r->edges.reset(new Block);
    r->edges->mem = s->pool.get(bytes, &r->edges->bytes);
    r->edges->pool = &s->pool;
    if (!r->edges->mem) return CVG_ERR_OUT_OF_MEMORY;
    std::memcpy(r->edges->mem, edges, bytes);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->table[r->id] = r;
        s->pending.push_back(r);
        *request_id = r->id;
    }
    s->cv_work.notify_one();
    return CVG_OK;
}

int cvg_server_wait(cvg_server *s, unsigned long long request_id, cvg_server_result *res)
{
    if (!s || !res) return CVG_ERR_INVALID_ARGUMENT;
    std::shared_ptr<Request> r;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->table.find(request_id);
    if (it == s->table.end()) return CVG_ERR_INVALID_ARGUMENT;
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    res->status = r->status;
    res->num_curvelets = r->rows;
    res->chain_width = r->width;
    res->num_edges = r->n;
    res->batch_maps = r->batch_maps;
    res->queue_ms = ms_between(r->t_submit, r->t_start);
    res->latency_ms = ms_between(r->t_submit, r->t_done);
    return r->status;
}

int cvg_server_fetch(cvg_server *s, unsigned long long request_id, unsigned *chains, double *info)
{
    if (!s) return CVG_ERR_INVALID_ARGUMENT;
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->table.find(request_id);
        if (it == s->table.end() || !it->second->done) return CVG_ERR_STATE;
        r = it->second;
        s->table.erase(it);
    }
    if (r->status != CVG_OK) return r->status;
    if (chains && r->rows) std::memcpy(chains, r->chains, static_cast<size_t>(r->rows) * r->width * sizeof(unsigned));
    if (info && r->rows) std::memcpy(info, r->info, static_cast<size_t>(r->rows) * 10 * sizeof(double));
    return CVG_OK;
}

int cvg_server_view(cvg_server *s, unsigned long long request_id, const unsigned **chains, const double **info)
{
    if (!s) return CVG_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->table.find(request_id);
    if (it == s->table.end() || !it->second->done) return CVG_ERR_STATE;
    if (it->second->status != CVG_OK) return it->second->status;
    if (chains) *chains = it->second->chains;
    if (info) *info = it->second->info;
    return CVG_OK;
}

int cvg_server_release(cvg_server *s, unsigned long long request_id)
{
    if (!s) return CVG_ERR_INVALID_ARGUMENT;
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->table.find(request_id);
        if (it == s->table.end() || !it->second->done) return CVG_ERR_STATE;
        r = it->second;
        s->table.erase(it);
    }
    return CVG_OK;
}

int cvg_server_counters(cvg_server *s, unsigned long long *batches, unsigned long long *maps,
                        unsigned long long *edges)
{
    if (!s) return CVG_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    if (batches) *batches = s->batches;
    if (maps) *maps = s->maps_done;
    if (edges) *edges = s->edges_done;
    return CVG_OK;
}

const char *cvg_server_last_error(cvg_server *s)
{
    static thread_local std::string copy;
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    copy = s->err;
    return copy.c_str();
}

void cvg_server_destroy(cvg_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    for (auto *c : s->ctxs) cvg_destroy(c);
    delete s;
}

}  // extern "C"
