// serving.cpp -- continuous-batching request loop over a device-resident
// PQIVF index.
//
// Requests (a few queries each, their own k and nprobe) are queued by
// submit(). A worker thread forms batches from the head of the queue and
// launches each on one of `slots` streams, each slot with its own scratch,
// pinned staging and completion event; up to `slots` batches are in flight
// on the device at once. A batch is admitted when a slot is free and its
// device buffers fit in what is left of the HBM budget; requests that arrive
// while batches run join the next batch as soon as a slot frees. A
// completion thread waits on the slots' events and hands results back.
//
// Mutations (BulkAdd, BulkDelete) queue in the same order: the worker lets
// every earlier batch finish, applies the mutation under the index's write
// lock, and only then launches later searches, which is the reader/writer
// rule of pqivf/index.go:37 (a search sees a write wholly or not at all).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pqivf.h"
#include "runtime.h"

namespace hann {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    int kind = 0;   // 0 search, 1 bulk add, 2 bulk delete
    int64_t nq = 0;
    int kmax = 0, pmax = 0;
    std::vector<float> q;
    std::vector<int32_t> k, np;
    std::vector<int64_t> ids;     // mutations
    std::vector<float> vecs;
    // results
    std::vector<int64_t> out_ids;
    std::vector<double> out_dist;
    std::vector<int32_t> out_cnt;
    int status = HANN_OK;
    std::string error;
    bool done = false;
    Clock::time_point t_submit, t_done;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done_ev = nullptr;
    Workspace ws;
    // pinned staging
    float *h_q = nullptr;
    int32_t *h_k = nullptr, *h_np = nullptr, *h_cnt = nullptr;
    float *h_key = nullptr;
    int64_t *h_id = nullptr;
    size_t cap_q = 0, cap_out = 0;
    DevBuf<float> d_q, d_key;
    DevBuf<int32_t> d_k, d_np, d_cnt;
    DevBuf<int64_t> d_id;
    std::vector<Request *> reqs;
    int64_t nq = 0;
    int kmax = 0;
    int64_t bytes = 0;       // the batch in flight's charge
    int64_t bound = 0;       // its scratch bound (the charge less the staging)
    int64_t ws_before = 0;   // the slot's workspace bytes when the batch started
    bool busy = false;
    // device bytes the slot holds: its workspace (grow-only) and buffers
    int64_t held() const {
        return (int64_t)ws.bytes() + (int64_t)(d_q.n * sizeof(float) + d_key.n * sizeof(float) +
                                               (d_k.n + d_np.n + d_cnt.n) * sizeof(int32_t) + d_id.n * sizeof(int64_t));
    }
};

struct Server {
    PQIVF *ix = nullptr;
    hann_server_options opt{};
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_slot;
    std::deque<Request *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> table;
    uint64_t next_ticket = 1;
    bool stop = false;
    int64_t budget_in_use = 0, budget_peak = 0;
    int64_t held_peak = 0;            // device bytes all slots held, at most
    int64_t scratch_over_bound = 0;   // most a batch grew its slot's workspace past its scratch bound (0: the bounds held)
    int64_t completed_queries = 0, completed_requests = 0, batches = 0, mutations = 0;
    int in_flight = 0;
    std::thread worker, completer;
    std::deque<Slot *> launched;   // in launch order
    std::string fatal;

    // a batch's charge against the budget: its query, parameter and result
    // buffers, and the search's workspace bound for its shape (list lengths
    // as of the last refresh of the index's shape)
    int64_t batch_bytes(int64_t nq, int kmax, int pmax) const {
        return staging_bytes(nq, kmax) + ix->search_scratch_bytes(nq, kmax, pmax, opt.params);
    }
    int64_t staging_bytes(int64_t nq, int kmax) const {
        return nq * (ix->dim * 4 + 12 + (int64_t)kmax * 12) + 4 * 256;
    }

    void ensure_slot_capacity(Slot &s, int64_t nq, int kmax) {
        const size_t need_q = (size_t)nq * ix->dim, need_o = (size_t)nq * kmax;
        if (need_q > s.cap_q || (size_t)nq > s.cap_q / std::max<int64_t>(1, ix->dim)) {
            if (s.h_q) (void)hipHostFree(s.h_q);
            if (s.h_k) (void)hipHostFree(s.h_k);
            if (s.h_np) (void)hipHostFree(s.h_np);
            if (s.h_cnt) (void)hipHostFree(s.h_cnt);
            HANN_HIP(hipHostMalloc((void **)&s.h_q, need_q * sizeof(float)));
            HANN_HIP(hipHostMalloc((void **)&s.h_k, (size_t)nq * sizeof(int32_t)));
            HANN_HIP(hipHostMalloc((void **)&s.h_np, (size_t)nq * sizeof(int32_t)));
            HANN_HIP(hipHostMalloc((void **)&s.h_cnt, (size_t)nq * sizeof(int32_t)));
            s.cap_q = need_q;
        }
        if (need_o > s.cap_out) {
            if (s.h_key) (void)hipHostFree(s.h_key);
            if (s.h_id) (void)hipHostFree(s.h_id);
            HANN_HIP(hipHostMalloc((void **)&s.h_key, need_o * sizeof(float)));
            HANN_HIP(hipHostMalloc((void **)&s.h_id, need_o * sizeof(int64_t)));
            s.cap_out = need_o;
        }
        s.d_q.resize(need_q);
        s.d_k.resize((size_t)nq);
        s.d_np.resize((size_t)nq);
        s.d_cnt.resize((size_t)nq);
        s.d_key.resize(need_o);
        s.d_id.resize(need_o);
    }

    // caller holds mu; returns with the batch launched (mu released meanwhile)
    void launch(Slot &s, std::vector<Request *> reqs, std::unique_lock<std::mutex> &lk) {
        int64_t nq = 0;
        int kmax = 1, pmax = 1;
        for (Request *r : reqs) {
            nq += r->nq;
            kmax = std::max(kmax, r->kmax);
            pmax = std::max(pmax, r->pmax);
        }
        s.reqs = std::move(reqs);
        s.nq = nq;
        s.kmax = kmax;
        s.bytes = batch_bytes(nq, kmax, pmax);
        s.bound = s.bytes - staging_bytes(nq, kmax);
        s.ws_before = (int64_t)s.ws.bytes();
        s.busy = true;
        budget_in_use += s.bytes;
        budget_peak = std::max(budget_peak, budget_in_use);
        ++in_flight;
        lk.unlock();
        try {
            ensure_slot_capacity(s, nq, kmax);
            int64_t o = 0;
            int pmin = INT32_MAX;
            for (Request *r : s.reqs) {
                std::memcpy(s.h_q + o * ix->dim, r->q.data(), r->q.size() * sizeof(float));
                std::memcpy(s.h_k + o, r->k.data(), r->nq * sizeof(int32_t));
                std::memcpy(s.h_np + o, r->np.data(), r->nq * sizeof(int32_t));
                for (int32_t v : r->np) pmin = std::min(pmin, v);
                o += r->nq;
            }
            HANN_HIP(hipMemcpyAsync(s.d_q.get(), s.h_q, (size_t)nq * ix->dim * sizeof(float), hipMemcpyHostToDevice,
                                    s.stream));
            HANN_HIP(hipMemcpyAsync(s.d_k.get(), s.h_k, (size_t)nq * sizeof(int32_t), hipMemcpyHostToDevice, s.stream));
            HANN_HIP(hipMemcpyAsync(s.d_np.get(), s.h_np, (size_t)nq * sizeof(int32_t), hipMemcpyHostToDevice, s.stream));
            {
                std::unique_lock<std::shared_mutex> ilk(ix->mu);
                ix->sync_device();
                ix->search_device(s.d_q.get(), nq, s.d_k.get(), s.d_np.get(), kmax, pmax, opt.params, s.d_key.get(),
                                  s.d_id.get(), s.d_cnt.get(), s.stream, pmin, &s.ws);
            }
            HANN_HIP(hipMemcpyAsync(s.h_key, s.d_key.get(), (size_t)nq * kmax * sizeof(float), hipMemcpyDeviceToHost,
                                    s.stream));
            HANN_HIP(hipMemcpyAsync(s.h_id, s.d_id.get(), (size_t)nq * kmax * sizeof(int64_t), hipMemcpyDeviceToHost,
                                    s.stream));
            HANN_HIP(hipMemcpyAsync(s.h_cnt, s.d_cnt.get(), (size_t)nq * sizeof(int32_t), hipMemcpyDeviceToHost,
                                    s.stream));
            HANN_HIP(hipEventRecord(s.done_ev, s.stream));
        } catch (const Error &e) {
            lk.lock();
            for (Request *r : s.reqs) {
                r->status = e.code;
                r->error = e.what();
            }
            lk.unlock();
            (void)hipStreamSynchronize(s.stream);
            (void)hipEventRecord(s.done_ev, s.stream);
        }
        lk.lock();
        ++batches;
        launched.push_back(&s);
        cv_done.notify_all();
    }

    void apply_mutation(Request *r) {
        try {
            std::unique_lock<std::shared_mutex> ilk(ix->mu);
            if (r->kind == 1)
                ix->bulk_add(r->ids.data(), r->vecs.data(), nullptr, r->ids.size());
            else
                ix->bulk_delete(r->ids.data(), r->ids.size());
            // later batches are charged for the lists as they now are
            ix->sync_device();
            ix->refresh_shape();
        } catch (const Error &e) {
            r->status = e.code;
            r->error = e.what();
        }
    }

    void worker_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (queue.empty() && stop) break;
            Request *head = queue.front();
            if (head->kind != 0) {
                // a write: every earlier batch completes first
                cv_slot.wait(lk, [&] { return in_flight == 0; });
                queue.pop_front();
                lk.unlock();
                apply_mutation(head);
                lk.lock();
                head->done = true;
                head->t_done = Clock::now();
                ++mutations;
                cv_done.notify_all();
                continue;
            }
            // a free slot and budget for at least the head request
            Slot *free_slot = nullptr;
            cv_slot.wait(lk, [&] {
                free_slot = nullptr;
                for (auto &s : slots)
                    if (!s->busy) {
                        free_slot = s.get();
                        break;
                    }
                if (!free_slot) return false;
                const int64_t need = batch_bytes(queue.front()->nq, queue.front()->kmax, queue.front()->pmax);
                return in_flight == 0 || budget_in_use + need <= opt.hbm_budget;
            });
            // gather the batch from the head of the queue
            std::vector<Request *> batch;
            int64_t nq = 0;
            int kmax = 1, pmax = 1;
            while (!queue.empty() && queue.front()->kind == 0) {
                Request *r = queue.front();
                const int64_t nnq = nq + r->nq;
                const int64_t need = batch_bytes(nnq, std::max(kmax, r->kmax), std::max(pmax, r->pmax));
                if (!batch.empty() && (nnq > opt.max_batch || (in_flight > 0 && budget_in_use + need > opt.hbm_budget)))
                    break;
                batch.push_back(r);
                queue.pop_front();
                nq = nnq;
                kmax = std::max(kmax, r->kmax);
                pmax = std::max(pmax, r->pmax);
            }
            launch(*free_slot, std::move(batch), lk);
        }
    }

    void completer_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_done.wait(lk, [&] { return (stop && launched.empty() && in_flight == 0) || !launched.empty(); });
            if (launched.empty()) {
                if (stop && in_flight == 0) break;
                continue;
            }
            Slot *s = launched.front();
            launched.pop_front();
            lk.unlock();
            (void)hipEventSynchronize(s->done_ev);
            const auto now = Clock::now();
            // unpack: FromRank is the float64 square root of the float32 rank
            int64_t o = 0;
            for (Request *r : s->reqs) {
                if (r->status == HANN_OK) {
                    r->out_ids.assign((size_t)r->nq * r->kmax, -1);
                    r->out_dist.assign((size_t)r->nq * r->kmax, INFINITY);
                    r->out_cnt.assign((size_t)r->nq, 0);
                    for (int64_t i = 0; i < r->nq; ++i) {
                        // a query never has more results than its own k
                        const int32_t c = std::min(s->h_cnt[o + i], r->k[i]);
                        r->out_cnt[i] = c;
                        for (int j = 0; j < c && j < r->kmax; ++j) {
                            r->out_ids[(size_t)i * r->kmax + j] = s->h_id[(o + i) * s->kmax + j];
                            r->out_dist[(size_t)i * r->kmax + j] = std::sqrt((double)s->h_key[(o + i) * s->kmax + j]);
                        }
                    }
                }
                o += r->nq;
            }
            lk.lock();
            for (Request *r : s->reqs) {
                r->done = true;
                r->t_done = now;
                completed_queries += r->nq;
                ++completed_requests;
            }
            s->reqs.clear();
            s->busy = false;
            budget_in_use -= s->bytes;
            // the workspace only grows, each slot to the largest need of
            // any batch it ran: what one batch adds is at most its bound
            scratch_over_bound = std::max(scratch_over_bound, (int64_t)s->ws.bytes() - s->ws_before - s->bound);
            int64_t held = 0;
            for (auto &sl : slots) held += sl->held();
            held_peak = std::max(held_peak, held);
            --in_flight;
            cv_slot.notify_all();
            cv_done.notify_all();
        }
    }
};

}  // namespace hann

using namespace hann;

struct hann_server {
    Server s;
};

extern "C" {

void hann_server_default_options(hann_server_options *o) {
    std::memset(o, 0, sizeof(*o));
    o->max_batch = 4096;
    o->slots = 2;
    o->hbm_budget = 1ll << 30;
    hann_search_default_params(&o->params);
    o->params.scan_mode = HANN_SCAN_ADC_F32;
    o->params.fused = 1;
}

int hann_server_create(hann_pqivf *idx, const hann_server_options *opts, hann_server **out) {
    HANN_API_BEGIN
    *out = nullptr;
    if (!idx) fail(HANN_ERR_INVALID_ARG, "index handle is null");
    auto h = std::make_unique<hann_server>();
    Server &s = h->s;
    s.ix = &idx->ix;
    hann_server_default_options(&s.opt);
    if (opts) s.opt = *opts;
    if (s.opt.max_batch <= 0) s.opt.max_batch = 4096;
    if (s.opt.slots <= 0) s.opt.slots = 2;
    if (s.opt.slots > 8) s.opt.slots = 8;
    if (s.opt.hbm_budget <= 0) s.opt.hbm_budget = 1ll << 30;
    {
        std::unique_lock<std::shared_mutex> ilk(s.ix->mu);
        if (!s.ix->trained) fail(HANN_ERR_NOT_TRAINED, "PQIVF index is not trained. Call Train() after adding data before searching");
        s.ix->ensure_stream();
        s.ix->sync_device();
        s.ix->refresh_shape();
    }
    for (int i = 0; i < s.opt.slots; ++i) {
        auto sl = std::make_unique<Slot>();
        HANN_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
        HANN_HIP(hipEventCreateWithFlags(&sl->done_ev, hipEventDisableTiming));
        s.slots.push_back(std::move(sl));
    }
    s.worker = std::thread([&s] { s.worker_loop(); });
    s.completer = std::thread([&s] { s.completer_loop(); });
    *out = h.release();
    HANN_API_END
}

void hann_server_free(hann_server *srv) {
    if (!srv) return;
    Server &s = srv->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.stop = true;
    }
    s.cv_work.notify_all();
    s.cv_slot.notify_all();
    s.cv_done.notify_all();
    if (s.worker.joinable()) s.worker.join();
    {
        std::unique_lock<std::mutex> lk(s.mu);
        s.cv_slot.wait(lk, [&] { return s.in_flight == 0; });
        s.cv_done.notify_all();
    }
    if (s.completer.joinable()) s.completer.join();
    for (auto &sl : s.slots) {
        (void)hipStreamSynchronize(sl->stream);
        for (void *p : {(void *)sl->h_q, (void *)sl->h_k, (void *)sl->h_np, (void *)sl->h_cnt, (void *)sl->h_key,
                        (void *)sl->h_id})
            if (p) (void)hipHostFree(p);
        (void)hipEventDestroy(sl->done_ev);
        (void)hipStreamDestroy(sl->stream);
    }
    delete srv;
}

int hann_server_submit(hann_server *srv, const float *queries, size_t nq, const int32_t *k, const int32_t *nprobe,
                       uint64_t *ticket) {
    HANN_API_BEGIN
    if (!srv) fail(HANN_ERR_INVALID_ARG, "server handle is null");
    Server &s = srv->s;
    if (nq == 0) fail(HANN_ERR_INVALID_ARG, "a request needs at least one query");
    if (!k) fail(HANN_ERR_INVALID_ARG, "k array is null");
    auto r = std::make_unique<Request>();
    r->kind = 0;
    r->nq = (int64_t)nq;
    r->q.assign(queries, queries + nq * s.ix->dim);
    r->k.assign(k, k + nq);
    r->np.resize(nq);
    for (size_t i = 0; i < nq; ++i) {
        if (k[i] <= 0) fail(HANN_ERR_INVALID_ARG, "k must be positive");
        r->np[i] = nprobe ? nprobe[i] : (int32_t)s.ix->nprobe;
        if (r->np[i] <= 0) fail(HANN_ERR_INVALID_ARG, "nprobe must be positive");
        r->kmax = std::max(r->kmax, k[i]);
        r->pmax = std::max(r->pmax, r->np[i]);
    }
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s.mu);
    r->ticket = s.next_ticket++;
    *ticket = r->ticket;
    s.queue.push_back(r.get());
    s.table[r->ticket] = std::move(r);
    s.cv_work.notify_one();
    HANN_API_END
}

static int submit_mutation(hann_server *srv, int kind, const int64_t *ids, const float *flat, size_t n,
                           uint64_t *ticket) {
    HANN_API_BEGIN
    if (!srv) fail(HANN_ERR_INVALID_ARG, "server handle is null");
    Server &s = srv->s;
    auto r = std::make_unique<Request>();
    r->kind = kind;
    r->ids.assign(ids, ids + n);
    if (flat) r->vecs.assign(flat, flat + n * s.ix->dim);
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s.mu);
    r->ticket = s.next_ticket++;
    *ticket = r->ticket;
    s.queue.push_back(r.get());
    s.table[r->ticket] = std::move(r);
    s.cv_work.notify_one();
    HANN_API_END
}

int hann_server_submit_bulk_add(hann_server *srv, const int64_t *ids, const float *flat, size_t n, uint64_t *ticket) {
    return submit_mutation(srv, 1, ids, flat, n, ticket);
}

int hann_server_submit_bulk_delete(hann_server *srv, const int64_t *ids, size_t n, uint64_t *ticket) {
    return submit_mutation(srv, 2, ids, nullptr, n, ticket);
}

int hann_server_poll(hann_server *srv, uint64_t ticket, int wait, int64_t *out_ids, double *out_dist,
                     int32_t *out_counts, double *latency_ms) {
    HANN_API_BEGIN
    if (!srv) fail(HANN_ERR_INVALID_ARG, "server handle is null");
    Server &s = srv->s;
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.table.find(ticket);
    if (it == s.table.end()) fail(HANN_ERR_NOT_FOUND, strf("ticket %llu not found", (unsigned long long)ticket));
    Request *r = it->second.get();
    if (!r->done) {
        if (!wait) return HANN_ERR_EMPTY;   // not ready yet
        s.cv_done.wait(lk, [&] { return r->done; });
    }
    std::unique_ptr<Request> owned = std::move(it->second);
    s.table.erase(it);
    lk.unlock();
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (r->status != HANN_OK) fail(r->status, r->error);
    if (r->kind == 0) {
        if (out_ids) std::memcpy(out_ids, r->out_ids.data(), r->out_ids.size() * sizeof(int64_t));
        if (out_dist) std::memcpy(out_dist, r->out_dist.data(), r->out_dist.size() * sizeof(double));
        if (out_counts) std::memcpy(out_counts, r->out_cnt.data(), r->out_cnt.size() * sizeof(int32_t));
    }
    HANN_API_END
}

int hann_server_stats(hann_server *srv, hann_server_counters *out) {
    HANN_API_BEGIN
    if (!srv) fail(HANN_ERR_INVALID_ARG, "server handle is null");
    Server &s = srv->s;
    std::lock_guard<std::mutex> lk(s.mu);
    out->completed_queries = s.completed_queries;
    out->completed_requests = s.completed_requests;
    out->batches = s.batches;
    out->mutations = s.mutations;
    out->queue_depth = (int64_t)s.queue.size();
    out->in_flight = s.in_flight;
    out->budget_in_use = s.budget_in_use;
    out->budget_peak = s.budget_peak;
    out->held_peak = s.held_peak;
    out->scratch_over_bound = s.scratch_over_bound;
    HANN_API_END
}

// Open-loop replay: request r is submitted at t0 + at_us[r] with queries
// [q_off[r], q_off[r] + q_cnt[r]) of `queries`; kind[r] 0 search, 1 bulk
// add of rows of `add_vecs` with ids add_ids (consuming q_cnt[r] of them in
// order), 2 bulk delete of ids del_ids (q_cnt[r] of them, in order).
// Writes each request's latency (ms, submit to completion) and the results
// of search requests ([q][kmax] at the request's query offset).
int hann_server_replay(hann_server *srv, size_t nreq, const double *at_us, const int32_t *kind,
                       const int64_t *q_off, const int32_t *q_cnt, const float *queries, const int32_t *k,
                       const int32_t *nprobe, const int64_t *add_ids, const float *add_vecs,
                       const int64_t *del_ids, int32_t kmax_out, int64_t *out_ids, double *out_dist,
                       double *latency_ms, double *wall_s) {
    HANN_API_BEGIN
    if (!srv) fail(HANN_ERR_INVALID_ARG, "server handle is null");
    Server &s = srv->s;
    const int64_t d = s.ix->dim;
    std::vector<uint64_t> tickets(nreq);
    std::thread poller;
    std::vector<int> status(nreq, HANN_OK);
    int64_t add_pos = 0, del_pos = 0;
    const auto t0 = Clock::now();
    // results are collected by a poller thread in submission order
    std::mutex tm;
    std::condition_variable tcv;
    size_t submitted = 0;
    poller = std::thread([&] {
        std::vector<int64_t> ids;
        std::vector<double> dist;
        std::vector<int32_t> cnt;
        for (size_t r = 0; r < nreq; ++r) {
            {
                std::unique_lock<std::mutex> lk(tm);
                tcv.wait(lk, [&] { return submitted > r; });
            }
            const int64_t nq = kind[r] == 0 ? q_cnt[r] : 0;
            int kr = 1;
            for (int64_t i = 0; i < nq; ++i) kr = std::max(kr, k[q_off[r] + i]);
            ids.assign((size_t)std::max<int64_t>(nq, 1) * kr, -1);
            dist.assign(ids.size(), INFINITY);
            cnt.assign((size_t)std::max<int64_t>(nq, 1), 0);
            double lat = 0;
            status[r] = hann_server_poll(srv, tickets[r], 1, ids.data(), dist.data(), cnt.data(), &lat);
            if (latency_ms) latency_ms[r] = lat;
            if (kind[r] == 0 && out_ids)
                for (int64_t i = 0; i < nq; ++i)
                    for (int j = 0; j < kmax_out; ++j) {
                        const bool ok = j < kr && j < cnt[i];
                        out_ids[(q_off[r] + i) * kmax_out + j] = ok ? ids[(size_t)i * kr + j] : -1;
                        if (out_dist) out_dist[(q_off[r] + i) * kmax_out + j] = ok ? dist[(size_t)i * kr + j] : INFINITY;
                    }
        }
    });
    for (size_t r = 0; r < nreq; ++r) {
        const auto due = t0 + std::chrono::nanoseconds((int64_t)(at_us[r] * 1000.0));
        std::this_thread::sleep_until(due);
        int rc;
        if (kind[r] == 0) {
            rc = hann_server_submit(srv, queries + q_off[r] * d, (size_t)q_cnt[r], k + q_off[r],
                                    nprobe ? nprobe + q_off[r] : nullptr, &tickets[r]);
        } else if (kind[r] == 1) {
            rc = hann_server_submit_bulk_add(srv, add_ids + add_pos, add_vecs + add_pos * d, (size_t)q_cnt[r],
                                             &tickets[r]);
            add_pos += q_cnt[r];
        } else {
            rc = hann_server_submit_bulk_delete(srv, del_ids + del_pos, (size_t)q_cnt[r], &tickets[r]);
            del_pos += q_cnt[r];
        }
        if (rc != HANN_OK) {
            {
                std::lock_guard<std::mutex> lk(tm);
                submitted = nreq;
            }
            tcv.notify_all();
            poller.join();
            fail(rc, hann_last_error());
        }
        {
            std::lock_guard<std::mutex> lk(tm);
            submitted = r + 1;
        }
        tcv.notify_all();
    }
    poller.join();
    if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    for (size_t r = 0; r < nreq; ++r)
        if (status[r] != HANN_OK) fail(status[r], strf("request %zu failed", r));
    HANN_API_END
}

}  // extern "C"
