// serve.cpp -- request server with continuous batching (include/coupe_hip.h).
//
// Every worker owns a context (its device work buffers) and a stream.  A free
// worker takes the queued requests of one dimension and weight type, in
// arrival order, that fit its batch limits, packs their points and weights
// into its own device buffers, runs them as one batch (rcb_batch) and writes
// each request's ids to its output.  Workers run concurrently, so batches
// overlap on the device while new requests keep arriving.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rcb_engine.h"

using coupe_hip::Context;
using coupe_hip::DevBuf;
using coupe_hip::HipFailure;
using Clock = std::chrono::steady_clock;

namespace {

struct Req {
    coupe_hip_request r;
    uint64_t ticket = 0;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    coupe_err err = COUPE_ERR_OK;
    uint32_t batch_requests = 0;
    uint64_t batch_points = 0;
};

struct Worker {
    Context cx;
    hipStream_t stream = nullptr;
    DevBuf pts, w, ids;
    std::thread th;
};

// Device bytes one point of a batch needs: packed input (points, weights),
// ids, and the engine's working arrays (3 coordinate slots, 2 weight slots),
// plus margin for the tables.
uint64_t bytes_per_point(uint64_t dim) { return 8 * dim + 8 + 8 + 12 * dim + 16 + 8; }

size_t weight_size(coupe_type t) { return t == COUPE_INT ? sizeof(int) : 8; }

}  // namespace

struct coupe_hip_server {
    coupe_hip_server_config cfg{};
    uint64_t slot_budget = 0;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> queue;
    std::unordered_map<uint64_t, Req *> live;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::vector<Worker *> workers;

    uint64_t max_points(uint64_t dim) const {
        uint64_t by_budget = slot_budget / bytes_per_point(dim);
        return cfg.max_batch_points < by_budget ? cfg.max_batch_points : by_budget;
    }
};

namespace {

void run_batch(coupe_hip_server *S, Worker *W, std::vector<Req *> &batch) {
    const Clock::time_point t0 = Clock::now();
    const uint64_t dim = batch[0]->r.dimension;
    const coupe_type wt = batch[0]->r.weight_type;
    const size_t ws = weight_size(wt);
    uint64_t total = 0;
    for (Req *q : batch) {
        q->t_start = t0;
        total += q->r.n;
    }
    coupe_err err = COUPE_ERR_OK;
    try {
        W->pts.ensure(total * dim * sizeof(double));
        W->w.ensure(total * ws);
        W->ids.ensure(total * sizeof(uint64_t));
        std::vector<uint64_t> counts;
        std::vector<size_t> levels;
        std::vector<double> tols;
        uint64_t off = 0;
        for (Req *q : batch) {
            const coupe_hip_request &r = q->r;
            const hipMemcpyKind k = r.device_input ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
            if (r.n) {
                hipError_t e = hipMemcpyAsync((char *)W->pts.ptr + off * dim * sizeof(double), r.points,
                                              r.n * dim * sizeof(double), k, W->stream);
                if (e == hipSuccess)
                    e = hipMemcpyAsync((char *)W->w.ptr + off * ws, r.weights, r.n * ws, k, W->stream);
                if (e != hipSuccess) throw HipFailure{e};
            }
            counts.push_back(r.n);
            levels.push_back(r.iter_count);
            tols.push_back(r.tolerance);
            off += r.n;
        }
        err = coupe_hip::rcb_batch(W->cx, W->stream, (uint64_t *)W->ids.ptr, dim, (const double *)W->pts.ptr, wt,
                                   W->w.ptr, counts.data(), levels.data(), tols.data(), counts.size());
        if (err == COUPE_ERR_OK) {
            off = 0;
            for (Req *q : batch) {
                const coupe_hip_request &r = q->r;
                if (r.n) {
                    hipError_t e = hipMemcpyAsync(r.partition, (uint64_t *)W->ids.ptr + off, r.n * sizeof(uint64_t),
                                                  r.device_output ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                                  W->stream);
                    if (e != hipSuccess) throw HipFailure{e};
                }
                off += r.n;
            }
            hipError_t e = hipStreamSynchronize(W->stream);
            if (e != hipSuccess) throw HipFailure{e};
        }
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        err = coupe_hip::map_hip_error(f.err);
    } catch (const std::bad_alloc &) {
        err = COUPE_ERR_ALLOC;
    }
    const Clock::time_point t1 = Clock::now();
    for (Req *q : batch) {
        q->t_done = t1;
        q->err = err;
        q->batch_requests = (uint32_t)batch.size();
        q->batch_points = total;
    }
}

void worker_main(coupe_hip_server *S, Worker *W) {
    (void)hipSetDevice(S->cfg.device);
    for (;;) {
        std::vector<Req *> batch;
        {
            std::unique_lock<std::mutex> lk(S->m);
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->queue.empty()) return;   // stopping and drained
            Req *first = S->queue.front();
            S->queue.pop_front();
            batch.push_back(first);
            const uint64_t cap = S->max_points(first->r.dimension);
            uint64_t pts = first->r.n;
            for (auto it = S->queue.begin(); it != S->queue.end() && batch.size() < S->cfg.max_batch_requests;) {
                Req *q = *it;
                if (q->r.dimension == first->r.dimension && q->r.weight_type == first->r.weight_type &&
                    pts + q->r.n <= cap) {
                    batch.push_back(q);
                    pts += q->r.n;
                    it = S->queue.erase(it);
                } else {
                    ++it;
                }
            }
        }
        run_batch(S, W, batch);
        {
            std::lock_guard<std::mutex> lk(S->m);
            for (Req *q : batch) q->done = true;
        }
        S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" coupe_hip_server *coupe_hip_server_create(const struct coupe_hip_server_config *config) {
    if (!config) return nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || config->device < 0 || config->device >= count) return nullptr;
    coupe_hip_server *S = new (std::nothrow) coupe_hip_server();
    if (!S) return nullptr;
    S->cfg = *config;
    if (S->cfg.slots == 0) S->cfg.slots = 1;
    if (S->cfg.max_batch_requests == 0) S->cfg.max_batch_requests = 4096;
    if (S->cfg.max_batch_points == 0) S->cfg.max_batch_points = 1ull << 27;
    S->slot_budget = S->cfg.memory_budget / S->cfg.slots;
    (void)hipSetDevice(config->device);
    for (uint32_t i = 0; i < S->cfg.slots; i++) {
        Worker *W = new (std::nothrow) Worker();
        if (!W || hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking) != hipSuccess) {
            delete W;
            coupe_hip_server_destroy(S);
            return nullptr;
        }
        W->cx.device = config->device;
        coupe_hip::configure_from_env(W->cx);
        S->workers.push_back(W);
    }
    for (Worker *W : S->workers) W->th = std::thread(worker_main, S, W);
    return S;
}

extern "C" void coupe_hip_server_destroy(coupe_hip_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (Worker *W : S->workers) {
        if (W->th.joinable()) W->th.join();
        (void)hipSetDevice(S->cfg.device);
        W->pts.release();
        W->w.release();
        W->ids.release();
        W->cx.release_all();
        if (W->stream) (void)hipStreamDestroy(W->stream);
        delete W;
    }
    for (auto &kv : S->live) delete kv.second;
    delete S;
}

extern "C" enum coupe_err coupe_hip_server_submit(coupe_hip_server *S, const struct coupe_hip_request *req,
                                                  uint64_t *ticket) {
    if (!S || !req || !ticket) return COUPE_ERR_CRASH;
    if (req->dimension != 2 && req->dimension != 3) return COUPE_ERR_BAD_DIMENSION;
    if (req->weight_type != COUPE_INT && req->weight_type != COUPE_INT64 && req->weight_type != COUPE_DOUBLE)
        return COUPE_ERR_BAD_TYPE;
    if (req->n > S->max_points(req->dimension)) return COUPE_ERR_ALLOC;   // cannot fit alone
    Req *q = new (std::nothrow) Req();
    if (!q) return COUPE_ERR_ALLOC;
    q->r = *req;
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (S->stop) {
            delete q;
            return COUPE_ERR_CRASH;
        }
        q->ticket = S->next_ticket++;
        S->live[q->ticket] = q;
        S->queue.push_back(q);
        *ticket = q->ticket;
    }
    S->cv_work.notify_one();
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_server_wait(coupe_hip_server *S, uint64_t ticket,
                                                struct coupe_hip_request_stats *stats) {
    if (!S) return COUPE_ERR_CRASH;
    Req *q = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->m);
        auto it = S->live.find(ticket);
        if (it == S->live.end()) return COUPE_ERR_NOT_FOUND;
        q = it->second;
        S->cv_done.wait(lk, [&] { return q->done; });
        S->live.erase(it);
    }
    if (stats) {
        auto ms = [](Clock::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
        stats->queue_ms = ms(q->t_start - q->t_submit);
        stats->run_ms = ms(q->t_done - q->t_start);
        stats->total_ms = ms(q->t_done - q->t_submit);
        stats->batch_requests = q->batch_requests;
        stats->batch_points = q->batch_points;
    }
    const coupe_err err = q->err;
    delete q;
    return err;
}
