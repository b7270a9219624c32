// serve.cpp -- request server with continuous batching and resident meshes
// (include/coupe_hip.h).
//
// Every worker owns a context (its device work buffers) and a stream.  A free
// worker takes the oldest queued request and, when it can share a batch, the
// queued requests of the same dimension and weight type that fit its batch
// limits; it packs their points and weights into its own device buffers, runs
// them as one batch (rcb_batch) and writes each request's ids to its output.
// A request on a resident mesh runs alone on the mesh's working copy (no
// input copy), starting from the mesh's own hints.  Workers run concurrently,
// so batches overlap on the device while new requests keep arriving.
//
// Device memory is admitted against the budget from live figures: the bytes
// every worker holds now (its context and packing buffers) plus the resident
// meshes.  A batch that does not fit waits; workers that are idle or done give
// their buffers back while some request waits for memory.
#include <hip/hip_runtime.h>

#include <algorithm>
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

constexpr uint64_t kDenseLimit = 24;   // deeper requests run alone (node-record descent, one instance)

size_t weight_size(coupe_type t) { return t == COUPE_INT ? sizeof(int) : 8; }

// A mesh resident in HBM: the working copy the engine starts from (f32 SoA
// coordinates, the caller's weights), its last ids and the grouping of its
// points by part (computed on request).
struct Mesh {
    uint64_t handle = 0;
    uintptr_t n = 0, dim = 0;
    coupe_type wt = COUPE_DOUBLE;
    DevBuf x[3], w, ids;
    DevBuf perm, offsets;             // points grouped by part (coupe_hip_mesh_parts)
    bool grouped = false;             // perm / offsets match ids
    bool partitioned = false;         // ids hold a partition
    uint64_t num_parts = 0;
    coupe_hip::Hints hints;           // what its last partition taught the engine
    uint32_t pending = 0;             // queued or running requests
    bool exclusive = false;           // an update or a grouping holds it (its requests wait)
    bool busy = false;                // one of its requests runs (they run one at a time)
    uint32_t holds = 0;               // threads waiting for exclusive use
    bool dying = false;               // destroyed: no longer in the server's map
    uint64_t bytes() const {
        return x[0].bytes + x[1].bytes + x[2].bytes + w.bytes + ids.bytes + perm.bytes + offsets.bytes;
    }
};

struct Req {
    coupe_hip_request r;
    Mesh *mesh = nullptr;
    uint64_t ticket = 0;
    hipEvent_t ready = nullptr;       // producer stream's work up to submit
    Clock::time_point t_submit, t_start, t_done;
    bool started = false, done = false;
    coupe_err err = COUPE_ERR_OK;
    uint32_t batch_requests = 0;
    uint64_t batch_points = 0;
};

struct Worker {
    Context cx;
    hipStream_t stream = nullptr;
    DevBuf pts, w, ids;
    std::thread th;
    uint64_t held = 0;                // device bytes this worker holds (accounted)
    uint64_t bytes() { return cx.device_bytes() + pts.bytes + w.bytes + ids.bytes; }
    void trim() {
        cx.release_all();
        pts.release();
        w.release();
        ids.release();
    }
};

// Upper bound of the device bytes a worker holds while running `n` points of
// dimension `dim` split into at most `parts` parts per level (all instances),
// packed (input copies and ids) or not (a resident mesh).
uint64_t batch_bytes(uint64_t n, uint64_t dim, uint64_t wsize, uint64_t parts, uint64_t levels, bool packed) {
    const uint64_t tiles = parts + n / 2048 + 1;
    uint64_t b = 0;
    b += n * (3 * 4 * dim);                     // coordinate slots
    b += n * (8 + 8 + 2 + 4);                   // weight slots, 16-bit copy, 32-bit copy
    b += parts * (3 * 160 + 256 * 16 + 64);     // part tables, bins, children tables
    b += tiles * (3 * 16 + 3 * 256 * 2 + 96);   // tile tables, per-tile bins and counts
    b += levels < 25 ? (2ull << levels) * 4 : parts * levels * 16;   // split tree
    b += 1ull << 20;                            // counters, scans, instance tables
    if (packed) b += n * (8 * dim + wsize + 8);
    return b + b / 8;
}

uint64_t parts_bound(uint64_t n, uint64_t levels) {
    uint64_t p = 1;
    for (uint64_t l = 1; l < levels && p < n; l++) p = std::min<uint64_t>(2 * p, n);
    return p;
}

}  // namespace

struct coupe_hip_server {
    coupe_hip_server_config cfg{};
    std::mutex m;
    std::condition_variable cv_work, cv_done, cv_idle;
    std::deque<Req *> queue;
    std::unordered_map<uint64_t, Req *> live;       // submitted, not yet taken by a wait
    std::unordered_map<uint64_t, Mesh *> meshes;
    uint64_t next_ticket = 1, next_mesh = 1;
    bool stop = false;
    uint32_t waiters = 0;                            // threads inside coupe_hip_server_wait
    bool memory_wanted = false;                      // a batch waits for device memory
    uint64_t mesh_bytes = 0, peak = 0, completed = 0, cancelled = 0;
    uint32_t running = 0;
    std::vector<Worker *> workers;
    std::vector<hipEvent_t> free_events;
    hipStream_t copy_stream = nullptr;               // mesh creation, updates, groupings
    std::mutex copy_lock;                            // one user of copy_stream at a time

    uint64_t held() const {
        uint64_t h = mesh_bytes;
        for (const Worker *W : workers) h += W->held;
        return h;
    }
    void note_peak() { peak = std::max(peak, held()); }
};

namespace {

uint64_t n_of(const Req *q) { return q->mesh ? q->mesh->n : q->r.n; }
uint64_t dim_of(const Req *q) { return q->mesh ? q->mesh->dim : q->r.dimension; }
uint64_t wsize_of(const Req *q) { return weight_size(q->mesh ? q->mesh->wt : q->r.weight_type); }

hipEvent_t take_event(coupe_hip_server *S) {   // under S->m
    if (!S->free_events.empty()) {
        hipEvent_t e = S->free_events.back();
        S->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Runs a packed batch, or one request on its resident mesh.
void run_batch(Worker *W, std::vector<Req *> &batch) {
    const Clock::time_point t0 = Clock::now();
    Req *first = batch[0];
    Mesh *M = first->mesh;
    const uint64_t dim = dim_of(first);
    const coupe_type wt = M ? M->wt : first->r.weight_type;
    const size_t ws = weight_size(wt);
    uint64_t total = 0;
    for (Req *q : batch) {
        q->t_start = t0;
        total += n_of(q);
    }
    coupe_err err = COUPE_ERR_OK;
    try {
        for (Req *q : batch)
            if (q->ready && hipStreamWaitEvent(W->stream, q->ready, 0) != hipSuccess) throw HipFailure{hipErrorUnknown};
        if (M) {
            // the mesh's working copy is the engine's starting point: no copy
            const size_t levels = first->r.iter_count;
            const uint64_t n = M->n;
            const double tol = first->r.tolerance;
            coupe_hip::ResidentSource src{{(const float *)M->x[0].ptr, (const float *)M->x[1].ptr,
                                           (const float *)M->x[2].ptr}};
            W->cx.load_hints(M->hints);
            err = coupe_hip::rcb_batch(W->cx, W->stream, (uint64_t *)M->ids.ptr, dim, nullptr, wt, M->w.ptr, &n,
                                       &levels, &tol, 1, nullptr, &src);
            W->cx.save_hints(M->hints);
            if (err == COUPE_ERR_OK) {
                M->partitioned = true;
                M->grouped = false;
                const coupe_hip_request &r = first->r;
                if (n && r.partition) {
                    hipError_t e = hipMemcpyAsync(r.partition, M->ids.ptr, n * sizeof(uint64_t),
                                                  r.device_output ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                                  W->stream);
                    if (e != hipSuccess) throw HipFailure{e};
                }
                hipError_t e = hipStreamSynchronize(W->stream);
                if (e != hipSuccess) throw HipFailure{e};
            }
        } else {
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
            err = coupe_hip::rcb_batch(W->cx, W->stream, (uint64_t *)W->ids.ptr, dim, (const double *)W->pts.ptr,
                                       wt, W->w.ptr, counts.data(), levels.data(), tols.data(), counts.size());
            if (err == COUPE_ERR_OK) {
                off = 0;
                for (Req *q : batch) {
                    const coupe_hip_request &r = q->r;
                    if (r.n) {
                        hipError_t e = hipMemcpyAsync(r.partition, (uint64_t *)W->ids.ptr + off,
                                                      r.n * sizeof(uint64_t),
                                                      r.device_output ? hipMemcpyDeviceToDevice
                                                                      : hipMemcpyDeviceToHost,
                                                      W->stream);
                        if (e != hipSuccess) throw HipFailure{e};
                    }
                    off += r.n;
                }
                hipError_t e = hipStreamSynchronize(W->stream);
                if (e != hipSuccess) throw HipFailure{e};
            }
        }
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(W->stream);
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

// A request that runs alone: on a resident mesh, or deeper than the dense
// split heap (a batch of several instances needs iter_count <= 24).
bool alone(const Req *q) { return q->mesh || q->r.iter_count > kDenseLimit; }

bool runnable(const Req *q) { return !(q->mesh && (q->mesh->exclusive || q->mesh->busy)); }

// Under S->m: picks the next batch for W (the oldest request that can run
// and the ones it can share a batch with), or returns false.
bool pick_batch(coupe_hip_server *S, Worker *W, std::vector<Req *> &batch) {
    auto head = std::find_if(S->queue.begin(), S->queue.end(), runnable);
    if (head == S->queue.end()) return false;
    Req *first = *head;
    const uint64_t dim = dim_of(first);
    const uint64_t wsize = wsize_of(first);
    const uint64_t room = S->cfg.memory_budget - std::min(S->cfg.memory_budget, S->mesh_bytes);
    uint64_t pts = n_of(first), parts = parts_bound(pts, first->r.iter_count);
    uint64_t levels = first->r.iter_count;
    std::vector<std::deque<Req *>::iterator> take;
    if (!alone(first)) {
        for (auto it = head + 1; it != S->queue.end() && take.size() + 1 < S->cfg.max_batch_requests; ++it) {
            Req *q = *it;
            if (alone(q) || q->r.dimension != dim || q->r.weight_type != first->r.weight_type ||
                pts + q->r.n > S->cfg.max_batch_points)
                continue;
            const uint64_t p2 = parts + parts_bound(q->r.n, q->r.iter_count);
            const uint64_t l2 = std::max<uint64_t>(levels, q->r.iter_count);
            if (batch_bytes(pts + q->r.n, dim, wsize, p2, l2, true) > room) continue;
            take.push_back(it);
            pts += q->r.n;
            parts = p2;
            levels = l2;
        }
    }
    // admission: what W will hold for this batch (at least what it holds now,
    // its buffers only grow), against what the others and the meshes hold now
    const uint64_t need = batch_bytes(pts, dim, wsize, parts, levels, !first->mesh);
    const uint64_t others = S->held() - W->held;
    if (others + std::max(need, W->held) > S->cfg.memory_budget) {
        S->memory_wanted = true;
        return false;
    }
    batch.push_back(first);
    for (auto it : take) batch.push_back(*it);
    for (Req *q : batch) q->started = true;
    S->queue.erase(std::remove_if(S->queue.begin(), S->queue.end(), [](const Req *q) { return q->started; }),
                   S->queue.end());
    if (first->mesh) first->mesh->busy = true;
    W->held = std::max(need, W->held);   // accounted until the batch's real figure is known
    S->note_peak();
    return true;
}

void worker_main(coupe_hip_server *S, Worker *W) {
    (void)hipSetDevice(S->cfg.device);
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        std::vector<Req *> batch;
        while (!pick_batch(S, W, batch)) {
            if (S->stop && S->queue.empty()) return;
            if (S->memory_wanted && W->held > 0) {
                // give the buffers back so a waiting batch can be admitted
                lk.unlock();
                W->trim();
                lk.lock();
                W->held = 0;
                S->cv_work.notify_all();
                continue;
            }
            S->cv_work.wait(lk);
        }
        S->memory_wanted = false;
        S->running += (uint32_t)batch.size();
        lk.unlock();
        run_batch(W, batch);
        if (batch[0]->err == COUPE_ERR_ALLOC) W->trim();   // nothing left behind by a failed growth
        const uint64_t real = W->bytes();
        lk.lock();
        W->held = real;
        S->note_peak();
        S->running -= (uint32_t)batch.size();
        for (Req *q : batch) {
            if (q->mesh) {
                q->mesh->pending--;
                q->mesh->busy = false;
            }
            if (q->ready) {
                S->free_events.push_back(q->ready);
                q->ready = nullptr;
            }
            S->completed++;
            q->done = true;   // from here on the request belongs to its waiter
        }
        S->cv_done.notify_all();
        S->cv_idle.notify_all();
        if (S->memory_wanted) S->cv_work.notify_all();
    }
}

Mesh *find_mesh(coupe_hip_server *S, uint64_t handle) {   // under S->m
    auto it = S->meshes.find(handle);
    return it == S->meshes.end() ? nullptr : it->second;
}

// Exclusive use of a mesh (an update, a grouping): waits under lk until none
// of its requests is queued or running and no other exclusive use is on,
// then takes it; false when the mesh was destroyed meanwhile.  Requests
// submitted meanwhile stay queued until release_mesh.
bool acquire_mesh(coupe_hip_server *S, std::unique_lock<std::mutex> &lk, Mesh *M) {
    M->holds++;
    S->cv_idle.wait(lk, [&] { return M->dying || (M->pending == 0 && !M->exclusive); });
    M->holds--;
    if (M->dying) {
        S->cv_idle.notify_all();
        return false;
    }
    M->exclusive = true;
    return true;
}

void release_mesh(coupe_hip_server *S, Mesh *M) {   // under S->m
    M->exclusive = false;
    S->cv_idle.notify_all();
    S->cv_work.notify_all();
}

// Copies points [first, first + count) and/or their weights (caller order,
// f64 AoS points) into the working copy of M on the server's copy stream,
// after the producer stream's work when `stream` is given.
coupe_err store_mesh(coupe_hip_server *S, Mesh *M, uintptr_t first, uintptr_t count, const double *points,
                     const void *weights, int device_input, void *stream) {
    if (count == 0 || (!points && !weights)) return COUPE_ERR_OK;
    std::lock_guard<std::mutex> g(S->copy_lock);
    hipStream_t cs = S->copy_stream;
    const size_t ws = weight_size(M->wt);
    DevBuf stage_p, stage_w;
    coupe_err err = COUPE_ERR_OK;
    try {
        if (device_input && stream) {
            hipEvent_t e = nullptr;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) throw HipFailure{hipErrorUnknown};
            hipError_t r = hipEventRecord(e, (hipStream_t)stream);
            if (r == hipSuccess) r = hipStreamWaitEvent(cs, e, 0);
            (void)hipEventDestroy(e);
            if (r != hipSuccess) throw HipFailure{r};
        }
        const double *dp = points;
        const void *dw = weights;
        if (!device_input) {   // host data: staged in device memory first
            if (points) {
                stage_p.ensure(count * M->dim * sizeof(double));
                hipError_t r = hipMemcpyAsync(stage_p.ptr, points, count * M->dim * sizeof(double),
                                              hipMemcpyHostToDevice, cs);
                if (r != hipSuccess) throw HipFailure{r};
                dp = (const double *)stage_p.ptr;
            }
            if (weights) {
                stage_w.ensure(count * ws);
                hipError_t r = hipMemcpyAsync(stage_w.ptr, weights, count * ws, hipMemcpyHostToDevice, cs);
                if (r != hipSuccess) throw HipFailure{r};
                dw = stage_w.ptr;
            }
        }
        float *x[3] = {(float *)M->x[0].ptr, (float *)M->x[1].ptr, (float *)M->x[2].ptr};
        coupe_hip::store_working_copy(cs, M->dim, M->wt, dp, dw, first, count, x, M->w.ptr);
        hipError_t r = hipGetLastError();
        if (r == hipSuccess) r = hipStreamSynchronize(cs);
        if (r != hipSuccess) throw HipFailure{r};
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(cs);
        err = coupe_hip::map_hip_error(f.err);
    }
    stage_p.release();
    stage_w.release();
    if (err == COUPE_ERR_OK) M->grouped = false;
    return err;
}

void free_mesh(Mesh *M) {
    for (DevBuf &b : M->x) b.release();
    M->w.release();
    M->ids.release();
    M->perm.release();
    M->offsets.release();
    delete M;
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
    (void)hipSetDevice(config->device);
    if (hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        return nullptr;
    }
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
    for (Worker *W : S->workers)
        if (W->th.joinable()) W->th.join();
    {
        // every request is done: wake the waiters and let them leave
        std::unique_lock<std::mutex> lk(S->m);
        S->cv_done.notify_all();
        S->cv_idle.wait(lk, [&] { return S->waiters == 0; });
    }
    (void)hipSetDevice(S->cfg.device);
    for (Worker *W : S->workers) {
        W->trim();
        if (W->stream) (void)hipStreamDestroy(W->stream);
        delete W;
    }
    for (auto &kv : S->live) {
        if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
        delete kv.second;
    }
    for (auto &kv : S->meshes) free_mesh(kv.second);
    for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
    if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
    delete S;
}

extern "C" enum coupe_err coupe_hip_server_submit(coupe_hip_server *S, const struct coupe_hip_request *req,
                                                  uint64_t *ticket) {
    if (!S || !req || !ticket) return COUPE_ERR_CRASH;
    Req *q = new (std::nothrow) Req();
    if (!q) return COUPE_ERR_ALLOC;
    q->r = *req;
    q->t_submit = Clock::now();
    std::unique_lock<std::mutex> lk(S->m);
    if (S->stop) {
        delete q;
        return COUPE_ERR_CRASH;
    }
    if (req->mesh) {
        q->mesh = find_mesh(S, req->mesh);
        if (!q->mesh) {
            delete q;
            return COUPE_ERR_NOT_FOUND;
        }
        q->r.n = q->mesh->n;
        q->r.dimension = q->mesh->dim;
        q->r.weight_type = q->mesh->wt;
        q->r.device_input = 1;
    } else {
        if (req->dimension != 2 && req->dimension != 3) {
            delete q;
            return COUPE_ERR_BAD_DIMENSION;
        }
        if (req->weight_type != COUPE_INT && req->weight_type != COUPE_INT64 && req->weight_type != COUPE_DOUBLE) {
            delete q;
            return COUPE_ERR_BAD_TYPE;
        }
    }
    // a request that cannot fit alone next to the resident meshes
    const uint64_t need = batch_bytes(q->r.n, q->r.dimension, weight_size(q->r.weight_type),
                                      parts_bound(q->r.n, q->r.iter_count), q->r.iter_count, !q->mesh);
    if (q->r.n > S->cfg.max_batch_points || need + S->mesh_bytes > S->cfg.memory_budget) {
        delete q;
        return COUPE_ERR_ALLOC;
    }
    if (req->stream && q->r.device_input) {
        q->ready = take_event(S);
        if (!q->ready || hipEventRecord(q->ready, (hipStream_t)req->stream) != hipSuccess) {
            if (q->ready) S->free_events.push_back(q->ready);
            delete q;
            return COUPE_ERR_CRASH;
        }
    }
    if (q->mesh) q->mesh->pending++;
    q->ticket = S->next_ticket++;
    S->live[q->ticket] = q;
    S->queue.push_back(q);
    *ticket = q->ticket;
    lk.unlock();
    S->cv_work.notify_all();
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
        S->live.erase(it);   // this wait owns the request: another wait finds nothing
        S->waiters++;
        S->cv_done.wait(lk, [&] { return q->done; });
        S->waiters--;
        if (S->waiters == 0) S->cv_idle.notify_all();
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

extern "C" enum coupe_err coupe_hip_server_poll(coupe_hip_server *S, uint64_t ticket, int *done) {
    if (!S || !done) return COUPE_ERR_CRASH;
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) return COUPE_ERR_NOT_FOUND;
    *done = it->second->done ? 1 : 0;
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_server_cancel(coupe_hip_server *S, uint64_t ticket, int *cancelled) {
    if (!S || !cancelled) return COUPE_ERR_CRASH;
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) return COUPE_ERR_NOT_FOUND;
    Req *q = it->second;
    *cancelled = 0;
    if (q->started) return COUPE_ERR_OK;   // running or done: it completes and is waited for as usual
    auto qi = std::find(S->queue.begin(), S->queue.end(), q);
    if (qi != S->queue.end()) S->queue.erase(qi);
    S->live.erase(it);
    if (q->mesh) {
        q->mesh->pending--;
        S->cv_idle.notify_all();
    }
    if (q->ready) S->free_events.push_back(q->ready);
    S->cancelled++;
    delete q;
    *cancelled = 1;
    lk.unlock();
    S->cv_work.notify_all();   // the requests behind it may fit now
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_server_stats(coupe_hip_server *S, struct coupe_hip_server_stats *out) {
    if (!S || !out) return COUPE_ERR_CRASH;
    std::lock_guard<std::mutex> lk(S->m);
    out->memory_budget = S->cfg.memory_budget;
    out->device_bytes = S->held();
    out->mesh_bytes = S->mesh_bytes;
    out->peak_device_bytes = S->peak;
    out->queued = (uint32_t)S->queue.size();
    out->running = S->running;
    out->meshes = (uint32_t)S->meshes.size();
    out->completed = S->completed;
    out->cancelled = S->cancelled;
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_mesh_create(coupe_hip_server *S, const double *points, const void *weights,
                                                enum coupe_type weight_type, uintptr_t n, uintptr_t dimension,
                                                int device_input, void *stream, uint64_t *mesh) {
    if (!S || !mesh || (n && (!points || !weights))) return COUPE_ERR_CRASH;
    if (dimension != 2 && dimension != 3) return COUPE_ERR_BAD_DIMENSION;
    if (weight_type != COUPE_INT && weight_type != COUPE_INT64 && weight_type != COUPE_DOUBLE)
        return COUPE_ERR_BAD_TYPE;
    if (n > 0xffffffffull) return COUPE_ERR_ALLOC;
    Mesh *M = new (std::nothrow) Mesh();
    if (!M) return COUPE_ERR_ALLOC;
    M->n = n;
    M->dim = dimension;
    M->wt = weight_type;
    const uint64_t bytes = std::max<uint64_t>(n, 1) * (4 * dimension + weight_size(weight_type) + 8) + 4096;
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (S->held() + bytes > S->cfg.memory_budget) {
            delete M;
            return COUPE_ERR_ALLOC;
        }
        S->mesh_bytes += bytes;   // reserved while the arrays are made
    }
    coupe_err err = COUPE_ERR_OK;
    try {
        (void)hipSetDevice(S->cfg.device);
        for (uintptr_t k = 0; k < dimension; k++) M->x[k].ensure(std::max<uint64_t>(n, 1) * sizeof(float));
        M->w.ensure(std::max<uint64_t>(n, 1) * weight_size(weight_type));
        M->ids.ensure(std::max<uint64_t>(n, 1) * sizeof(uint64_t));
        err = store_mesh(S, M, 0, n, points, weights, device_input, stream);
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        err = coupe_hip::map_hip_error(f.err);
    }
    std::lock_guard<std::mutex> lk(S->m);
    S->mesh_bytes -= bytes;
    if (err != COUPE_ERR_OK) {
        free_mesh(M);
        return err;
    }
    M->handle = S->next_mesh++;
    S->mesh_bytes += M->bytes();
    S->meshes[M->handle] = M;
    S->note_peak();
    *mesh = M->handle;
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_mesh_update(coupe_hip_server *S, uint64_t mesh, uintptr_t first,
                                                uintptr_t count, const double *points, const void *weights,
                                                int device_input, void *stream) {
    if (!S) return COUPE_ERR_CRASH;
    std::unique_lock<std::mutex> lk(S->m);
    Mesh *M = find_mesh(S, mesh);
    if (!M) return COUPE_ERR_NOT_FOUND;
    if (first > M->n || count > M->n - first) return COUPE_ERR_LEN_MISMATCH;
    // after the mesh's queued and running requests, before later ones
    if (!acquire_mesh(S, lk, M)) return COUPE_ERR_NOT_FOUND;
    lk.unlock();
    (void)hipSetDevice(S->cfg.device);
    const coupe_err err = store_mesh(S, M, first, count, points, weights, device_input, stream);
    lk.lock();
    release_mesh(S, M);
    return err;
}

// Points grouped by part: a stable sort of the ids (mesh_parts.hip).
namespace coupe_hip {
enum coupe_err group_by_part(hipStream_t st, const uint64_t *ids, uint64_t n, uint32_t *perm, DevBuf &offsets,
                             uint64_t *num_parts, DevBuf &scratch);
}

extern "C" enum coupe_err coupe_hip_mesh_parts(coupe_hip_server *S, uint64_t mesh, uint32_t *perm,
                                               uint64_t *offsets, uintptr_t offsets_len, uint64_t *num_parts,
                                               int device_output) {
    if (!S || !num_parts) return COUPE_ERR_CRASH;
    std::unique_lock<std::mutex> lk(S->m);
    Mesh *M = find_mesh(S, mesh);
    if (!M) return COUPE_ERR_NOT_FOUND;
    if (!acquire_mesh(S, lk, M)) return COUPE_ERR_NOT_FOUND;
    if (!M->partitioned) {
        release_mesh(S, M);
        return COUPE_ERR_NOT_FOUND;
    }
    lk.unlock();
    (void)hipSetDevice(S->cfg.device);
    coupe_err err = COUPE_ERR_OK;
    {
        std::lock_guard<std::mutex> g(S->copy_lock);
        hipStream_t cs = S->copy_stream;
        try {
            if (!M->grouped) {
                M->perm.ensure(std::max<uint64_t>(M->n, 1) * sizeof(uint32_t));
                DevBuf scratch;
                err = coupe_hip::group_by_part(cs, (const uint64_t *)M->ids.ptr, M->n, (uint32_t *)M->perm.ptr,
                                               M->offsets, &M->num_parts, scratch);
                scratch.release();
                M->grouped = err == COUPE_ERR_OK;
            }
            *num_parts = M->num_parts;
            if (err == COUPE_ERR_OK && (perm || offsets)) {
                if (offsets && offsets_len < M->num_parts + 1) {
                    err = COUPE_ERR_LEN_MISMATCH;
                } else {
                    const hipMemcpyKind k = device_output ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
                    hipError_t r = hipSuccess;
                    if (perm && M->n) r = hipMemcpyAsync(perm, M->perm.ptr, M->n * sizeof(uint32_t), k, cs);
                    if (r == hipSuccess && offsets)
                        r = hipMemcpyAsync(offsets, M->offsets.ptr, (M->num_parts + 1) * sizeof(uint64_t), k, cs);
                    if (r == hipSuccess) r = hipStreamSynchronize(cs);
                    if (r != hipSuccess) throw HipFailure{r};
                }
            }
        } catch (const HipFailure &f) {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(cs);
            err = coupe_hip::map_hip_error(f.err);
        }
    }
    lk.lock();
    S->mesh_bytes = 0;
    for (auto &kv : S->meshes) S->mesh_bytes += kv.second->bytes();
    S->note_peak();
    release_mesh(S, M);
    return err;
}

extern "C" enum coupe_err coupe_hip_mesh_destroy(coupe_hip_server *S, uint64_t mesh) {
    if (!S) return COUPE_ERR_CRASH;
    std::unique_lock<std::mutex> lk(S->m);
    Mesh *M = find_mesh(S, mesh);
    if (!M) return COUPE_ERR_NOT_FOUND;
    // no new request or exclusive use finds it; the queued and running ones end first
    S->meshes.erase(mesh);
    M->dying = true;
    S->cv_idle.notify_all();
    S->cv_idle.wait(lk, [&] { return M->pending == 0 && !M->exclusive && M->holds == 0; });
    S->mesh_bytes -= std::min(S->mesh_bytes, M->bytes());
    lk.unlock();
    (void)hipSetDevice(S->cfg.device);
    free_mesh(M);
    S->cv_work.notify_all();
    return COUPE_ERR_OK;
}
