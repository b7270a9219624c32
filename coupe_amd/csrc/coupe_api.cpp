// coupe_api.cpp -- the C ABI of include/coupe.h and include/coupe_hip.h.
//
// coupe_rcb keeps coupe's host-memory contract (coupe-ffi/src/lib.rs:272-304):
// the data sets are read on the host, staged to HBM, partitioned on the
// current HIP device, and the ids copied back into the caller's array.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "rcb_engine.h"

using coupe_hip::Context;

// Data sets (coupe-ffi/src/data.rs:11-174): borrowed array, repeated value,
// or per-index callback.
struct coupe_data {
    enum Kind { Array, Constant, Fn } kind;
    uintptr_t len;
    enum coupe_type type;
    const void *ptr;   // array or value
    const void *context;
    const void *(*i_th)(const void *, uintptr_t);
};

extern "C" {

const char *coupe_strerror(enum coupe_err err) {
    // Same strings as coupe-ffi/src/lib.rs:69-110.
    switch (err) {
        case COUPE_ERR_OK: return "success";
        case COUPE_ERR_ALLOC: return "allocation failed";
        case COUPE_ERR_CRASH: return "coupe encountered a bug and crashed";
        case COUPE_ERR_BAD_DIMENSION: return "this algorithm does not support the given mesh dimension";
        case COUPE_ERR_BAD_TYPE: return "this algorithm does not support the given type";
        case COUPE_ERR_BIPART_ONLY: return "this algorithm does not support k-way partitioning";
        case COUPE_ERR_LEN_MISMATCH: return "input iters (e.g. weights and points) don't have the same length";
        case COUPE_ERR_NOT_FOUND: return "no partition has been found for the given constraints";
        case COUPE_ERR_NEG_VALUES: return "this algorithm does not support negative values";
    }
    return "unknown error";
}

void coupe_data_free(coupe_data *data) { delete data; }

coupe_data *coupe_data_array(uintptr_t len, enum coupe_type type, const void *data) {
    return new (std::nothrow) coupe_data{coupe_data::Array, len, type, data, nullptr, nullptr};
}

coupe_data *coupe_data_constant(uintptr_t len, enum coupe_type type, const void *value) {
    return new (std::nothrow) coupe_data{coupe_data::Constant, len, type, value, nullptr, nullptr};
}

coupe_data *coupe_data_fn(const void *context, uintptr_t len, enum coupe_type type,
                          const void *(*i_th)(const void *, uintptr_t)) {
    return new (std::nothrow) coupe_data{coupe_data::Fn, len, type, nullptr, context, i_th};
}

}  // extern "C"

namespace {

size_t type_size(enum coupe_type t) { return t == COUPE_INT ? sizeof(int) : 8; }

// Host side of the host-memory entry: the data sets are produced (copied,
// repeated or called for, on several threads, as coupe's rayon workers read
// them, coupe-ffi/src/data.rs:231-256) into pinned chunks whose DMA to HBM
// overlaps the production of the next chunk; the ids come back the same way.
constexpr size_t kStageChunk = 32ull << 20;   // bytes per pinned chunk
constexpr size_t kStageDirect = 4ull << 20;   // below this, one plain copy

unsigned host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return h ? std::min(h, 8u) : 4u;
}

// f(lo, hi) over [0, n), split across up to `nt` threads (the caller's included).
template <typename F>
void parallel_for(size_t n, unsigned nt, F &&f) {
    if (nt <= 1 || n < (1u << 16)) {
        f((size_t)0, n);
        return;
    }
    const size_t step = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned k = 1; k < nt; k++) {
        const size_t lo = k * step, hi = std::min(n, lo + step);
        if (lo < hi) th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    f((size_t)0, std::min(n, step));
    for (auto &t : th) t.join();
}

// Elements [first, first + count) of a data set (`elem` bytes each) into out.
void produce(const coupe_data *d, size_t elem, uintptr_t first, uintptr_t count, unsigned char *out, unsigned nt) {
    parallel_for(count, nt, [&](size_t lo, size_t hi) {
        if (d->kind == coupe_data::Array) {
            memcpy(out + lo * elem, (const unsigned char *)d->ptr + (first + lo) * elem, (hi - lo) * elem);
        } else if (d->kind == coupe_data::Constant) {
            for (size_t i = lo; i < hi; i++) memcpy(out + i * elem, d->ptr, elem);
        } else {
            for (size_t i = lo; i < hi; i++) memcpy(out + i * elem, d->i_th(d->context, first + i), elem);
        }
    });
}

// The context's pinned chunks and their events (made on first use, kept).
void stage_init(Context &cx) {
    if (cx.stage_buf[0]) return;
    for (int b = 0; b < Context::kStageBufs; b++) {
        hipError_t e = hipHostMalloc(&cx.stage_buf[b], kStageChunk, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&cx.stage_ev[b], hipEventDisableTiming);
        if (e != hipSuccess) throw coupe_hip::HipFailure{e};
    }
}

// A data set of n elements into device memory at dst, on st.
void stage_in(Context &cx, hipStream_t st, const coupe_data *d, size_t elem, uintptr_t n, void *dst, unsigned nt) {
    const size_t bytes = n * elem;
    if (bytes <= kStageDirect) {   // small: one copy (from the caller's array, or produced)
        std::vector<unsigned char> tmp;
        const void *src = d->ptr;
        if (d->kind != coupe_data::Array) {
            tmp.resize(bytes);
            produce(d, elem, 0, n, tmp.data(), 1);
            src = tmp.data();
        }
        const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) (void)hipStreamSynchronize(st);
        if (e != hipSuccess) throw coupe_hip::HipFailure{e};
        return;
    }
    stage_init(cx);
    const uintptr_t per = kStageChunk / elem;
    for (uintptr_t c = 0, first = 0; first < n; c++, first += per) {
        const int b = (int)(c % Context::kStageBufs);
        const uintptr_t count = std::min<uintptr_t>(per, n - first);
        // the chunk's previous copy to the device has run
        hipError_t e = hipEventSynchronize(cx.stage_ev[b]);
        if (e != hipSuccess) throw coupe_hip::HipFailure{e};
        produce(d, elem, first, count, (unsigned char *)cx.stage_buf[b], nt);
        e = hipMemcpyAsync((unsigned char *)dst + first * elem, cx.stage_buf[b], count * elem, hipMemcpyHostToDevice,
                           st);
        if (e == hipSuccess) e = hipEventRecord(cx.stage_ev[b], st);
        if (e != hipSuccess) throw coupe_hip::HipFailure{e};
    }
}

// n ids from device memory at src into the caller's array, on st.
void stage_out(Context &cx, hipStream_t st, const uint64_t *src, uintptr_t n, uintptr_t *out, unsigned nt) {
    const size_t bytes = n * sizeof(uint64_t);
    if (bytes <= kStageDirect) {
        hipError_t e = hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) throw coupe_hip::HipFailure{e};
        return;
    }
    stage_init(cx);
    const uintptr_t per = kStageChunk / sizeof(uint64_t);
    const uintptr_t chunks = (n + per - 1) / per;
    auto copy_out = [&](uintptr_t c) {   // chunk c, once its copy from the device has run
        const int b = (int)(c % Context::kStageBufs);
        const hipError_t e = hipEventSynchronize(cx.stage_ev[b]);
        if (e != hipSuccess) throw coupe_hip::HipFailure{e};
        const uintptr_t first = c * per, count = std::min<uintptr_t>(per, n - first);
        parallel_for(count * sizeof(uint64_t), nt, [&](size_t lo, size_t hi) {
            memcpy((unsigned char *)(out + first) + lo, (const unsigned char *)cx.stage_buf[b] + lo, hi - lo);
        });
    };
    for (uintptr_t c = 0; c < chunks; c++) {
        const int b = (int)(c % Context::kStageBufs);
        const uintptr_t first = c * per, count = std::min<uintptr_t>(per, n - first);
        if (c >= (uintptr_t)Context::kStageBufs) copy_out(c - Context::kStageBufs);   // frees chunk b
        hipError_t e = hipMemcpyAsync(cx.stage_buf[b], src + first, count * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                      st);
        if (e == hipSuccess) e = hipEventRecord(cx.stage_ev[b], st);
        if (e != hipSuccess) throw coupe_hip::HipFailure{e};
    }
    for (uintptr_t c = chunks > (uintptr_t)Context::kStageBufs ? chunks - Context::kStageBufs : 0; c < chunks; c++)
        copy_out(c);
}

std::mutex g_default_lock;
Context *g_default[64];

Context *default_context() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> g(g_default_lock);
    if (!g_default[dev]) {
        Context *c = new (std::nothrow) Context();
        if (!c) return nullptr;
        c->device = dev;
        coupe_hip::configure_from_env(*c);
        g_default[dev] = c;
    }
    return g_default[dev];
}

}  // namespace

namespace {

// Host-memory contract shared by coupe_rcb and coupe_rib
// (coupe-ffi/src/lib.rs:272-304, :331-364).
enum coupe_err host_partition(uintptr_t *partition, uintptr_t dimension, const coupe_data *points,
                              const coupe_data *weights, uintptr_t iter_count, double tolerance, bool rib) {
    const uintptr_t n = points->len;
    if (n != weights->len) return COUPE_ERR_LEN_MISMATCH;   // lib.rs:285-288
    if (dimension != 2 && dimension != 3) return COUPE_ERR_BAD_DIMENSION;   // lib.rs:297-301
    if (n == 0) return COUPE_ERR_OK;
    if (weights->type != COUPE_INT && weights->type != COUPE_INT64 && weights->type != COUPE_DOUBLE)
        return COUPE_ERR_BAD_TYPE;
    Context *cx = default_context();
    if (!cx) return COUPE_ERR_CRASH;
    std::lock_guard<std::mutex> g(cx->lock);
    try {
        const size_t wsz = type_size(weights->type);
        const size_t pbytes = n * dimension * sizeof(double), wbytes = n * wsz, obytes = n * sizeof(uint64_t);
        cx->in_points.ensure(pbytes);
        cx->in_weights.ensure(wbytes);
        cx->out_ids.ensure(obytes);
        if (!cx->stage_st && hipStreamCreateWithFlags(&cx->stage_st, hipStreamNonBlocking) != hipSuccess)
            return COUPE_ERR_CRASH;
        hipStream_t st = cx->stage_st;
        const unsigned nt = host_threads();
        stage_in(*cx, st, points, dimension * sizeof(double), n, cx->in_points.ptr, nt);
        stage_in(*cx, st, weights, wsz, n, cx->in_weights.ptr, nt);
        auto run = rib ? coupe_hip::rib_device : coupe_hip::rcb_device;
        enum coupe_err r = run(*cx, st, (uint64_t *)cx->out_ids.ptr, n, dimension,
                               (const double *)cx->in_points.ptr, weights->type, cx->in_weights.ptr, iter_count,
                               tolerance);
        if (r != COUPE_ERR_OK) return r;
        stage_out(*cx, st, (const uint64_t *)cx->out_ids.ptr, n, partition, nt);
    } catch (const coupe_hip::HipFailure &f) {
        (void)hipGetLastError();
        return coupe_hip::map_hip_error(f.err);
    } catch (const std::bad_alloc &) {
        return COUPE_ERR_ALLOC;
    } catch (...) {
        return COUPE_ERR_ALLOC;   // DevBuf::ensure failures
    }
    return COUPE_ERR_OK;
}

}  // namespace

extern "C" enum coupe_err coupe_rcb(uintptr_t *partition, uintptr_t dimension, const coupe_data *points,
                                    const coupe_data *weights, uintptr_t iter_count, double tolerance) {
    return host_partition(partition, dimension, points, weights, iter_count, tolerance, false);
}

extern "C" enum coupe_err coupe_rib(uintptr_t *partition, uintptr_t dimension, const coupe_data *points,
                                    const coupe_data *weights, uintptr_t iter_count, double tolerance) {
    return host_partition(partition, dimension, points, weights, iter_count, tolerance, true);
}

extern "C" coupe_hip_ctx *coupe_hip_ctx_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return nullptr;
    Context *c = new (std::nothrow) Context();
    if (!c) return nullptr;
    c->device = device;
    coupe_hip::configure_from_env(*c);
    return reinterpret_cast<coupe_hip_ctx *>(c);
}

extern "C" void coupe_hip_ctx_destroy(coupe_hip_ctx *ctx) {
    if (!ctx) return;
    Context *c = reinterpret_cast<Context *>(ctx);
    (void)hipSetDevice(c->device);
    c->release_all();
    delete c;
}

extern "C" enum coupe_err coupe_hip_rcb_device(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition, uintptr_t n,
                                               uintptr_t dimension, const double *d_points,
                                               enum coupe_type weight_type, const void *d_weights,
                                               uintptr_t iter_count, double tolerance) {
    if (!ctx) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    return coupe_hip::rcb_device(*c, (hipStream_t)stream, d_partition, n, dimension, d_points, weight_type,
                                 d_weights, iter_count, tolerance);
}

extern "C" enum coupe_err coupe_hip_rib_device(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition, uintptr_t n,
                                               uintptr_t dimension, const double *d_points,
                                               enum coupe_type weight_type, const void *d_weights,
                                               uintptr_t iter_count, double tolerance) {
    if (!ctx) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    return coupe_hip::rib_device(*c, (hipStream_t)stream, d_partition, n, dimension, d_points, weight_type,
                                 d_weights, iter_count, tolerance);
}

extern "C" enum coupe_err coupe_hip_last_rib(coupe_hip_ctx *ctx, struct coupe_hip_rib_info *info) {
    if (!ctx || !info) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    *info = c->rib;
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_rcb_batch(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition,
                                              uintptr_t dimension, const double *d_points,
                                              enum coupe_type weight_type, const void *d_weights,
                                              const struct coupe_hip_instance *instances, uintptr_t n_instances) {
    if (!ctx) return COUPE_ERR_CRASH;
    if (n_instances > 0 && !instances) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    try {
        std::vector<uint64_t> counts(n_instances);
        std::vector<size_t> levels(n_instances);
        std::vector<double> tols(n_instances);
        for (uintptr_t q = 0; q < n_instances; q++) {
            counts[q] = instances[q].n;
            levels[q] = instances[q].iter_count;
            tols[q] = instances[q].tolerance;
        }
        return coupe_hip::rcb_batch(*c, (hipStream_t)stream, d_partition, dimension, d_points, weight_type,
                                    d_weights, counts.data(), levels.data(), tols.data(), n_instances);
    } catch (const std::bad_alloc &) {
        return COUPE_ERR_ALLOC;
    }
}

extern "C" enum coupe_err coupe_hip_rcb_batch_mixed(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition,
                                                    const double *d_points, enum coupe_type weight_type,
                                                    const void *d_weights,
                                                    const struct coupe_hip_instance_nd *instances,
                                                    uintptr_t n_instances) {
    if (!ctx) return COUPE_ERR_CRASH;
    if (n_instances > 0 && !instances) return COUPE_ERR_CRASH;
    if (weight_type != COUPE_INT && weight_type != COUPE_INT64 && weight_type != COUPE_DOUBLE)
        return COUPE_ERR_BAD_TYPE;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    try {
        std::vector<uint64_t> counts(n_instances);
        std::vector<size_t> dims(n_instances), levels(n_instances);
        std::vector<double> tols(n_instances);
        for (uintptr_t q = 0; q < n_instances; q++) {
            counts[q] = instances[q].n;
            dims[q] = instances[q].dimension;
            levels[q] = instances[q].iter_count;
            tols[q] = instances[q].tolerance;
        }
        return coupe_hip::rcb_batch_mixed(*c, (hipStream_t)stream, d_partition, d_points, weight_type, d_weights,
                                          counts.data(), dims.data(), levels.data(), tols.data(), n_instances);
    } catch (const std::bad_alloc &) {
        return COUPE_ERR_ALLOC;
    }
}

extern "C" enum coupe_err coupe_hip_autotune(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition,
                                             const double *d_points, enum coupe_type weight_type,
                                             const void *d_weights, const struct coupe_hip_instance_nd *instances,
                                             uintptr_t n_instances, uint32_t trials, struct coupe_hip_tuning *chosen,
                                             double *best_ms) {
    if (!ctx) return COUPE_ERR_CRASH;
    if (n_instances > 0 && !instances) return COUPE_ERR_CRASH;
    if (weight_type != COUPE_INT && weight_type != COUPE_INT64 && weight_type != COUPE_DOUBLE)
        return COUPE_ERR_BAD_TYPE;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    try {
        std::vector<uint64_t> counts(n_instances);
        std::vector<size_t> dims(n_instances), levels(n_instances);
        std::vector<double> tols(n_instances);
        for (uintptr_t q = 0; q < n_instances; q++) {
            counts[q] = instances[q].n;
            dims[q] = instances[q].dimension;
            levels[q] = instances[q].iter_count;
            tols[q] = instances[q].tolerance;
        }
        return coupe_hip::autotune(*c, (hipStream_t)stream, d_partition, d_points, weight_type, d_weights,
                                   counts.data(), dims.data(), levels.data(), tols.data(), n_instances, trials,
                                   chosen, best_ms);
    } catch (const std::bad_alloc &) {
        return COUPE_ERR_ALLOC;
    }
}

extern "C" enum coupe_err coupe_hip_last_stats(const coupe_hip_ctx *ctx, struct coupe_hip_stats *out) {
    if (!ctx || !out) return COUPE_ERR_CRASH;
    *out = reinterpret_cast<const Context *>(ctx)->stats;
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_set_profiling(coupe_hip_ctx *ctx, int enable) {
    if (!ctx) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    c->prof.enabled = enable != 0;
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_kernel_times(const coupe_hip_ctx *ctx, struct coupe_hip_kernel_time *out,
                                                 int count) {
    if (!ctx || !out) return COUPE_ERR_CRASH;
    const Context *c = reinterpret_cast<const Context *>(ctx);
    for (int k = 0; k < count && k < COUPE_HIP_K_COUNT; k++) out[k] = c->prof.out[k];
    return COUPE_ERR_OK;
}

extern "C" enum coupe_err coupe_hip_parts_load(coupe_hip_ctx *ctx, void *stream, const uint64_t *d_partition,
                                               uintptr_t n, enum coupe_type weight_type, const void *d_weights,
                                               uintptr_t num_parts, void *loads) {
    if (!ctx || (num_parts && !loads)) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    return coupe_hip::parts_load(*c, (hipStream_t)stream, d_partition, n, weight_type, d_weights, num_parts, loads);
}

extern "C" enum coupe_err coupe_hip_imbalance(coupe_hip_ctx *ctx, void *stream, const uint64_t *d_partition,
                                              uintptr_t n, enum coupe_type weight_type, const void *d_weights,
                                              uintptr_t num_parts, double *imbalance) {
    if (!ctx || !imbalance) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    return coupe_hip::imbalance(*c, (hipStream_t)stream, d_partition, n, weight_type, d_weights, num_parts,
                                imbalance);
}

extern "C" enum coupe_err coupe_hip_barycentres(coupe_hip_ctx *ctx, void *stream, const double *d_coords,
                                                uintptr_t n_nodes, uintptr_t dimension,
                                                const int64_t *d_element_nodes, uintptr_t n_elements,
                                                uintptr_t nodes_per_element, double *d_out) {
    if (!ctx) return COUPE_ERR_CRASH;
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    return coupe_hip::barycentres(*c, (hipStream_t)stream, d_coords, n_nodes, dimension, d_element_nodes,
                                  n_elements, nodes_per_element, d_out);
}

extern "C" enum coupe_err coupe_hip_grid_rcb(coupe_hip_ctx *ctx, void *stream, uintptr_t dimension,
                                             const uintptr_t *size, enum coupe_type weight_type,
                                             const void *d_weights, uintptr_t iter_count, uint32_t threads,
                                             uint64_t *d_partition) {
    if (!ctx || !size) return COUPE_ERR_CRASH;
    if (dimension != 2 && dimension != 3) return COUPE_ERR_BAD_DIMENSION;
    if (!d_weights || !d_partition) return COUPE_ERR_CRASH;   // every grid has a cell
    Context *c = reinterpret_cast<Context *>(ctx);
    std::lock_guard<std::mutex> g(c->lock);
    const uint64_t sz[3] = {size[0], size[1], dimension == 3 ? size[2] : 1};
    return coupe_hip::grid_rcb(*c, (hipStream_t)stream, dimension, sz, weight_type, d_weights, iter_count, threads,
                               d_partition);
}
