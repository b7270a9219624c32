// rcb_engine.h -- context and entry points of the RCB engine (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "../../include/coupe.h"
#include "../../include/coupe_hip.h"
#include "rcb_common.h"

namespace coupe_hip {

struct Rotation;   // rcb_common.h

constexpr size_t kHostScratch = 65536;   // pinned readback area of a context

// A HIP failure inside the engine; entry points map it onto coupe's codes.
struct HipFailure {
    hipError_t err;
};

// A region of device memory that DevBuf::ensure carves buffers from instead
// of allocating, while it is installed on the calling thread (ArenaScope).
struct Arena {
    char *base = nullptr;
    size_t size = 0;
    size_t used = 0;
};

// Installs an arena for the lifetime of the scope (calling thread only).
struct ArenaScope {
    Arena *prev;
    explicit ArenaScope(Arena *a);
    ~ArenaScope();
};

// A grow-only device allocation (or, under an ArenaScope, a carve of the
// arena: `borrowed` buffers are never freed).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    void ensure(size_t need);   // throws HipFailure on allocation failure
    void release();
    void forget();              // drops a borrowed buffer without freeing anything
};

// Opt-in per-launch timing with HIP events on the launch stream.
struct Profiler {
    bool enabled = false;
    struct Rec {
        int kind;
        int ev;          // index of the start event; stop is ev + 1
        double bytes;    // algorithmic bytes of this launch
    };
    std::vector<hipEvent_t> events;
    std::vector<Rec> recs;
    coupe_hip_kernel_time out[COUPE_HIP_K_COUNT]{};
    void reset();
    int begin(hipStream_t st, int kind, double bytes);
    void end(hipStream_t st, int rec);
    void collect();      // after the stream is synchronised
    void release();
};

// A repartition's starting point: the resident working copy of a mesh (SoA
// f32 coordinates of its points in input order), which replaces the
// prologue's AoS -> SoA conversion; the weights are the call's d_weights.
struct ResidentSource {
    const float *x[3];
};

// What a context learns from a call about the next call of the same shape
// (the same instance table and dimension): the weight mode, and per level
// the decisions the split search needed and the parts still searching after
// the first decision.
struct Hints {
    uint64_t shape = 0;                 // hash of the instance table and the dimension (0: none)
    int weight_hint = -1;               // weight mode (0 u16, 1 i32, 2 i64, 3 f64)
    int passes_hint = 0;                // most decisions one level's search needed
    std::vector<uint8_t> level_passes;  // decisions each level's search needed
    std::vector<uint32_t> level_open;   // parts of each level still searching after the first decision
    uint64_t graph_key = 0;             // graph key of its last call (a key is captured when it repeats)
    int moves_coherent = -1;            // its moved tiles went to few destinations (-1: not known)
};

constexpr size_t kHintShapes = 16;      // shapes a context remembers

// What a call queued and must check once the stream has run it: the
// readback of its counters (split-search passes per level, errors) and of
// its prologue flags (the weight mode the data needs).
struct PendingCall {
    size_t levels = 0;
    int spec = 0;                       // search passes queued per level (0: host-checked)
    std::vector<int> depth;             // queued passes of each level
    uint64_t shape = 0;
    double ws_bytes = 0.0;              // bytes of one stored weight
    int mode = -1;                      // weight mode the call ran in
    int wkind = 0;                      // caller's weight type
};

// A captured call: the graph of every launch and copy of one call of a
// shape, replayed while the context's buffers, the call's arrays and the
// shape's hints are those it was captured with.
struct GraphEntry {
    uint64_t key = 0;                   // shape, arrays, hints and switches
    uint64_t layout = 0;                // the context's buffers when captured
    hipGraphExec_t exec = nullptr;
    PendingCall pending;
    coupe_hip_stats stats{};            // the counters queuing the call set
};

constexpr size_t kGraphs = 8;           // graphs a context keeps

struct Context {
    int device = 0;
    std::mutex lock;                 // one call at a time per context
    DevBuf cx[3][3];                 // SoA f32 coordinates: input order, then two sides
    DevBuf w[2];                     // stored weights, two sides
    DevBuf w16;                      // the prologue's 16-bit copy of the weights
    DevBuf parts[3], tiles[3];       // part / tile tables (two levels per move uses three)
    DevBuf bcnt, bw, bmin;           // per-part histogram bins
    DevBuf xcnt, xw, xmin;           // per-XCD copies of the bins (levels of few parts)
    DevBuf nchild, ntiles;           // scan inputs / outputs (children, tiles per part)
    DevBuf bsum;                     // scan block sums
    DevBuf nodes, leaf;              // split tree of all levels, part ids of the leaves
    DevBuf heap_split;               // split values by heap index (shallow trees)
    DevBuf instances;                // instance table of the current call
    DevBuf tiles0;                   // level-0 tiles (prologue, first level, descent)
    DevBuf below, tile_left;         // per-tile left counts (fused / counted) and their scan
    DevBuf tile_bins2, tile4;        // two levels per move: per-(tile, side) bins, 4 grandchild counts
    DevBuf pleft[2], pbelow2[2];     // per tile: left count, per side the count below the children's candidate
    DevBuf fwd_cnt, fwd_w, fwd_mn;   // children's first pass forwarded by the move, per (part, side)
    int search_passes = 3;           // histogram passes queued per level without a host check
    bool pairs = true;               // move the points once per two levels
    bool forward = true;             // the move forwards the next children's first pass
    bool weights16 = true;           // integer weights in [0, 65535] stored in 16 bits
    bool full_first = true;          // children's first pass with every candidate where forwarding mostly fails
    uint64_t scan_direct = 2048;     // scans of more blocks scan their block totals first
    bool graphs = false;             // replay a call of a known shape as one captured graph
    uint32_t copy_parts = 256;       // levels of at most this many parts flush into per-XCD copies
    uint32_t search_chunk = 2;       // consecutive tiles a workgroup takes at a time in the search passes
    double move_nt_bytes = 2.5e9;    // moves and id descents of more bytes than this go non-temporal
    int max_bins = kMaxBins;         // bins per part of the later search passes (16..256, a power of two)
    int wg_cap = 0;                  // workgroups per CU of the persistent kernels (0: as many as fit)
    // tuned switches per batch class (tune_key), chosen by coupe_hip_autotune
    std::vector<std::pair<uint64_t, coupe_hip_tuning>> tuned;
    bool tuning_now = false;         // the autotuner runs the call: the table is not applied
    bool capturing = false;          // the running call is being captured
    PendingCall pending;             // what the running call queued
    std::vector<GraphEntry> graph_cache;
    hipStream_t gstream = nullptr;   // where graphs are captured and launched
    hipEvent_t gevent = nullptr;     // orders gstream after the caller's stream
    uint64_t layout() ;              // hash of every buffer's address and size
    void drop_graphs();
    int weight_hint = -1;            // weight mode of the last call, for a shape not seen before
    std::vector<Hints> hints;        // per shape, the most recent last
    const Hints *hints_for(uint64_t shape) const {
        for (const Hints &h : hints)
            if (h.shape == shape) return &h;
        return nullptr;
    }
    Hints &remember(uint64_t shape) {   // the shape's entry, made the most recent
        for (size_t i = 0; i < hints.size(); i++)
            if (hints[i].shape == shape) {
                Hints h = std::move(hints[i]);
                hints.erase(hints.begin() + (long)i);
                hints.push_back(std::move(h));
                return hints.back();
            }
        if (hints.size() >= kHintShapes) hints.erase(hints.begin());
        hints.emplace_back();
        hints.back().shape = shape;
        return hints.back();
    }
    DevBuf loads;                    // per-part loads (imbalance)
    DevBuf tile_bins;                // per-tile bin counts of the histogram passes
    DevBuf gparts[2], gaxis, gchunk, gcount, gctr;   // Grid::rcb: subgrids, axis sums, chunk offsets
    DevBuf gsat, gcol, grow, gheap, gscan;   // Grid::rcb summed-area path: table, carries, heap of splits, chunk sums
    bool grid_clean = false;         // Grid::rcb axis sums are all zero
    DevBuf rib_moments;              // Rib: block partials, centroid, inertia
    coupe_hip_rib_info rib{};        // Rib: rotation of the last call
    const Rotation *rot = nullptr;   // change of basis for the running call
    DevBuf counters;
    DevBuf ired;                     // per-XCD copies of the prologue's instance reductions
    DevBuf in_points, in_weights, out_ids;   // staging for the host-memory ABI
    static constexpr int kStageBufs = 3;
    void *stage_buf[kStageBufs] = {};        // pinned chunks of the host-memory ABI's copies
    hipEvent_t stage_ev[kStageBufs] = {};    // each chunk's last copy
    hipStream_t stage_st = nullptr;          // the host-memory ABI's stream
    void *host_scratch = nullptr;    // pinned, for small readbacks
    bool bins_clean = false;
    int cu_count = 0;                // compute units of the device
    std::vector<std::pair<const void *, int>> occupancy;   // resident workgroups per CU, per kernel
    double swept_bytes = 0.0;        // HIST algorithmic bytes of the last call
    coupe_hip_stats stats{};
    Profiler prof;
    bool borrowed_host = false;      // host_scratch is not the context's own pinned allocation
    const float *slot0[3] = {nullptr, nullptr, nullptr};   // resident working copy of the running call, if any
    const float *coords0(int k) const { return slot0[0] ? slot0[k] : (const float *)cx[0][k].ptr; }
    // the hints of the context's most recent call (a resident mesh keeps its own)
    void save_hints(Hints &h) const {
        if (!hints.empty()) h = hints.back();
    }
    void load_hints(const Hints &h) {
        if (!h.shape) return;
        remember(h.shape) = h;
        weight_hint = h.weight_hint;
    }
    uint64_t device_bytes() const;
    void release_all();
    void forget_all();               // every buffer dropped without freeing (arena contexts)
    template <typename F>
    void for_each_buf(F &&f);
};

enum coupe_err map_hip_error(hipError_t e);

// COUPE_HIP_PAIRS=0: move the points at every level instead of every other.
inline bool pairs_from_env() {
    const char *v = getenv("COUPE_HIP_PAIRS");
    return !(v && v[0] == '0');
}

// COUPE_HIP_FORWARD=0: the move does not forward the children's first pass
// of the level it fills (every such pass reads the points).
inline bool forward_from_env() {
    const char *v = getenv("COUPE_HIP_FORWARD");
    return !(v && v[0] == '0');
}

// COUPE_HIP_WEIGHTS16=0: integer weights are stored in 32 bits even when
// they all fit in 16.
inline bool weights16_from_env() {
    const char *v = getenv("COUPE_HIP_WEIGHTS16");
    return !(v && v[0] == '0');
}

// COUPE_HIP_FULL_FIRST=0: the children's first pass always takes the bins
// of their first candidate, even where the context's last call of the same
// shape saw most of their parents still searching after the first decision.
inline bool full_first_from_env() {
    const char *v = getenv("COUPE_HIP_FULL_FIRST");
    return !(v && v[0] == '0');
}

// COUPE_HIP_BIN_COPIES=k: the search passes of levels of at most k parts
// flush into one copy of the bins per XCD (default 256; 0: never).
inline uint32_t copy_parts_from_env() {
    const char *v = getenv("COUPE_HIP_BIN_COPIES");
    if (!v || !*v) return 256;
    const long long k = atoll(v);
    return k < 0 ? 0u : (uint32_t)std::min<long long>(k, 1 << 20);
}

// COUPE_HIP_MOVE_NT_BYTES=b: moves of more than b bytes (read and written)
// store non-temporally (default 2.5e9) when the shape's last call moved its
// tiles to few destinations.  Measured per move: cached stores faster on
// jittered grids in grid order from 0.47 to 1.9 GB moved (by 6 to 21%),
// non-temporal ones at 3.8 GB (by 5%); on points in random order (Gaussian
// mixtures, shuffled grids, the batch) cached stores faster at every size
// measured, up to 7.8 GB (by 11 to 21%).
inline double move_nt_bytes_from_env() {
    const char *v = getenv("COUPE_HIP_MOVE_NT_BYTES");
    if (!v || !*v) return 2.5e9;
    const double b = atof(v);
    return b < 0 ? 0.0 : b;
}

// COUPE_HIP_SEARCH_CHUNK=k: the later search passes hand k consecutive
// tiles at a time to each workgroup (default 2, at most 64).
inline uint32_t search_chunk_from_env() {
    const char *v = getenv("COUPE_HIP_SEARCH_CHUNK");
    const int k = v && *v ? atoi(v) : 2;
    return (uint32_t)(k < 1 ? 1 : (k > 64 ? 64 : k));
}

// COUPE_HIP_GRAPHS=1: calls of a shape seen before replay a captured graph
// (default 0: kernel by kernel, unless the autotuner chose graphs for the
// call's class).
inline bool graphs_from_env() {
    const char *v = getenv("COUPE_HIP_GRAPHS");
    return v && v[0] == '1';
}

// COUPE_HIP_BINS=b: bins per part of the later search passes (16..256,
// rounded up to a power of two; default 256).
inline int bins_from_env() {
    const char *v = getenv("COUPE_HIP_BINS");
    int b = v && *v ? atoi(v) : 256;
    int p = 16;
    while (p < 256 && p < b) p *= 2;
    return p;
}

// COUPE_HIP_WG_PER_CU=k: at most k workgroups per CU for the persistent
// kernels (default 0: as many as fit).
inline int wg_cap_from_env() {
    const char *v = getenv("COUPE_HIP_WG_PER_CU");
    const int k = v && *v ? atoi(v) : 0;
    return k < 0 ? 0 : (k > 64 ? 64 : k);
}

// COUPE_HIP_SCAN_DIRECT=b: a scan of more than b blocks scans the block
// totals instead of re-summing them per block (default 2048, at least 1).
inline uint64_t scan_direct_from_env() {
    const char *v = getenv("COUPE_HIP_SCAN_DIRECT");
    if (!v || !*v) return 2048;
    const long long k = atoll(v);
    return k < 1 ? 1 : (uint64_t)k;
}

// COUPE_HIP_SEARCH_PASSES=k: histogram passes queued per level before the
// host looks (default 3; 0 = the host checks after every pass).  A context
// queues fewer, never under 2, when its last call needed fewer.
inline int search_passes_from_env() {
    const char *v = getenv("COUPE_HIP_SEARCH_PASSES");
    if (!v || !*v) return 3;
    const int k = atoi(v);
    return k < 0 ? 0 : (k > 16 ? 16 : k);
}

// A context's switches from the environment (the COUPE_HIP_* variables above).
inline void configure_from_env(Context &c) {
    c.search_passes = search_passes_from_env();
    c.pairs = pairs_from_env();
    c.forward = forward_from_env();
    c.weights16 = weights16_from_env();
    c.full_first = full_first_from_env();
    c.scan_direct = scan_direct_from_env();
    c.graphs = graphs_from_env();
    c.copy_parts = copy_parts_from_env();
    c.search_chunk = search_chunk_from_env();
    c.move_nt_bytes = move_nt_bytes_from_env();
    c.max_bins = bins_from_env();
    c.wg_cap = wg_cap_from_env();
}

// The context's switches as a tuning record, and back.
coupe_hip_tuning current_tuning(const Context &cx);
void apply_tuning(Context &cx, const coupe_hip_tuning &t);

// Autotuner of a batch of both dimensions (see coupe_hip_autotune).
enum coupe_err autotune(Context &cx, hipStream_t st, uint64_t *d_out, const double *d_points, int wkind,
                        const void *d_weights, const uint64_t *counts, const size_t *dims, const size_t *levels,
                        const double *tols, size_t n_inst, uint32_t trials, coupe_hip_tuning *chosen,
                        double *best_ms);

// RCB of a batch of independent instances packed in input order (see
// coupe_hip_rcb_batch).  counts / levels / tols: one entry per instance.
// With `rot`, every point is first mapped by the rotation (Rib); the call
// then adds to the profile the caller started instead of starting one.
enum coupe_err rcb_batch(Context &cx, hipStream_t st, uint64_t *d_out, size_t dim, const double *d_points,
                         int wkind, const void *d_weights, const uint64_t *counts, const size_t *levels,
                         const double *tols, size_t n_inst, const Rotation *rot = nullptr,
                         const ResidentSource *src = nullptr, const uint64_t *src_idx = nullptr,
                         const uint64_t *src_pt = nullptr);

// A batch of instances of dimension 2 and 3 in one call: dims[q] per
// instance, the caller's arrays packed in the given order (instance q's
// points are counts[q] * dims[q] doubles).  The instances of each dimension
// run as one batch, both on `st`.
enum coupe_err rcb_batch_mixed(Context &cx, hipStream_t st, uint64_t *d_out, const double *d_points, int wkind,
                               const void *d_weights, const uint64_t *counts, const size_t *dims, const size_t *levels,
                               const double *tols, size_t n_inst);

// Exclusive scan of n_dev (read on the device, at most n_ub; n_ub itself when
// n_dev is null) uint64 values in place; `total` (device, may be null)
// receives their sum.
void scan_u64(Context &cx, hipStream_t st, uint64_t *data, const uint32_t *n_dev, uint64_t n_ub, uint64_t *total);

// Grid::rcb on a device-resident weight array (see coupe_hip_grid_rcb).
enum coupe_err grid_rcb(Context &cx, hipStream_t st, size_t dim, const uint64_t *size, int wkind, const void *w,
                        size_t iter_count, uint32_t threads, uint64_t *out);

// coupe::Rib on device-resident inputs (see coupe_hip_rib_device).
enum coupe_err rib_device(Context &cx, hipStream_t st, uint64_t *d_out, uint64_t n, size_t dim,
                          const double *d_points, int wkind, const void *d_weights, size_t iter_count, double tol);

// Element barycentres (see coupe_hip_barycentres).
enum coupe_err barycentres(Context &cx, hipStream_t st, const double *coords, uint64_t n_nodes, size_t dim,
                           const int64_t *nodes, uint64_t n_el, size_t k, double *out);

// Per-part loads and imbalance of a partition (coupe/src/imbalance.rs).
enum coupe_err parts_load(Context &cx, hipStream_t st, const uint64_t *ids, uint64_t n, int wkind, const void *w,
                          uint64_t num_parts, void *host_loads);
enum coupe_err imbalance(Context &cx, hipStream_t st, const uint64_t *ids, uint64_t n, int wkind, const void *w,
                         uint64_t num_parts, double *out);

// The working copy of a resident mesh, points [first, first + count): f32
// SoA coordinates x[k] and weights mw (of kind wkind) from f64 AoS points and
// weights given for that range (device pointers; either may be null).
// Queued on `st`; the caller checks hipGetLastError and synchronises.
void store_working_copy(hipStream_t st, size_t dim, int wkind, const double *pts, const void *w, uint64_t first,
                        uint64_t count, float *const x[3], void *mw);

// RCB on device-resident inputs (see coupe_hip_rcb_device).
enum coupe_err rcb_device(Context &cx, hipStream_t st, uint64_t *d_out, uint64_t n, size_t dim,
                          const double *d_points, int wkind, const void *d_weights, size_t iter_count,
                          double tol);

}  // namespace coupe_hip
