// rcb_engine.hip -- host side of the MI355X RCB engine: device buffers of a
// context, weight-mode selection, and the level loop that drives the kernels
// of rcb_kernels.h.  Entry point: coupe_hip_rcb_device (include/coupe_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "rcb_engine.h"
#include "rcb_kernels.h"
#include "rcb_pairs.h"

namespace coupe_hip {

// ---------------------------------------------------------------------------
// Errors: any HIP failure unwinds to the entry point, which maps it onto
// coupe's codes (out of memory -> COUPE_ERR_ALLOC, anything else -> CRASH).
// ---------------------------------------------------------------------------
static void check(hipError_t e) {
    if (e != hipSuccess) throw HipFailure{e};
}

enum coupe_err map_hip_error(hipError_t e) {
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? COUPE_ERR_ALLOC : COUPE_ERR_CRASH;
}

static thread_local Arena *t_arena = nullptr;
static thread_local bool t_capturing = false;   // a stream capture is running on this thread

ArenaScope::ArenaScope(Arena *a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

void DevBuf::ensure(size_t need) {
    if (need <= bytes) return;
    // a capture records no allocation: the caller gives it up and runs eagerly
    if (t_capturing) throw HipFailure{hipErrorStreamCaptureUnsupported};
    const size_t sz = need < 256 ? 256 : need;
    if (t_arena) {
        // carve: 256-byte aligned, never freed (the old carve stays unused)
        const size_t off = (t_arena->used + 255) & ~(size_t)255;
        if (off > t_arena->size || sz > t_arena->size - off) throw HipFailure{hipErrorOutOfMemory};
        ptr = t_arena->base + off;
        bytes = sz;
        borrowed = true;
        t_arena->used = off + sz;
        return;
    }
    if (ptr && !borrowed) check(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
    check(hipMalloc(&ptr, sz));
    bytes = sz;
}

void DevBuf::release() {
    if (ptr && !borrowed) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
}

void DevBuf::forget() {
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
}

void Profiler::reset() {
    recs.clear();
    for (auto &o : out) o = coupe_hip_kernel_time{};
}

int Profiler::begin(hipStream_t st, int kind, double bytes) {
    if (!enabled) return -1;
    int ev = (int)recs.size() * 2;
    while ((int)events.size() < ev + 2) {
        hipEvent_t e;
        check(hipEventCreate(&e));
        events.push_back(e);
    }
    check(hipEventRecord(events[ev], st));
    recs.push_back(Rec{kind, ev, bytes});
    return (int)recs.size() - 1;
}

void Profiler::end(hipStream_t st, int rec) {
    if (rec < 0) return;
    check(hipEventRecord(events[recs[rec].ev + 1], st));
}

void Profiler::collect() {
    for (auto &o : out) o = coupe_hip_kernel_time{};
    if (!enabled) return;
    for (const Rec &r : recs) {
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, events[r.ev], events[r.ev + 1]));
        out[r.kind].launches++;
        out[r.kind].total_ms += ms;
        out[r.kind].algo_bytes += r.bytes;
    }
}

void Profiler::release() {
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    events.clear();
    recs.clear();
}

// Trace ranges (roctx, seen by rocprofv3 --marker-trace): one per call, per
// bisection level (coupe's spans per split, recursive_bisection.rs:467-468,
// and per leaf pass, :590-595) and per kernel launch, named by kernel class.
static const char *const kKernelRange[COUPE_HIP_K_COUNT] = {"prologue", "hist", "decide", "partition",
                                                            "ids",      "other", "moments"};

struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
};

static const char *level_range(size_t level) {
    static const char *const names[] = {"level 0",  "level 1",  "level 2",  "level 3",  "level 4",  "level 5",
                                        "level 6",  "level 7",  "level 8",  "level 9",  "level 10", "level 11",
                                        "level 12", "level 13", "level 14", "level 15", "level 16", "level 17",
                                        "level 18", "level 19", "level 20", "level 21", "level 22", "level 23"};
    return level < sizeof(names) / sizeof(names[0]) ? names[level] : "level 24+";
}

// Runs `launch` between two profiling events when profiling is on.
template <typename F>
static void timed(Context &cx, hipStream_t st, int kind, double bytes, F &&launch) {
    TraceRange range(kKernelRange[kind]);
    int r = cx.prof.begin(st, kind, bytes);
    launch();
    check(hipGetLastError());
    cx.prof.end(st, r);
}

// Small device counters shared by the kernels of one call.

constexpr int kSpecSlots = 1024;   // (level, pass) slots of the speculative search
// The counters that many workgroups bump hold one copy per XCD (the kernels
// add into copy blockIdx mod kCopies); readers sum the copies.
struct Counters {
    uint32_t n_parts[3];
    uint32_t n_tiles[3];
    uint32_t searching[kCopies];                // parts still searching (host-checked search)
    uint64_t parts_total;
    uint64_t tiles_total;
    uint64_t min_id;
    // read back at the end of the call, from here
    unsigned long long swept[kCopies];          // points swept by standalone passes
    unsigned long long swept_fused[kCopies];    // points binned inside partitions
    unsigned long long first_read[kCopies];     // points read by the children's first passes
    unsigned long long move_runs[2 * kCopies];  // destinations of the moved tiles, then the moved tiles
    Prologue pro[kCopies];                      // the prologue's weight flags
    uint32_t forwarded[kCopies];                // parts whose children's first pass the move forwarded
    uint32_t unfinished;                        // a speculative search did not end in its passes
    uint32_t error;                             // the node-record descent found an inconsistent tree
    uint32_t spec[kSpecSlots * kCopies];        // parts still searching after pass p of level l (slot's copies)
    uint32_t end;                               // end of the call's readback
};

template <typename T>
static T sum_copies(const T *c) {
    T s = 0;
    for (int k = 0; k < kCopies; k++) s += c[k];
    return s;
}

// The call's counters, read back into host_scratch once the stream has run
// it: errors, the bytes its search passes read, and per level the decisions
// its search needed (speculative searches), which become the shape's hints.
// Returns true when a speculative search did not end in its queued passes.
static bool finish_levels(Context &cx) {
    const PendingCall &P = cx.pending;
    const size_t rb_off = offsetof(Counters, swept);
    const char *h_rb = (const char *)cx.host_scratch + 256;
    const Counters *h = (const Counters *)(h_rb - rb_off);   // only the copied fields are read
    if (h->error) throw HipFailure{hipErrorUnknown};
    cx.swept_bytes = (double)sum_copies(h->swept) * (4.0 + P.ws_bytes) +
                     (double)sum_copies(h->first_read) * (8.0 + P.ws_bytes);
    cx.stats.first_read = sum_copies(h->first_read);
    cx.stats.forwarded = sum_copies(h->forwarded);
    {   // moved tiles going to at most 2.5 destinations on average (of four)
        const unsigned long long r = sum_copies(h->move_runs), t = sum_copies(h->move_runs + kCopies);
        if (t) cx.remember(P.shape).moves_coherent = 2 * r <= 5 * t ? 1 : 0;
    }
    cx.bins_clean = true;
    if (P.spec == 0) return false;
    // passes per level: the first decision plus every pass whose gate was open
    std::vector<uint8_t> used(P.levels, 0);
    std::vector<uint32_t> open1(P.levels, 0);
    for (size_t l = 0; l < P.levels; l++) {
        uint32_t passes = 1;
        for (int p = 1; p <= P.depth[l]; p++)
            if (sum_copies(&h->spec[(l * (P.spec + 1) + p - 1) * kCopies])) passes++;
        cx.stats.passes += passes;
        cx.stats.max_passes = std::max(cx.stats.max_passes, passes);
        used[l] = (uint8_t)passes;
        open1[l] = sum_copies(&h->spec[l * (P.spec + 1) * kCopies]);
    }
    if (h->unfinished == 0) {   // (an unfinished run keeps the last call's figures for its rerun)
        Hints &e = cx.remember(P.shape);
        e.level_passes = used;
        e.level_open = open1;
    }
    return h->unfinished != 0;
}

static inline unsigned grid_for(uint64_t items, unsigned per_block) {
    uint64_t g = (items + per_block - 1) / per_block;
    return (unsigned)(g < 1 ? 1 : g);
}

// One level of the exclusive scan below: block totals at bsum, and the
// space past them for the levels under it.
static void scan_level(hipStream_t st, uint64_t *data, const uint32_t *n_dev, uint64_t n_ub, uint64_t *total,
                       uint64_t *bsum, uint64_t direct) {
    const uint64_t nblocks = std::max<uint64_t>((n_ub + kScanTile - 1) / kScanTile, 1);
    if (nblocks == 1) {
        k_scan_local<<<1, kBlock, 0, st>>>(data, data, bsum, n_dev, n_ub, total ? total : bsum + 1);
        return;
    }
    k_scan_local<<<(unsigned)nblocks, kBlock, 0, st>>>(data, data, bsum, n_dev, n_ub, nullptr);
    if (nblocks <= direct) {
        k_scan_add<<<(unsigned)nblocks, kBlock, 0, st>>>(data, bsum, n_dev, n_ub, total);
        return;
    }
    // many blocks: scan their totals (zero past the elements), then add
    scan_level(st, bsum, nullptr, nblocks, total, bsum + nblocks + 1, direct);
    k_scan_add_scanned<<<(unsigned)nblocks, kBlock, 0, st>>>(data, bsum, n_dev, n_ub);
}

// Exclusive scan of `n_dev` (device) uint64 values, in place, upper bound n_ub.
void scan_u64(Context &cx, hipStream_t st, uint64_t *data, const uint32_t *n_dev, uint64_t n_ub,
                     uint64_t *total) {
    // block totals of every level of the recursion, each followed by a spare entry
    uint64_t need = 2, m = std::max<uint64_t>(n_ub, 1);
    while (m > 1) {
        m = (m + kScanTile - 1) / kScanTile;
        need += m + 1;
    }
    cx.bsum.ensure(need * sizeof(uint64_t));
    timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
        scan_level(st, data, n_dev, n_ub, total, (uint64_t *)cx.bsum.ptr, cx.scan_direct);
    });
}

// Search passes to queue at a level whose search took `decisions` decisions
// (its first one and one per pass) in the context's last call of the same
// shape: none when the first decision ended every search there (the split
// at the first candidate, the usual case on even data), else the passes it
// took plus one to spare, at most spec.
static int queued_passes(int spec, uint32_t decisions) {
    if (decisions <= 1) return 0;
    return std::min(spec, (int)decisions);
}

// Bins per part of the standalone LDS passes, bounded by the bin buffers.
static int choose_bins(uint64_t parts_ub, int max_bins) {
    const uint64_t cap = 1ull << 23;   // entries of each bin array
    int nb = max_bins;
    while (nb > kFirstBins && parts_ub * nb > cap) nb >>= 1;
    return nb;
}

// Resident workgroups of a persistent kernel: CUs x blocks per CU.
template <typename K>
static unsigned persistent_grid(Context &cx, K kernel, uint64_t work_items) {
    // the occupancy query is a runtime call: asked once per kernel and
    // context, so that queuing a launch stays cheaper than running it
    const void *key = reinterpret_cast<const void *>(kernel);
    int per_cu = 0;
    for (const auto &e : cx.occupancy)
        if (e.first == key) { per_cu = e.second; break; }
    if (per_cu == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu < 1)
            per_cu = 2;
        cx.occupancy.emplace_back(key, per_cu);
    }
    if (cx.wg_cap > 0 && per_cu > cx.wg_cap) per_cu = cx.wg_cap;
    uint64_t g = (uint64_t)cx.cu_count * (uint64_t)per_cu;
    if (g > work_items) g = work_items;
    return (unsigned)(g < 1 ? 1 : g);
}

// Host view of a batch: the instance table and the bounds derived from it.
struct BatchPlan {
    std::vector<Instance> inst;   // begin, count, levels, tile_begin, tol, heap_base
    uint64_t n = 0;               // points of all instances
    uint32_t max_levels = 0;
    uint64_t tiles0 = 0;          // level-0 tiles
    std::vector<uint64_t> parts_ub;   // parts at each level, upper bound
    uint64_t parts_cap = 1;
    uint64_t tiles_cap = 1;
    uint64_t heap_total = 0;      // dense split heap entries
    bool dense = true;            // every instance has iter_count <= kDenseLevels
    uint64_t shape = 0;           // hash of the instance table and the dimension (hints)
};

// The instances in the order given, packed one after the other in the
// working arrays.  src / src_pt (optional): where each instance's points
// start in the caller's weights and ids, and in the caller's points (doubles);
// by default the caller packs them the same way.
static BatchPlan plan_batch(const uint64_t *counts, const size_t *levels, const double *tols, size_t n_inst,
                            size_t dim, const uint64_t *src = nullptr, const uint64_t *src_pt = nullptr) {
    BatchPlan b;
    b.inst.resize(n_inst);
    uint64_t begin = 0, tile = 0;
    uint64_t h = 0xcbf29ce484222325ull ^ (uint64_t)dim;   // FNV-1a over the table
    auto mix = [&](uint64_t v) {
        for (int k = 0; k < 8; k++) {
            h ^= (v >> (8 * k)) & 0xff;
            h *= 0x100000001b3ull;
        }
    };
    for (size_t q = 0; q < n_inst; q++) {
        Instance &I = b.inst[q];
        memset(&I, 0, sizeof I);
        I.begin = begin;
        I.count = counts[q];
        I.levels = (uint32_t)std::min<size_t>(levels[q], 0xffffffffu);
        I.tile_begin = (uint32_t)tile;
        I.tol = tols[q];
        I.src = src ? src[q] : begin;
        I.src_pt = src_pt ? src_pt[q] : begin * dim;
        mix(I.count);
        mix(I.levels);
        mix(__builtin_bit_cast(uint64_t, I.tol));
        begin += counts[q];
        tile += (counts[q] + kTile - 1) / kTile;
        if (counts[q] > 0) b.max_levels = std::max<uint32_t>(b.max_levels, I.levels);
        if (I.levels > (uint32_t)kDenseLevels) b.dense = false;
    }
    b.n = begin;
    b.tiles0 = tile;
    // parts at level l: one root per instance at level 0, then at most
    // min(2^l, n_i) for the instances still splitting
    b.parts_ub.assign((size_t)b.max_levels + 1, 0);
    b.parts_ub[0] = n_inst;
    for (size_t q = 0; q < n_inst; q++) {
        const uint64_t c = counts[q];
        uint64_t p = 1;
        for (uint32_t l = 1; l < b.inst[q].levels && c > 0; l++) {
            p = std::min<uint64_t>(2 * p, c);
            b.parts_ub[l] += p;
        }
    }
    for (uint64_t v : b.parts_ub) b.parts_cap = std::max<uint64_t>(b.parts_cap, v);
    b.tiles_cap = b.parts_cap + (b.n + kTile - 1) / kTile;
    if (b.dense) {
        for (auto &I : b.inst) {
            I.heap_base = b.heap_total;
            b.heap_total += (I.levels ? (1ull << I.levels) - 1 : 0);
        }
    }
    b.shape = h ? h : 1;
    return b;
}

// Runs the levels of a batch.  With spec_passes > 0 the split search of a
// level is queued as its first decision plus spec_passes further
// histogram+decision passes, each skipped on the device when no part is left
// searching, so the host never waits inside the level loop; returns true when
// some part needed more passes (the result is then invalid and the caller
// runs the batch again with spec_passes = 0, which checks every pass).
template <int D, typename WS, typename Acc>
static bool run_levels(Context &cx, hipStream_t st, const BatchPlan &B, int wrap32, uint64_t *d_out, WS *w_first,
                       int spec_passes) {
    Counters *ctr = (Counters *)cx.counters.ptr;
    uint32_t *h_search = (uint32_t *)cx.host_scratch;
    Instance *insts = (Instance *)cx.instances.ptr;
    const uint32_t n_inst = (uint32_t)B.inst.size();
    const uint64_t n = B.n;
    const size_t L = B.max_levels;
    const uint64_t parts_cap = B.parts_cap, tiles_cap = B.tiles_cap;
    if (tiles_cap > 0xffffffffull || parts_cap > 0x7fffffffull) throw HipFailure{hipErrorOutOfMemory};
    if ((uint64_t)L * (uint64_t)(spec_passes + 1) > (uint64_t)kSpecSlots) spec_passes = 0;
    // passes queued per level: spec_passes, or, when the context's last call
    // had this shape, queued_passes() of what that level's search needed
    // then (a search that needs more is closed by its last queued decision,
    // the first one when none is queued, and the batch runs again with host
    // checks)
    const uint64_t shape = B.shape;
    const Hints *seen = cx.hints_for(shape);
    std::vector<int> depth(L, spec_passes);
    if (spec_passes > 0 && seen && seen->level_passes.size() == L)
        for (size_t l = 0; l < L; l++) depth[l] = queued_passes(spec_passes, seen->level_passes[l]);
    std::vector<uint8_t> used(L, 0);   // decisions per level in this call
    std::vector<uint32_t> open1(L, 0);  // parts still searching after each level's first decision
    const std::vector<uint32_t> last_open = seen && seen->level_open.size() == L ? seen->level_open
                                                                                  : std::vector<uint32_t>();
    const bool same_shape = !last_open.empty();

    // Two levels per move (dense trees, prefix partition): three part / tile
    // tables (a level, its children searched in place, its grandchildren).
    const bool pairs = cx.pairs && B.dense;
    for (int b = 0; b < (pairs ? 3 : 2); b++) {
        cx.parts[b].ensure(parts_cap * sizeof(Part));
        cx.tiles[b].ensure(tiles_cap * sizeof(Tile));
    }
    cx.nchild.ensure(parts_cap * sizeof(uint64_t));
    cx.ntiles.ensure(parts_cap * sizeof(uint64_t));
    // Two levels per move: per-tile counts (two buffers, alternating per
    // move) and the children's first passes forwarded by the moves (and, at
    // level 0, by the roots' first pass).
    const bool root_fwd = pairs && cx.forward && L > 1;
    if (pairs) {
        for (int b = 0; b < 2; b++) {
            cx.pleft[b].ensure(tiles_cap * sizeof(uint64_t));
            cx.pbelow2[b].ensure(tiles_cap * 2 * sizeof(uint32_t));
        }
        cx.fwd_cnt.ensure(parts_cap * 2 * sizeof(uint32_t));
        cx.fwd_w.ensure(parts_cap * 2 * sizeof(Acc));
        cx.fwd_mn.ensure(parts_cap * 2 * sizeof(uint32_t));
        cx.tile_bins2.ensure(tiles_cap * 2 * kMaxBins * sizeof(uint16_t));
        cx.tile4.ensure(tiles_cap * 4 * sizeof(uint64_t));
    }
    const Forward<Acc> none{nullptr, nullptr, nullptr, nullptr, nullptr};
    auto fwd_at = [&](int b) {
        return Forward<Acc>{(uint32_t *)cx.fwd_cnt.ptr, (Acc *)cx.fwd_w.ptr, (uint32_t *)cx.fwd_mn.ptr,
                            (uint64_t *)cx.pleft[b].ptr, (uint32_t *)cx.pbelow2[b].ptr};
    };
    // Split tree for the descent: a dense heap per instance for shallow
    // trees; node records with child links otherwise (single instance).
    const bool dense = B.dense;
    std::vector<uint64_t> node_base(L + 1, 0);
    uint64_t n_nodes = 0;
    if (dense) {
        cx.heap_split.ensure(std::max<uint64_t>(B.heap_total, 1) * sizeof(float));
    } else {
        for (size_t l = 0; l < L; l++) node_base[l + 1] = node_base[l] + B.parts_ub[l];
        n_nodes = node_base[L];
        if (n_nodes > 0x7fffffffull) throw HipFailure{hipErrorOutOfMemory};
        cx.nodes.ensure(n_nodes * sizeof(Node));
        cx.leaf.ensure(2 * parts_cap * sizeof(uint64_t));
    }
    float *heap_split = dense ? (float *)cx.heap_split.ptr : nullptr;
    Node *nodes = dense ? nullptr : (Node *)cx.nodes.ptr;
    // Level 0 has its own tile table: the prologue and the final descent
    // read it, while the two ping-pong tables serve levels 1, 2, ...
    Tile *tiles0 = (Tile *)cx.tiles0.ptr;

    // Partition: per-tile left counts and their scan.
    cx.tile_left.ensure(tiles_cap * sizeof(uint64_t));
    cx.below.ensure(tiles_cap * sizeof(uint32_t));
    // Per-tile bin counts of the histogram passes (prefix partition): the
    // partition's per-tile left counts of a part decided in such a pass.
    cx.tile_bins.ensure(tiles_cap * kMaxBins * sizeof(uint16_t));
    uint16_t *tile_bins = (uint16_t *)cx.tile_bins.ptr;
    const uint32_t hist_counts = kCountsBins;
    // The per-tile counts start at zero: each partition's fused pass adds to
    // the entries of the next level's tiles and k_tile_left clears what it
    // reads, but the last level's counts are never read (no partition), so
    // a call cannot rely on the previous one leaving them clear.
    if (!pairs) k_fill_u32<<<grid_for(tiles_cap, kBlock), kBlock, 0, st>>>((uint32_t *)cx.below.ptr, tiles_cap, 0u);
    // Bins: parts x bins of the widest pass of any level.
    uint64_t bins = 1;
    for (size_t l = 0; l < L; l++)
        bins = std::max<uint64_t>(bins, std::max<uint64_t>(B.parts_ub[l], 1) * (uint64_t)choose_bins(B.parts_ub[l], cx.max_bins));
    // Copies 1..7 of the bins, one per XCD, for the levels of at most
    // cx.copy_parts parts (the first levels of a call of few instances).
    uint64_t xparts = 0;
    for (size_t l = 0; l < L; l++)
        if (B.parts_ub[l] <= cx.copy_parts) xparts = std::max<uint64_t>(xparts, B.parts_ub[l]);
    const bool copies = xparts > 0;
    const uint64_t xbins = (uint64_t)(kCopies - 1) * xparts * kMaxBins;
    const bool fresh_bins = cx.bcnt.bytes < bins * 4 || cx.bw.bytes < bins * 8 || cx.bmin.bytes < bins * 4 ||
                            cx.xcnt.bytes < xbins * 4 || cx.xw.bytes < xbins * 8 || cx.xmin.bytes < xbins * 4;
    cx.bcnt.ensure(bins * 4);
    cx.bw.ensure(bins * 8);
    cx.bmin.ensure(bins * 4);
    if (copies) {
        cx.xcnt.ensure(xbins * 4);
        cx.xw.ensure(xbins * 8);
        cx.xmin.ensure(xbins * 4);
    }
    if (fresh_bins || !cx.bins_clean) {
        // k_decide leaves bins zeroed after reading them; only new or dirty
        // buffers need a clear (never in a captured call: it starts clean).
        if (cx.capturing) throw HipFailure{hipErrorStreamCaptureUnsupported};
        check(hipMemsetAsync(cx.bcnt.ptr, 0, cx.bcnt.bytes, st));
        check(hipMemsetAsync(cx.bw.ptr, 0, cx.bw.bytes, st));
        check(hipMemsetAsync(cx.bmin.ptr, 0xff, cx.bmin.bytes, st));
        if (cx.xcnt.bytes) {
            check(hipMemsetAsync(cx.xcnt.ptr, 0, cx.xcnt.bytes, st));
            check(hipMemsetAsync(cx.xw.ptr, 0, cx.xw.bytes, st));
            check(hipMemsetAsync(cx.xmin.ptr, 0xff, cx.xmin.bytes, st));
        }
    }
    cx.bins_clean = false;   // until this call completes
    uint32_t *bcnt = (uint32_t *)cx.bcnt.ptr;
    Acc *bw = (Acc *)cx.bw.ptr;
    uint32_t *bmin = (uint32_t *)cx.bmin.ptr;
    // the copies a level of pub parts uses (none: nullptr)
    auto copies_for = [&](uint64_t pub) {
        if (!copies || pub > xparts) return BinCopies<Acc>{nullptr, nullptr, nullptr, 0u};
        return BinCopies<Acc>{(uint32_t *)cx.xcnt.ptr, (Acc *)cx.xw.ptr, (uint32_t *)cx.xmin.ptr, (uint32_t)xparts};
    };

    // Working arrays.  Coordinates: slot 0 is the prologue output in input
    // order and is kept for the final descent; levels 1, 3, ... use slot 1,
    // levels 2, 4, ... slot 2.  Weights alternate between two slots.
    float *cs[3][3];
    WS *ws[2];
    for (int b = 0; b < 3; b++)
        for (int k = 0; k < 3; k++)
            cs[b][k] = k < D ? (b == 0 ? (float *)cx.coords0(k) : (float *)cx.cx[b][k].ptr) : nullptr;
    ws[0] = w_first;
    ws[1] = (WS *)cx.w[1].ptr;
    auto cslot = [](size_t level) { return level == 0 ? 0 : (level % 2 == 1 ? 1 : 2); };
    auto wslot = [](size_t level) { return (int)(level % 2); };

    const int real_mode = std::is_same<Acc, double>::value;
    const double payload = 4.0 * D + sizeof(WS);   // bytes of one point in the working arrays
    timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
        // (the counters were cleared by k_init_roots, before the prologue)
        k_roots_finalize<<<grid_for(n_inst, kBlock), kBlock, 0, st>>>(
            insts, n_inst, (Part *)cx.parts[0].ptr, D, real_mode, root_fwd ? (uint32_t *)cx.fwd_cnt.ptr : nullptr,
            (uint64_t *)cx.fwd_w.ptr, (uint32_t *)cx.fwd_mn.ptr, &ctr->n_parts[0], &ctr->n_tiles[0],
            (uint32_t)B.tiles0, (const InstReduce *)cx.ired.ptr);
    });

    cx.stats.levels = 0;
    cx.stats.passes = 0;
    cx.stats.max_passes = 0;
    cx.stats.full_first = 0;
    cx.stats.unit_points = kTile;
    cx.stats.bins = kFirstBins;

    // Split search of the parts of table `sl` (level `level`) after a first
    // pass filled their first bins (kFirstBins or 2).  `hist(gate)` launches
    // one histogram pass over the parts' points.  With spec_passes > 0 the
    // passes are queued without host checks (see run_levels' comment).
    auto search_gen = [&](int sl, size_t level, uint64_t parts_ub, int first_nb, unsigned long long *swept_first,
                          uint32_t counts_first, uint32_t counts_hist, auto &&hist) {
        Part *parts = (Part *)cx.parts[sl].ptr;
        uint32_t *np = &ctr->n_parts[sl];
        const int nb = choose_bins(parts_ub, cx.max_bins);
        const BinCopies<Acc> xc = copies_for(parts_ub);
        cx.stats.bins = std::max<uint32_t>(cx.stats.bins, (uint32_t)nb);
        // a decision over nbd bins: a lane per part up to kFirstBins bins
        // without bin copies (levels of many parts), a wavefront per part up
        // to 64 bins, else a workgroup (measured faster than a wavefront with
        // four bins per lane on 2^24 Gaussian points, no slower on the batch)
        auto decide = [&](int nbd, uint32_t *searching, unsigned long long *swept, uint32_t counts,
                          const uint32_t *gate, uint32_t *unfin) {
            if (nbd <= kFirstBins && !xc.cnt)
                k_decide_lanes<Acc><<<grid_for(parts_ub, kBlock), kBlock, 0, st>>>(
                    parts, np, bcnt, bw, bmin, nbd, wrap32, searching, swept, counts, gate, unfin, xc);
            else if (nbd <= 64)
                k_decide_wave<Acc><<<(unsigned)((parts_ub + kBlock / 64 - 1) / (kBlock / 64)), kBlock, 0, st>>>(
                    parts, np, bcnt, bw, bmin, nbd, wrap32, searching, swept, counts, gate, unfin, xc);
            else
                k_decide<Acc><<<(unsigned)parts_ub, kBlock, 0, st>>>(parts, np, bcnt, bw, bmin, nbd, wrap32, searching,
                                                                   swept, counts, gate, unfin, xc);
        };
        if (spec_passes > 0) {
            // slot p counts the parts still searching after decision p; pass
            // p + 1 runs only if it is non-zero
            uint32_t *slot = &ctr->spec[level * (spec_passes + 1) * kCopies];   // slot p: slot + p * kCopies
            const int queued = depth[level];
            timed(cx, st, COUPE_HIP_K_DECIDE, 0.0, [&] {
                // with no pass queued this decision closes the searches it leaves open
                decide(first_nb, slot, swept_first, counts_first, nullptr,
                       queued == 0 ? &ctr->unfinished : nullptr);
            });
            for (int p = 1; p <= queued; p++) {
                timed(cx, st, COUPE_HIP_K_HIST, 0.0, [&] { hist(nb, slot + (p - 1) * kCopies); });
                // the last queued decision also closes the searches it leaves open
                timed(cx, st, COUPE_HIP_K_DECIDE, 0.0, [&] {
                    decide(nb, slot + p * kCopies, ctr->swept, counts_hist, slot + (p - 1) * kCopies,
                           p == queued ? &ctr->unfinished : nullptr);
                });
            }
            return;
        }
        check(hipMemsetAsync(ctr->searching, 0, sizeof(ctr->searching), st));
        timed(cx, st, COUPE_HIP_K_DECIDE, 0.0, [&] {
            decide(first_nb, ctr->searching, swept_first, counts_first, nullptr, nullptr);
        });
        uint32_t passes = 1;
        for (;;) {
            check(hipMemcpyAsync(h_search, ctr->searching, sizeof(ctr->searching), hipMemcpyDeviceToHost, st));
            check(hipStreamSynchronize(st));
            const uint32_t searching = sum_copies(h_search);
            if (passes == 1) open1[level] = searching;
            if (searching == 0) break;
            // f32 bisection ends within a few hundred candidates; each pass takes 8
            if (passes >= 4096) throw HipFailure{hipErrorUnknown};
            check(hipMemsetAsync(ctr->searching, 0, sizeof(ctr->searching), st));
            timed(cx, st, COUPE_HIP_K_HIST, 0.0, [&] { hist(nb, nullptr); });
            timed(cx, st, COUPE_HIP_K_DECIDE, 0.0, [&] {
                decide(nb, ctr->searching, ctr->swept, counts_hist, nullptr, nullptr);
            });
            passes++;
        }
        cx.stats.passes += passes;
        cx.stats.max_passes = std::max(cx.stats.max_passes, passes);
        used[level] = (uint8_t)std::min<uint32_t>(passes, 255);
    };

    // Search of parts whose points are contiguous in `tiles` (coordinates cl,
    // weights wl).
    auto search_at = [&](int sl, size_t level, uint64_t parts_ub, uint64_t tiles_ub, bool first_standalone,
                         Tile *tiles, float *const *cl, WS *wl, uint32_t counts_first) {
        Part *parts = (Part *)cx.parts[sl].ptr;
        uint32_t *ntl = &ctr->n_tiles[sl];
        const unsigned g = persistent_grid(cx, k_hist<WS, Acc>, tiles_ub);
        search_gen(sl, level, parts_ub, first_standalone ? kFirstBins : 2,
                   first_standalone ? ctr->swept : ctr->swept_fused, counts_first, hist_counts,
                   [&](int nb, const uint32_t *gate) {
                       k_hist<WS, Acc><<<g, kBlock, 0, st>>>(parts, tiles, ntl, cl[0], cl[1], cl[2], wl, nb, bcnt,
                                                             bw, bmin, gate, tile_bins, copies_for(parts_ub),
                                                             cx.search_chunk);
                   });
    };
    auto search = [&](int sl, size_t level, uint64_t parts_ub, uint64_t tiles_ub, uint64_t n_parts_host,
                      bool first_standalone) {
        (void)n_parts_host;
        search_at(sl, level, parts_ub, tiles_ub, first_standalone, level == 0 ? tiles0 : (Tile *)cx.tiles[sl].ptr,
                  cs[cslot(level)], ws[wslot(level)],
                  first_standalone ? kCountsRead : kCountsBelow);
    };

    // Level 0: the root parts' tables come from k_init_roots (in prologue());
    // register sweep of the roots, then their split search.
    uint64_t parts_ub = n_inst;
    uint64_t tiles_ub = std::max<uint64_t>(B.tiles0, 1);
    if (L > 0 && root_fwd) {
        const unsigned g = persistent_grid(cx, k_root_first<D, WS, Acc>, tiles_ub);
        // the roots' first pass with their children's: two axes and the weight
        timed(cx, st, COUPE_HIP_K_HIST, (8.0 + sizeof(WS)) * (double)n, [&] {
            k_root_first<D, WS, Acc><<<g, kBlock, 0, st>>>((const Part *)cx.parts[0].ptr, tiles0,
                                                          &ctr->n_tiles[0], cs[0][0], cs[0][1], cs[0][2], ws[0],
                                                          bcnt, bw, bmin, fwd_at(0));
        });
        search(0, 0, parts_ub, tiles_ub, n_inst, true);
    } else if (L > 0) {
        const unsigned g = persistent_grid(cx, k_sweep<WS, Acc>, tiles_ub);
        // the roots' first pass: split-axis coordinate and weight of every point
        timed(cx, st, COUPE_HIP_K_HIST, (4.0 + sizeof(WS)) * (double)n, [&] {
            k_sweep<WS, Acc><<<g, kBlock, 0, st>>>((const Part *)cx.parts[0].ptr, tiles0,
                                                  &ctr->n_tiles[0], cs[0][0], cs[0][1], cs[0][2], ws[0], bcnt, bw,
                                                  bmin);
        });
        search(0, 0, parts_ub, tiles_ub, n_inst, true);
    }

    int cur = 0;
    if (pairs) {
        // Two levels per move.  Real levels l = 0, 2, 4, ...: parts with
        // contiguous points (table ia, tiles tl, coordinates cl, weights wl).
        // Their children (table ib) are searched in place over the parent's
        // tiles; the grandchildren (table ic) receive the points.
        uint16_t *tile_bins2 = (uint16_t *)cx.tile_bins2.ptr;
        // Forwarding: the move to level l + 2 also sums, for each new part,
        // its children's first pass (per side of its first candidate) and
        // each of its tiles' counts; a part whose split ends at its first
        // candidate then needs no first pass of its children.  Per-tile
        // counts alternate between two buffers (par: the current level's).
        int par = 0;
        uint64_t *q4[4];
        int ia = 0, csl = 0, wsl = 0;
        Tile *tl = tiles0;
        uint64_t tl_ub = std::max<uint64_t>(B.tiles0, 1);
        auto children_of = [&](int from, int to, size_t level, Tile *next_tiles, const Forward<Acc> &fsrc,
                               const Forward<Acc> &fclr) {
            Part *pf = (Part *)cx.parts[from].ptr;
            const uint64_t pub = std::max<uint64_t>(B.parts_ub[level], 1);
            const bool self_plan = pub <= kSelfPlanParts;   // the write kernel scans the offsets itself
            if (!self_plan && pub <= kPlanParts) {
                timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                    k_children_plan<<<1, kPlanThreads, 0, st>>>(pf, &ctr->n_parts[from], (uint64_t *)cx.nchild.ptr,
                                                               (uint64_t *)cx.ntiles.ptr, &ctr->parts_total,
                                                               &ctr->tiles_total, &ctr->n_parts[to], &ctr->n_tiles[to]);
                });
            } else if (!self_plan) {
                timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                    k_children_count<<<grid_for(pub, kBlock), kBlock, 0, st>>>(
                        pf, &ctr->n_parts[from], (uint64_t *)cx.nchild.ptr, (uint64_t *)cx.ntiles.ptr);
                });
                scan_u64(cx, st, (uint64_t *)cx.nchild.ptr, &ctr->n_parts[from], pub, &ctr->parts_total);
                scan_u64(cx, st, (uint64_t *)cx.ntiles.ptr, &ctr->n_parts[from], pub, &ctr->tiles_total);
            }
            timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                k_children_write<Acc><<<(unsigned)pub, kBlock, 0, st>>>(
                    pf, &ctr->n_parts[from], D, self_plan ? nullptr : (const uint64_t *)cx.nchild.ptr,
                    self_plan ? nullptr : (const uint64_t *)cx.ntiles.ptr, (Part *)cx.parts[to].ptr, next_tiles,
                    nodes, 0, 0, heap_split, insts, fsrc, bcnt, bw, bmin, fclr, ctr->forwarded,
                    &ctr->n_parts[to], &ctr->n_tiles[to]);
                if (pub > kPlanParts)
                    k_set_sizes<<<1, 64, 0, st>>>(&ctr->parts_total, &ctr->tiles_total, &ctr->n_parts[to],
                                                  &ctr->n_tiles[to]);
            });
        };
        for (size_t level = 0; level < L; level += 2) {
            TraceRange range(level_range(level));   // a pair of levels: this range covers both
            const int ib = (ia + 1) % 3, ic = (ia + 2) % 3;
            Part *pa = (Part *)cx.parts[ia].ptr;
            Part *pb = (Part *)cx.parts[ib].ptr;
            uint32_t *ntl = &ctr->n_tiles[ia];
            float *const *cl = cs[csl];
            WS *wl = ws[wsl];
            cx.stats.levels++;
            // The children's first pass: the bins of their first candidate
            // (forwarded by the move that made this level, or read by
            // k_child_first for the parents whose split is not there, which
            // also counts per tile the points left of the parent's split for
            // the next move).  For the last pair of levels (no move follows),
            // when in the context's last call of this shape most parents of
            // this level were still searching after their first decision (so
            // most children were not forwarded and needed later passes): one
            // pass with every candidate over all the parents' tiles instead.
            const uint64_t pub0 = std::max<uint64_t>(B.parts_ub[level], 1);
            const bool full_first = cx.full_first && same_shape && level + 2 == L &&
                                    2 * (uint64_t)last_open[level] > pub0;
            // the move that made this level forwarded its children's first pass
            const bool fwd_in = cx.forward && (level >= 2 || root_fwd) && level + 1 < L && !full_first;
            uint64_t *pl = (uint64_t *)cx.pleft[par].ptr;
            uint32_t *pb2 = (uint32_t *)cx.pbelow2[par].ptr;
            children_of(ia, ib, level, nullptr, fwd_in ? fwd_at(par) : none, none);   // in place: no tiles
            if (level + 1 == L) break;
            cx.stats.levels++;
            const uint64_t pub1 = std::max<uint64_t>(B.parts_ub[level + 1], 1);
            const unsigned gv = persistent_grid(cx, k_hist_virtual<D, WS, Acc>, tl_ub);
            auto hist_v = [&](int nb, const uint32_t *gate) {
                k_hist_virtual<D, WS, Acc><<<gv, kBlock, 0, st>>>(pa, tl, ntl, pb, cl[0], cl[1], cl[2], wl, nb, bcnt,
                                                                 bw, bmin, gate, tile_bins2, copies_for(pub1),
                                                                 cx.search_chunk);
            };
            if (full_first) {
                cx.stats.full_first++;
                // every child searching: the pass reads both axes and the weight
                // of every point (counted with the first passes by the decision)
                const int nbf = choose_bins(pub1, cx.max_bins);
                timed(cx, st, COUPE_HIP_K_HIST, 0.0, [&] { hist_v(nbf, nullptr); });
                search_gen(ib, level + 1, pub1, nbf, ctr->first_read, kCountsBins, kCountsBins, hist_v);
            } else {
                const unsigned g = persistent_grid(cx, k_child_first<D, WS, Acc>, tl_ub);
                // bytes: parent-axis and child-axis coordinates and the weight
                // of every point read (counted on the device, added below)
                timed(cx, st, COUPE_HIP_K_HIST, 0.0, [&] {
                    k_child_first<D, WS, Acc><<<g, kBlock, 0, st>>>(pa, tl, ntl, cl[0], cl[1], cl[2], wl, bcnt, bw,
                                                                   bmin, pb2, pl, fwd_in ? 1u : 0u,
                                                                   ctr->first_read);
                });
                search_gen(ib, level + 1, pub1, 2, ctr->swept_fused, kCountsBelow, kCountsBins, hist_v);
            }
            // grandchildren, with tiles
            Tile *tn = (Tile *)cx.tiles[ic].ptr;
            // the next move forwards when the grandchildren's children are searched
            const bool fwd_out = cx.forward && level + 3 < L;
            children_of(ib, ic, level + 1, tn, none, fwd_out ? fwd_at(par ^ 1) : none);
            if (level + 2 == L) break;
            // the move.  The per-tile counts of the 4 grandchild slots lie end
            // to end (stride tl_ub) and are scanned as one array: the move only
            // takes differences within one slot's tiles of one part, which the
            // sums of the preceding slots (and the unused entries past the
            // tile count) do not change.
            for (int j = 0; j < 4; j++) q4[j] = (uint64_t *)cx.tile4.ptr + (uint64_t)j * tl_ub;
            timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                const unsigned g = grid_for(tl_ub, kBlock);   // one thread per tile
                k_tile_count4<D><<<g, kBlock, 0, st>>>(pa, tl, ntl, pb, cl[0], cl[1], cl[2],
                                                      pl, pb2, tile_bins2, q4[0],
                                                      q4[1], q4[2], q4[3], ctr->move_runs);
            });
            scan_u64(cx, st, q4[0], nullptr, 4 * tl_ub, nullptr);
            const int csn = csl == 1 ? 2 : 1, wsn = wsl ^ 1;
            float *const *co = cs[csn];
            // level l splits axis l mod D; after this move only levels
            // level + 2 and level + 3 are split (or fewer): in 3D the axis of
            // level + 1, which would be next split at level + 4, is not
            // written (the ids come from the prologue's copy)
            const uint32_t skip_axis = (D == 3 && level + 4 >= L) ? (uint32_t)((level + 1) % D) : 3u;
            {
                auto move = [&](auto kernel) {
                    const unsigned g = persistent_grid(cx, kernel, tl_ub);
                    // bytes: every point's payload read, and written less the skipped axis
                    const double moved = 2.0 * payload - (skip_axis < 3 ? 4.0 : 0.0);
                    timed(cx, st, COUPE_HIP_K_PARTITION, moved * (double)n, [&] {
                        kernel<<<g, kBlock, 0, st>>>(pa, pb, (const Part *)cx.parts[ic].ptr, tl, ntl, q4[0], q4[1],
                                                     q4[2], q4[3], cl[0], cl[1], cl[2], wl, co[0], co[1], co[2],
                                                     ws[wsn], n, bcnt, bw, bmin, fwd_out ? fwd_at(par ^ 1) : none,
                                                     skip_axis);
                    });
                };
                // the stores are non-temporal for moves larger than the switch whose
                // tiles went to few destinations in the shape's last call
                const bool nt = (2.0 * payload - (skip_axis < 3 ? 4.0 : 0.0)) * (double)n > cx.move_nt_bytes &&
                                seen && seen->moves_coherent == 1;
                if (fwd_out) nt ? move(k_partition4<D, WS, Acc, true, true>) : move(k_partition4<D, WS, Acc, true, false>);
                else nt ? move(k_partition4<D, WS, Acc, false, true>) : move(k_partition4<D, WS, Acc, false, false>);
            }
            // the grandchildren's search over their own tiles
            const uint64_t pub2 = std::max<uint64_t>(B.parts_ub[level + 2], 1);
            const uint64_t tub2 = std::min<uint64_t>(tiles_cap, pub2 + (n + kTile - 1) / kTile);
            search_at(ic, level + 2, pub2, tub2, false, tn, co, ws[wsn], kCountsRead);
            ia = ic;
            csl = csn;
            wsl = wsn;
            tl = tn;
            tl_ub = tub2;
            par ^= 1;
        }
        parts_ub = std::max<uint64_t>(B.parts_ub[L - 1], 1);
    } else {
        for (size_t level = 0; level < L; level++) {
            TraceRange range(level_range(level));
            cx.stats.levels++;
            const int nxt = cur ^ 1;
            Part *parts = (Part *)cx.parts[cur].ptr;
            Tile *tiles = level == 0 ? tiles0 : (Tile *)cx.tiles[cur].ptr;
            uint32_t *np = &ctr->n_parts[cur];
            uint32_t *ntl = &ctr->n_tiles[cur];

            // Children (next level's parts and tiles), split records, id offsets.
            const bool plan1 = parts_ub <= kPlanParts;
            const bool self_plan = parts_ub <= kSelfPlanParts;   // the write kernel scans the offsets itself
            if (!self_plan && plan1) {
                timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                    k_children_plan<<<1, kPlanThreads, 0, st>>>(parts, np, (uint64_t *)cx.nchild.ptr,
                                                               (uint64_t *)cx.ntiles.ptr, &ctr->parts_total,
                                                               &ctr->tiles_total, &ctr->n_parts[nxt], &ctr->n_tiles[nxt]);
                });
            } else if (!self_plan) {
                timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                    k_children_count<<<grid_for(parts_ub, kBlock), kBlock, 0, st>>>(
                        parts, np, (uint64_t *)cx.nchild.ptr, (uint64_t *)cx.ntiles.ptr);
                });
                scan_u64(cx, st, (uint64_t *)cx.nchild.ptr, np, parts_ub, &ctr->parts_total);
                scan_u64(cx, st, (uint64_t *)cx.ntiles.ptr, np, parts_ub, &ctr->tiles_total);
            }
            timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                k_children_write<Acc><<<(unsigned)parts_ub, kBlock, 0, st>>>(
                    parts, np, D, self_plan ? nullptr : (const uint64_t *)cx.nchild.ptr,
                    self_plan ? nullptr : (const uint64_t *)cx.ntiles.ptr, (Part *)cx.parts[nxt].ptr,
                    (Tile *)cx.tiles[nxt].ptr, nodes, node_base[level], node_base[std::min(level + 1, L)], heap_split,
                    insts, Forward<Acc>{}, bcnt, bw, bmin, Forward<Acc>{}, ctr->forwarded, &ctr->n_parts[nxt],
                    &ctr->n_tiles[nxt]);
                if (!plan1)
                    k_set_sizes<<<1, 64, 0, st>>>(&ctr->parts_total, &ctr->tiles_total, &ctr->n_parts[nxt],
                                                  &ctr->n_tiles[nxt]);
            });
            if (level + 1 == L) break;   // every part left is at its last level

            // Partition of this level, fused with the children's first pass.
            float *const *ci = cs[cslot(level)];
            float *const *co = cs[cslot(level + 1)];
            {
                const unsigned gt = persistent_grid(cx, k_tile_left, (tiles_ub + kBlock / 64 - 1) / (kBlock / 64));
                timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                    k_tile_left<<<gt, kBlock, 0, st>>>(parts, tiles, ntl, ci[0], ci[1], ci[2], (uint32_t *)cx.below.ptr,
                                                      tile_bins, (uint64_t *)cx.tile_left.ptr);
                });
                scan_u64(cx, st, (uint64_t *)cx.tile_left.ptr, ntl, tiles_ub, nullptr);
                const unsigned g = persistent_grid(cx, k_partition_tiles<D, WS, Acc>, tiles_ub);
                timed(cx, st, COUPE_HIP_K_PARTITION, 2.0 * payload * (double)n, [&] {
                    k_partition_tiles<D, WS, Acc><<<g, kBlock, 0, st>>>(
                        parts, (const Part *)cx.parts[nxt].ptr, tiles, ntl, (const uint64_t *)cx.tile_left.ptr, ci[0],
                        ci[1], ci[2], ws[wslot(level)], co[0], co[1], co[2], ws[wslot(level + 1)], n, bcnt, bw, bmin,
                        (uint32_t *)cx.below.ptr);
                });
            }

            cur = nxt;
            parts_ub = std::max<uint64_t>(B.parts_ub[level + 1], 1);
            tiles_ub = std::min<uint64_t>(tiles_cap, parts_ub + (n + kTile - 1) / kTile);
            search(cur, level + 1, parts_ub, tiles_ub, 0, false);
        }
    }

    // Part ids in input order (:589-601, :698-702).
    if (dense) {
        // reads the input-order f32 coordinates, writes the 8-byte id;
        // non-temporal above the moves' size threshold (2^27 points 3D:
        // 0.521 -> 0.495 ms, and the next call's first pass 0.604 -> 0.578;
        // 2^24 points: 0.060 -> 0.063 ms, so cached there)
        const double bytes = (4.0 * D + 8.0) * (double)n;
        auto descend = [&](auto kernel) {
            const unsigned g = persistent_grid(cx, kernel, B.tiles0);
            timed(cx, st, COUPE_HIP_K_IDS, bytes, [&] {
                kernel<<<g, kBlock, 0, st>>>(tiles0, (uint32_t)B.tiles0, insts, heap_split, cs[0][0], cs[0][1],
                                             cs[0][2], d_out);
            });
        };
        if (bytes > cx.move_nt_bytes) descend(k_descend_dense<D, true>);
        else descend(k_descend_dense<D, false>);
    } else {
        // single instance, deep tree: node records and leaf ids
        Part *parts = (Part *)cx.parts[cur].ptr;
        timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
            k_last_nodes<<<grid_for(parts_ub, kBlock), kBlock, 0, st>>>(
                parts, &ctr->n_parts[cur], node_base[L - 1], &insts[0].min_id, nodes, (uint64_t *)cx.leaf.ptr,
                nullptr);
        });
        const unsigned g = persistent_grid(cx, k_descend<D>, (n + kBlock - 1) / kBlock);
        timed(cx, st, COUPE_HIP_K_IDS, (4.0 * D + 8.0) * (double)n, [&] {
            k_descend<D><<<g, kBlock, 0, st>>>(nodes, n_nodes, (uint32_t)L, node_base[L - 1],
                                               (const uint64_t *)cx.leaf.ptr, parts_ub, cs[0][0], cs[0][1],
                                               cs[0][2], n, d_out, &ctr->error);
        });
    }

    // Standalone-pass bytes: swept points times (axis coordinate + stored weight).
    // One readback for the whole call: the counters from `swept` to the end
    // of the speculative slots.
    const size_t rb_off = offsetof(Counters, swept);
    // up to the last speculative slot this call uses
    const size_t rb_len = offsetof(Counters, spec) - rb_off +
                          (spec_passes > 0 ? L * (size_t)(spec_passes + 1) * kCopies * sizeof(uint32_t) : 0);
    static_assert(offsetof(Counters, end) - offsetof(Counters, swept) <= kHostScratch - 256, "readback");
    char *h_rb = (char *)cx.host_scratch + 256;
    check(hipMemcpyAsync(h_rb, (char *)ctr + rb_off, rb_len, hipMemcpyDeviceToHost, st));
    cx.pending.levels = L;
    cx.pending.spec = spec_passes;
    cx.pending.depth = depth;
    cx.pending.shape = shape;
    cx.pending.ws_bytes = sizeof(WS);
    if (cx.capturing) return false;   // checked once the graph has run
    check(hipStreamSynchronize(st));
    // the host-checked search counted its passes itself
    if (spec_passes == 0) {
        Hints &e = cx.remember(shape);
        e.level_passes = used;
        e.level_open = open1;
    }
    return finish_levels(cx);
}

// Weight mode of a call's data from its prologue flags.  Integer kinds and
// integer-valued doubles whose partial sums stay exact in f64 (|sum| <= 2^53)
// use exact int64 sums: the same bits as coupe's sums in any order; they are
// stored in 16 bits when all are in [0, 65535], else in 32 or 64.  Other
// doubles use f64 sums.  Modes: 0 u16, 1 i32, 2 i64, 3 f64.
static int weight_mode_of(const Prologue &p, int wkind, bool weights16) {
    const bool integral = wkind != COUPE_DOUBLE || (!p.non_integral && p.sum_abs <= 9007199254740992.0 * (1.0 - 1e-6));
    if (!integral) return 3;
    if (weights16 && !p.not_u16) return 0;
    return p.not_i32 ? 2 : 1;
}

// The prologue's flags are kept in one copy per XCD, in the call's
// counters; the host reads the copies with the call's last counters (or,
// when it must wait for them, right after the prologue, into the early
// area) and combines them.
constexpr size_t kPrologueEarly = 49152;   // the early prologue readback, past the counters' readback
static Prologue *prologue_early(Context &cx) {
    static_assert(kHostScratch >= kPrologueEarly + kCopies * sizeof(Prologue), "prologue readback");
    static_assert(256 + offsetof(Counters, end) - offsetof(Counters, swept) <= kPrologueEarly, "readback areas");
    return (Prologue *)((char *)cx.host_scratch + kPrologueEarly);
}
static const Prologue *prologue_late(Context &cx) {   // in the counters' readback area
    return (const Prologue *)((const char *)cx.host_scratch + 256 + offsetof(Counters, pro) -
                              offsetof(Counters, swept));
}
static Prologue combine_flags(const Prologue *c) {
    Prologue p = c[0];
    for (int k = 1; k < kCopies; k++) {
        p.sum_abs += c[k].sum_abs;
        p.max_abs_bits = std::max(p.max_abs_bits, c[k].max_abs_bits);   // non-negative doubles order as integers
        p.non_integral |= c[k].non_integral;
        p.not_i32 |= c[k].not_i32;
        p.not_u16 |= c[k].not_u16;
    }
    return p;
}

// Once the call has run: the weight mode its data needs becomes the hint;
// false when the mode the call ran in cannot hold that data (an integer
// mode holds any narrower one; f64 and the integer modes differ), and the
// call must run again.
static bool mode_held(Context &cx) {
    const int best = weight_mode_of(combine_flags(prologue_late(cx)), cx.pending.wkind, cx.weights16);
    cx.weight_hint = best;
    cx.remember(cx.pending.shape).weight_hint = best;
    const int mode = cx.pending.mode;
    return (mode == 3) == (best == 3) && mode >= best;
}

template <int D, typename WIN>
static void prologue(Context &cx, hipStream_t st, const BatchPlan &B, const double *pts, const WIN *win,
                     Prologue *h, bool wait, int hint) {
    const uint32_t n_inst = (uint32_t)B.inst.size();
    Prologue *flags = (Prologue *)((char *)cx.counters.ptr + offsetof(Counters, pro));
    // the instance table (a captured call's is copied before each launch)
    if (!cx.capturing)
        check(hipMemcpyAsync(cx.instances.ptr, B.inst.data(), n_inst * sizeof(Instance), hipMemcpyHostToDevice, st));
    timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
        // blocks (instance, share of its tiles): about 4 tiles per thread;
        // the same launch clears the call's counters and the prologue's
        // reductions (bounding boxes start empty, flags clear)
        uint64_t max_nt = 1;
        for (const Instance &I : B.inst) max_nt = std::max<uint64_t>(max_nt, (I.count + kTile - 1) / kTile);
        const unsigned gy = (unsigned)std::min<uint64_t>((max_nt + 4 * kBlock - 1) / (4 * kBlock), 1024);
        static_assert(sizeof(Counters) % 4 == 0, "counters are cleared as 32-bit words");
        k_init_roots<<<dim3(n_inst, gy), kBlock, 0, st>>>(
            (const Instance *)cx.instances.ptr, n_inst, (Part *)cx.parts[0].ptr, (Tile *)cx.tiles0.ptr,
            (uint32_t *)cx.counters.ptr, (uint32_t)(sizeof(Counters) / 4), (InstReduce *)cx.ired.ptr);
    });
    const uint32_t nt = (uint32_t)B.tiles0;
    if (nt > 0) {
        // reads AoS f64 points and the input weights; writes SoA f32, int32
        // weights and (optionally) their 16-bit copy
        // the weight copies the expected mode reads (both without a hint;
        // none for the 64-bit modes, which store their own)
        int32_t *w32 = hint < 0 || hint == 1 ? (int32_t *)cx.w[0].ptr : nullptr;
        uint16_t *w16 = nullptr;
        if (cx.weights16 && hint <= 0) {
            cx.w16.ensure(std::max<uint64_t>(B.n, 1) * sizeof(uint16_t));
            w16 = (uint16_t *)cx.w16.ptr;
        }
        const double bytes =
            (double)B.n * (8.0 * D + sizeof(WIN) + 4.0 * D + (w32 ? 4.0 : 0.0) + (w16 ? 2.0 : 0.0));
        Rotation rot;
        memset(&rot, 0, sizeof rot);
        if (cx.rot) rot = *cx.rot;
        auto launch = [&](auto kernel) {
            const unsigned g = persistent_grid(cx, kernel, nt);
            timed(cx, st, COUPE_HIP_K_PROLOGUE, bytes, [&] {
                kernel<<<g, kBlock, 0, st>>>((const Tile *)cx.tiles0.ptr, nt, pts, win, (float *)cx.cx[0][0].ptr,
                                             (float *)cx.cx[0][1].ptr, (float *)(D > 2 ? cx.cx[0][2].ptr : nullptr),
                                             w32, w16, (Instance *)cx.instances.ptr,
                                             flags, rot, (InstReduce *)cx.ired.ptr, n_inst);
            });
        };
        if (cx.slot0[0]) {
            // a resident working copy: read its f32 coordinates, write none
            const double sbytes = (double)B.n * (4.0 * D + sizeof(WIN) + (w32 ? 4.0 : 0.0) + (w16 ? 2.0 : 0.0));
            const unsigned g = persistent_grid(cx, k_prologue_soa<D, WIN>, nt);
            timed(cx, st, COUPE_HIP_K_PROLOGUE, sbytes, [&] {
                k_prologue_soa<D, WIN><<<g, kBlock, 0, st>>>((const Tile *)cx.tiles0.ptr, nt, cx.coords0(0),
                                                             cx.coords0(1), D > 2 ? cx.coords0(2) : nullptr, win,
                                                             w32, w16, (Instance *)cx.instances.ptr,
                                                             flags, (InstReduce *)cx.ired.ptr,
                                                             n_inst);
            });
        } else if (cx.rot) {
            launch(k_prologue<D, WIN, true>);
        } else {
            launch(k_prologue<D, WIN, false>);
        }
    }
    // without `wait` the flags come back with the call's last counters
    if (wait) {
        check(hipMemcpyAsync(h, flags, kCopies * sizeof(Prologue), hipMemcpyDeviceToHost, st));
        check(hipStreamSynchronize(st));
    }
}

template <int D, typename WIN>
static bool run_typed(Context &cx, hipStream_t st, uint64_t *d_out, const BatchPlan &B, const double *pts,
                      const WIN *win, int wkind, int spec_passes) {
    const uint64_t n = B.n;
    // Coordinate slots: 0 (prologue output, input order), 1 and 2 (levels).
    for (int b = cx.slot0[0] ? 1 : 0; b < 3; b++)
        for (int k = 0; k < D; k++) cx.cx[b][k].ensure(std::max<uint64_t>(n, 1) * sizeof(float));
    cx.w[0].ensure(std::max<uint64_t>(n, 1) * sizeof(int32_t));
    cx.ired.ensure(kCopies * std::max<size_t>(B.inst.size(), 1) * sizeof(InstReduce));
    cx.counters.ensure(sizeof(Counters));
    cx.instances.ensure(std::max<size_t>(B.inst.size(), 1) * sizeof(Instance));
    cx.parts[0].ensure(std::max<uint64_t>(B.parts_cap, B.inst.size()) * sizeof(Part));
    cx.tiles0.ensure(std::max<uint64_t>(B.tiles0, 1) * sizeof(Tile));

    // Weight mode.  Integer kinds and integer-valued doubles whose partial
    // sums stay exact in f64 (|sum| <= 2^53) use exact int64 sums: the same
    // bits as coupe's sums in any order; they are stored in 16 bits when all
    // are in [0, 65535], else in 32 or 64.  Other doubles use f64 sums.
    // Modes: 0 u16, 1 i32, 2 i64, 3 f64.
    const bool wrap32 = wkind == COUPE_INT;
    auto best_mode = [&](const Prologue &p) { return weight_mode_of(p, wkind, cx.weights16); };
    auto weight_bytes = [](int mode) { return mode == 0 ? 2u : (mode == 1 ? 4u : 8u); };
    const unsigned tgrid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(B.tiles0, 1), 4096);
    auto run_mode = [&](int mode) {
        cx.stats.weight_mode = mode == 3 ? 1 : 0;
        cx.stats.weight_bytes = weight_bytes(mode);
        cx.w[1].ensure(std::max<uint64_t>(n, 1) * weight_bytes(mode));
        if (mode == 0)
            return run_levels<D, uint16_t, uint64_t>(cx, st, B, wrap32, d_out, (uint16_t *)cx.w16.ptr, spec_passes);
        if (mode == 1)
            return run_levels<D, int32_t, uint64_t>(cx, st, B, wrap32, d_out, (int32_t *)cx.w[0].ptr, spec_passes);
        if (mode == 2) {
            timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
                k_store_weights<WIN, int64_t><<<tgrid, kBlock, 0, st>>>(
                    (const Tile *)cx.tiles0.ptr, (uint32_t)B.tiles0, (const Instance *)cx.instances.ptr, win,
                    (int64_t *)cx.w[0].ptr);
            });
            return run_levels<D, int64_t, uint64_t>(cx, st, B, wrap32, d_out, (int64_t *)cx.w[0].ptr, spec_passes);
        }
        timed(cx, st, COUPE_HIP_K_OTHER, 0.0, [&] {
            k_store_weights<WIN, double><<<tgrid, kBlock, 0, st>>>(
                (const Tile *)cx.tiles0.ptr, (uint32_t)B.tiles0, (const Instance *)cx.instances.ptr, win,
                (double *)cx.w[0].ptr);
        });
        return run_levels<D, double, double>(cx, st, B, 0, d_out, (double *)cx.w[0].ptr, spec_passes);
    };
    // The mode of this context's last call is taken as this call's: the
    // prologue's flags are then read with the call's last counters, without a
    // wait after the prologue, and a call whose data does not fit that mode
    // runs again in its own.  A new context waits for the flags first.
    Prologue *hp = prologue_early(cx);
    for (;;) {
        const Hints *seen = cx.hints_for(B.shape);
        const int hint = seen ? seen->weight_hint : cx.weight_hint;
        int mode = hint;
        // buffers for the expected mode before anything is queued (a buffer
        // that grows is freed and allocated again)
        cx.w[0].ensure(std::max<uint64_t>(n, 1) * (hint >= 2 ? 8u : 4u));
        if (hint >= 0) cx.w[1].ensure(std::max<uint64_t>(n, 1) * weight_bytes(hint));
        prologue<D, WIN>(cx, st, B, pts, win, hp, hint < 0, hint);
        if (hint < 0) {
            mode = best_mode(combine_flags(hp));
            if (mode >= 2) cx.w[0].ensure(std::max<uint64_t>(n, 1) * 8u);
        }
        cx.pending.mode = mode;
        cx.pending.wkind = wkind;
        const bool unfinished = run_mode(mode);
        if (cx.capturing) return false;   // checked once the graph has run
        if (!mode_held(cx)) continue;     // complete: run_levels waited for the stream
        return unfinished;
    }
}

// Returns true when a speculative search did not end (see run_levels).
template <int D>
static bool run_dim(Context &cx, hipStream_t st, uint64_t *d_out, const BatchPlan &B, const double *pts, int wkind,
                    const void *w, int spec_passes) {
    switch (wkind) {
        case COUPE_INT: return run_typed<D, int32_t>(cx, st, d_out, B, pts, (const int32_t *)w, wkind, spec_passes);
        case COUPE_INT64: return run_typed<D, int64_t>(cx, st, d_out, B, pts, (const int64_t *)w, wkind, spec_passes);
        default: return run_typed<D, double>(cx, st, d_out, B, pts, (const double *)w, wkind, spec_passes);
    }
}

coupe_hip_tuning current_tuning(const Context &cx) {
    return coupe_hip_tuning{(uint32_t)cx.max_bins, (uint32_t)cx.wg_cap, (uint32_t)cx.search_passes,
                            cx.full_first ? 1u : 0u, cx.graphs ? 1u : 0u, cx.search_chunk};
}

void apply_tuning(Context &cx, const coupe_hip_tuning &t) {
    int b = kFirstBins * 4;   // a power of two in [16, kMaxBins]
    while (b < kMaxBins && (uint32_t)b < t.bins) b *= 2;
    cx.max_bins = b;
    cx.wg_cap = (int)std::min<uint32_t>(t.workgroups_per_cu, 64);
    cx.search_passes = (int)std::max<uint32_t>(2, std::min<uint32_t>(t.search_passes, 8));
    cx.full_first = t.full_first != 0;
    cx.graphs = t.graphs != 0;
    cx.search_chunk = std::max<uint32_t>(1, std::min<uint32_t>(t.search_chunk, 64));
}

static uint64_t tune_key(const BatchPlan &B, size_t dim) {
    const uint64_t lg_n = B.n ? 63 - __builtin_clzll(B.n) : 0;
    uint64_t lg4_inst = 0;
    for (uint64_t k = B.inst.size(); k > 1; k >>= 2) lg4_inst++;
    return (uint64_t)dim | lg_n << 8 | (uint64_t)B.max_levels << 16 | lg4_inst << 32;
}

uint64_t Context::layout() {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 0x100000001b3ull;
        h ^= h >> 29;
    };
    for_each_buf([&](DevBuf &d) {
        mix((uint64_t)(uintptr_t)d.ptr);
        mix(d.bytes);
    });
    mix((uint64_t)(uintptr_t)host_scratch);
    return h;
}

void Context::drop_graphs() {
    for (GraphEntry &g : graph_cache)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graph_cache.clear();
    if (gstream) (void)hipStreamDestroy(gstream);
    if (gevent) (void)hipEventDestroy(gevent);
    gstream = nullptr;
    gevent = nullptr;
}

// The graph key of a call: its shape, arrays, the launch sequence its hints
// select (queued passes per level, the last pair's first pass) and the
// context's switches.
static uint64_t graph_key(const Context &cx, const BatchPlan &B, size_t dim, int wkind, const void *out,
                          const void *pts, const void *w, int spec) {
    uint64_t h = 0x84222325cbf29ce4ull;
    auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 0x100000001b3ull;
        h ^= h >> 31;
    };
    mix(B.shape);
    mix(dim);
    mix((uint64_t)wkind);
    mix((uint64_t)(uintptr_t)out);
    mix((uint64_t)(uintptr_t)pts);
    mix((uint64_t)(uintptr_t)w);
    for (int k = 0; k < 3; k++) mix((uint64_t)(uintptr_t)cx.slot0[k]);
    mix((uint64_t)spec);
    mix((uint64_t)cx.pairs | (uint64_t)cx.forward << 1 | (uint64_t)cx.weights16 << 2 | (uint64_t)cx.full_first << 3 |
        (uint64_t)cx.copy_parts << 4 | (uint64_t)cx.search_chunk << 40);
    mix((uint64_t)cx.max_bins | (uint64_t)cx.wg_cap << 16);
    mix(cx.scan_direct);
    const Hints *e = cx.hints_for(B.shape);
    mix((uint64_t)(int64_t)(e ? e->weight_hint : cx.weight_hint));
    const size_t L = B.max_levels;
    for (size_t l = 0; l < L; l++) {
        const int d = e && e->level_passes.size() == L ? queued_passes(spec, e->level_passes[l]) : spec;
        const bool ff = e && e->level_open.size() == L &&
                        2 * (uint64_t)e->level_open[l] > std::max<uint64_t>(B.parts_ub[l], 1);
        mix((uint64_t)d | (uint64_t)ff << 16);
    }
    return h;
}

// One call of a known shape as a graph: captured on its first such call
// (every launch and readback of the eager path, none of its waits), then
// replayed.  The instance table is copied before each launch.  Returns 0
// when the results stand, 1 when a search outran its queued passes (rerun
// host-checked), 2 when the graph could not be made or run, or the data
// needs another weight mode (rerun eagerly), 3 when the key is new (run
// eagerly; captured if the next call has it too).
template <typename Run>
static int run_captured(Context &cx, hipStream_t caller, const BatchPlan &B, size_t dim, int wkind, const void *out,
                        const void *pts, const void *w, int spec, Run &&run) {
    // graphs are captured and launched on the context's own stream (a
    // capture cannot use the null stream), after the caller's work so far
    if (!cx.gstream) {
        check(hipStreamCreateWithFlags(&cx.gstream, hipStreamNonBlocking));
        check(hipEventCreateWithFlags(&cx.gevent, hipEventDisableTiming));
    }
    hipStream_t st = cx.gstream;
    check(hipEventRecord(cx.gevent, caller));
    check(hipStreamWaitEvent(st, cx.gevent, 0));
    const uint64_t key = graph_key(cx, B, dim, wkind, out, pts, w, spec);
    const uint64_t layout = cx.layout();
    GraphEntry *G = nullptr;
    for (GraphEntry &g : cx.graph_cache)
        if (g.key == key && g.layout == layout) G = &g;
    // a key is captured when it repeats (the same shape, arrays and hints
    // twice in a row): calls whose hints keep changing stay eager
    if (!G) {
        Hints &e = cx.remember(B.shape);
        if (key != e.graph_key) {
            e.graph_key = key;
            return 3;
        }
    }
    if (!G) {
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return 2;
        }
        cx.capturing = t_capturing = true;
        bool queued = true;
        try {
            run(spec, st);
        } catch (const HipFailure &) {
            queued = false;
        }
        cx.capturing = t_capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        hipGraphExec_t exec = nullptr;
        const bool made = queued && e == hipSuccess && graph &&
                          hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!made) {
            (void)hipGetLastError();
            return 2;
        }
        if (cx.graph_cache.size() >= kGraphs) {
            (void)hipGraphExecDestroy(cx.graph_cache.front().exec);
            cx.graph_cache.erase(cx.graph_cache.begin());
        }
        GraphEntry g;
        g.key = key;
        g.layout = layout;
        g.exec = exec;
        g.pending = cx.pending;
        g.stats = cx.stats;   // what queuing the call set (levels, bins, weight mode, ...)
        cx.graph_cache.push_back(std::move(g));
        G = &cx.graph_cache.back();
        cx.stats.graph = 2;
    } else {
        cx.stats = G->stats;
        cx.stats.graph = 1;
    }
    cx.pending = G->pending;
    check(hipMemcpyAsync(cx.instances.ptr, B.inst.data(), B.inst.size() * sizeof(Instance), hipMemcpyHostToDevice,
                         st));
    check(hipGraphLaunch(G->exec, st));
    check(hipStreamSynchronize(st));
    const bool unfinished = finish_levels(cx);
    if (!mode_held(cx)) return 2;
    return unfinished ? 1 : 0;
}

enum coupe_err rcb_batch(Context &cx, hipStream_t st, uint64_t *d_out, size_t dim, const double *d_points,
                         int wkind, const void *d_weights, const uint64_t *counts, const size_t *levels,
                         const double *tols, size_t n_inst, const Rotation *rot, const ResidentSource *src,
                         const uint64_t *src_idx, const uint64_t *src_pt) {
    if (dim != 2 && dim != 3) return COUPE_ERR_BAD_DIMENSION;
    if (wkind != COUPE_INT && wkind != COUPE_INT64 && wkind != COUPE_DOUBLE) return COUPE_ERR_BAD_TYPE;
    if (rot && src) return COUPE_ERR_CRASH;   // a rotation needs the points themselves
    if (n_inst && !counts) return COUPE_ERR_CRASH;
    uint64_t total = 0;
    for (size_t q = 0; q < n_inst; q++) total += counts[q];
    // a null array with points to read or ids to write (a resident mesh
    // brings its own coordinates)
    if (total && (!d_out || !d_weights || (!src && !d_points))) return COUPE_ERR_CRASH;
    TraceRange range(n_inst == 1 ? "coupe_hip rcb" : "coupe_hip rcb batch");
    memset(&cx.stats, 0, sizeof cx.stats);
    if (!rot) cx.prof.reset();
    struct CallScope {
        Context &c;
        CallScope(Context &c_, const Rotation *r, const ResidentSource *s) : c(c_) {
            c.rot = r;
            for (int k = 0; k < 3; k++) c.slot0[k] = s ? s->x[k] : nullptr;
        }
        ~CallScope() {
            c.rot = nullptr;
            for (int k = 0; k < 3; k++) c.slot0[k] = nullptr;
        }
    } call_scope(cx, rot, src);
    cx.swept_bytes = 0.0;
    // switches of the call: the autotuner's for its class, if any
    struct TuneScope {   // the context's own switches come back after the call
        Context &c;
        int bins, wg, passes;
        bool full, graphs;
        uint32_t chunk;
        TuneScope(Context &c_)
            : c(c_), bins(c_.max_bins), wg(c_.wg_cap), passes(c_.search_passes), full(c_.full_first),
              graphs(c_.graphs), chunk(c_.search_chunk) {}
        ~TuneScope() {
            c.max_bins = bins;
            c.wg_cap = wg;
            c.search_passes = passes;
            c.full_first = full;
            c.graphs = graphs;
            c.search_chunk = chunk;
        }
    } tune_scope(cx);
    try {
        BatchPlan B = plan_batch(counts, levels, tols, n_inst, dim, src_idx, src_pt);
        bool tuned = false;
        if (!cx.tuning_now) {
            const uint64_t tk = tune_key(B, dim);
            for (const auto &e : cx.tuned)
                if (e.first == tk) {
                    apply_tuning(cx, e.second);
                    tuned = true;
                }
        }
        cx.stats.points = B.n;
        if (B.n == 0) return COUPE_ERR_OK;            // BoundingBox::from_points is None (:685-688)
        if (B.n > 0xffffffffull) return COUPE_ERR_ALLOC;
        if (B.inst.size() > 1 && !B.dense) return COUPE_ERR_BAD_TYPE;
        check(hipSetDevice(cx.device));
        if (!cx.host_scratch) check(hipHostMalloc(&cx.host_scratch, kHostScratch, hipHostMallocDefault));
        if (cx.cu_count <= 0) {
            check(hipDeviceGetAttribute(&cx.cu_count, hipDeviceAttributeMultiprocessorCount, cx.device));
            if (cx.cu_count <= 0) cx.cu_count = 1;
        }
        auto run = [&](int spec, hipStream_t s) {
            return dim == 2 ? run_dim<2>(cx, s, d_out, B, d_points, wkind, d_weights, spec)
                            : run_dim<3>(cx, s, d_out, B, d_points, wkind, d_weights, spec);
        };
        // Passes queued per level: the configured depth, or fewer (never
        // under 2) when this context's last call needed fewer (its deepest
        // search's decisions); each queued pass costs two launches even when
        // its gate is closed.
        // A shape run before queues as many as its deepest search needed then
        // (2..16; a level queues no more than it needed).
        int spec = cx.search_passes;
        const Hints *seen = cx.hints_for(B.shape);
        const int passes_hint = seen ? seen->passes_hint : 0;
        if (passes_hint > 0 && spec > 0) spec = std::max(2, std::min(16, passes_hint));
        // A shape this context has run (hints known, buffers sized) runs as
        // one captured graph: 0 done, 1 a search outran its queued passes,
        // 2 not run that way (or its data needs another weight mode).
        int how = 2;
        if (cx.graphs && !rot && !cx.prof.enabled && !t_arena && spec > 0 && cx.bins_clean && seen &&
            seen->weight_hint >= 0 && seen->level_passes.size() == B.max_levels)
            how = run_captured(cx, st, B, dim, wkind, d_out, d_points, d_weights, spec, run);
        if (how == 3) how = 2;   // first sight of the graph key: eager
        const bool unfinished = how == 1 || (how == 2 && run(spec, st));
        if (unfinished) {
            // a split search needed more passes than were queued: again,
            // checking every pass on the host
            cx.stats = coupe_hip_stats{};
            cx.stats.points = B.n;
            cx.stats.reruns = 1;
            run(0, st);
        }
        cx.remember(B.shape).passes_hint = (int)cx.stats.max_passes;
        cx.stats.tuned = tuned ? 1u : 0u;
        cx.stats.tuning = current_tuning(cx);
        cx.prof.collect();
        cx.prof.out[COUPE_HIP_K_HIST].algo_bytes += cx.swept_bytes;   // first passes + standalone passes
    } catch (const HipFailure &f) {
        cx.bins_clean = false;
        (void)hipGetLastError();
        return map_hip_error(f.err);
    } catch (const std::bad_alloc &) {
        cx.bins_clean = false;
        return COUPE_ERR_ALLOC;
    } catch (const std::length_error &) {
        cx.bins_clean = false;
        return COUPE_ERR_ALLOC;
    }
    cx.stats.device_bytes = cx.device_bytes();
    return COUPE_ERR_OK;
}

enum coupe_err rcb_batch_mixed(Context &cx, hipStream_t st, uint64_t *d_out, const double *d_points, int wkind,
                               const void *d_weights, const uint64_t *counts, const size_t *dims, const size_t *levels,
                               const double *tols, size_t n_inst) {
    std::vector<uint64_t> cnt[2], src[2], src_pt[2];
    std::vector<size_t> lev[2];
    std::vector<double> tol[2];
    uint64_t idx = 0, pt = 0;
    for (size_t q = 0; q < n_inst; q++) {
        if (dims[q] != 2 && dims[q] != 3) return COUPE_ERR_BAD_DIMENSION;
        // a group of one instance may be deep (node-record descent); it then
        // writes its ids at its working positions, so only the whole call may be
        if (n_inst > 1 && levels[q] > (size_t)kDenseLevels) return COUPE_ERR_BAD_TYPE;
        const int g = (int)dims[q] - 2;
        cnt[g].push_back(counts[q]);
        lev[g].push_back(levels[q]);
        tol[g].push_back(tols[q]);
        src[g].push_back(idx);
        src_pt[g].push_back(pt);
        idx += counts[q];
        pt += counts[q] * dims[q];
    }
    for (int g = 0; g < 2; g++) {
        if (cnt[g].empty()) continue;
        const enum coupe_err e = rcb_batch(cx, st, d_out, (size_t)g + 2, d_points, wkind, d_weights, cnt[g].data(),
                                           lev[g].data(), tol[g].data(), cnt[g].size(), nullptr, nullptr,
                                           src[g].data(), src_pt[g].data());
        if (e != COUPE_ERR_OK) return e;
    }
    return COUPE_ERR_OK;
}

enum coupe_err autotune(Context &cx, hipStream_t st, uint64_t *d_out, const double *d_points, int wkind,
                        const void *d_weights, const uint64_t *counts, const size_t *dims, const size_t *levels,
                        const double *tols, size_t n_inst, uint32_t trials, coupe_hip_tuning *chosen,
                        double *best_ms) {
    if (trials < 1) trials = 1;
    // the classes of the call's groups (one per dimension present)
    std::vector<uint64_t> keys;
    for (size_t d = 2; d <= 3; d++) {
        std::vector<uint64_t> c;
        std::vector<size_t> l;
        std::vector<double> t;
        for (size_t q = 0; q < n_inst; q++)
            if (dims[q] == d) {
                c.push_back(counts[q]);
                l.push_back(levels[q]);
                t.push_back(tols[q]);
            }
        if (!c.empty()) keys.push_back(tune_key(plan_batch(c.data(), l.data(), t.data(), c.size(), d), d));
    }
    struct Restore {   // the context's own switches come back after the tuning
        Context &c;
        int bins, wg, passes;
        bool full, graphs;
        uint32_t chunk;
        ~Restore() {
            c.max_bins = bins;
            c.wg_cap = wg;
            c.search_passes = passes;
            c.full_first = full;
            c.graphs = graphs;
            c.search_chunk = chunk;
            c.tuning_now = false;
        }
    } restore{cx, cx.max_bins, cx.wg_cap, cx.search_passes, cx.full_first, cx.graphs, cx.search_chunk};
    cx.tuning_now = true;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return COUPE_ERR_CRASH;
    }
    enum coupe_err err = COUPE_ERR_OK;
    // mean ms of `trials` calls under t, after a warm-up call (hints, graph)
    auto measure = [&](const coupe_hip_tuning &t, double &ms) {
        apply_tuning(cx, t);
        err = rcb_batch_mixed(cx, st, d_out, d_points, wkind, d_weights, counts, dims, levels, tols, n_inst);
        if (err != COUPE_ERR_OK) return false;
        if (hipEventRecord(e0, st) != hipSuccess) return false;
        for (uint32_t k = 0; k < trials && err == COUPE_ERR_OK; k++)
            err = rcb_batch_mixed(cx, st, d_out, d_points, wkind, d_weights, counts, dims, levels, tols, n_inst);
        if (err != COUPE_ERR_OK || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return false;
        float f = 0.f;
        if (hipEventElapsedTime(&f, e0, e1) != hipSuccess) return false;
        ms = f / trials;
        return true;
    };
    coupe_hip_tuning best = current_tuning(cx);
    double best_t = 0.0;
    bool ok = measure(best, best_t);
    // one switch at a time from the best so far; a candidate must win by 1%
    auto try_set = [&](auto field, std::initializer_list<uint32_t> values) {
        for (uint32_t v : values) {
            if (!ok) return;
            coupe_hip_tuning t = best;
            t.*field = v;
            if (t.*field == best.*field) continue;
            double m = 0.0;
            ok = measure(t, m);
            if (ok && m < 0.99 * best_t) {
                best = t;
                best_t = m;
            }
        }
    };
    try_set(&coupe_hip_tuning::graphs, {1, 0});
    try_set(&coupe_hip_tuning::bins, {256, 128, 64, 32});
    try_set(&coupe_hip_tuning::workgroups_per_cu, {0, 4, 2, 1});
    try_set(&coupe_hip_tuning::full_first, {1, 0});
    try_set(&coupe_hip_tuning::search_chunk, {2, 1, 4});
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) return err != COUPE_ERR_OK ? err : COUPE_ERR_CRASH;
    for (uint64_t k : keys) {
        bool found = false;
        for (auto &e : cx.tuned)
            if (e.first == k) {
                e.second = best;
                found = true;
            }
        if (!found) cx.tuned.emplace_back(k, best);
    }
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    // the ids of the chosen switches (the last call ran a candidate)
    cx.tuning_now = false;
    return rcb_batch_mixed(cx, st, d_out, d_points, wkind, d_weights, counts, dims, levels, tols, n_inst);
}

enum coupe_err rcb_device(Context &cx, hipStream_t st, uint64_t *d_out, uint64_t n, size_t dim,
                          const double *d_points, int wkind, const void *d_weights, size_t iter_count,
                          double tol) {
    return rcb_batch(cx, st, d_out, dim, d_points, wkind, d_weights, &n, &iter_count, &tol, 1);
}

template <typename F>
void Context::for_each_buf(F &&f) {
    DevBuf *all[] = {&cx[0][0], &cx[0][1], &cx[0][2], &cx[1][0], &cx[1][1], &cx[1][2], &cx[2][0], &cx[2][1],
                     &cx[2][2], &w[0], &w[1], &w16, &parts[0], &parts[1], &parts[2], &tiles[0], &tiles[1],
                     &tiles[2], &pleft[0], &pleft[1], &pbelow2[0], &pbelow2[1], &fwd_cnt, &fwd_w, &fwd_mn,
                     &tile_bins2, &tile4, &bcnt, &bw, &bmin, &xcnt, &xw, &xmin, &nchild, &ntiles, &bsum, &nodes, &leaf,
                     &heap_split, &instances, &tiles0, &below, &tile_left, &loads, &rib_moments, &tile_bins,
                     &gparts[0], &gparts[1], &gaxis, &gchunk, &gcount, &gctr, &gsat, &gcol, &grow, &gheap, &gscan,
                     &ired, &counters, &in_points, &in_weights, &out_ids};
    for (DevBuf *d : all) f(*d);
}

void store_working_copy(hipStream_t st, size_t dim, int wkind, const double *pts, const void *w, uint64_t first,
                        uint64_t count, float *const x[3], void *mw) {
    if (count == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((count + kBlock - 1) / kBlock, 4096);
    auto launch = [&](auto wtag) {
        using WT = decltype(wtag);
        if (dim == 2)
            k_mesh_store<2, WT><<<g, kBlock, 0, st>>>(pts, (const WT *)w, first, count, x[0], x[1], x[2], (WT *)mw);
        else
            k_mesh_store<3, WT><<<g, kBlock, 0, st>>>(pts, (const WT *)w, first, count, x[0], x[1], x[2], (WT *)mw);
    };
    if (wkind == COUPE_INT) launch(int32_t{});
    else if (wkind == COUPE_INT64) launch(int64_t{});
    else launch(double{});
}

uint64_t Context::device_bytes() const {
    uint64_t b = 0;
    const_cast<Context *>(this)->for_each_buf([&](DevBuf &d) { b += d.bytes; });
    return b;
}

void Context::release_all() {
    drop_graphs();
    for (int b = 0; b < kStageBufs; b++) {
        if (stage_buf[b]) (void)hipHostFree(stage_buf[b]);
        if (stage_ev[b]) (void)hipEventDestroy(stage_ev[b]);
        stage_buf[b] = nullptr;
        stage_ev[b] = nullptr;
    }
    if (stage_st) (void)hipStreamDestroy(stage_st);
    stage_st = nullptr;
    for_each_buf([](DevBuf &d) { d.release(); });
    if (host_scratch && !borrowed_host) (void)hipHostFree(host_scratch);
    host_scratch = nullptr;
    prof.release();
}

void Context::forget_all() {
    for_each_buf([](DevBuf &d) { d.forget(); });
    bins_clean = false;
}

}  // namespace coupe_hip

// ---------------------------------------------------------------------------
// coupe_hip_rcb_device_noalloc (include/coupe_hip.h): coupe_hip_rcb_device
// with every work buffer carved from one static device array per GPU, for
// callers that cannot allocate.  The host name `scratch` is only a stand-in:
// its device address comes from hipGetSymbolAddress, once per device, and
// only that address reaches the kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[COUPE_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNoAllocDevices = 64;

struct NoAllocDevice {
    std::mutex lock;                  // calls on one device run one at a time
    char *scratch_ptr = nullptr;      // device address of `scratch` on this device
    coupe_hip::Context cx;            // hints and bookkeeping; every buffer is a carve of scratch
    alignas(256) char host[coupe_hip::kHostScratch];   // readbacks (pageable; no pinned allocation)
};

NoAllocDevice g_noalloc[kNoAllocDevices];

}  // namespace

extern "C" enum coupe_err coupe_hip_rcb_device_noalloc(void *stream, uint64_t *d_partition, uintptr_t n,
                                                       uintptr_t dimension, const double *d_points,
                                                       enum coupe_type weight_type, const void *d_weights,
                                                       uintptr_t iter_count, double tolerance) {
    if (n > COUPE_HIP_NOALLOC_MAX_POINTS || iter_count > COUPE_HIP_NOALLOC_MAX_ITER_COUNT) return COUPE_ERR_ALLOC;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) return COUPE_ERR_CRASH;
    NoAllocDevice &D = g_noalloc[dev];
    std::lock_guard<std::mutex> g(D.lock);
    if (!D.scratch_ptr) {   // first call on this device, which is current
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return COUPE_ERR_CRASH;
        }
        D.scratch_ptr = (char *)p;
        D.cx.device = dev;
        coupe_hip::configure_from_env(D.cx);
        D.cx.host_scratch = D.host;
        D.cx.borrowed_host = true;
    }
    // the previous call's carves are dropped; this call carves from offset 0
    D.cx.forget_all();
    coupe_hip::Arena arena{D.scratch_ptr, (size_t)COUPE_HIP_NOALLOC_SCRATCH_BYTES, 0};
    coupe_hip::ArenaScope scope(&arena);
    return coupe_hip::rcb_device(D.cx, (hipStream_t)stream, d_partition, n, dimension, d_points, weight_type,
                                 d_weights, iter_count, tolerance);
}
