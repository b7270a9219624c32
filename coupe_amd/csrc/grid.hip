// grid.hip -- RCB on cartesian grids (coupe::Grid::rcb,
// coupe/src/cartesian/mod.rs:127-180, coupe/src/cartesian/rcb.rs).
//
// The grid is split level by level.  At level l every non-empty subgrid
// (a box of cells) is cut across axis (1 + l) mod D at the weighted median
// of its axis sums (the sums of its weights over the cross-section at each
// index of that axis):
//   k_grid_counts + scans  chunks of 4096 cells and axis-sum slots per subgrid
//   k_grid_axis            one pass over the cells of every subgrid: axis sums
//                          in LDS, flushed with atomics (exact for integers)
//   k_grid_median          one wavefront per subgrid replays weighted_median
//                          (rcb.rs:52-99) with the rayon pool size as parameter
//   k_grid_children        the two halves (empty ones dropped), their weights
//                          and id prefixes (left 2p, right 2p + 1)
// and at the end k_grid_fill writes each leaf subgrid's id into its cells.
// Integer weights (and integer-valued doubles whose sums stay below 2^53)
// give the reference's ids bit for bit; other doubles sum in another order.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "rcb_device.h"
#include "rcb_engine.h"

namespace coupe_hip {

namespace {

constexpr uint32_t kGChunk = 4096;     // cells per work unit
constexpr uint32_t kGMaxThreads = 1024;
constexpr double kGridTolerance = 0.01;   // rcb.rs:44

void ck(hipError_t e) {
    if (e != hipSuccess) throw HipFailure{e};
}

struct GridPart {
    uint32_t off[3], size[3];
    uint64_t cells;
    uint64_t sum;     // weight (W bits: int64 value, or f64 bits)
    uint64_t pid;     // part id prefix
    uint64_t left;    // left weight of the split (W bits)
    uint32_t split;   // absolute split index along the level's axis
    uint32_t known;   // sum is valid (false for the root: taken from its axis sums)
};

// W arithmetic in its 64-bit carrier (K = 0 int32 wrapping, 1 int64, 2 f64).
template <int K> struct GW {
    __host__ __device__ static uint64_t add(uint64_t a, uint64_t b) {
        if constexpr (K == 2) return __builtin_bit_cast(uint64_t, __builtin_bit_cast(double, a) + __builtin_bit_cast(double, b));
        else if constexpr (K == 0) return (uint64_t)(int64_t)(int32_t)(uint32_t)(a + b);
        else return a + b;
    }
    __host__ __device__ static uint64_t sub(uint64_t a, uint64_t b) {
        if constexpr (K == 2) return __builtin_bit_cast(uint64_t, __builtin_bit_cast(double, a) - __builtin_bit_cast(double, b));
        else if constexpr (K == 0) return (uint64_t)(int64_t)(int32_t)(uint32_t)(a - b);
        else return a - b;
    }
    __host__ __device__ static bool lt(uint64_t a, uint64_t b) {
        if constexpr (K == 2) return __builtin_bit_cast(double, a) < __builtin_bit_cast(double, b);
        else return (int64_t)a < (int64_t)b;
    }
    __host__ __device__ static double f64(uint64_t a) {
        if constexpr (K == 2) return __builtin_bit_cast(double, a);
        else return (double)(int64_t)a;
    }
    // Rust's `as` from f64: truncation, saturation, NaN -> 0
    __host__ __device__ static uint64_t from_f64(double x) {
        if constexpr (K == 2) return __builtin_bit_cast(uint64_t, x);
        const double lo = K == 0 ? -2147483648.0 : -9223372036854775808.0;
        const double hi = K == 0 ? 2147483647.0 : 9223372036854775807.0;
        int64_t r;
        if (x != x) r = 0;
        else if (x <= lo) r = K == 0 ? (int64_t)INT32_MIN : INT64_MIN;
        else if (x >= hi) r = K == 0 ? (int64_t)INT32_MAX : INT64_MAX;
        else r = (int64_t)x;
        return (uint64_t)r;
    }
};

__device__ inline uint32_t upper_bound_u64(const uint64_t *a, uint32_t n, uint64_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Position of local cell c of a box in memory order (x fastest).
__device__ inline void local_pos(const GridPart &P, uint64_t c, uint32_t (&p)[3]) {
    p[0] = (uint32_t)(c % P.size[0]);
    const uint64_t r = c / P.size[0];
    p[1] = (uint32_t)(r % P.size[1]);
    p[2] = (uint32_t)(r / P.size[1]);
}

__global__ void __launch_bounds__(kBlock)
k_grid_counts(const GridPart *__restrict__ parts, const uint32_t *n_parts, int coord, uint64_t *__restrict__ chunks,
              uint64_t *__restrict__ axis) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_parts) return;
    const GridPart &P = parts[s];
    chunks[s] = (P.cells + kGChunk - 1) / kGChunk;
    if (axis) axis[s] = coord >= 0 ? P.size[coord] : 0;
}

// Axis sums of every subgrid along `coord`.  Accumulated in LDS per chunk
// (slot = index along the axis minus the chunk's first index), flushed with
// 64-bit atomics: integer sums are exact in any order.
template <int K, typename WIN>
__global__ void __launch_bounds__(kBlock)
k_grid_axis(const GridPart *__restrict__ parts, const uint32_t *n_parts, const uint64_t *__restrict__ chunk_off,
            const uint64_t *chunk_total, const uint64_t *__restrict__ axis_off, int coord, uint64_t gx, uint64_t gy,
            const WIN *__restrict__ w, uint64_t *__restrict__ axis) {
    using Acc = typename std::conditional<K == 2, double, unsigned long long>::type;
    __shared__ Acc sb[kGChunk];
    const uint32_t np = *n_parts;
    const uint64_t nc = *chunk_total;
    for (uint64_t c = blockIdx.x; c < nc; c += gridDim.x) {
        const uint32_t s = upper_bound_u64(chunk_off, np, c) - 1;
        const GridPart P = parts[s];
        const uint64_t c0 = (c - chunk_off[s]) * kGChunk;
        const uint64_t c1 = c0 + kGChunk < P.cells ? c0 + kGChunk : P.cells;
        uint32_t p0[3], p1[3];
        local_pos(P, c0, p0);
        local_pos(P, c1 - 1, p1);
        // Slots: x for axis 0; for axis 1 the y range of the chunk when it
        // stays in one z slab (y then only grows), else every y; for axis 2
        // the z range (z only grows along the chunk).
        uint32_t first, range;
        if (coord == 0 || (coord == 1 && p0[2] != p1[2])) {
            first = 0;
            range = P.size[coord];
        } else {
            first = p0[coord];
            range = p1[coord] - p0[coord] + 1;
        }
        const bool lds = range <= kGChunk;
        Acc *g = reinterpret_cast<Acc *>(axis + axis_off[s]);
        if (lds)
            for (uint32_t j = threadIdx.x; j < range; j += kBlock) sb[j] = (Acc)0;
        __syncthreads();
        for (uint64_t cc = c0 + threadIdx.x; cc < c1; cc += kBlock) {
            uint32_t p[3];
            local_pos(P, cc, p);
            const uint64_t x = P.off[0] + p[0], y = P.off[1] + p[1], z = P.off[2] + p[2];
            const uint64_t idx = x + gx * (y + gy * z);
            Acc v;
            if constexpr (K == 2) v = (double)w[idx];
            else v = (Acc)(int64_t)w[idx];
            const uint32_t b = p[coord];
            if (lds) atomicAdd(&sb[b - first], v);
            else atomicAdd(&g[b], v);
        }
        __syncthreads();
        if (lds)
            for (uint32_t j = threadIdx.x; j < range; j += kBlock)
                if (sb[j] != (Acc)0) atomicAdd(&g[first + j], sb[j]);
        __syncthreads();
    }
}

// weighted_median (rcb.rs:52-99), one wavefront per subgrid: chunk sums in
// the lanes (each sequential over its chunk, as fold_chunks), the prefix over
// the chunks and the decisions in lane 0, in the reference's order.  The
// axis sums read are zeroed for the next level.
template <int K>
__global__ void __launch_bounds__(64)
k_grid_median(GridPart *__restrict__ parts, const uint32_t *n_parts, int coord, const uint64_t *__restrict__ axis_off,
              uint64_t *__restrict__ axis, uint32_t threads) {
    using Wt = GW<K>;
    __shared__ uint64_t cs_sum[2 * kGMaxThreads + 2];
    __shared__ uint64_t st[4];   // min, max, left, done
    const uint32_t s = blockIdx.x;
    if (s >= *n_parts) return;
    GridPart *P = &parts[s];
    const int lane = threadIdx.x;
    uint64_t *a = axis + axis_off[s];
    const uint32_t m = P->size[coord];
    // the carrier of the sums: integers as wrapped int64 / int32, f64 bits
    auto norm = [&](uint64_t v) -> uint64_t {
        if constexpr (K == 0) return (uint64_t)(int64_t)(int32_t)(uint32_t)v;
        else return v;
    };
    uint64_t total = P->sum;
    if (!P->known) {   // the root: its weight from its axis sums
        uint64_t t = 0;
        if constexpr (K == 2) {
            double d = 0.0;
            for (uint32_t i = lane; i < m; i += 64) d += __builtin_bit_cast(double, a[i]);
            d = wave_sum(d);
            t = __builtin_bit_cast(uint64_t, d);
        } else {
            uint64_t u = 0;
            for (uint32_t i = lane; i < m; i += 64) u += a[i];
            t = norm(wave_sum(u));
        }
        total = t;
        if (lane == 0) { P->sum = total; P->known = 1; }
    }
    const double ideal = Wt::f64(total) / 2.0;
    const uint64_t minp = Wt::from_f64(ideal * (1.0 - kGridTolerance));
    const uint64_t maxp = Wt::from_f64(ideal * (1.0 + kGridTolerance));
    uint64_t mn = 0, mx = m, left = 0;
    if (lane == 0) { st[0] = mn; st[1] = mx; st[2] = left; st[3] = 0; }
    __syncthreads();
    uint64_t pos_out = 0, left_out = 0;
    for (;;) {
        mn = st[0]; mx = st[1]; left = st[2];
        uint64_t cs = (mx - mn) / threads;
        if (cs < 1) cs = 1;
        const uint64_t nch = (mx - mn + cs - 1) / cs;
        for (uint64_t k = lane; k < nch; k += 64) {
            const uint64_t b = mn + k * cs, e = b + cs < mx ? b + cs : mx;
            uint64_t acc = 0;
            for (uint64_t i = b; i < e; i++) acc = Wt::add(acc, norm(a[i]));
            cs_sum[k] = acc;
        }
        __syncthreads();
        if (lane == 0) {
            uint64_t prefix = 0, nmin = mn, nmax = mx, nleft = left;
            int done = 0;
            for (uint64_t k = 0; k < nch; k++) {
                const uint64_t start = mn + k * cs;
                const uint64_t pcw = Wt::add(left, prefix);
                prefix = Wt::add(prefix, cs_sum[k]);
                if (Wt::lt(pcw, minp)) {
                    nmin = start;
                    nleft = pcw;
                } else if (Wt::lt(maxp, pcw)) {
                    nmax = start;
                    break;
                } else {
                    pos_out = start;
                    left_out = pcw;
                    done = 1;
                    break;
                }
            }
            if (!done && (nmin + 1 >= nmax || (nmin == mn && nmax == mx))) {
                pos_out = nmin;
                left_out = nleft;
                done = 1;
            }
            st[0] = nmin; st[1] = nmax; st[2] = nleft; st[3] = (uint64_t)done;
            if (done) {
                P->split = (uint32_t)(P->off[coord] + pos_out);
                P->left = left_out;
            }
        }
        __syncthreads();
        if (st[3]) break;
    }
    for (uint32_t i = lane; i < m; i += 64) a[i] = 0;   // clean for the next level
}

// Children of every subgrid: counts (0, 1 or 2 non-empty halves) for the
// scan, then the halves themselves.
__global__ void __launch_bounds__(kBlock)
k_grid_nchild(const GridPart *__restrict__ parts, const uint32_t *n_parts, int coord, uint64_t *__restrict__ nchild) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_parts) return;
    const GridPart &P = parts[s];
    const uint64_t lsz = P.split - P.off[coord], rsz = P.size[coord] - lsz;
    const uint64_t per = P.cells / P.size[coord];
    nchild[s] = (lsz * per > 0) + (rsz * per > 0);
}

template <int K>
__global__ void __launch_bounds__(kBlock)
k_grid_children(const GridPart *__restrict__ parts, const uint32_t *n_parts, int coord,
                const uint64_t *__restrict__ child_off, const uint64_t *total, GridPart *__restrict__ next,
                uint32_t *next_n) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s == 0) *next_n = (uint32_t)*total;
    if (s >= *n_parts) return;
    const GridPart &P = parts[s];
    const uint32_t lsz = P.split - P.off[coord], rsz = P.size[coord] - lsz;
    const uint64_t per = P.cells / P.size[coord];
    uint64_t o = child_off[s];
    for (int side = 0; side < 2; side++) {
        const uint32_t sz = side == 0 ? lsz : rsz;
        if ((uint64_t)sz * per == 0) continue;
        GridPart C = P;
        C.size[coord] = sz;
        if (side == 1) C.off[coord] = P.split;
        C.cells = (uint64_t)sz * per;
        C.sum = side == 0 ? P.left : GW<K>::sub(P.sum, P.left);
        C.pid = 2 * P.pid + (uint64_t)side;
        C.left = 0;
        C.split = 0;
        C.known = 1;
        next[o++] = C;
    }
}

// Each leaf subgrid's id into its cells.
__global__ void __launch_bounds__(kBlock)
k_grid_fill(const GridPart *__restrict__ parts, const uint32_t *n_parts, const uint64_t *__restrict__ chunk_off,
            const uint64_t *chunk_total, uint64_t gx, uint64_t gy, uint64_t *__restrict__ out) {
    const uint32_t np = *n_parts;
    const uint64_t nc = *chunk_total;
    for (uint64_t c = blockIdx.x; c < nc; c += gridDim.x) {
        const uint32_t s = upper_bound_u64(chunk_off, np, c) - 1;
        const GridPart P = parts[s];
        const uint64_t c0 = (c - chunk_off[s]) * kGChunk;
        const uint64_t c1 = c0 + kGChunk < P.cells ? c0 + kGChunk : P.cells;
        for (uint64_t cc = c0 + threadIdx.x; cc < c1; cc += kBlock) {
            uint32_t p[3];
            local_pos(P, cc, p);
            const uint64_t x = P.off[0] + p[0], y = P.off[1] + p[1], z = P.off[2] + p[2];
            out[x + gx * (y + gy * z)] = P.pid;
        }
    }
}

// The rayon pool size coupe would run with: RAYON_NUM_THREADS, else the
// number of logical CPUs.
uint32_t default_threads() {
    const char *v = getenv("RAYON_NUM_THREADS");
    if (v && *v) {
        const long k = strtol(v, nullptr, 10);
        if (k > 0) return (uint32_t)k;
    }
    const unsigned h = std::thread::hardware_concurrency();
    return h ? h : 1;
}

struct GridCounters {
    uint32_t n_parts[2];
    uint32_t pad[2];
    uint64_t chunk_total;
    uint64_t child_total;
};

template <int K, typename WIN>
void grid_run(Context &cx, hipStream_t st, size_t dim, const uint64_t *size, const WIN *w, size_t L, uint32_t T,
              uint64_t *out) {
    uint64_t sz[3] = {size[0], size[1], dim == 3 ? size[2] : 1};
    const uint64_t cells = sz[0] * sz[1] * sz[2];
    // bounds: parts per level, axis slots per level, chunks per level
    uint64_t parts_cap = 1, axis_cap = 1;
    for (size_t l = 0; l <= L; l++) {
        const uint64_t p = l < 63 ? std::min<uint64_t>((uint64_t)1 << l, cells) : cells;
        parts_cap = std::max(parts_cap, p);
        if (l < L) axis_cap = std::max(axis_cap, std::min<uint64_t>(cells, p * sz[(1 + l) % dim]));
    }
    if (parts_cap > 0x7fffffffull) throw HipFailure{hipErrorOutOfMemory};
    cx.gparts[0].ensure(parts_cap * sizeof(GridPart));
    cx.gparts[1].ensure(parts_cap * sizeof(GridPart));
    const bool fresh = cx.gaxis.bytes < axis_cap * 8;
    cx.gaxis.ensure(axis_cap * 8);
    if (fresh || !cx.grid_clean) ck(hipMemsetAsync(cx.gaxis.ptr, 0, cx.gaxis.bytes, st));
    cx.gchunk.ensure(parts_cap * 8);
    cx.gcount.ensure(parts_cap * 8);
    cx.gctr.ensure(sizeof(GridCounters));
    GridCounters *ctr = (GridCounters *)cx.gctr.ptr;
    ck(hipMemsetAsync(ctr, 0, sizeof(GridCounters), st));
    cx.grid_clean = false;   // until the medians have cleared every axis sum again
    // root subgrid
    GridPart root;
    memset(&root, 0, sizeof root);
    for (int k = 0; k < 3; k++) root.size[k] = (uint32_t)sz[k];
    root.cells = cells;
    root.known = 0;
    ck(hipMemcpyAsync(cx.gparts[0].ptr, &root, sizeof root, hipMemcpyHostToDevice, st));
    const uint32_t one = 1;
    ck(hipMemcpyAsync(&ctr->n_parts[0], &one, sizeof one, hipMemcpyHostToDevice, st));

    uint64_t *chunk_off = (uint64_t *)cx.gchunk.ptr, *aux = (uint64_t *)cx.gcount.ptr;
    uint64_t *axis = (uint64_t *)cx.gaxis.ptr;
    int cur = 0;
    uint64_t parts_ub = 1;
    auto plan_chunks = [&](GridPart *parts, uint32_t *np, int coord, uint64_t *axis_off) {
        k_grid_counts<<<(unsigned)((parts_ub + kBlock - 1) / kBlock), kBlock, 0, st>>>(parts, np, coord, chunk_off,
                                                                                      axis_off);
        ck(hipGetLastError());
        scan_u64(cx, st, chunk_off, np, parts_ub, &ctr->chunk_total);
        if (axis_off) scan_u64(cx, st, axis_off, np, parts_ub, nullptr);
    };
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_grid_axis<K, WIN>, kBlock, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 2;
    for (size_t l = 0; l < L; l++) {
        const int coord = (int)((1 + l) % dim);
        GridPart *parts = (GridPart *)cx.gparts[cur].ptr;
        GridPart *next = (GridPart *)cx.gparts[cur ^ 1].ptr;
        uint32_t *np = &ctr->n_parts[cur];
        uint64_t *axis_off = aux;
        plan_chunks(parts, np, coord, axis_off);
        const uint64_t chunks_ub = cells / kGChunk + parts_ub + 1;
        const unsigned g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(chunks_ub, (uint64_t)cx.cu_count * per_cu));
        int r = cx.prof.begin(st, COUPE_HIP_K_HIST, (double)cells * sizeof(WIN));
        k_grid_axis<K, WIN><<<g, kBlock, 0, st>>>(parts, np, chunk_off, &ctr->chunk_total, axis_off, coord, sz[0],
                                                  sz[1], w, axis);
        ck(hipGetLastError());
        cx.prof.end(st, r);
        r = cx.prof.begin(st, COUPE_HIP_K_DECIDE, 0.0);
        k_grid_median<K><<<(unsigned)parts_ub, 64, 0, st>>>(parts, np, coord, axis_off, axis, T);
        ck(hipGetLastError());
        cx.prof.end(st, r);
        r = cx.prof.begin(st, COUPE_HIP_K_OTHER, 0.0);
        k_grid_nchild<<<(unsigned)((parts_ub + kBlock - 1) / kBlock), kBlock, 0, st>>>(parts, np, coord, aux);
        ck(hipGetLastError());
        scan_u64(cx, st, aux, np, parts_ub, &ctr->child_total);
        k_grid_children<K><<<(unsigned)((parts_ub + kBlock - 1) / kBlock), kBlock, 0, st>>>(
            parts, np, coord, aux, &ctr->child_total, next, &ctr->n_parts[cur ^ 1]);
        ck(hipGetLastError());
        cx.prof.end(st, r);
        cur ^= 1;
        parts_ub = std::min<uint64_t>(2 * parts_ub, std::min<uint64_t>(cells, parts_cap));
    }
    GridPart *leaves = (GridPart *)cx.gparts[cur].ptr;
    plan_chunks(leaves, &ctr->n_parts[cur], -1, nullptr);
    const uint64_t chunks_ub = cells / kGChunk + parts_ub + 1;
    const unsigned g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(chunks_ub, (uint64_t)cx.cu_count * 8));
    const int r = cx.prof.begin(st, COUPE_HIP_K_IDS, (double)cells * 8.0);
    k_grid_fill<<<g, kBlock, 0, st>>>(leaves, &ctr->n_parts[cur], chunk_off, &ctr->chunk_total, sz[0], sz[1], out);
    ck(hipGetLastError());
    cx.prof.end(st, r);
    ck(hipStreamSynchronize(st));
    cx.grid_clean = true;
}

}  // namespace

enum coupe_err grid_rcb(Context &cx, hipStream_t st, size_t dim, const uint64_t *size, int wkind, const void *w,
                        size_t iter_count, uint32_t threads, uint64_t *out) {
    if (dim != 2 && dim != 3) return COUPE_ERR_BAD_DIMENSION;
    if (wkind != COUPE_INT && wkind != COUPE_INT64 && wkind != COUPE_DOUBLE) return COUPE_ERR_BAD_TYPE;
    for (size_t k = 0; k < dim; k++)
        if (size[k] == 0 || size[k] > 0xffffffffull) return COUPE_ERR_CRASH;   // NonZeroUsize, u32 positions
    const uint64_t cells = size[0] * size[1] * (dim == 3 ? size[2] : 1);
    if (iter_count > 62) return COUPE_ERR_CRASH;
    const uint32_t T = std::min<uint32_t>(threads ? threads : default_threads(), kGMaxThreads);
    memset(&cx.stats, 0, sizeof cx.stats);
    cx.stats.points = cells;
    cx.stats.levels = (uint32_t)iter_count;
    cx.prof.reset();
    try {
        ck(hipSetDevice(cx.device));
        if (cx.cu_count <= 0) {
            ck(hipDeviceGetAttribute(&cx.cu_count, hipDeviceAttributeMultiprocessorCount, cx.device));
            if (cx.cu_count <= 0) cx.cu_count = 1;
        }
        if (wkind == COUPE_INT) grid_run<0, int32_t>(cx, st, dim, size, (const int32_t *)w, iter_count, T, out);
        else if (wkind == COUPE_INT64) grid_run<1, int64_t>(cx, st, dim, size, (const int64_t *)w, iter_count, T, out);
        else grid_run<2, double>(cx, st, dim, size, (const double *)w, iter_count, T, out);
        cx.prof.collect();
    } catch (const HipFailure &f) {
        cx.grid_clean = false;
        (void)hipGetLastError();
        return map_hip_error(f.err);
    } catch (const std::bad_alloc &) {
        cx.grid_clean = false;
        return COUPE_ERR_ALLOC;
    }
    cx.stats.device_bytes = cx.device_bytes();
    return COUPE_ERR_OK;
}

}  // namespace coupe_hip
