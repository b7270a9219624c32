// grid.hip -- RCB on cartesian grids (coupe::Grid::rcb,
// coupe/src/cartesian/mod.rs:127-180, coupe/src/cartesian/rcb.rs).
//
// The grid is split level by level.  At level l every non-empty subgrid
// (a box of cells) is cut across axis (1 + l) mod D at the weighted median
// of its axis sums (the sums of its weights over the cross-section at each
// index of that axis).  Two paths: the summed-area path further down (calls
// of at most 20 levels whose weights are exactly summable) and the
// per-level path, which reads the weights at every level:
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

// ---------------------------------------------------------------------------
// Summed-area path.  Every axis sum weighted_median asks for is a box sum of
// the weights (the subgrid's cross-sections from its first index to the
// chunk's start), and a box sum is 2^D lookups into the table of inclusive
// prefix sums S(x, y, z) = sum of w over [0, x] x [0, y] x [0, z].  Integer
// weights are summed in 64-bit two's complement (a ring: any order gives the
// wrapping sums of coupe's i32 / i64 loops); integer-valued doubles whose
// absolute sum is at most 2^53 have exact sums in any order, so their
// table, kept in f64, gives coupe's f64 sums bit for bit.  Other doubles
// take the per-level path above.  The table costs two reads of the weights
// and one write of 8 bytes per cell (2D; 3D adds a scan along z), where the
// per-level path reads the weights once per level; each level is then one
// wavefront per subgrid doing the median's rounds on lookups:
//   k_sat_sums        per tile of 16 rows x 512 columns (one wavefront):
//                     column and row sums
//   k_sat_colscan     exclusive scans over the tiles' rows / columns (tall
//   k_sat_chunk*      arrays in chunks of rows: sums, scan, apply)
//   k_sat_linescan    inclusive scans along rows of the row carries
//   k_sat_tiles       per tile: the table from the weights and the carries
//   k_sat_zsum/zapply 3D: the scan along z, in bands of 16 planes
//   k_grid_level      one wavefront per subgrid of a level: weighted_median
//                     on lookups, the split into a dense heap, the children
//   k_grid_fill_heap  each cell walks the heap to its part id (part_of)
// ---------------------------------------------------------------------------
constexpr uint32_t kSatRows = 16;                 // rows per tile
constexpr uint32_t kSatPer = 8;                   // columns per lane of a tile row
constexpr uint32_t kSatCols = 64 * kSatPer;       // columns per tile (one wavefront)
constexpr uint32_t kLineChunk = kBlock * kSatPer;  // values per step of a line scan
constexpr uint32_t kSatGroup = 4;                 // rows whose loads are issued together
constexpr uint32_t kSatZBand = 16;                // planes per band of the z scan
constexpr uint32_t kDenseGridLevels = 20;         // deeper calls take the per-level path
constexpr uint32_t kFillTop = 12;                 // heap levels the fill keeps in LDS
constexpr uint32_t kFillCells = 16;               // cells per thread of the fill

struct SatGeom {
    uint64_t gx, gy, gz;
    uint32_t nb, nseg;     // tiles per plane: nb bands of rows x nseg segments of columns
    uint32_t vec_w;        // every row of the weights starts at a multiple of two weights' size
    uint32_t vec_s;        // every row of the table starts 16-byte aligned
};

struct GridSatFlags {
    double abs_sum;        // sum of |w| (f64 weights)
    uint32_t bad;          // a weight that is not an integer below 2^53 in magnitude
    uint32_t pad;
};

// One subgrid of a level of the summed-area path; any size 0: no cells.
struct GSlot {
    uint32_t off[3], size[3];
    uint64_t sum;          // weight, in the carrier of GW<K>
};

// The table's accumulator: 64-bit two's complement for integer weights,
// f64 for f64 weights (exact while every partial sum is an integer of
// magnitude at most 2^53, which the host checks).
template <int K> using SatAcc = typename std::conditional<K == 2, double, uint64_t>::type;

template <int K, typename WIN>
__device__ inline SatAcc<K> sat_of(WIN v) {
    if constexpr (K == 2) return (double)v;
    else return (uint64_t)(int64_t)v;
}

__device__ inline uint64_t readlane64(uint64_t v, int l) {
    return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l) |
           (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l) << 32;
}
template <typename Acc>
__device__ inline Acc readlane_acc(Acc v, int l) {
    return __builtin_bit_cast(Acc, readlane64(__builtin_bit_cast(uint64_t, v), l));
}

// A tile is kSatRows rows x kSatCols columns, one wavefront's work (the
// wavefronts of a workgroup take tiles independently).  Layout of a tile row
// in the lanes: kSatPer / 2 chunks of 128 columns from the tile's first
// column xw; lane l holds columns xw + 128 c + 2 l and + 1 of chunk c
// (element e = 2 c + {0, 1}), so every load or store instruction of the
// wavefront covers one contiguous run of 128 values.
constexpr uint32_t kSatChunks = kSatPer / 2;
struct SatTile {
    uint64_t z, y0;
    uint64_t xl;           // this lane's first column: xw + 2 * lane
    uint32_t b, s, nr;
    bool whole;            // every column of the wavefront lies in the grid
};
__device__ inline SatTile sat_tile(const SatGeom &g, uint64_t t) {
    const uint64_t per_plane = (uint64_t)g.nb * g.nseg;
    SatTile T;
    T.z = t / per_plane;
    const uint32_t r = (uint32_t)(t - T.z * per_plane);
    T.b = r / g.nseg;
    T.s = r - T.b * g.nseg;
    const uint64_t xw = (uint64_t)T.s * kSatCols;
    T.xl = xw + 2 * (threadIdx.x & 63);
    T.whole = xw + 64 * kSatPer <= g.gx;
    T.y0 = (uint64_t)T.b * kSatRows;
    T.nr = (uint32_t)std::min<uint64_t>(kSatRows, g.gy - T.y0);
    return T;
}
__device__ inline uint64_t sat_col(const SatTile &T, uint32_t e) { return T.xl + (e >> 1) * 128 + (e & 1); }

// The kSatPer values of this lane from row pointer p (columns past gx read
// as 0); pairs as one load when the wavefront's columns are all in the grid
// and rows start at a multiple of the pair's size.
template <bool NT = false, typename T>
__device__ inline void load_lane(const T *p, const SatTile &Tl, uint64_t gx, bool vec, T (&v)[kSatPer]) {
    if (vec && Tl.whole) {
        using V = __attribute__((ext_vector_type(2 * sizeof(T) / 4))) uint32_t;
#pragma unroll
        for (uint32_t c = 0; c < kSatChunks; c++) {
            const V *q_at = reinterpret_cast<const V *>(p + Tl.xl + c * 128);
            const V q = NT ? __builtin_nontemporal_load(q_at) : *q_at;
            __builtin_memcpy(&v[2 * c], &q, sizeof(V));
        }
    } else {
#pragma unroll
        for (uint32_t e = 0; e < kSatPer; e++) {
            const uint64_t x = sat_col(Tl, e);
            v[e] = x < gx ? p[x] : (T)0;
        }
    }
}
template <typename Acc>
__device__ inline void store_lane(Acc *p, const SatTile &Tl, uint64_t gx, bool vec, const Acc (&v)[kSatPer]) {
    if (vec && Tl.whole) {
        using V = __attribute__((ext_vector_type(4))) uint32_t;
#pragma unroll
        for (uint32_t c = 0; c < kSatChunks; c++) {
            V q;
            __builtin_memcpy(&q, &v[2 * c], 16);
            *reinterpret_cast<V *>(p + Tl.xl + c * 128) = q;
        }
    } else {
#pragma unroll
        for (uint32_t e = 0; e < kSatPer; e++) {
            const uint64_t x = sat_col(Tl, e);
            if (x < gx) p[x] = v[e];
        }
    }
}

// Rows r0 .. r0 + B - 1 of a tile (rows past the tile read its last row
// again; the caller masks them), all loads issued before any is used;
// non-temporal when NT.
template <int B, bool NT = false, typename WIN>
__device__ inline void load_rows(const WIN *__restrict__ w, const SatGeom &g, const SatTile &T, uint32_t r0,
                                 WIN (&v)[B][kSatPer]) {
#pragma unroll
    for (int b = 0; b < B; b++) {
        const uint32_t row = std::min<uint32_t>(r0 + b, T.nr - 1);
        load_lane<NT>(w + (T.z * g.gy + T.y0 + row) * g.gx, T, g.gx, g.vec_w, v[b]);
    }
}

// Inclusive scan over the wavefront (every lane active): DPP row shifts
// within rows of 16 lanes (lanes shifted in read 0), then the earlier rows'
// totals by readlane.
template <typename Acc>
__device__ inline Acc wave_scan(Acc v) {
    auto shr = [](Acc a, auto ctrl) {
        return __builtin_bit_cast(Acc, dpp_row64<decltype(ctrl)::value>(__builtin_bit_cast(uint64_t, a)));
    };
    v += shr(v, std::integral_constant<int, 0x111>{});
    v += shr(v, std::integral_constant<int, 0x112>{});
    v += shr(v, std::integral_constant<int, 0x114>{});
    v += shr(v, std::integral_constant<int, 0x118>{});
    auto last = [&](int r) { return __builtin_bit_cast(Acc, rows_last_lane64(__builtin_bit_cast(uint64_t, v), r)); };
    const Acc t0 = last(0), t1 = last(1), t2 = last(2);
    const int row = (threadIdx.x & 63) >> 4;
    return v + (row >= 1 ? t0 : (Acc)0) + (row >= 2 ? t1 : (Acc)0) + (row >= 3 ? t2 : (Acc)0);
}

// Inclusive prefix of a tile row over the wavefront's columns, in place:
// per chunk the lanes' pair sums scanned across the wavefront, after the
// earlier chunks' totals.  Returns the wavefront's total.
template <typename Acc>
__device__ inline Acc wave_row_prefix(Acc (&v)[kSatPer]) {
    Acc carry = 0;
#pragma unroll
    for (uint32_t c = 0; c < kSatChunks; c++) {
        v[2 * c + 1] += v[2 * c];
        const Acc inc = wave_scan(v[2 * c + 1]);
        const Acc ex = carry + (inc - v[2 * c + 1]);
        v[2 * c] += ex;
        v[2 * c + 1] += ex;
        carry += readlane_acc(inc, 63);
    }
    return carry;
}

// Exclusive scan of one value per thread over the workgroup; `total` gets
// the workgroup's sum.  Every thread calls it.
template <typename Acc>
__device__ inline Acc block_excl(Acc v, Acc &total, Acc *lds4) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const Acc inc = wave_scan(v);
    if (lane == 63) lds4[wave] = inc;
    __syncthreads();
    Acc off = 0;
    for (int k = 0; k < wave; k++) off += lds4[k];
    total = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return off + (inc - v);
}

// Per tile: the sums over its rows of each column, and over its columns of
// each row; for f64 weights, whether each is an integer below 2^53 and the
// sum of their magnitudes (into the copy of the workgroup's XCD).
template <int K, typename WIN>
__global__ void __launch_bounds__(kBlock)
k_sat_sums(const WIN *__restrict__ w, SatGeom g, SatAcc<K> *__restrict__ colsum, SatAcc<K> *__restrict__ rowsum,
           GridSatFlags *flags) {
    using Acc = SatAcc<K>;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t n_tiles = (uint64_t)g.nb * g.nseg * g.gz;
    double abs_sum = 0.0;
    uint32_t bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * (kBlock / 64);
    for (uint64_t t = (uint64_t)blockIdx.x * (kBlock / 64) + wave; t < n_tiles; t += stride) {
        const SatTile T = sat_tile(g, t);
        Acc col[kSatPer];
#pragma unroll
        for (uint32_t j = 0; j < kSatPer; j++) col[j] = 0;
        Acc mine = 0;   // lane i: the sum of row i of the tile
#pragma unroll
        for (uint32_t i0 = 0; i0 < kSatRows; i0 += kSatGroup) {
            WIN v[kSatGroup][kSatPer];
            load_rows<kSatGroup, true>(w, g, T, i0, v);   // non-temporal: measured faster here, not in k_sat_tiles
#pragma unroll
            for (uint32_t b = 0; b < kSatGroup; b++) {
                const uint32_t i = i0 + b;
                if (i >= T.nr) continue;
                Acc rowp = 0;
#pragma unroll
                for (uint32_t j = 0; j < kSatPer; j++) {
                    const Acc u = sat_of<K>(v[b][j]);
                    col[j] += u;
                    rowp += u;
                    if constexpr (K == 2) {
                        const double d = (double)v[b][j];
                        abs_sum += fabs(d);
                        bad |= !(d == trunc(d) && fabs(d) < 0x1p53);
                    }
                }
                const Acc r = wave_total(rowp);
                if (lane == (int)i) mine = r;
            }
        }
        store_lane(colsum + (T.z * g.nb + T.b) * g.gx, T, g.gx, g.vec_s, col);
        if (lane < (int)T.nr) rowsum[(T.z * g.nseg + T.s) * g.gy + T.y0 + lane] = mine;
    }
    if constexpr (K == 2) {
        abs_sum = wave_total(abs_sum);
        bad = wave_total(bad);
        GridSatFlags *f = flags + blockIdx.x % kCopies;   // one copy per XCD
        if (lane == 0) {
            if (abs_sum != 0.0) atomicAdd(&f->abs_sum, abs_sum);
            if (bad) atomicOr(&f->bad, 1u);
        }
    }
}

// Exclusive scan over the rows of each column of `batch` row-major arrays of
// `rows` x W, in place: a workgroup takes 64 columns, each of its
// kScanWaves wavefronts a contiguous share of the rows (the columns are few
// next to the lanes of the device: the rows are split to keep more loads in
// flight).
constexpr int kScanWaves = 8;
template <typename Acc>
__global__ void __launch_bounds__(64 * kScanWaves)
k_sat_colscan(Acc *__restrict__ a, uint64_t W, uint32_t rows, uint32_t batch) {
    __shared__ Acc part[kScanWaves][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t ncg = (W + 63) / 64;
    const uint32_t q = (rows + kScanWaves - 1) / kScanWaves;
    const uint32_t r0 = std::min(rows, wave * q), r1 = std::min(rows, r0 + q);
    for (uint64_t blk = blockIdx.x; blk < ncg * batch; blk += gridDim.x) {
        const uint64_t p = blk / ncg, cg = blk - p * ncg;
        const uint64_t col = cg * 64 + lane;
        const bool valid = col < W;
        Acc *base = a + p * rows * W + col;
        Acc sum = 0;
        if (valid) {
#pragma unroll 16
            for (uint32_t r = r0; r < r1; r++) sum += base[(uint64_t)r * W];
        }
        part[wave][lane] = sum;
        __syncthreads();
        Acc carry = 0;
        for (int k = 0; k < wave; k++) carry += part[k][lane];
        if (valid) {
#pragma unroll 16
            for (uint32_t r = r0; r < r1; r++) {
                const Acc v = base[(uint64_t)r * W];
                base[(uint64_t)r * W] = carry;
                carry += v;
            }
        }
        __syncthreads();
    }
}

// The same scan for tall arrays (few columns, many rows), in three steps:
// the sums of chunks of kScanChunk rows (k_sat_chunksum), their exclusive
// scan over the chunks (k_sat_colscan), then each chunk's scan from its
// carry (k_sat_chunkapply); a workgroup per chunk and 64 columns.
constexpr uint32_t kScanChunk = 64 * kScanWaves;
template <typename Acc>
__global__ void __launch_bounds__(64 * kScanWaves)
k_sat_chunksum(const Acc *__restrict__ a, uint64_t W, uint32_t rows, uint32_t batch, Acc *__restrict__ part_sums) {
    __shared__ Acc part[kScanWaves][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t ncg = (W + 63) / 64;
    const uint32_t nch = (rows + kScanChunk - 1) / kScanChunk;
    for (uint64_t blk = blockIdx.x; blk < ncg * nch * batch; blk += gridDim.x) {
        const uint64_t p = blk / (ncg * nch), r = blk - p * ncg * nch, ch = r / ncg, cg = r - ch * ncg;
        const uint64_t col = cg * 64 + lane;
        const uint32_t r0 = (uint32_t)ch * kScanChunk + wave * (kScanChunk / kScanWaves);
        const uint32_t r1 = std::min<uint32_t>(rows, r0 + kScanChunk / kScanWaves);
        Acc sum = 0;
        if (col < W) {
            const Acc *base = a + p * rows * W + col;
#pragma unroll 16
            for (uint32_t q = r0; q < r1; q++) sum += base[(uint64_t)q * W];
        }
        part[wave][lane] = sum;
        __syncthreads();
        if (wave == 0 && col < W) {
            Acc t = 0;
            for (int k = 0; k < kScanWaves; k++) t += part[k][lane];
            part_sums[(p * nch + ch) * W + col] = t;
        }
        __syncthreads();
    }
}
template <typename Acc>
__global__ void __launch_bounds__(64 * kScanWaves)
k_sat_chunkapply(Acc *__restrict__ a, uint64_t W, uint32_t rows, uint32_t batch, const Acc *__restrict__ carries) {
    __shared__ Acc part[kScanWaves][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t ncg = (W + 63) / 64;
    const uint32_t nch = (rows + kScanChunk - 1) / kScanChunk;
    for (uint64_t blk = blockIdx.x; blk < ncg * nch * batch; blk += gridDim.x) {
        const uint64_t p = blk / (ncg * nch), r = blk - p * ncg * nch, ch = r / ncg, cg = r - ch * ncg;
        const uint64_t col = cg * 64 + lane;
        const bool valid = col < W;
        const uint32_t r0 = (uint32_t)ch * kScanChunk + wave * (kScanChunk / kScanWaves);
        const uint32_t r1 = std::min<uint32_t>(rows, r0 + kScanChunk / kScanWaves);
        Acc *base = a + p * rows * W + col;
        Acc sum = 0;
        if (valid) {
#pragma unroll 16
            for (uint32_t q = r0; q < r1; q++) sum += base[(uint64_t)q * W];
        }
        part[wave][lane] = sum;
        __syncthreads();
        Acc carry = valid ? carries[(p * nch + ch) * W + col] : (Acc)0;
        for (int k = 0; k < wave; k++) carry += part[k][lane];
        if (valid) {
#pragma unroll 16
            for (uint32_t q = r0; q < r1; q++) {
                const Acc v = base[(uint64_t)q * W];
                base[(uint64_t)q * W] = carry;
                carry += v;
            }
        }
        __syncthreads();
    }
}

// Inclusive scan along each of `lines` contiguous lines of n values, in
// place; one workgroup per line.
template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_sat_linescan(Acc *__restrict__ a, uint64_t n, uint64_t lines) {
    __shared__ Acc lds4[kBlock / 64];
    for (uint64_t line = blockIdx.x; line < lines; line += gridDim.x) {
        Acc *p = a + line * n;
        Acc carry = 0;
        for (uint64_t x0 = 0; x0 < n; x0 += kLineChunk) {
            const uint64_t x = x0 + (uint64_t)threadIdx.x * kSatPer;
            const uint32_t nv = x < n ? (uint32_t)std::min<uint64_t>(kSatPer, n - x) : 0;
            Acc v[kSatPer];
#pragma unroll
            for (uint32_t j = 0; j < kSatPer; j++) v[j] = j < nv ? p[x + j] : (Acc)0;
#pragma unroll
            for (uint32_t j = 1; j < kSatPer; j++) v[j] += v[j - 1];
            Acc total;
            const Acc off = carry + block_excl(v[kSatPer - 1], total, lds4);
#pragma unroll
            for (uint32_t j = 0; j < kSatPer; j++)
                if (j < nv) p[x + j] = v[j] + off;
            carry += total;
        }
    }
}

// Per tile (rows y0.., columns c0..): S(y, x) = la(y) + above(x) + local(y, x)
// where la(y) is the sum of w over rows <= y left of c0, above(x) over rows
// < y0 and columns c0..x (prefix of the column carries cex), local(y, x)
// over rows y0..y and columns c0..x.
template <int K, typename WIN>
__global__ void __launch_bounds__(kBlock)
k_sat_tiles(const WIN *__restrict__ w, SatGeom g, const SatAcc<K> *__restrict__ cex,
            const SatAcc<K> *__restrict__ la, SatAcc<K> *__restrict__ S) {
    using Acc = SatAcc<K>;
    const int wave = threadIdx.x >> 6;
    const uint64_t n_tiles = (uint64_t)g.nb * g.nseg * g.gz;
    const uint64_t stride = (uint64_t)gridDim.x * (kBlock / 64);
    for (uint64_t t = (uint64_t)blockIdx.x * (kBlock / 64) + wave; t < n_tiles; t += stride) {
        const SatTile T = sat_tile(g, t);
        Acc abv[kSatPer];
        load_lane(cex + (T.z * g.nb + T.b) * g.gx, T, g.gx, g.vec_s, abv);
        wave_row_prefix(abv);
        Acc run[kSatPer];
#pragma unroll
        for (uint32_t j = 0; j < kSatPer; j++) run[j] = 0;
        const Acc *lap = la ? la + ((T.z * g.nseg + T.s) * g.gy + T.y0) : nullptr;
        for (uint32_t g0 = 0; g0 < T.nr; g0 += kSatGroup) {
            Acc rv[kSatGroup][kSatPer];
            {
                WIN v[kSatGroup][kSatPer];
                load_rows<kSatGroup>(w, g, T, g0, v);
#pragma unroll
                for (uint32_t i = 0; i < kSatGroup; i++) {
                    const bool in = g0 + i < T.nr;
#pragma unroll
                    for (uint32_t j = 0; j < kSatPer; j++) rv[i][j] = in ? sat_of<K>(v[i][j]) : (Acc)0;
                }
            }
#pragma unroll
            for (uint32_t i = 0; i < kSatGroup; i++) {
                if (g0 + i >= T.nr) break;
                wave_row_prefix(rv[i]);
                const Acc lay = lap ? lap[g0 + i] : (Acc)0;
                Acc o[kSatPer];
#pragma unroll
                for (uint32_t j = 0; j < kSatPer; j++) {
                    run[j] += rv[i][j];
                    o[j] = lay + abv[j] + run[j];
                }
                store_lane(S + (T.z * g.gy + T.y0 + g0 + i) * g.gx, T, g.gx, g.vec_s, o);
            }
        }
    }
}

// 3D: the band sums over z of the per-plane tables (W = gx * gy cells per
// plane), then, after their exclusive scan, each band's running sums.
template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_sat_zsum(const Acc *__restrict__ S, uint64_t W, uint64_t gz, uint32_t nbz, Acc *__restrict__ bz) {
    const uint64_t nwb = (W + kBlock - 1) / kBlock;
    for (uint64_t blk = blockIdx.x; blk < nwb * nbz; blk += gridDim.x) {
        const uint64_t band = blk / nwb, j = (blk - band * nwb) * kBlock + threadIdx.x;
        if (j >= W) continue;
        const uint64_t z0 = band * kSatZBand, z1 = std::min<uint64_t>(gz, z0 + kSatZBand);
        Acc s = 0;
#pragma unroll 8
        for (uint64_t z = z0; z < z1; z++) s += S[z * W + j];
        bz[band * W + j] = s;
    }
}
template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_sat_zapply(Acc *__restrict__ S, uint64_t W, uint64_t gz, uint32_t nbz, const Acc *__restrict__ bz) {
    const uint64_t nwb = (W + kBlock - 1) / kBlock;
    for (uint64_t blk = blockIdx.x; blk < nwb * nbz; blk += gridDim.x) {
        const uint64_t band = blk / nwb, j = (blk - band * nwb) * kBlock + threadIdx.x;
        if (j >= W) continue;
        const uint64_t z0 = band * kSatZBand, z1 = std::min<uint64_t>(gz, z0 + kSatZBand);
        Acc c = bz[band * W + j];
#pragma unroll 8
        for (uint64_t z = z0; z < z1; z++) {
            c += S[z * W + j];
            S[z * W + j] = c;
        }
    }
}

// Sum of w over the box [lo, hi) (hi exclusive) from the inclusive table.
// In f64 every partial result is a sum over cells with coefficients -1, 0
// or 1, so at most the sum of |w| in magnitude: exact under the host's check.
template <int D, typename Acc>
__device__ inline Acc sat_box(const Acc *__restrict__ S, uint64_t gx, uint64_t gy, const uint64_t (&lo)[3],
                              const uint64_t (&hi)[3]) {
    auto at = [&](uint64_t X, uint64_t Y, uint64_t Z) -> Acc {
        return (X == 0 || Y == 0 || Z == 0) ? (Acc)0 : S[((Z - 1) * gy + (Y - 1)) * gx + (X - 1)];
    };
    Acc r = at(hi[0], hi[1], hi[2]) - at(lo[0], hi[1], hi[2]) - at(hi[0], lo[1], hi[2]) + at(lo[0], lo[1], hi[2]);
    if constexpr (D == 3) {
        r = r - at(hi[0], hi[1], lo[2]);
        r = r + at(lo[0], hi[1], lo[2]);
        r = r + at(hi[0], lo[1], lo[2]);
        r = r - at(lo[0], lo[1], lo[2]);
    }
    return r;
}

// Sum of w over the cells whose index along `coord` is below t and whose
// other indices lie in [lo, hi): the box's lower face along coord at 0, so
// 2^(D-1) lookups.  A subgrid's axis prefix is face(off + pos) - face(off).
template <int D, typename Acc>
__device__ inline Acc sat_face(const Acc *__restrict__ S, uint64_t gx, uint64_t gy, const uint64_t (&lo)[3],
                               const uint64_t (&hi)[3], int coord, uint64_t t) {
    auto at = [&](const uint64_t (&q)[3]) -> Acc {
        return (q[0] == 0 || q[1] == 0 || q[2] == 0) ? (Acc)0 : S[((q[2] - 1) * gy + (q[1] - 1)) * gx + (q[0] - 1)];
    };
    uint64_t a[3] = {hi[0], hi[1], hi[2]};
    a[coord] = t;
    const int o1 = coord == 0 ? 1 : 0;
    uint64_t b[3] = {a[0], a[1], a[2]};
    b[o1] = lo[o1];
    Acc r = at(a) - at(b);
    if constexpr (D == 3) {
        const int o2 = 3 - coord - o1;
        uint64_t c[3] = {a[0], a[1], a[2]}, d[3] = {b[0], b[1], b[2]};
        c[o2] = lo[o2];
        d[o2] = lo[o2];
        r = r - at(c);
        r = r + at(d);
    }
    return r;
}

// The carrier of GW<K> of an exact sum of the table.
template <int K>
__device__ inline uint64_t sat_carrier(SatAcc<K> v) {
    if constexpr (K == 0) return (uint64_t)(int64_t)(int32_t)(uint32_t)v;
    else if constexpr (K == 1) return v;
    else return __builtin_bit_cast(uint64_t, v);
}

// One wavefront per subgrid h of level l (2^l of them): weighted_median
// (rcb.rs:52-99) with every chunk's prefix weight a lookup (the weight of
// the subgrid's cross-sections before the chunk's start: coupe's
// left_weight + prefix, the same integer; the first chunk's is the left
// weight itself), the decisions by ballot in
// coupe's order; then the split into the heap and the two halves into
// slots 2h and 2h + 1 of the next level.
template <int K, int D>
__global__ void __launch_bounds__(kBlock)
k_grid_level(const GSlot *__restrict__ in, GSlot *__restrict__ out, uint32_t *__restrict__ heap, uint32_t l,
             const SatAcc<K> *__restrict__ S, uint64_t gx, uint64_t gy, uint64_t gz, uint32_t threads) {
    using Wt = GW<K>;
    const int lane = threadIdx.x & 63;
    const uint64_t n = (uint64_t)1 << l;
    const uint64_t h = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (h >= n) return;
    const int coord = (int)((1 + l) % D);
    GSlot P;   // level 0: the whole grid
    if (l == 0) P = GSlot{{0, 0, 0}, {(uint32_t)gx, (uint32_t)gy, (uint32_t)gz}, 0};
    else P = in[h];
    if (P.size[0] == 0 || P.size[1] == 0 || P.size[2] == 0) {   // no cells: nothing below reaches a cell
        if (lane == 0) {
            out[2 * h] = P;
            out[2 * h + 1] = P;
            heap[n - 1 + h] = 0;
        }
        return;
    }
    uint64_t lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = P.off[k];
        hi[k] = (uint64_t)P.off[k] + P.size[k];
    }
    const uint64_t total = l == 0 ? sat_carrier<K>(sat_box<D>(S, gx, gy, lo, hi)) : P.sum;
    const double ideal = Wt::f64(total) / 2.0;
    const uint64_t minp = Wt::from_f64(ideal * (1.0 - kGridTolerance));
    const uint64_t maxp = Wt::from_f64(ideal * (1.0 + kGridTolerance));
    const uint64_t m = P.size[coord];
    const SatAcc<K> f0 = sat_face<D>(S, gx, gy, lo, hi, coord, lo[coord]);
    uint64_t mn = 0, mx = m, left = 0, pos_out = 0, left_out = 0;
    for (;;) {
        uint64_t cs = (mx - mn) / threads;
        if (cs < 1) cs = 1;
        const uint64_t nch = (mx - mn + cs - 1) / cs;
        uint64_t nmin = mn, nmax = mx, nleft = left;
        bool done = false;
        for (uint64_t base = 0; base < nch; base += 64) {
            const uint64_t k = base + lane;
            const bool valid = k < nch;
            const uint64_t pos = mn + k * cs;
            uint64_t c = left;   // the first chunk starts at the left weight's position
            if (valid && k > 0) c = sat_carrier<K>(sat_face<D>(S, gx, gy, lo, hi, coord, lo[coord] + pos) - f0);
            const bool stop = valid && !Wt::lt(c, minp);
            const uint64_t mask = __ballot(stop);
            if (mask == 0) {   // every chunk of this batch below the minimum
                const int last = (int)std::min<uint64_t>(63, nch - 1 - base);
                nmin = readlane64(pos, last);
                nleft = readlane64(c, last);
                continue;
            }
            const int f = __builtin_ctzll(mask);
            if (f > 0) {
                nmin = readlane64(pos, f - 1);
                nleft = readlane64(c, f - 1);
            }
            const uint64_t cf = readlane64(c, f), pf = readlane64(pos, f);
            if (Wt::lt(maxp, cf)) {
                nmax = pf;
            } else {
                pos_out = pf;
                left_out = cf;
                done = true;
            }
            break;
        }
        if (!done && (nmin + 1 >= nmax || (nmin == mn && nmax == mx))) {
            pos_out = nmin;
            left_out = nleft;
            done = true;
        }
        if (done) break;
        mn = nmin;
        mx = nmax;
        left = nleft;
    }
    if (lane == 0) {
        GSlot A = P, B = P;
        A.size[coord] = (uint32_t)pos_out;
        A.sum = left_out;
        B.off[coord] = P.off[coord] + (uint32_t)pos_out;
        B.size[coord] = (uint32_t)(m - pos_out);
        B.sum = Wt::sub(total, left_out);
        out[2 * h] = A;
        out[2 * h + 1] = B;
        heap[n - 1 + h] = P.off[coord] + (uint32_t)pos_out;
    }
}

// part_of (rcb.rs:21-42) from the heap of splits: at level l the cell goes
// right when its index along axis (1 + l) mod D is at or past the split.
// Cells in memory order: a workgroup takes kBlock * kFillCells consecutive
// cells, each wavefront stores kFillCells windows of 64 consecutive cells.
// A leaf is a box, so the cells of a row that lie in one leaf are
// consecutive: when a window lies in one row and its first and last cells
// walk to the same leaf, so do all its cells.  The wavefront walks the
// 2 * kFillCells window ends at once, one per lane, and walks every cell only
// of the windows that cross a leaf boundary or a row.  Rows and columns come
// from f64 reciprocals of the sizes, corrected by one step (exact for
// indices below 2^52).  The first kFillTop levels of the heap sit in LDS.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_grid_fill_heap(const uint32_t *__restrict__ heap, uint32_t L, uint64_t gx, uint64_t gy, uint64_t cells,
                 double inv_gx, double inv_gy, uint64_t *__restrict__ out) {
    static_assert(2 * kFillCells <= 64, "window ends per wavefront");
    __shared__ uint32_t top[(1u << kFillTop) - 1];
    const uint32_t lt = L < kFillTop ? L : kFillTop;
    const uint32_t ntop = (1u << lt) - 1;
    for (uint32_t i = threadIdx.x; i < ntop; i += kBlock) top[i] = heap[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // the leaf (index within the last level) of cell (x, y, z); levels D at
    // a time, so that each one's axis is known at compile time
    auto leaf_of = [&](uint32_t x, uint32_t y, uint32_t z) -> uint32_t {
        const uint32_t q[3] = {x, y, z};
        uint32_t n = 0;   // heap node: children 2n + 1, 2n + 2
        uint32_t l = 0;
        for (; l + D <= lt; l += D) {
#pragma unroll
            for (int k = 0; k < D; k++) n = 2 * n + 1 + (q[(1 + k) % D] >= top[n]);
        }
        for (; l < lt; l++) n = 2 * n + 1 + (q[(1 + l) % D] >= top[n]);
        for (; l < L; l++) n = 2 * n + 1 + (q[(1 + l) % D] >= heap[n]);
        return n - ((1u << L) - 1);
    };
    // a / b and a % b for a < 2^52
    auto divmod = [](uint64_t a, uint64_t b, double inv_b, uint64_t &r) -> uint64_t {
        uint64_t qt = (uint64_t)((double)a * inv_b);
        int64_t rr = (int64_t)(a - qt * b);
        if (rr < 0) {
            qt--;
            rr += (int64_t)b;
        } else if ((uint64_t)rr >= b) {
            qt++;
            rr -= (int64_t)b;
        }
        r = (uint64_t)rr;
        return qt;
    };
    // the leaf of cell f of row `row` (the index of its run of gx cells)
    auto leaf_in_row = [&](uint64_t f, uint64_t row) -> uint32_t {
        const uint64_t x = f - row * gx;
        uint64_t y = row, z = 0;
        if constexpr (D == 3) z = divmod(row, gy, inv_gy, y);
        return leaf_of((uint32_t)x, (uint32_t)y, (uint32_t)z);
    };
    // the row and the leaf of cell f
    auto cell_leaf = [&](uint64_t f, uint64_t &row) -> uint32_t {
        uint64_t x;
        row = divmod(f, gx, inv_gx, x);
        return leaf_in_row(f, row);
    };
    const uint64_t chunks = (cells + kBlock * kFillCells - 1) / (kBlock * kFillCells);
    for (uint64_t blk = blockIdx.x; blk < chunks; blk += gridDim.x) {
        // window i of this wavefront starts at w0 + i * kBlock
        const uint64_t w0 = blk * (kBlock * kFillCells) + (threadIdx.x & ~63u);
        const uint32_t ei = (uint32_t)(lane >> 1) % kFillCells;
        const uint64_t fe = std::min<uint64_t>(w0 + (uint64_t)ei * kBlock + ((lane & 1) ? 63 : 0), cells - 1);
        uint64_t end_row;
        const uint32_t end_leaf = cell_leaf(fe, end_row);
#pragma unroll
        for (uint32_t i = 0; i < kFillCells; i++) {
            const uint64_t ws = w0 + (uint64_t)i * kBlock;
            if (ws >= cells) break;
            const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)end_leaf, 2 * i);
            const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)end_leaf, 2 * i + 1);
            const uint64_t ra = readlane64(end_row, 2 * i), rb = readlane64(end_row, 2 * i + 1);
            const uint64_t f = ws + lane;
            uint32_t id = a;
            if (ra != rb) {   // the window crosses a row
                uint64_t row;
                id = cell_leaf(f < cells ? f : cells - 1, row);
            } else if (a != b) {   // one row, more than one leaf
                id = leaf_in_row(f < cells ? f : cells - 1, ra);
            }
            if (f < cells) out[f] = id;
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

// Blocks for a grid-stride kernel: enough to fill the device, at most `work`.
// The occupancy query is a runtime call: asked once per kernel and context.
template <typename Kern>
unsigned fill_grid(Context &cx, Kern kern, uint64_t work) {
    const void *key = reinterpret_cast<const void *>(kern);
    int per_cu = 0;
    for (const auto &e : cx.occupancy)
        if (e.first == key) {
            per_cu = e.second;
            break;
        }
    if (per_cu == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess || per_cu < 1)
            per_cu = 2;
        cx.occupancy.emplace_back(key, per_cu);
    }
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(work, (uint64_t)cx.cu_count * per_cu));
}

// The same for the column scans' workgroups of 64 * kScanWaves threads.
template <typename Kern>
unsigned scan_grid(Context &cx, Kern kern, uint64_t work) {
    const void *key = reinterpret_cast<const void *>(kern);
    int per_cu = 0;
    for (const auto &e : cx.occupancy)
        if (e.first == key) {
            per_cu = e.second;
            break;
        }
    if (per_cu == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * kScanWaves, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        cx.occupancy.emplace_back(key, per_cu);
    }
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(work, (uint64_t)cx.cu_count * per_cu));
}

// Exclusive scan over the rows of each column of `batch` arrays of rows x W
// (in place).  Arrays too narrow to give the device enough workgroups of 64
// columns are scanned in chunks of rows (`tmp`: room for their sums).
template <typename Acc>
void colscan(Context &cx, hipStream_t st, Acc *a, uint64_t W, uint32_t rows, uint32_t batch, Acc *tmp) {
    const uint64_t ncg = (W + 63) / 64;
    const uint32_t nch = (rows + kScanChunk - 1) / kScanChunk;
    if (ncg * batch >= 2 * (uint64_t)cx.cu_count || nch < 2 || !tmp) {
        k_sat_colscan<Acc><<<scan_grid(cx, k_sat_colscan<Acc>, ncg * batch), 64 * kScanWaves, 0, st>>>(a, W, rows,
                                                                                                     batch);
        ck(hipGetLastError());
        return;
    }
    const uint64_t work = ncg * nch * batch;
    k_sat_chunksum<Acc><<<scan_grid(cx, k_sat_chunksum<Acc>, work), 64 * kScanWaves, 0, st>>>(a, W, rows, batch, tmp);
    ck(hipGetLastError());
    colscan<Acc>(cx, st, tmp, W, nch, batch, nullptr);
    k_sat_chunkapply<Acc><<<scan_grid(cx, k_sat_chunkapply<Acc>, work), 64 * kScanWaves, 0, st>>>(a, W, rows, batch,
                                                                                                  tmp);
    ck(hipGetLastError());
}

// The summed-area path.  Returns false when f64 weights are not exactly
// summable (then nothing it wrote is a result and the caller runs the
// per-level path).
template <int K, typename WIN>
bool grid_run_sat(Context &cx, hipStream_t st, size_t dim, const uint64_t *size, const WIN *w, size_t L, uint32_t T,
                  uint64_t *out) {
    SatGeom g;
    g.gx = size[0];
    g.gy = size[1];
    g.gz = dim == 3 ? size[2] : 1;
    const uint64_t cells = g.gx * g.gy * g.gz;
    g.nb = (uint32_t)((g.gy + kSatRows - 1) / kSatRows);
    g.nseg = (uint32_t)((g.gx + kSatCols - 1) / kSatCols);
    g.vec_w = g.gx % 2 == 0 && (uintptr_t)w % (2 * sizeof(WIN)) == 0;
    g.vec_s = g.gx % 2 == 0;
    const uint64_t n_tiles = (uint64_t)g.nb * g.nseg * g.gz;
    const uint32_t nbz = (uint32_t)((g.gz + kSatZBand - 1) / kSatZBand);
    const uint64_t plane = g.gx * g.gy;
    cx.gsat.ensure(cells * 8);
    cx.gcol.ensure(std::max<uint64_t>(g.gz * g.nb * g.gx, dim == 3 ? nbz * plane : 0) * 8);
    cx.grow.ensure(g.gz * g.nseg * g.gy * 8);
    auto chunks = [](uint64_t rows) { return (rows + kScanChunk - 1) / kScanChunk; };
    cx.gscan.ensure(std::max({chunks(g.nb) * g.gx * g.gz, chunks(g.nseg) * g.gy * g.gz,
                              dim == 3 ? chunks(nbz) * plane : (uint64_t)0, (uint64_t)1}) * 8);
    const uint64_t slots = (uint64_t)1 << L;
    cx.gparts[0].ensure(slots * sizeof(GSlot));
    cx.gparts[1].ensure(slots * sizeof(GSlot));
    cx.gheap.ensure(slots * 4);
    cx.gctr.ensure(kCopies * sizeof(GridSatFlags));
    GridSatFlags *flags = (GridSatFlags *)cx.gctr.ptr;
    using Acc = SatAcc<K>;
    Acc *S = (Acc *)cx.gsat.ptr, *col = (Acc *)cx.gcol.ptr, *row = (Acc *)cx.grow.ptr, *tmp = (Acc *)cx.gscan.ptr;
    uint32_t *heap = (uint32_t *)cx.gheap.ptr;
    if (K == 2) ck(hipMemsetAsync(flags, 0, kCopies * sizeof(GridSatFlags), st));

    int r = cx.prof.begin(st, COUPE_HIP_K_PROLOGUE, L ? (double)cells * 2.0 * sizeof(WIN) + (double)cells * 8.0 : 0.0);
    if (L) {   // no table for a call without levels
        const unsigned tile_grid = (unsigned)std::min<uint64_t>((n_tiles + 3) / 4, 1u << 24);   // wavefront per tile
        k_sat_sums<K, WIN><<<tile_grid, kBlock, 0, st>>>(w, g, col, row, flags);
        ck(hipGetLastError());
        colscan<Acc>(cx, st, col, g.gx, g.nb, (uint32_t)g.gz, tmp);
        if (g.nseg > 1) {
            colscan<Acc>(cx, st, row, g.gy, g.nseg, (uint32_t)g.gz, tmp);
            k_sat_linescan<Acc><<<fill_grid(cx, k_sat_linescan<Acc>, g.gz * g.nseg), kBlock, 0, st>>>(row, g.gy,
                                                                                                    g.gz * g.nseg);
            ck(hipGetLastError());
        }
        k_sat_tiles<K, WIN><<<tile_grid, kBlock, 0, st>>>(
            w, g, col, g.nseg > 1 ? row : nullptr, S);
        ck(hipGetLastError());
        if (dim == 3) {
            const uint64_t work = (plane + kBlock - 1) / kBlock * nbz;
            k_sat_zsum<Acc><<<fill_grid(cx, k_sat_zsum<Acc>, work), kBlock, 0, st>>>(S, plane, g.gz, nbz, col);
            ck(hipGetLastError());
            colscan<Acc>(cx, st, col, plane, nbz, 1, tmp);
            k_sat_zapply<Acc><<<fill_grid(cx, k_sat_zapply<Acc>, work), kBlock, 0, st>>>(S, plane, g.gz, nbz, col);
            ck(hipGetLastError());
        }
    }
    cx.prof.end(st, r);

    int cur = 0;
    r = cx.prof.begin(st, COUPE_HIP_K_DECIDE, 0.0);
    for (size_t l = 0; l < L; l++) {
        const GSlot *in = (const GSlot *)cx.gparts[cur].ptr;
        GSlot *nx = (GSlot *)cx.gparts[cur ^ 1].ptr;
        const unsigned blocks = (unsigned)((((uint64_t)1 << l) + kBlock / 64 - 1) / (kBlock / 64));
        if (dim == 2)
            k_grid_level<K, 2><<<blocks, kBlock, 0, st>>>(in, nx, heap, (uint32_t)l, S, g.gx, g.gy, g.gz, T);
        else
            k_grid_level<K, 3><<<blocks, kBlock, 0, st>>>(in, nx, heap, (uint32_t)l, S, g.gx, g.gy, g.gz, T);
        ck(hipGetLastError());
        cur ^= 1;
    }
    cx.prof.end(st, r);

    const uint64_t fill_chunks = (cells + kBlock * kFillCells - 1) / (kBlock * kFillCells);
    const double inv_gx = 1.0 / (double)g.gx, inv_gy = 1.0 / (double)g.gy;
    r = cx.prof.begin(st, COUPE_HIP_K_IDS, (double)cells * 8.0);
    if (dim == 2)
        k_grid_fill_heap<2><<<fill_grid(cx, k_grid_fill_heap<2>, fill_chunks), kBlock, 0, st>>>(heap, (uint32_t)L, g.gx, g.gy,
                                                                                        cells, inv_gx, inv_gy, out);
    else
        k_grid_fill_heap<3><<<fill_grid(cx, k_grid_fill_heap<3>, fill_chunks), kBlock, 0, st>>>(heap, (uint32_t)L, g.gx, g.gy,
                                                                                        cells, inv_gx, inv_gy, out);
    ck(hipGetLastError());
    cx.prof.end(st, r);
    if (K != 2) {
        ck(hipStreamSynchronize(st));
        return true;
    }
    if (!cx.host_scratch) ck(hipHostMalloc(&cx.host_scratch, kHostScratch, hipHostMallocDefault));
    GridSatFlags *hf = (GridSatFlags *)cx.host_scratch;
    ck(hipMemcpyAsync(hf, flags, kCopies * sizeof(GridSatFlags), hipMemcpyDeviceToHost, st));
    ck(hipStreamSynchronize(st));
    double abs_sum = 0.0;
    uint32_t bad = 0;
    for (int c = 0; c < kCopies; c++) {
        abs_sum += hf[c].abs_sum;
        bad |= hf[c].bad;
    }
    // Every partial sum is an integer of magnitude at most sum |w|; the f64
    // sum of the magnitudes, over at most 2^36 terms, is within a factor
    // 1 + 2^-17 of the exact one, so this bound keeps the exact sum <= 2^53.
    return bad == 0 && cells <= (1ull << 36) && abs_sum <= 0x1p53 * (1.0 - 0x1p-15);
}

}  // namespace

enum coupe_err grid_rcb(Context &cx, hipStream_t st, size_t dim, const uint64_t *size, int wkind, const void *w,
                        size_t iter_count, uint32_t threads, uint64_t *out) {
    if (dim != 2 && dim != 3) return COUPE_ERR_BAD_DIMENSION;
    if (wkind != COUPE_INT && wkind != COUPE_INT64 && wkind != COUPE_DOUBLE) return COUPE_ERR_BAD_TYPE;
    for (size_t k = 0; k < dim; k++)
        if (size[k] == 0 || size[k] > 0xffffffffull) return COUPE_ERR_CRASH;   // NonZeroUsize, u32 positions
    uint64_t cells = size[0];
    for (size_t k = 1; k < dim; k++)
        if (__builtin_mul_overflow(cells, size[k], &cells)) return COUPE_ERR_CRASH;   // no memory holds such a grid
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
        // the summed-area path up to kDenseGridLevels levels (COUPE_HIP_GRID_SAT=0: never)
        const char *e = getenv("COUPE_HIP_GRID_SAT");
        const bool sat = iter_count <= kDenseGridLevels && !(e && e[0] == '0');
        if (wkind == COUPE_INT) {
            if (!sat || !grid_run_sat<0, int32_t>(cx, st, dim, size, (const int32_t *)w, iter_count, T, out))
                grid_run<0, int32_t>(cx, st, dim, size, (const int32_t *)w, iter_count, T, out);
        } else if (wkind == COUPE_INT64) {
            if (!sat || !grid_run_sat<1, int64_t>(cx, st, dim, size, (const int64_t *)w, iter_count, T, out))
                grid_run<1, int64_t>(cx, st, dim, size, (const int64_t *)w, iter_count, T, out);
        } else {
            if (!sat || !grid_run_sat<2, double>(cx, st, dim, size, (const double *)w, iter_count, T, out))
                grid_run<2, double>(cx, st, dim, size, (const double *)w, iter_count, T, out);
        }
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
