// rcb_common.h -- shared types of the MI355X RCB engine (device and host).
//
// Vocabulary follows coupe: a *part* is a node of the bisection tree, i.e. a
// contiguous range of the working arrays; a *level* is one depth of the tree;
// a *pass* is one sweep that evaluates several split candidates of every part
// still searching; a *tile* is a range of at most kTile points of one part,
// the unit of work of a workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace coupe_hip {

constexpr int kMaxDim = 3;
constexpr int kBlock = 256;          // threads per workgroup of the sweeps
constexpr int kItems = 8;            // points per thread per tile
constexpr int kTile = kBlock * kItems;
constexpr int kMaxBins = 256;        // histogram bins per part per pass
constexpr int kFirstBins = 4;        // bins of the register sweep (2 bisection steps)
constexpr int kCopies = 8;           // copies of a search pass's bins, one per XCD

// Part status during a level.
enum : uint32_t {
    kSearching = 0,   // split position not found yet
    kSplitAtT = 1,    // split found: left = {c < t}
    kAllLeft = 2,     // no point at or right of t twice: every point goes left
    kDone = 3,        // nothing to split (iter_count 0 or no point)
};

// One part of the current level.  `sum` and `wl` hold the bits of the
// accumulator type (uint64 two's complement or double).
struct Part {
    uint64_t begin;        // first position in the working arrays
    uint64_t count;        // points in the part
    uint64_t heap_id;      // iter_id of coupe's recursion (root 0, children 2i+1 / 2i+2)
    uint64_t sum;          // weight of the part, accumulator bits
    float lo[kMaxDim];     // inherited box, f32 (coupe casts its f64 box at use)
    float hi[kMaxDim];
    // split search state (par_rcb_split's min, max, prev_count_left)
    float blo, bhi;
    uint64_t prev;
    uint32_t axis;
    uint32_t status;
    float split;           // split_pos
    int32_t kstar;         // in-order index of the stopping candidate in the last pass
    uint64_t wl;           // weight_left, accumulator bits
    uint64_t cnt_left;     // points going left
    uint32_t tile_begin;   // first tile of this part in the tile table
    uint32_t tile_count;
    int32_t child[2];      // indices of the left / right child in the next level (-1: empty)
    uint32_t inst;         // instance of a batch (0 for a single call)
    uint32_t depth;        // level of this part in its instance's tree
    uint32_t levels;       // iter_count of its instance
    uint32_t counts_ready;      // where the partition finds each tile's left count (kCounts*)
    double tol;            // tolerance of its instance
};

// One independent RCB problem of a batch: its points are the range
// [begin, begin + count) of the packed arrays.
struct Instance {
    uint64_t begin;
    uint64_t count;
    uint32_t levels;       // iter_count
    uint32_t tile_begin;   // its first tile at level 0
    double tol;
    uint64_t heap_base;    // its dense split heap: heap_split[heap_base + h]
    uint64_t src;          // index of its first point in the caller's weights and ids
    uint64_t src_pt;       // offset of its first point's first coordinate in the caller's points
    // filled on the device
    uint64_t bmin[kMaxDim];     // ordered keys of the f64 bounding box
    uint64_t bmax[kMaxDim];
    uint64_t sum_u64;           // wrapping integer sum of the weights
    double sum_f64;             // f64 sum (real weights)
    uint64_t min_id;            // smallest leaf heap id (the id offset)
};

// A tile: the range [begin, begin + count) of part `part`, count <= kTile.
struct Tile {
    uint64_t begin;
    uint32_t count;
    uint32_t part;
};

// One split of the bisection tree, kept for the final descent that gives
// every point its part id in input order.
constexpr uint32_t kNodeAllLeft = 1u << 8;
struct Node {
    float split;           // split_pos
    uint32_t meta;         // axis | kNodeAllLeft
    int32_t child[2];      // node index of the left / right child (-1: empty or leaf level)
};

// Change of basis applied to every point in the prologue (Rib): row-major
// D x D, p -> m p with the sum over columns taken left to right.
struct Rotation {
    double m[9];
};

// Part::counts_ready: each tile's count of points left of the split is
// counted by reading the tile (kCountsRead), was counted by the partition
// that made the part (kCountsBelow: split at the first candidate), or is the
// sum of the tile's bins in the histogram pass that decided (kCountsBins).
constexpr uint32_t kCountsRead = 0, kCountsBelow = 1, kCountsBins = 2;

// One copy of an instance's prologue reductions (ordered keys of the f64
// box, the wrapping integer and the f64 weight sums); the prologue's
// workgroups on XCD c reduce into copy c and the roots' set-up combines them.
struct InstReduce {
    uint64_t bmin[kMaxDim];
    uint64_t bmax[kMaxDim];
    uint64_t sum_u64;
    double sum_f64;
};

// Flags and guards of the prologue over the whole batch, one copy per XCD;
// the host combines the copies.
struct Prologue {
    double sum_abs;             // sum of |w| (exactness guard)
    uint64_t max_abs_bits;      // bits of max |w| (non-negative doubles order as integers)
    uint32_t non_integral;      // any weight with a fractional part or |w| > 2^53
    uint32_t not_i32;           // any weight outside int32
    uint32_t not_u16;           // any weight that is not an integer in [0, 65535]
    uint32_t pad;
};

__host__ __device__ inline uint32_t f32_key(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float key_f32(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __builtin_bit_cast(float, u);
}
__host__ __device__ inline uint64_t f64_key(double x) {
    uint64_t u = __builtin_bit_cast(uint64_t, x);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double key_f64(uint64_t k) {
    uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, u);
}

}  // namespace coupe_hip
