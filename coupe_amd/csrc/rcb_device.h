// rcb_device.h -- device helpers shared by the HIP kernels: accumulator
// types, wavefront reductions, axis select and candidate bins.
#pragma once

#include <type_traits>

#include "rcb_common.h"

namespace coupe_hip {

// ---------------------------------------------------------------------------
// Accumulator helpers.  Acc = uint64_t: integer weights, two's complement sums
// (exact; i32 semantics are recovered by truncation when wrap32 is set, as
// Rust's wrapping i32 arithmetic is a ring morphism mod 2^32).  Acc = double:
// non-integer f64 weights.
// ---------------------------------------------------------------------------
template <typename Acc> struct AccOps;

template <> struct AccOps<uint64_t> {
    __device__ static uint64_t from_bits(uint64_t b) { return b; }
    __device__ static uint64_t to_bits(uint64_t v) { return v; }
    template <typename W> __device__ static uint64_t of(W w) { return (uint64_t)(int64_t)w; }
    // RcbWeight::to_f64 (num-traits): i32 / i64 -> f64, round to nearest.
    __device__ static double f64(uint64_t v, int wrap32) {
        return wrap32 ? (double)(int32_t)(uint32_t)v : (double)(int64_t)v;
    }
    __device__ static bool lt(uint64_t a, uint64_t b, int wrap32) {
        return wrap32 ? (int32_t)(uint32_t)a < (int32_t)(uint32_t)b : (int64_t)a < (int64_t)b;
    }
    __device__ static uint64_t sub(uint64_t a, uint64_t b) { return a - b; }
};

template <> struct AccOps<double> {
    __device__ static double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
    __device__ static uint64_t to_bits(double v) { return __builtin_bit_cast(uint64_t, v); }
    template <typename W> __device__ static double of(W w) { return (double)w; }
    __device__ static double f64(double v, int) { return v; }
    __device__ static bool lt(double a, double b, int) { return a < b; }
    __device__ static double sub(double a, double b) { return a - b; }
};

__device__ inline void atomic_add_acc(uint64_t *p, uint64_t v) {
    atomicAdd((unsigned long long *)p, (unsigned long long)v);
}
__device__ inline void atomic_add_acc(double *p, double v) { atomicAdd(p, v); }

__device__ inline int lane_id() { return (int)__lane_id(); }

template <typename T> __device__ inline T wave_incl_add(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T n = __shfl_up(v, o, 64);
        if (lane >= o) v += n;
    }
    return v;
}
template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t x = __shfl_xor(v, o, 64);
        v = x < v ? x : v;
    }
    return v;
}

// Wavefront totals without LDS: DPP row shifts leave each row's total in the
// row's last lane (an inclusive scan within rows of 16), four readlanes
// combine the rows.  Called with every lane of the wavefront active; the
// result is uniform.
template <int Ctrl, bool Bound>
__device__ __forceinline__ uint32_t dpp_row(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, Ctrl, 0xf, 0xf, Bound);
}
template <int Ctrl>
__device__ __forceinline__ uint64_t dpp_row64(uint64_t v) {   // shifted-in lanes read 0
    return (uint64_t)dpp_row<Ctrl, true>(0u, (uint32_t)v) | (uint64_t)dpp_row<Ctrl, true>(0u, (uint32_t)(v >> 32)) << 32;
}
__device__ __forceinline__ uint32_t rows_last_lane(uint32_t v, int r) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 16 * r + 15);
}
__device__ __forceinline__ uint64_t rows_last_lane64(uint64_t v, int r) {
    return (uint64_t)rows_last_lane((uint32_t)v, r) | (uint64_t)rows_last_lane((uint32_t)(v >> 32), r) << 32;
}
__device__ __forceinline__ uint32_t wave_total_min(uint32_t v) {
    constexpr uint32_t kId = 0xffffffffu;
    v = min(v, dpp_row<0x111, false>(kId, v));   // row_shr:1 (lanes shifted past the row's start keep kId)
    v = min(v, dpp_row<0x112, false>(kId, v));   // row_shr:2
    v = min(v, dpp_row<0x114, false>(kId, v));   // row_shr:4
    v = min(v, dpp_row<0x118, false>(kId, v));   // row_shr:8
    return min(min(rows_last_lane(v, 0), rows_last_lane(v, 1)), min(rows_last_lane(v, 2), rows_last_lane(v, 3)));
}
__device__ __forceinline__ uint32_t wave_total(uint32_t v) {
    v += dpp_row<0x111, true>(0u, v);
    v += dpp_row<0x112, true>(0u, v);
    v += dpp_row<0x114, true>(0u, v);
    v += dpp_row<0x118, true>(0u, v);
    return rows_last_lane(v, 0) + rows_last_lane(v, 1) + rows_last_lane(v, 2) + rows_last_lane(v, 3);
}
__device__ __forceinline__ uint64_t wave_total(uint64_t v) {
    v += dpp_row64<0x111>(v);
    v += dpp_row64<0x112>(v);
    v += dpp_row64<0x114>(v);
    v += dpp_row64<0x118>(v);
    return rows_last_lane64(v, 0) + rows_last_lane64(v, 1) + rows_last_lane64(v, 2) + rows_last_lane64(v, 3);
}
__device__ __forceinline__ double wave_total(double v) {   // 0 bits are 0.0
    auto f = [](uint64_t b) { return __builtin_bit_cast(double, b); };
    auto u = [](double d) { return __builtin_bit_cast(uint64_t, d); };
    v += f(dpp_row64<0x111>(u(v)));
    v += f(dpp_row64<0x112>(u(v)));
    v += f(dpp_row64<0x114>(u(v)));
    v += f(dpp_row64<0x118>(u(v)));
    return (f(rows_last_lane64(u(v), 0)) + f(rows_last_lane64(u(v), 1))) +
           (f(rows_last_lane64(u(v), 2)) + f(rows_last_lane64(u(v), 3)));
}

// The call's counters that many workgroups bump (parts still searching,
// points swept) are kept in one copy per XCD (copy blockIdx mod kCopies):
// an atomic on one line shared by every XCD moves the line between their
// L2s at each update.  A gate (the parts still searching after a decision)
// is the sum of its copies.
__device__ __forceinline__ uint32_t xcd_copy() { return blockIdx.x % kCopies; }
__device__ __forceinline__ uint32_t gate_count(const uint32_t *g) {
    uint32_t s = 0;
#pragma unroll
    for (int c = 0; c < kCopies; c++) s += g[c];
    return s;
}

// Copies 1..7 of a search pass's bins (copy 0 is the bins themselves).
// Workgroups go to the XCDs in turn (blockIdx mod 8), each XCD with its own
// L2; with few parts, every workgroup of the grid flushes into the same few
// lines, which then move between the L2s at each atomic.  The workgroups of
// XCD c flush into copy c instead and the decision sums the copies (the
// atomics stay device-wide: any placement is correct, only slower).
// cnt == nullptr: no copies; parts: the copies' stride in parts.
template <typename Acc>
struct BinCopies {
    uint32_t *cnt;
    Acc *w;
    uint32_t *mn;
    uint32_t parts;
};
template <typename Acc>
__device__ __forceinline__ uint64_t copy_index(const BinCopies<Acc> &xc, int c, uint32_t part, int j) {
    return ((uint64_t)(c - 1) * xc.parts + part) * kMaxBins + j;
}

// v[a] for a uniform runtime axis a, as bit masks the compiler cannot turn
// into an indexed (scratch) access to a register array.
__device__ __forceinline__ float sel3(int a, float v0, float v1, float v2) {
    uint32_t m0 = a == 0 ? 0xffffffffu : 0u, m1 = a == 1 ? 0xffffffffu : 0u, m2 = a == 2 ? 0xffffffffu : 0u;
    asm volatile("" : "+v"(m0), "+v"(m1), "+v"(m2));
    return __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, v0) & m0) |
                                         (__builtin_bit_cast(uint32_t, v1) & m1) |
                                         (__builtin_bit_cast(uint32_t, v2) & m2));
}

// Candidate `p` (in-order index among nb-1) of the bisection tree rooted at
// (lo, hi): the midpoint sequence par_rcb_split computes (:472), following
// the path to node p.  Going left sets max = t, going right min = t.
__device__ inline float candidate(float lo, float hi, int p, int nb) {
    int s = nb >> 1;
    int q = s - 1;
    for (;;) {
        float t = (lo + hi) / 2.0f;
        if (p == q) return t;
        s >>= 1;
        if (p < q) { hi = t; q -= s; } else { lo = t; q += s; }
    }
}

// Number of candidates <= x (sorted ascending): the bin of x.  Bin j holds
// cand[j-1] <= c < cand[j], so count(c < cand[q]) = sum of bins 0..q.
// A NaN coordinate lands in bin 0, like !(c >= t) in the partition.
__device__ inline int bin_of(const float *cand, int nb, float x) {
    int pos = 0;
    for (int step = nb >> 1; step >= 1; step >>= 1)
        if (x >= cand[pos + step - 1]) pos += step;
    return pos;
}

// bin_of by an estimate and comparisons: the candidates of a pass lie close
// to lo + (hi - lo) k / nb, so (x - lo) * nb / (hi - lo) is within a bin or
// two of x's bin; the result is the bin where cand[pos - 1] <= x < cand[pos]
// (the candidates are sorted), exactly bin_of's.  A search whose candidates
// have collapsed (a few ulps wide) falls back to bin_of.  scale = nb / (hi - lo).
__device__ inline int bin_near(const float *cand, int nb, float x, float lo, float scale) {
    const float e = (x - lo) * scale;
    int pos = e >= 0.0f ? (e < (float)(nb - 1) ? (int)e : nb - 1) : 0;   // NaN: 0
#pragma unroll
    for (int it = 0; it < 3; it++) {
        const bool down = pos > 0 && x < cand[pos - 1];
        const bool up = pos < nb - 1 && x >= cand[pos];
        if (!down && !up) return pos;
        pos += up ? 1 : -1;
    }
    return bin_of(cand, nb, x);
}
// Same for the 3 register candidates of the first pass.
__device__ inline int bin4(float c0, float c1, float c2, float x) {
    int b = (x >= c1) ? 2 : 0;
    b += (x >= (b ? c2 : c0)) ? 1 : 0;
    return b;
}
// The tiles [t0, t1) of this workgroup: a contiguous share of the nt tiles,
// so per-part state (destination bases, register bins flushed at a part
// change) changes seldom within a workgroup.  Used by the kernels where every
// tile costs the same (the moves, the children's first pass); the later
// search passes keep the strided order because tiles of finished parts are
// skipped there, and contiguous shares of those would leave workgroups idle.
__device__ __forceinline__ void tile_range(uint32_t nt, uint32_t &t0, uint32_t &t1) {
    const uint64_t per = ((uint64_t)nt + gridDim.x - 1) / gridDim.x;
    const uint64_t a = (uint64_t)blockIdx.x * per;
    t0 = (uint32_t)(a < nt ? a : nt);
    t1 = (uint32_t)(a + per < nt ? a + per : nt);
}

}  // namespace coupe_hip
