// rcb_kernels.h -- HIP kernels (gfx950) of the RCB engine.
//
// Level structure (one level = one depth of coupe's recursion, for all parts
// of that depth at once):
//   k_hist    one pass: every workgroup sweeps one unit of a searching part
//             and bins its axis coordinate against the 2^b - 1 candidates
//             that coupe's split loop can reach in its next b steps
//             (par_rcb_split, recursive_bisection.rs:471-520); bins hold
//             count, weight sum and smallest coordinate, in LDS, then are
//             added into the part's bins with one atomic per bin.
//   k_decide  one wavefront per part: prefix sums / suffix minima over the
//             bins, then lane 0 replays coupe's decision loop exactly
//             (:522-571) for up to b steps.
//   k_unit_left + scans: where each unit's left/right points start.
//   k_partition  stable split of every part into {c < t} and {c >= t}:
//             wave ballots give ranks, each wave store covers two contiguous
//             runs.  Replaces reorder_split (:437-453, :122-181).
//   k_children   builds the next level's parts (:603-641) and units.
//   k_write_ids  last level: part ids through the original index (:589-601,
//             :698-702) without moving the data again.
#pragma once

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
__device__ inline int bin_of(const float *cand, int nb, float x) {
    int pos = 0;
    for (int step = nb >> 1; step >= 1; step >>= 1)
        if (x >= cand[pos + step - 1]) pos += step;
    return pos;
}

// ---------------------------------------------------------------------------
// Prologue: AoS f64 -> SoA f32 (`as f32`, :674-679), original index, bounding
// box (geometry.rs:42-76), weight sums (:684) and the weight-kind checks that
// select exact integer accumulation.
// ---------------------------------------------------------------------------
template <int D, typename WIN>
__global__ void __launch_bounds__(kBlock)
k_prologue(const double *__restrict__ pts, const WIN *__restrict__ win, uint64_t n,
           float *__restrict__ c0, float *__restrict__ c1, float *__restrict__ c2,
           uint32_t *__restrict__ ix, int32_t *__restrict__ w32, Prologue *out) {
    double mn[D], mx[D];
#pragma unroll
    for (int k = 0; k < D; k++) { mn[k] = 1.7976931348623157e308; mx[k] = -1.7976931348623157e308; }
    uint64_t su = 0;
    double sf = 0.0, sa = 0.0, ma = 0.0;
    uint32_t nonint = 0, not32 = 0;
    float *cs[3] = {c0, c1, c2};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            double v = pts[i * D + k];
            if (v < mn[k]) mn[k] = v;
            if (mx[k] < v) mx[k] = v;
            cs[k][i] = (float)v;
        }
        ix[i] = (uint32_t)i;
        WIN w = win[i];
        if constexpr (sizeof(WIN) == 8 && ((WIN)0.5) != 0) {   // double
            double a = fabs(w);
            bool integral = (w == trunc(w)) && a <= 9007199254740992.0;
            nonint |= !integral;
            sf += w;
            sa += a;
            ma = a > ma ? a : ma;
            if (integral) {
                su += (uint64_t)(int64_t)w;
                bool fits = w >= -2147483648.0 && w <= 2147483647.0;
                not32 |= !fits;
                w32[i] = fits ? (int32_t)w : 0;
            } else {
                not32 = 1;
            }
        } else {
            int64_t wi = (int64_t)w;
            su += (uint64_t)wi;
            double a = fabs((double)wi);
            sf += (double)wi;
            sa += a;
            ma = a > ma ? a : ma;
            bool fits = wi >= INT32_MIN && wi <= INT32_MAX;
            not32 |= !fits;
            w32[i] = (int32_t)wi;
        }
    }
    // wave reductions
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            double a = __shfl_xor(mn[k], o, 64);
            double b = __shfl_xor(mx[k], o, 64);
            mn[k] = a < mn[k] ? a : mn[k];
            mx[k] = mx[k] < b ? b : mx[k];
        }
        su += __shfl_xor(su, o, 64);
        sf += __shfl_xor(sf, o, 64);
        sa += __shfl_xor(sa, o, 64);
        double m2 = __shfl_xor(ma, o, 64);
        ma = m2 > ma ? m2 : ma;
        nonint |= __shfl_xor(nonint, o, 64);
        not32 |= __shfl_xor(not32, o, 64);
    }
    if (lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            atomicMin((unsigned long long *)&out->bmin[k], (unsigned long long)f64_key(mn[k]));
            atomicMax((unsigned long long *)&out->bmax[k], (unsigned long long)f64_key(mx[k]));
        }
        atomicAdd((unsigned long long *)&out->sum_u64, (unsigned long long)su);
        atomicAdd(&out->sum_f64, sf);
        atomicAdd(&out->sum_abs, sa);
        atomicMax((unsigned long long *)&out->max_abs_bits,
                  (unsigned long long)__builtin_bit_cast(uint64_t, ma));
        if (nonint) atomicOr(&out->non_integral, 1u);
        if (not32) atomicOr(&out->not_i32, 1u);
    }
}

// Weights stored as int64 (integer-valued, beyond int32) or kept as double.
template <typename WIN, typename WS>
__global__ void __launch_bounds__(kBlock)
k_store_weights(const WIN *__restrict__ win, WS *__restrict__ ws, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        ws[i] = (WS)win[i];
}

// ---------------------------------------------------------------------------
// Histogram pass over all searching parts.
// ---------------------------------------------------------------------------
template <typename WS, typename Acc>
__global__ void __launch_bounds__(kBlock)
k_hist(const Part *__restrict__ parts, const Unit *__restrict__ units, const uint32_t *n_units,
       const float *__restrict__ c0, const float *__restrict__ c1, const float *__restrict__ c2,
       const WS *__restrict__ w, int nb, uint32_t *__restrict__ bcnt, Acc *__restrict__ bw,
       uint32_t *__restrict__ bmin, uint32_t *__restrict__ ucnt) {
    __shared__ float cand[kMaxBins];
    __shared__ uint32_t scnt[kMaxBins];
    __shared__ Acc sw[kMaxBins];
    __shared__ uint32_t smin[kMaxBins];
    const uint32_t u = blockIdx.x;
    if (u >= *n_units) return;
    const Unit un = units[u];
    const Part &P = parts[un.part];
    if (P.status != kSearching) return;
    const int tid = threadIdx.x;
    const float blo = P.blo, bhi = P.bhi;
    for (int p = tid; p < nb - 1; p += kBlock) cand[p] = candidate(blo, bhi, p, nb);
    for (int j = tid; j < nb; j += kBlock) { scnt[j] = 0; sw[j] = (Acc)0; smin[j] = 0xffffffffu; }
    __syncthreads();
    const float *c = P.axis == 0 ? c0 : (P.axis == 1 ? c1 : c2);
    const uint64_t end = un.begin + un.count;
    for (uint64_t i = un.begin + tid; i < end; i += kBlock) {
        float x = c[i];
        Acc wv = AccOps<Acc>::of(w[i]);
        int b = bin_of(cand, nb, x);
        atomicAdd(&scnt[b], 1u);
        atomicAdd(&sw[b], wv);
        atomicMin(&smin[b], f32_key(x));
    }
    __syncthreads();
    const uint64_t base = (uint64_t)un.part * nb;
    for (int j = tid; j < nb; j += kBlock) {
        uint32_t k = scnt[j];
        ucnt[(uint64_t)u * nb + j] = k;
        if (k) {
            atomicAdd(&bcnt[base + j], k);
            atomicAdd(&bw[base + j], sw[j]);
            atomicMin(&bmin[base + j], smin[j]);
        }
    }
}

// ---------------------------------------------------------------------------
// Decision replay: one wavefront per searching part.
// ---------------------------------------------------------------------------
template <typename Acc>
__global__ void __launch_bounds__(64)
k_decide(Part *__restrict__ parts, const uint32_t *n_parts, uint32_t *__restrict__ bcnt,
         Acc *__restrict__ bw, uint32_t *__restrict__ bmin, int nb, double tol, int wrap32,
         uint32_t *n_searching, unsigned long long *points_swept) {
    __shared__ uint64_t pc[kMaxBins];
    __shared__ Acc pw[kMaxBins];
    __shared__ uint32_t sm[kMaxBins + 1];
    const uint32_t s = blockIdx.x;
    if (s >= *n_parts) return;
    Part *P = &parts[s];
    if (P->status != kSearching) return;
    const int lane = threadIdx.x;
    if (lane == 0) atomicAdd(points_swept, (unsigned long long)P->count);   // swept by the last pass
    const int per = nb >= 64 ? nb / 64 : 1;
    const int j0 = lane * per;
    const bool mine = j0 < nb;
    const uint64_t base = (uint64_t)s * nb;
    uint64_t c[4];
    Acc w[4];
    uint32_t m[4];
    uint64_t lc = 0;
    Acc lw = (Acc)0;
    uint32_t lm = 0xffffffffu;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k] = 0; w[k] = (Acc)0; m[k] = 0xffffffffu;
        if (k < per && mine) {
            uint64_t j = base + j0 + k;
            c[k] = bcnt[j]; w[k] = bw[j]; m[k] = bmin[j];
            bcnt[j] = 0; bw[j] = (Acc)0; bmin[j] = 0xffffffffu;   // ready for the next pass
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) { lc += c[k]; lw += w[k]; }
    // exclusive prefix of lane totals
    uint64_t ic = wave_incl_add(lc);
    Acc iw = wave_incl_add(lw);
    uint64_t ec = ic - lc;
    Acc ew = iw - lw;
    if (mine) {
        uint64_t rc = ec;
        Acc rw = ew;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < per) {
                rc += c[k]; rw += w[k];
                pc[j0 + k] = rc; pw[j0 + k] = rw;
            }
        }
    }
    // suffix minima: lane-local, then across lanes from the right
#pragma unroll
    for (int k = 3; k >= 0; k--) lm = m[k] < lm ? m[k] : lm;
    uint32_t suf = lm;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t v = __shfl_down(suf, o, 64);
        if (lane + o < 64) suf = v < suf ? v : suf;
    }
    uint32_t after = __shfl_down(suf, 1, 64);   // min over lanes > lane
    if (lane == 63) after = 0xffffffffu;
    if (mine) {
        uint32_t r = after;
        for (int k = per - 1; k >= 0; k--) {
            r = m[k] < r ? m[k] : r;
            sm[j0 + k] = r;
        }
    }
    if (lane == 0) sm[nb] = 0xffffffffu;
    __syncthreads();
    if (lane != 0) return;

    // Replay of par_rcb_split (:470-571) over the candidate tree.
    using Ops = AccOps<Acc>;
    const Acc S = Ops::from_bits(P->sum);
    const uint64_t n = P->count;
    float lo = P->blo, hi = P->bhi;
    uint64_t prev = P->prev;
    int step = nb >> 1;
    int q = step - 1;
    for (;;) {
        const float t = (lo + hi) / 2.0f;
        const uint64_t cnt = pc[q];
        const Acc wl = pw[q];
        bool left;
        if (cnt == n) {                                   // no point at or right of t
            if (prev == cnt) {                            // :528-539
                P->status = kAllLeft; P->split = hi; P->wl = Ops::to_bits(S);
                P->cnt_left = n; P->kstar = -1;
                return;
            }
            hi = t; prev = cnt;                           // :540-544
            left = true;
        } else {
            const float mnear = key_f32(sm[q + 1]);       // smallest c >= t
            const float d = mnear - t;                    // nearest_distance
            const double ideal = Ops::f64(S, wrap32) / 2.0;
            const double imb = fabs((Ops::f64(wl, wrap32) - ideal) / ideal);
            if (cnt == prev || hi <= t + d || imb <= tol) {   // :552-563
                P->status = kSplitAtT; P->split = t; P->wl = Ops::to_bits(wl);
                P->cnt_left = cnt; P->kstar = q;
                return;
            }
            prev = cnt;                                   // :564
            if (Ops::lt(wl, Ops::sub(S, wl), wrap32)) { lo = t; left = false; }
            else { hi = t; left = true; }
        }
        step >>= 1;
        if (step == 0) break;
        q = left ? q - step : q + step;
    }
    P->blo = lo; P->bhi = hi; P->prev = prev;             // next pass continues here
    atomicAdd(n_searching, 1u);
}

// Points of each unit that go left (all of them for kAllLeft parts).
__global__ void __launch_bounds__(kBlock)
k_unit_left(const Part *__restrict__ parts, const Unit *__restrict__ units, const uint32_t *n_units,
            const uint32_t *__restrict__ ucnt, int nb, uint64_t *__restrict__ left) {
    const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= *n_units) return;
    const Unit un = units[u];
    const Part &P = parts[un.part];
    uint64_t l;
    if (P.status == kAllLeft) {
        l = un.count;
    } else {
        l = 0;
        const uint32_t *row = ucnt + (uint64_t)u * nb;
        for (int j = 0; j <= P.kstar; j++) l += row[j];
    }
    left[u] = l;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of uint64 (3 kernels).  n is read on the device;
// the launch covers an upper bound.  `total` receives the sum.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t *wsum, uint64_t *total) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    uint64_t inc = wave_incl_add(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
        for (int i = 0; i < kBlock / 64; i++) { uint64_t t = wsum[i]; wsum[i] = run; run += t; }
        *total = run;
    }
    __syncthreads();
    uint64_t r = wsum[wid] + inc - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kBlock)
k_scan_local(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
             uint64_t *__restrict__ bsum, const uint32_t *n_dev) {
    __shared__ uint64_t wsum[kBlock / 64];
    __shared__ uint64_t tot;
    const uint64_t n = *n_dev;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    if ((uint64_t)blockIdx.x * kScanTile >= n) return;
    uint64_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    uint64_t e = block_excl_scan(s, wsum, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = e;
        e += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock)
k_scan_top(uint64_t *__restrict__ bsum, const uint32_t *n_dev, uint64_t *total) {
    __shared__ uint64_t wsum[kBlock / 64];
    __shared__ uint64_t tot;
    const uint64_t nb = (*n_dev + kScanTile - 1) / kScanTile;
    uint64_t carry = 0;
    for (uint64_t b0 = 0; b0 < nb; b0 += kBlock) {
        uint64_t i = b0 + threadIdx.x;
        uint64_t v = i < nb ? bsum[i] : 0;
        uint64_t e = block_excl_scan(v, wsum, &tot);
        if (i < nb) bsum[i] = carry + e;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(kBlock)
k_scan_add(uint64_t *__restrict__ out, const uint64_t *__restrict__ bsum, const uint32_t *n_dev) {
    const uint64_t n = *n_dev;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    if (base >= n) return;
    const uint64_t add = bsum[blockIdx.x];
    for (int k = threadIdx.x; k < kScanTile; k += kBlock)
        if (base + k < n) out[base + k] += add;
}

// ---------------------------------------------------------------------------
// Stable partition of every part: left = {c < t} (or all), right = the rest.
// ---------------------------------------------------------------------------
template <int D, typename WS>
__global__ void __launch_bounds__(kBlock)
k_partition(const Part *__restrict__ parts, const Unit *__restrict__ units, const uint32_t *n_units,
            const uint64_t *__restrict__ uleft,
            const float *__restrict__ i0, const float *__restrict__ i1, const float *__restrict__ i2,
            const WS *__restrict__ iw, const uint32_t *__restrict__ iix,
            float *__restrict__ o0, float *__restrict__ o1, float *__restrict__ o2,
            WS *__restrict__ ow, uint32_t *__restrict__ oix, uint64_t n_total) {
    __shared__ uint32_t wtot[kItems][kBlock / 64];
    __shared__ uint32_t tile_left;
    const uint32_t u = blockIdx.x;
    if (u >= *n_units) return;
    const Unit un = units[u];
    const Part &P = parts[un.part];
    const int tid = threadIdx.x, lane = lane_id(), wid = tid >> 6;
    const bool all_left = P.status == kAllLeft;
    const float t = P.split;
    const uint64_t lprefix = uleft[u] - uleft[P.unit_begin];
    const uint64_t lbase = P.begin + lprefix;
    const uint64_t rbase = P.begin + P.cnt_left + ((un.begin - P.begin) - lprefix);
    const float *ins[3] = {i0, i1, i2};
    float *outs[3] = {o0, o1, o2};
    const float *ca = ins[P.axis];
    uint64_t runL = 0, runR = 0;
    const uint64_t end = un.begin + un.count;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint64_t t0 = un.begin; t0 < end; t0 += kTile) {
        bool f[kItems];
        uint32_t pre[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            uint64_t p = t0 + (uint64_t)k * kBlock + tid;
            bool valid = p < end;
            // !(c >= t) is c < t for numbers and sends NaN left, as bin_of does
            f[k] = valid && (all_left || !(ca[p] >= t));
            uint64_t bal = __ballot(f[k]);
            pre[k] = __popcll(bal & lt_mask);
            if (lane == 0) wtot[k][wid] = __popcll(bal);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t run = 0;
            for (int k = 0; k < kItems; k++)
                for (int wv = 0; wv < kBlock / 64; wv++) { uint32_t x = wtot[k][wv]; wtot[k][wv] = run; run += x; }
            tile_left = run;
        }
        __syncthreads();
        uint64_t dst[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            uint64_t off = (uint64_t)k * kBlock + tid;   // position in tile
            uint32_t rl = wtot[k][wid] + pre[k];          // left items before it
            dst[k] = f[k] ? lbase + runL + rl : rbase + runR + (off - rl);
        }
        const uint64_t valid_n = (end - t0) < (uint64_t)kTile ? (end - t0) : (uint64_t)kTile;
        const uint32_t tl = tile_left;
#pragma unroll
        for (int a = 0; a < D; a++) {
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                uint64_t p = t0 + (uint64_t)k * kBlock + tid;
                if (p < end && dst[k] < n_total) outs[a][dst[k]] = ins[a][p];
            }
        }
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            uint64_t p = t0 + (uint64_t)k * kBlock + tid;
            if (p < end && dst[k] < n_total) { ow[dst[k]] = iw[p]; oix[dst[k]] = iix[p]; }
        }
        runL += tl;
        runR += valid_n - tl;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Next level's parts.  Children of part s: left (2i+1, box hi = split, sum wl)
// and right (2i+2, box lo = split, sum S - wl), kept when non-empty
// (:586-588, :603-641).
// ---------------------------------------------------------------------------
__device__ inline uint32_t units_for(uint64_t count, uint64_t unit) {
    return (uint32_t)((count + unit - 1) / unit);
}

template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_children_count(const Part *__restrict__ parts, const uint32_t *n_parts, uint64_t unit,
                 uint64_t *__restrict__ nchild, uint64_t *__restrict__ nunits) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_parts) return;
    const Part &P = parts[s];
    const uint64_t nl = P.cnt_left, nr = P.count - P.cnt_left;
    nchild[s] = (nl > 0) + (nr > 0);
    nunits[s] = units_for(nl, unit) + units_for(nr, unit);
}

template <typename Acc>
__global__ void __launch_bounds__(64)
k_children_write(const Part *__restrict__ parts, const uint32_t *n_parts, int dim, uint64_t unit,
                 const uint64_t *__restrict__ child_off, const uint64_t *__restrict__ unit_off,
                 Part *__restrict__ next, Unit *__restrict__ next_units) {
    const uint32_t s = blockIdx.x;
    if (s >= *n_parts) return;
    const Part P = parts[s];
    const uint64_t nl = P.cnt_left, nr = P.count - P.cnt_left;
    const uint32_t a = P.axis;
    const uint32_t na = (a + 1) % dim;
    const Acc S = AccOps<Acc>::from_bits(P.sum);
    const Acc wl = AccOps<Acc>::from_bits(P.wl);
    uint64_t co = child_off[s];
    uint64_t uo = unit_off[s];
    for (int side = 0; side < 2; side++) {
        const uint64_t cnt = side == 0 ? nl : nr;
        if (cnt == 0) continue;
        const uint64_t b = side == 0 ? P.begin : P.begin + nl;
        const uint32_t nu = units_for(cnt, unit);
        if (threadIdx.x == 0) {
            Part C;
            C.begin = b;
            C.count = cnt;
            C.heap_id = 2 * P.heap_id + 1 + side;
            C.sum = side == 0 ? AccOps<Acc>::to_bits(wl) : AccOps<Acc>::to_bits(AccOps<Acc>::sub(S, wl));
            for (int k = 0; k < kMaxDim; k++) { C.lo[k] = P.lo[k]; C.hi[k] = P.hi[k]; }
            if (side == 0) C.hi[a] = P.split; else C.lo[a] = P.split;
            C.axis = na;
            C.blo = C.lo[na];
            C.bhi = C.hi[na];
            C.prev = ~0ull;
            C.status = kSearching;
            C.split = 0.0f;
            C.kstar = -1;
            C.wl = 0;
            C.cnt_left = 0;
            C.unit_begin = (uint32_t)uo;
            C.unit_count = nu;
            next[co] = C;
        }
        for (uint32_t k = threadIdx.x; k < nu; k += 64) {
            Unit U;
            U.begin = b + (uint64_t)k * unit;
            uint64_t rem = cnt - (uint64_t)k * unit;
            U.count = (uint32_t)(rem < unit ? rem : unit);
            U.part = (uint32_t)co;
            next_units[uo + k] = U;
        }
        co++;
        uo += nu;
    }
}

// Smallest leaf heap id (:698-699).
__global__ void __launch_bounds__(kBlock)
k_leaf_min(const Part *__restrict__ parts, const uint32_t *n_parts, uint64_t *min_id) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    uint64_t v = ~0ull;
    if (s < *n_parts) {
        const Part &P = parts[s];
        v = P.cnt_left > 0 ? 2 * P.heap_id + 1 : 2 * P.heap_id + 2;
        if (P.cnt_left < P.count) {
            uint64_t r = 2 * P.heap_id + 2;
            v = v < r ? v : r;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint64_t x = __shfl_xor(v, o, 64);
        v = x < v ? x : v;
    }
    if (lane_id() == 0 && v != ~0ull) atomicMin((unsigned long long *)min_id, (unsigned long long)v);
}

// Last level: part id of every point (:589-601) minus the offset (:700-702).
__global__ void __launch_bounds__(kBlock)
k_write_ids(const Part *__restrict__ parts, const Unit *__restrict__ units, const uint32_t *n_units,
            const float *__restrict__ c0, const float *__restrict__ c1, const float *__restrict__ c2,
            const uint32_t *__restrict__ ix, const uint64_t *min_id, uint64_t *__restrict__ out,
            uint64_t n_total) {
    const uint32_t u = blockIdx.x;
    if (u >= *n_units) return;
    const Unit un = units[u];
    const Part &P = parts[un.part];
    const float *c = P.axis == 0 ? c0 : (P.axis == 1 ? c1 : c2);
    const bool all_left = P.status == kAllLeft;
    const float t = P.split;
    const uint64_t off = *min_id;
    const uint64_t idl = 2 * P.heap_id + 1 - off, idr = 2 * P.heap_id + 2 - off;
    const uint64_t end = un.begin + un.count;
    for (uint64_t i = un.begin + threadIdx.x; i < end; i += kBlock) {
        bool l = all_left || !(c[i] >= t);
        const uint32_t o = ix[i];
        if (o < n_total) out[o] = l ? idl : idr;
    }
}

// iter_count == 0: every point is in part 0.
__global__ void __launch_bounds__(kBlock) k_fill_u64(uint64_t *out, uint64_t n, uint64_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = v;
}

// Root part: whole array, box from the prologue, sum in the accumulator.
template <typename Acc>
__global__ void k_init_root(Part *parts, uint32_t *n_parts, Unit *units, uint32_t *n_units,
                            const Prologue *pro, uint64_t n, int dim, uint64_t unit, int real_mode) {
    const uint32_t nu = units_for(n, unit);
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        Part R;
        R.begin = 0;
        R.count = n;
        R.heap_id = 0;
        R.sum = real_mode ? __builtin_bit_cast(uint64_t, pro->sum_f64) : pro->sum_u64;
        for (int k = 0; k < kMaxDim; k++) {
            R.lo[k] = k < dim ? (float)key_f64(pro->bmin[k]) : 0.0f;   // `as f32`, :604-605
            R.hi[k] = k < dim ? (float)key_f64(pro->bmax[k]) : 0.0f;
        }
        R.axis = 0;
        R.blo = R.lo[0];
        R.bhi = R.hi[0];
        R.prev = ~0ull;
        R.status = kSearching;
        R.split = 0.0f;
        R.kstar = -1;
        R.wl = 0;
        R.cnt_left = 0;
        R.unit_begin = 0;
        R.unit_count = nu;
        parts[0] = R;
        *n_parts = 1;
        *n_units = nu;
    }
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nu;
         k += (uint64_t)gridDim.x * blockDim.x) {
        Unit U;
        U.begin = k * unit;
        uint64_t rem = n - k * unit;
        U.count = (uint32_t)(rem < unit ? rem : unit);
        U.part = 0;
        units[k] = U;
    }
}

// Copies the 64-bit totals of the scans into the 32-bit table sizes.
__global__ void k_set_sizes(const uint64_t *parts_total, const uint64_t *units_total,
                            uint32_t *n_parts, uint32_t *n_units) {
    if (threadIdx.x == 0) {
        *n_parts = (uint32_t)*parts_total;
        *n_units = (uint32_t)*units_total;
    }
}

}  // namespace coupe_hip
