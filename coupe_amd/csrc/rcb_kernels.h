// rcb_kernels.h -- HIP kernels (gfx950) of the RCB engine.
//
// One level = one depth of coupe's recursion, processed for all its parts at
// once.  The working set is SoA: f32 coordinates per axis, stored weights, and
// the u32 original index, in two copies (each partition writes the other one).
//
//   k_prologue   AoS f64 -> SoA f32, bounding box, weight sums (:674-688)
//   k_sweep      first split pass of the root: 3 candidates (2 bisection steps),
//                bins accumulated in registers, no atomics per point
//   k_hist       further passes for parts still searching: 255 candidates
//                (8 steps), bins in LDS
//   k_decide     one wavefront per searching part: prefix sums over the bins,
//                then lane 0 replays coupe's split loop exactly (:470-571)
//   k_children   next level's parts and tiles (:603-641)
//   k_partition_tiles  stable split of every part into {c < t} | {c >= t}
//                from per-tile left counts and their scan, with LDS-staged
//                contiguous stores; in the same sweep it bins the children's
//                next-axis coordinates (their first split pass)
//   k_descend    part ids in input order by walking the split tree (:589-601,
//                :698-702)
#pragma once

#include <type_traits>

#include "rcb_device.h"

namespace coupe_hip {


// ---------------------------------------------------------------------------
// Prologue: AoS f64 -> SoA f32 (`as f32`, :674-679), bounding box
// (geometry.rs:42-76), weight sums (:684) and the weight-kind checks that
// select exact integer accumulation.  Points are read through LDS so every
// global load is contiguous.
// ---------------------------------------------------------------------------
constexpr int kProPoints = kBlock * 4;   // points per prologue round (4 per thread)

// Per-instance bounding box and weight sums, reduced over the workgroup and
// added to the instance with one atomic per field.
template <int D>
__device__ __forceinline__ void flush_instance(InstReduce *I, double (&mn)[D], double (&mx)[D], uint64_t &su,
                                               double &sf, double *s_d, uint64_t *s_u) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double a = __shfl_xor(mn[k], o, 64);
            const double b = __shfl_xor(mx[k], o, 64);
            mn[k] = a < mn[k] ? a : mn[k];
            mx[k] = mx[k] < b ? b : mx[k];
        }
        su += __shfl_xor(su, o, 64);
        sf += __shfl_xor(sf, o, 64);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < D; k++) { s_d[wid * 8 + k] = mn[k]; s_d[wid * 8 + 3 + k] = mx[k]; }
        s_d[wid * 8 + 6] = sf;
        s_u[wid] = su;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < kBlock / 64; v++) {
#pragma unroll
            for (int k = 0; k < D; k++) {
                const double a = s_d[v * 8 + k], b = s_d[v * 8 + 3 + k];
                s_d[k] = a < s_d[k] ? a : s_d[k];
                s_d[3 + k] = s_d[3 + k] < b ? b : s_d[3 + k];
            }
            s_d[6] += s_d[v * 8 + 6];
            s_u[0] += s_u[v];
        }
#pragma unroll
        for (int k = 0; k < D; k++) {
            atomicMin((unsigned long long *)&I->bmin[k], (unsigned long long)f64_key(s_d[k]));
            atomicMax((unsigned long long *)&I->bmax[k], (unsigned long long)f64_key(s_d[3 + k]));
        }
        atomicAdd((unsigned long long *)&I->sum_u64, (unsigned long long)s_u[0]);
        atomicAdd(&I->sum_f64, s_d[6]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < D; k++) { mn[k] = 1.7976931348623157e308; mx[k] = -1.7976931348623157e308; }
    su = 0;
    sf = 0.0;
}

// Per-thread weight sums and the weight-kind flags of the prologue; w32 /
// w16 (each optional) receive the weight as int32 and as a 16-bit unsigned
// integer (used when every weight is an integer in [0, 65535]).
struct WeightScan {
    uint64_t su = 0;    // integer sum (two's complement, exact in any order)
    double sf = 0.0;    // f64 sum
    double sa = 0.0;    // sum of |w|
    double ma = 0.0;    // max |w|
    uint32_t nonint = 0, not32 = 0, not16 = 0;

    template <typename WIN>
    __device__ __forceinline__ void add(WIN w, uint64_t i, int32_t *__restrict__ w32, uint16_t *__restrict__ w16) {
        if constexpr (std::is_same<WIN, double>::value) {
            const double aw = fabs(w);
            const bool integral = (w == trunc(w)) && aw <= 9007199254740992.0;
            nonint |= !integral;
            sf += w;
            sa += aw;
            ma = aw > ma ? aw : ma;
            if (integral) {
                su += (uint64_t)(int64_t)w;
                const bool fits = w >= -2147483648.0 && w <= 2147483647.0;
                not32 |= !fits;
                if (w32) __builtin_nontemporal_store(fits ? (int32_t)w : 0, &w32[i]);
                const bool f16 = w >= 0.0 && w <= 65535.0;
                not16 |= !f16;
                if (w16) __builtin_nontemporal_store(f16 ? (uint16_t)w : (uint16_t)0, &w16[i]);
            } else {
                not32 = 1;
                not16 = 1;
            }
        } else {
            const int64_t wi = (int64_t)w;
            su += (uint64_t)wi;
            const double aw = fabs((double)wi);
            sf += (double)wi;
            sa += aw;
            ma = aw > ma ? aw : ma;
            const bool fits = wi >= INT32_MIN && wi <= INT32_MAX;
            not32 |= !fits;
            if (w32) __builtin_nontemporal_store((int32_t)wi, &w32[i]);
            const bool f16 = wi >= 0 && wi <= 65535;
            not16 |= !f16;
            if (w16) __builtin_nontemporal_store((uint16_t)wi, &w16[i]);
        }
    }

    // the flags and |w| statistics of the whole call: the workgroup's, into
    // its XCD's copy of the flags (blockIdx mod kCopies; the host combines
    // the copies).  Every thread; s_d holds kBlock / 64 * 8 doubles.
    __device__ __forceinline__ void flush_flags(Prologue *out, double *s_d) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sa += __shfl_xor(sa, o, 64);
            const double m2 = __shfl_xor(ma, o, 64);
            ma = m2 > ma ? m2 : ma;
            nonint |= __shfl_xor(nonint, o, 64);
            not32 |= __shfl_xor(not32, o, 64);
            not16 |= __shfl_xor(not16, o, 64);
        }
        __syncthreads();   // s_d free
        const int wid = threadIdx.x >> 6;
        if (lane_id() == 0) {
            s_d[wid * 4] = sa;
            s_d[wid * 4 + 1] = ma;
            s_d[wid * 4 + 2] = (double)(nonint | not32 << 1 | not16 << 2);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, m = 0.0;
            uint32_t f = 0;
            for (int v = 0; v < kBlock / 64; v++) {
                a += s_d[v * 4];
                m = s_d[v * 4 + 1] > m ? s_d[v * 4 + 1] : m;
                f |= (uint32_t)s_d[v * 4 + 2];
            }
            Prologue *c = &out[blockIdx.x % kCopies];
            atomicAdd(&c->sum_abs, a);
            atomicMax((unsigned long long *)&c->max_abs_bits, (unsigned long long)__builtin_bit_cast(uint64_t, m));
            if (f & 1u) atomicOr(&c->non_integral, 1u);
            if (f & 2u) atomicOr(&c->not_i32, 1u);
            if (f & 4u) atomicOr(&c->not_u16, 1u);
        }
    }
};

// Workgroups stride over the level-0 tiles (each inside one instance); each
// lane reads its own points' AoS doubles (a wavefront's three loads per point
// cover whole lines between them, so no LDS staging is needed).
// With ROT every point p is replaced by rot.m p (Rib's obb_to_aabb map,
// recursive_bisection.rs:850) before its coordinates are used.
template <int D, typename WIN, bool ROT>
__global__ void __launch_bounds__(kBlock)
k_prologue(const Tile *__restrict__ tiles, uint32_t n_tiles, const double *__restrict__ pts,
           const WIN *__restrict__ win, float *__restrict__ c0, float *__restrict__ c1, float *__restrict__ c2,
           int32_t *__restrict__ w32, uint16_t *__restrict__ w16, Instance *__restrict__ insts, Prologue *out,
           Rotation rot, InstReduce *red, uint32_t n_inst) {
    // w32, w16 (each optional): the weights as int32 and as 16-bit unsigned
    // integers; the latter are used when every weight is an integer in
    // [0, 65535] (not_u16 stays 0)
    __shared__ double s_d[kBlock / 64 * 8];
    __shared__ uint64_t s_u[kBlock / 64];
    const int tid = threadIdx.x;
    double mn[D], mx[D];
#pragma unroll
    for (int k = 0; k < D; k++) { mn[k] = 1.7976931348623157e308; mx[k] = -1.7976931348623157e308; }
    WeightScan ws;
    uint32_t cur = 0xffffffffu;
    // the instance's points and weights in the caller's arrays
    const double *ipts = pts;
    const WIN *iwin = win;
    uint64_t ibeg = 0;
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const Tile T = tiles[t];
        if (T.part != cur) {
            if (cur != 0xffffffffu) flush_instance<D>(&red[(blockIdx.x % kCopies) * n_inst + cur], mn, mx, ws.su, ws.sf, s_d, s_u);
            cur = T.part;
            ipts = pts + insts[cur].src_pt;
            iwin = win + insts[cur].src;
            ibeg = insts[cur].begin;
        }
        for (uint32_t r0 = 0; r0 < T.count; r0 += kProPoints) {
            const uint64_t base = T.begin + r0;
            const uint32_t cnt = T.count - r0 < (uint32_t)kProPoints ? T.count - r0 : (uint32_t)kProPoints;
            const double *src = ipts + (base - ibeg) * D;
            const WIN *wsrc = iwin + (base - ibeg);
            // every load of the round issued back to back (clamped to the
            // round's last point).  The caller's arrays are read once and the
            // working arrays written once, each far larger than the caches:
            // non-temporal loads and stores (2^27 points 3D: 1.23 -> 1.20 ms)
            double qv[4][D];
            WIN wv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t pc = min((uint32_t)(k * kBlock + tid), cnt - 1u);
#pragma unroll
                for (int a = 0; a < D; a++) qv[k][a] = __builtin_nontemporal_load(&src[(uint64_t)pc * D + a]);
                wv[k] = __builtin_nontemporal_load(&wsrc[pc]);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t p = k * kBlock + tid;
                if (p < cnt) {
                    const uint64_t i = base + p;
                    double q[D];
#pragma unroll
                    for (int a = 0; a < D; a++) q[a] = qv[k][a];
                    if constexpr (ROT) {
                        double r[D];
#pragma unroll
                        for (int a = 0; a < D; a++) {
                            r[a] = rot.m[a * D] * q[0];
#pragma unroll
                            for (int b = 1; b < D; b++) r[a] = r[a] + rot.m[a * D + b] * q[b];
                        }
#pragma unroll
                        for (int a = 0; a < D; a++) q[a] = r[a];
                    }
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        const double v = q[a];
                        if (v < mn[a]) mn[a] = v;
                        if (mx[a] < v) mx[a] = v;
                        float *cs = a == 0 ? c0 : (a == 1 ? c1 : c2);
                        __builtin_nontemporal_store((float)v, &cs[i]);
                    }
                    ws.add(wv[k], i, w32, w16);
                }
            }
        }
    }
    if (cur != 0xffffffffu) flush_instance<D>(&red[(blockIdx.x % kCopies) * n_inst + cur], mn, mx, ws.su, ws.sf, s_d, s_u);
    ws.flush_flags(out, s_d);
}

// The prologue of a repartition that starts from a resident working copy
// (SoA f32 coordinates in input order, weights as given): the per-instance
// bounding box of the f32 coordinates (the f32 cast is monotone, so its
// extremes are the f32 casts of the f64 extremes coupe takes,
// geometry.rs:42-76), the weight sums and flags, and the weight copies the
// expected mode reads.  Nothing is written to the coordinates.
template <int D, typename WIN>
__global__ void __launch_bounds__(kBlock)
k_prologue_soa(const Tile *__restrict__ tiles, uint32_t n_tiles, const float *__restrict__ c0,
               const float *__restrict__ c1, const float *__restrict__ c2, const WIN *__restrict__ win,
               int32_t *__restrict__ w32, uint16_t *__restrict__ w16, Instance *__restrict__ insts, Prologue *out,
               InstReduce *red, uint32_t n_inst) {
    __shared__ double s_d[kBlock / 64 * 8];
    __shared__ uint64_t s_u[kBlock / 64];
    double mn[D], mx[D];
#pragma unroll
    for (int k = 0; k < D; k++) { mn[k] = 1.7976931348623157e308; mx[k] = -1.7976931348623157e308; }
    WeightScan ws;
    uint32_t cur = 0xffffffffu;
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const Tile T = tiles[t];
        if (T.part != cur) {
            if (cur != 0xffffffffu) flush_instance<D>(&red[(blockIdx.x % kCopies) * n_inst + cur], mn, mx, ws.su, ws.sf, s_d, s_u);
            cur = T.part;
        }
        for (uint32_t p = threadIdx.x; p < T.count; p += kBlock) {
            const uint64_t i = T.begin + p;
            const float x[3] = {c0[i], c1[i], D > 2 ? c2[i] : 0.0f};
#pragma unroll
            for (int a = 0; a < D; a++) {
                const double v = (double)x[a];
                if (v < mn[a]) mn[a] = v;
                if (mx[a] < v) mx[a] = v;
            }
            ws.add(win[i], i, w32, w16);
        }
    }
    if (cur != 0xffffffffu) flush_instance<D>(&red[(blockIdx.x % kCopies) * n_inst + cur], mn, mx, ws.su, ws.sf, s_d, s_u);
    ws.flush_flags(out, s_d);
}

// Working copy of a resident mesh: points [first, first + count) as f32 SoA
// (`as f32`, :677) and their weights, from f64 AoS points and weights given
// for that range.  Null `pts` or `w` leaves that array unchanged.
template <int D, typename WIN>
__global__ void __launch_bounds__(kBlock)
k_mesh_store(const double *__restrict__ pts, const WIN *__restrict__ w, uint64_t first, uint64_t count,
             float *__restrict__ c0, float *__restrict__ c1, float *__restrict__ c2, WIN *__restrict__ mw) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < count; j += stride) {
        const uint64_t i = first + j;
        if (pts) {
            c0[i] = (float)pts[j * D];
            c1[i] = (float)pts[j * D + 1];
            if (D > 2) c2[i] = (float)pts[j * D + 2];
        }
        if (w) mw[i] = w[j];
    }
}

// Weights stored as int64 (integer-valued, beyond int32) or kept as double,
// at the working positions of the level-0 tiles (from each instance's range
// of the caller's weights).
template <typename WIN, typename WS>
__global__ void __launch_bounds__(kBlock)
k_store_weights(const Tile *__restrict__ tiles, uint32_t n_tiles, const Instance *__restrict__ insts,
                const WIN *__restrict__ win, WS *__restrict__ ws) {
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const Tile T = tiles[t];
        const Instance &I = insts[T.part];
        const WIN *src = win + I.src + (T.begin - I.begin);
        for (uint32_t p = threadIdx.x; p < T.count; p += kBlock) ws[T.begin + p] = (WS)src[p];
    }
}

// ---------------------------------------------------------------------------
// Register bins: block reduction of per-thread bins, then one global atomic
// per non-empty bin.  Bins [0, NB) go to part slot `slot0`, [NB, 2NB) to
// `slot1` (the two children of a part); a negative slot drops the bins.
// ---------------------------------------------------------------------------
template <int NR, typename Acc>
struct RegBins {
    uint32_t cnt[NR];
    Acc w[NR];
    uint32_t mn[NR];
    __device__ __forceinline__ void reset() {
#pragma unroll
        for (int j = 0; j < NR; j++) { cnt[j] = 0; w[j] = (Acc)0; mn[j] = 0xffffffffu; }
    }
    __device__ __forceinline__ void add(int b, Acc wv, uint32_t key) {
#pragma unroll
        for (int j = 0; j < NR; j++) {
            const bool s = b == j;
            cnt[j] += s ? 1u : 0u;
            w[j] += s ? wv : (Acc)0;
            const uint32_t k = s ? key : 0xffffffffu;
            mn[j] = k < mn[j] ? k : mn[j];
        }
    }
};

template <int NR, typename Acc>
struct RegBinsShared {
    uint32_t cnt[kBlock / 64][NR];
    Acc w[kBlock / 64][NR];
    uint32_t mn[kBlock / 64][NR];
};

template <int NR, int NB, typename Acc>
__device__ __forceinline__ void flush_bins(RegBins<NR, Acc> &r, RegBinsShared<NR, Acc> &sh, int64_t slot0, int64_t slot1,
                           uint32_t *__restrict__ gcnt, Acc *__restrict__ gw, uint32_t *__restrict__ gmn) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NR; j++) {
        uint32_t c = wave_sum(r.cnt[j]);
        Acc x = wave_sum(r.w[j]);
        uint32_t m = wave_min(r.mn[j]);
        if (lane == 0) { sh.cnt[wid][j] = c; sh.w[wid][j] = x; sh.mn[wid][j] = m; }
    }
    __syncthreads();
    const int j = threadIdx.x;
    if (j < NR) {
        uint32_t c = 0, m = 0xffffffffu;
        Acc x = (Acc)0;
#pragma unroll
        for (int v = 0; v < kBlock / 64; v++) {
            c += sh.cnt[v][j];
            x += sh.w[v][j];
            m = sh.mn[v][j] < m ? sh.mn[v][j] : m;
        }
        const int64_t slot = j < NB ? slot0 : slot1;
        if (c && slot >= 0) {
            const uint64_t g = (uint64_t)slot * NB + (j % NB);
            atomicAdd(&gcnt[g], c);
            atomic_add_acc(&gw[g], x);
            atomicMin(&gmn[g], m);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// First pass of the root part: the 3 candidates of 2 bisection steps, bins in
// registers.  Workgroups stride over the tile table and flush when the part
// changes.
// ---------------------------------------------------------------------------
template <typename WS, typename Acc>
__global__ void __launch_bounds__(kBlock)
k_sweep(const Part *__restrict__ parts, const Tile *__restrict__ tiles, const uint32_t *n_tiles,
        const float *__restrict__ c0, const float *__restrict__ c1, const float *__restrict__ c2,
        const WS *__restrict__ w, uint32_t *__restrict__ bcnt, Acc *__restrict__ bw, uint32_t *__restrict__ bmin) {
    __shared__ RegBinsShared<kFirstBins, Acc> sh;
    RegBins<kFirstBins, Acc> r;
    r.reset();
    const uint32_t nt = *n_tiles;
    uint32_t cur = 0xffffffffu;
    bool active = false;
    float e0 = 0.f, e1 = 0.f, e2 = 0.f;
    const float *c = c0;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile T = tiles[t];
        if (T.part != cur) {
            if (active) { flush_bins<kFirstBins, kFirstBins>(r, sh, cur, -1, bcnt, bw, bmin); r.reset(); }
            cur = T.part;
            const Part &P = parts[cur];
            active = P.status == kSearching;
            e0 = candidate(P.blo, P.bhi, 0, kFirstBins);
            e1 = candidate(P.blo, P.bhi, 1, kFirstBins);
            e2 = candidate(P.blo, P.bhi, 2, kFirstBins);
            c = P.axis == 0 ? c0 : (P.axis == 1 ? c1 : c2);
        }
        if (!active) continue;
        const uint64_t end = T.begin + T.count;
        for (uint64_t i = T.begin + threadIdx.x; i < end; i += kBlock) {
            const float x = c[i];
            r.add(bin4(e0, e1, e2, x), AccOps<Acc>::of(w[i]), f32_key(x));
        }
    }
    if (active) flush_bins<kFirstBins, kFirstBins>(r, sh, cur, -1, bcnt, bw, bmin);
}

// ---------------------------------------------------------------------------
// Later passes: nb - 1 candidates (nb <= 256) binned in LDS.
// ---------------------------------------------------------------------------
// LDS bins of a search pass: count, weight and smallest key per bin.  With
// 16-bit weights and integer sums, a part whose points a workgroup holds in
// its bins stay under 2^24 (`packed`) keeps count and weight in one 64-bit
// word (count << 40 | weight: at most 2^24 x 65535 < 2^40), one LDS atomic
// instead of two.  The smallest key is taken with an atomic only when the
// point's key is below the bin's current one (the bin's minimum only falls).
constexpr int kPackShift = 40;

template <typename WS, typename Acc>
__device__ __forceinline__ constexpr bool packable() {
    return std::is_same<WS, uint16_t>::value && std::is_same<Acc, uint64_t>::value;
}

// Whether a workgroup of a grid of `grid` workgroups striding over the tiles
// of a part of `count` points can hold its share in packed bins: the part
// spans at most count/kTile + 2 tiles, a workgroup takes at most
// ceil(tiles/grid) of them, so it bins at most count/grid + 2 kTile points.
__device__ __forceinline__ bool pack_fits(uint64_t count, uint32_t grid) {
    return count < (uint64_t)grid * ((1u << 24) - 2 * kTile);
}

template <typename WS, typename Acc>
__device__ __forceinline__ void bin_add(bool packed, uint32_t *scnt, Acc *sw, uint32_t *smin, int b, WS wv, float x) {
    if constexpr (packable<WS, Acc>()) {
        if (packed) {
            atomicAdd((unsigned long long *)&sw[b], (1ull << kPackShift) | (unsigned long long)wv);
        } else {
            atomicAdd(&scnt[b], 1u);
            atomic_add_acc(&sw[b], AccOps<Acc>::of(wv));
        }
    } else {
        atomicAdd(&scnt[b], 1u);
        atomic_add_acc(&sw[b], AccOps<Acc>::of(wv));
    }
    const uint32_t key = f32_key(x);
    if (key < __atomic_load_n(&smin[b], __ATOMIC_RELAXED)) atomicMin(&smin[b], key);
}

// The first and last bins of a pass hold every point below and above the
// search window: in a refining pass most of them, so that the LDS atomics of
// a wavefront would meet at two addresses.  Each lane sums its points of
// those two bins in registers; once per tile the wavefront's totals go to
// the bins, one atomic per bin (lane 0).
template <typename WS, typename Acc>
struct EdgeBins {
    uint32_t c[2] = {0u, 0u};
    Acc w[2] = {(Acc)0, (Acc)0};
    uint32_t m[2] = {0xffffffffu, 0xffffffffu};

    // Takes the point when bin b is the first or the last.
    __device__ __forceinline__ bool take(int b, int nb, WS wv, float x) {
        if (b != 0 && b != nb - 1) return false;
        const uint32_t key = f32_key(x);
        if (b == 0) { c[0]++; w[0] += AccOps<Acc>::of(wv); m[0] = min(m[0], key); }
        else        { c[1]++; w[1] += AccOps<Acc>::of(wv); m[1] = min(m[1], key); }
        return true;
    }

    // The whole wavefront, every lane active.
    __device__ __forceinline__ void flush(bool packed, int nb, uint32_t *scnt, Acc *sw, uint32_t *smin) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            if (__ballot(c[e] != 0) == 0) continue;   // wave-uniform
            const int b = e ? nb - 1 : 0;
            const uint32_t mk = wave_total_min(m[e]);
            bool done = false;
            if constexpr (packable<WS, Acc>()) {
                if (packed) {
                    const uint64_t v = wave_total((uint64_t)c[e] << kPackShift | w[e]);
                    if (lane_id() == 0) atomicAdd((unsigned long long *)&sw[b], (unsigned long long)v);
                    done = true;
                }
            }
            if (!done) {
                const uint32_t cn = wave_total(c[e]);
                const Acc ws = wave_total(w[e]);
                if (lane_id() == 0) { atomicAdd(&scnt[b], cn); atomic_add_acc(&sw[b], ws); }
            }
            if (lane_id() == 0 && mk < smin[b]) atomicMin(&smin[b], mk);
            c[e] = 0u;
            w[e] = (Acc)0;
            m[e] = 0xffffffffu;
        }
    }
};

// Where a workgroup flushes the bins of part `part`: the bins, or its XCD's
// copy while the level uses copies.
template <typename Acc>
struct BinsDest {
    uint32_t *cnt;
    Acc *w;
    uint32_t *mn;
    uint64_t base;
};
template <typename Acc>
__device__ __forceinline__ BinsDest<Acc> bins_dest(const BinCopies<Acc> &xc, uint32_t part, int nb, uint32_t *bcnt,
                                                   Acc *bw, uint32_t *bmin) {
    const int c = (int)(blockIdx.x % kCopies);
    if (xc.cnt && c) return {xc.cnt, xc.w, xc.mn, copy_index(xc, c, part, 0)};
    return {bcnt, bw, bmin, (uint64_t)part * nb};
}

template <typename Acc>
__device__ __forceinline__ uint32_t bin_count(bool packed, const uint32_t *scnt, const Acc *sw, int j) {
    if constexpr (std::is_same<Acc, uint64_t>::value)
        if (packed) return (uint32_t)(sw[j] >> kPackShift);
    return scnt[j];
}

template <typename Acc>
__device__ inline void flush_lds_bins(uint64_t base, int nb, const uint32_t *scnt, const Acc *sw,
                                      const uint32_t *smin, uint32_t *bcnt, Acc *bw, uint32_t *bmin,
                                      bool packed = false) {
    for (int j = threadIdx.x; j < nb; j += kBlock) {
        const uint32_t c = bin_count(packed, scnt, sw, j);
        if (c) {
            Acc wsum = sw[j];
            if constexpr (std::is_same<Acc, uint64_t>::value)
                if (packed) wsum &= (1ull << kPackShift) - 1;
            atomicAdd(&bcnt[base + j], c);
            atomic_add_acc(&bw[base + j], wsum);
            atomicMin(&bmin[base + j], smin[j]);
        }
    }
}

template <typename WS, typename Acc>
__global__ void __launch_bounds__(kBlock)
k_hist(const Part *__restrict__ parts, const Tile *__restrict__ tiles, const uint32_t *n_tiles,
       const float *__restrict__ c0, const float *__restrict__ c1, const float *__restrict__ c2,
       const WS *__restrict__ w, int nb, uint32_t *__restrict__ bcnt, Acc *__restrict__ bw,
       uint32_t *__restrict__ bmin, const uint32_t *gate, uint16_t *__restrict__ tile_bins, BinCopies<Acc> xc,
       uint32_t chunk) {
    if (gate && gate_count(gate) == 0) return;   // no part left searching after the previous decision
    __shared__ float cand[kMaxBins];
    __shared__ uint32_t scnt[kMaxBins];
    __shared__ Acc sw[kMaxBins];
    __shared__ uint32_t smin[kMaxBins];
    const int tid = threadIdx.x;
    const uint32_t nt = *n_tiles;
    uint32_t cur = 0xffffffffu;
    bool active = false;
    uint32_t before = 0;   // this thread's bin count of the part before the tile (tile_bins)
    const float *c = c0;
    float lo = 0.0f, scale = 0.0f;
    bool packed = false;
    auto flush = [&]() {
        const BinsDest<Acc> d = bins_dest(xc, cur, nb, bcnt, bw, bmin);
        flush_lds_bins(d.base, nb, scnt, sw, smin, d.cnt, d.w, d.mn, packed);
    };
    // chunks of `chunk` consecutive tiles go to the workgroups in turn: a
    // workgroup stays on one part over a chunk (one bin setup and flush)
    for (uint32_t t0 = blockIdx.x * chunk; t0 < nt; t0 += gridDim.x * chunk)
    for (uint32_t t = t0; t < min(nt, t0 + chunk); t++) {
        const Tile T = tiles[t];
        if (T.part != cur) {
            before = 0;
            __syncthreads();
            if (active) flush();
            __syncthreads();
            cur = T.part;
            const Part &P = parts[cur];
            active = P.status == kSearching;
            packed = packable<WS, Acc>() && pack_fits(P.count, gridDim.x);
            if (active) {
                for (int p = tid; p < nb - 1; p += kBlock) cand[p] = candidate(P.blo, P.bhi, p, nb);
                for (int j = tid; j < nb; j += kBlock) { scnt[j] = 0; sw[j] = (Acc)0; smin[j] = 0xffffffffu; }
                lo = P.blo;
                scale = (float)nb / (P.bhi - P.blo);
            }
            c = P.axis == 0 ? c0 : (P.axis == 1 ? c1 : c2);
            __syncthreads();
        }
        if (!active) continue;
        // every load of the tile issued before the binning (clamped, unconditional)
        const uint32_t cnt = T.count;
        float xs[kItems];
        WS wk[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint32_t p = min((uint32_t)(k * kBlock + tid), cnt - 1u);
            xs[k] = c[T.begin + p];
            wk[k] = w[T.begin + p];
        }
        EdgeBins<WS, Acc> edge;
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            if ((uint32_t)(k * kBlock + tid) >= cnt) continue;
            const float x = xs[k];
            const int b = bin_near(cand, nb, x, lo, scale);
            if (!edge.take(b, nb, wk[k], x)) bin_add<WS, Acc>(packed, scnt, sw, smin, b, wk[k], x);
        }
        edge.flush(packed, nb, scnt, sw, smin);
        if (tile_bins) {
            // each tile's own bin counts (the growth of the part's counts
            // over the tile) are kept for the partition's offsets
            __syncthreads();
            const uint32_t after = tid < nb ? bin_count(packed, scnt, sw, tid) : 0u;
            if (tid < nb) tile_bins[(uint64_t)t * kMaxBins + tid] = (uint16_t)(after - before);
            before = after;
            __syncthreads();
        }
    }
    __syncthreads();
    if (active) flush();
}

// ---------------------------------------------------------------------------
// Decision replay of one part from its pass's bins: pc, pw the inclusive
// prefix counts and weights over the bins, sm[j] the smallest key of bins
// j.. (sm[nb]: none), as arrays in LDS or in the thread's registers.  Run by
// one thread per part; true when the part is still searching after it (the
// caller counts those).
// ---------------------------------------------------------------------------
template <typename Acc, typename PC, typename PW, typename SM>
__device__ bool decide_replay(Part *P, const PC &pc, const PW &pw, const SM &sm, int nb, int wrap32,
                              uint32_t counts_ready, uint32_t *unfinished) {
    // Replay of par_rcb_split (:470-571) over the candidate tree.
    using Ops = AccOps<Acc>;
    const Acc S = Ops::from_bits(P->sum);
    const uint64_t n = P->count;
    const double tol = P->tol;
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
                return false;
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
                P->counts_ready = counts_ready;   // per-tile left counts already known
                return false;
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
    if (unfinished) {
        P->status = kAllLeft;
        P->split = hi;
        P->cnt_left = n;
        P->wl = P->sum;
        P->kstar = -1;
        atomicOr(unfinished, 1u);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Decision replay: one workgroup per searching part (passes of more than 64
// bins; k_decide_wave and k_decide_lanes take the others).
// ---------------------------------------------------------------------------
template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_decide(Part *__restrict__ parts, const uint32_t *n_parts, uint32_t *__restrict__ bcnt,
         Acc *__restrict__ bw, uint32_t *__restrict__ bmin, int nb, int wrap32,
         uint32_t *n_searching, unsigned long long *points_swept, uint32_t counts_ready,
         const uint32_t *gate, uint32_t *unfinished, BinCopies<Acc> xc) {
    // unfinished (non-null): the last queued decision of a speculative
    // search; a part still searching after it is closed as "all left" (a
    // consistent tree for the kernels that follow) and the call is flagged,
    // so the host runs it again with a search that checks every pass
    if (gate && gate_count(gate) == 0) return;   // no part left searching after the previous decision
    __shared__ uint64_t pc[kMaxBins];
    __shared__ Acc pw[kMaxBins];
    __shared__ uint32_t sm[kMaxBins + 1];
    __shared__ uint32_t gc[kMaxBins];   // the pass's bins, gathered by the workgroup
    __shared__ Acc gw[kMaxBins];
    __shared__ uint32_t gm[kMaxBins];
    const uint32_t s = blockIdx.x;
    if (s >= *n_parts) return;
    Part *P = &parts[s];
    if (P->status != kSearching) return;
    {
        // one bin per thread: the bins and their copies read (all loads in
        // flight at once) and left cleared for the next pass
        const int j = threadIdx.x;
        if (j < nb) {
            const uint64_t g = (uint64_t)s * nb + j;
            uint32_t cn = bcnt[g];
            Acc wt = bw[g];
            uint32_t mn = bmin[g];
            bcnt[g] = 0; bw[g] = (Acc)0; bmin[g] = 0xffffffffu;   // ready for the next pass
            if (xc.cnt) {
#pragma unroll
                for (int cp = 1; cp < kCopies; cp++) {
                    const uint64_t x = copy_index(xc, cp, s, j);
                    cn += xc.cnt[x]; wt += xc.w[x]; mn = min(mn, xc.mn[x]);
                }
#pragma unroll
                for (int cp = 1; cp < kCopies; cp++) {
                    const uint64_t x = copy_index(xc, cp, s, j);
                    xc.cnt[x] = 0; xc.w[x] = (Acc)0; xc.mn[x] = 0xffffffffu;
                }
            }
            gc[j] = cn; gw[j] = wt; gm[j] = mn;
        }
        __syncthreads();
    }
    // the first wavefront scans the bins (the others hold no bins: j0 >= nb)
    // and its lane 0 replays the decisions
    const int lane = threadIdx.x;
    if (lane == 0) atomicAdd(&points_swept[xcd_copy()], (unsigned long long)P->count);   // swept by the last pass
    const int per = nb >= 64 ? nb / 64 : 1;
    const int j0 = lane * per;
    const bool mine = j0 < nb;
    uint64_t c[4];
    Acc w[4];
    uint32_t m[4];
    uint64_t lc = 0;
    Acc lw = (Acc)0;
    uint32_t lm = 0xffffffffu;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k] = 0; w[k] = (Acc)0; m[k] = 0xffffffffu;
        if (k < per && mine) { c[k] = gc[j0 + k]; w[k] = gw[j0 + k]; m[k] = gm[j0 + k]; }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) { lc += c[k]; lw += w[k]; }
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
    if (decide_replay<Acc>(P, pc, pw, sm, nb, wrap32, counts_ready, unfinished))
        atomicAdd(&n_searching[xcd_copy()], 1u);
}

// The same decision, one wavefront per part: four parts per workgroup,
// each lane up to four consecutive bins, no workgroup barrier (the engine
// takes it for passes of at most 64 bins).
template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_decide_wave(Part *__restrict__ parts, const uint32_t *n_parts, uint32_t *__restrict__ bcnt,
              Acc *__restrict__ bw, uint32_t *__restrict__ bmin, int nb, int wrap32,
              uint32_t *n_searching, unsigned long long *points_swept, uint32_t counts_ready,
              const uint32_t *gate, uint32_t *unfinished, BinCopies<Acc> xc) {
    if (gate && gate_count(gate) == 0) return;   // no part left searching after the previous decision
    static_assert(kMaxBins <= 4 * 64, "at most four bins per lane");
    __shared__ uint64_t pc_all[kBlock / 64][kMaxBins];
    __shared__ Acc pw_all[kBlock / 64][kMaxBins];
    __shared__ uint32_t sm_all[kBlock / 64][kMaxBins + 1];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t s = blockIdx.x * (kBlock / 64) + wv;
    if (s >= *n_parts) return;
    Part *P = &parts[s];
    if (P->status != kSearching) return;
    // lane l holds bins l * per .. l * per + per - 1 (nb is a power of two)
    const int per = nb > 64 ? nb / 64 : 1;
    const int j0 = lane * per;
    uint64_t c[4];
    Acc w[4];
    uint32_t m[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k] = 0; w[k] = (Acc)0; m[k] = 0xffffffffu;
        if (k < per && j0 + k < nb) {   // the bin and its copies read, and left cleared for the next pass
            const int j = j0 + k;
            const uint64_t g = (uint64_t)s * nb + j;
            uint32_t cn = bcnt[g], mn = bmin[g];
            Acc wt = bw[g];
            bcnt[g] = 0; bw[g] = (Acc)0; bmin[g] = 0xffffffffu;
            if (xc.cnt) {
#pragma unroll
                for (int cp = 1; cp < kCopies; cp++) {
                    const uint64_t x = copy_index(xc, cp, s, j);
                    cn += xc.cnt[x]; wt += xc.w[x]; mn = min(mn, xc.mn[x]);
                }
#pragma unroll
                for (int cp = 1; cp < kCopies; cp++) {
                    const uint64_t x = copy_index(xc, cp, s, j);
                    xc.cnt[x] = 0; xc.w[x] = (Acc)0; xc.mn[x] = 0xffffffffu;
                }
            }
            c[k] = cn; w[k] = wt; m[k] = mn;
        }
    }
    uint64_t lc = 0;
    Acc lw = (Acc)0;
    uint32_t lm = 0xffffffffu;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        lc += c[k];
        lw += w[k];
        lm = m[k] < lm ? m[k] : lm;
    }
    const uint64_t ic = wave_incl_add(lc);
    const Acc iw = wave_incl_add(lw);
    uint32_t suf = lm;   // the smallest key of the bins from this lane's on
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_down(suf, o, 64);
        if (lane + o < 64) suf = v < suf ? v : suf;
    }
    uint32_t after = __shfl_down(suf, 1, 64);   // the smallest key of the later lanes' bins
    if (lane == 63) after = 0xffffffffu;
    if (j0 < nb) {
        uint64_t rc = ic - lc;
        Acc rw = iw - lw;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < per) {
                rc += c[k];
                rw += w[k];
                pc_all[wv][j0 + k] = rc;
                pw_all[wv][j0 + k] = rw;
            }
        uint32_t r = after;
#pragma unroll
        for (int k = 3; k >= 0; k--)
            if (k < per) {
                r = m[k] < r ? m[k] : r;
                sm_all[wv][j0 + k] = r;
            }
    }
    if (lane == 0) sm_all[wv][nb] = 0xffffffffu;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane != 0) return;
    atomicAdd(&points_swept[xcd_copy()], (unsigned long long)P->count);   // swept by the last pass
    if (decide_replay<Acc>(P, pc_all[wv], pw_all[wv], sm_all[wv], nb, wrap32, counts_ready, unfinished))
        atomicAdd(&n_searching[xcd_copy()], 1u);
}

// The same decision, one lane per part, for passes of at most kFirstBins
// bins (every first decision): each lane gathers its part's bins, and a
// wavefront counts its parts still searching and their points with one
// atomic each.
template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_decide_lanes(Part *__restrict__ parts, const uint32_t *n_parts, uint32_t *__restrict__ bcnt,
               Acc *__restrict__ bw, uint32_t *__restrict__ bmin, int nb, int wrap32,
               uint32_t *n_searching, unsigned long long *points_swept, uint32_t counts_ready,
               const uint32_t *gate, uint32_t *unfinished, BinCopies<Acc> xc) {
    if (gate && gate_count(gate) == 0) return;   // no part left searching after the previous decision
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    bool searching = false;
    uint64_t swept = 0;
    if (s < *n_parts && parts[s].status == kSearching) {
        Part *P = &parts[s];
        uint64_t pc[kFirstBins];
        Acc pw[kFirstBins];
        uint32_t mk[kFirstBins], sm[kFirstBins + 1];
        uint64_t c = 0;
        Acc w = (Acc)0;
#pragma unroll
        for (int j = 0; j < kFirstBins; j++) {
            uint32_t cn = 0, mn = 0xffffffffu;
            Acc wt = (Acc)0;
            if (j < nb) {   // the bin and its copies read, and left cleared for the next pass
                const uint64_t g = (uint64_t)s * nb + j;
                cn = bcnt[g];
                wt = bw[g];
                mn = bmin[g];
                bcnt[g] = 0; bw[g] = (Acc)0; bmin[g] = 0xffffffffu;
                if (xc.cnt) {
                    for (int cp = 1; cp < kCopies; cp++) {
                        const uint64_t x = copy_index(xc, cp, s, j);
                        cn += xc.cnt[x]; wt += xc.w[x]; mn = min(mn, xc.mn[x]);
                        xc.cnt[x] = 0; xc.w[x] = (Acc)0; xc.mn[x] = 0xffffffffu;
                    }
                }
            }
            c += cn;
            w += wt;
            pc[j] = c;
            pw[j] = w;
            mk[j] = mn;
        }
        sm[kFirstBins] = 0xffffffffu;
#pragma unroll
        for (int j = kFirstBins - 1; j >= 0; j--) sm[j] = min(mk[j], sm[j + 1]);
        swept = P->count;
        searching = decide_replay<Acc>(P, pc, pw, sm, nb, wrap32, counts_ready, unfinished);
    }
    const uint64_t sw = wave_sum(swept);
    const uint32_t ns = (uint32_t)__popcll(__ballot(searching));
    if ((threadIdx.x & 63) == 0) {
        if (sw) atomicAdd(&points_swept[xcd_copy()], (unsigned long long)sw);
        if (ns) atomicAdd(&n_searching[xcd_copy()], ns);
    }
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of uint64 (3 kernels).  n is read on the device;
// the launch covers an upper bound.  `total` receives the sum.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr uint64_t kScanAddDirect = 2048;   // blocks whose totals k_scan_add re-sums per block

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

// With `single` (a launch of one block covering every element) the block's
// total is the scan's total and no further kernel is needed.  The element
// count is *n_dev, or n_host when n_dev is null.
__global__ void __launch_bounds__(kBlock)
k_scan_local(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
             uint64_t *__restrict__ bsum, const uint32_t *n_dev, uint64_t n_host, uint64_t *single) {
    __shared__ uint64_t wsum[kBlock / 64];
    __shared__ uint64_t tot;
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    if ((uint64_t)blockIdx.x * kScanTile >= n) {
        // past the elements: a zero total, so a scan of the totals over the
        // launch's blocks stays exact
        if (threadIdx.x == 0) {
            bsum[blockIdx.x] = 0;
            if (single) *single = 0;
        }
        return;
    }
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
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = tot;
        if (single) *single = tot;
    }
}

// The second half of a multi-block scan of at most kScanAddDirect blocks:
// block b adds the totals of blocks 0..b-1 (read from bsum, the block totals
// k_scan_local wrote) to its elements; block 0 writes the scan's total.  Each
// block sums the preceding totals itself (at most kScanAddDirect loads per
// block), so no separate scan of the totals is launched.
__global__ void __launch_bounds__(kBlock)
k_scan_add(uint64_t *__restrict__ out, const uint64_t *__restrict__ bsum, const uint32_t *n_dev,
           uint64_t n_host, uint64_t *total) {
    __shared__ uint64_t wsum[kBlock / 64];
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    if (base >= n) {   // past the elements (block 0 only when there are none)
        if (blockIdx.x == 0 && threadIdx.x == 0 && total) *total = 0;
        return;
    }
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    const uint64_t upto = blockIdx.x == 0 ? nb : blockIdx.x;   // block 0: every total
    uint64_t s = 0;
    for (uint64_t i = threadIdx.x; i < upto; i += kBlock) s += bsum[i];
    s = wave_sum(s);
    if (lane_id() == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    s = 0;
#pragma unroll
    for (int v = 0; v < kBlock / 64; v++) s += wsum[v];
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0 && total) *total = s;
        return;   // block 0's elements need no carry
    }
    for (int k = threadIdx.x; k < kScanTile; k += kBlock)
        if (base + k < n) out[base + k] += s;
}

// The second half of a scan of more blocks: the block totals were scanned
// (exclusive, in place); block b adds bsum[b] to its elements.
__global__ void __launch_bounds__(kBlock)
k_scan_add_scanned(uint64_t *__restrict__ out, const uint64_t *__restrict__ bsum, const uint32_t *n_dev,
                   uint64_t n_host) {
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    if (base >= n || blockIdx.x == 0) return;
    const uint64_t s = bsum[blockIdx.x];
    for (int k = threadIdx.x; k < kScanTile; k += kBlock)
        if (base + k < n) out[base + k] += s;
}

// ---------------------------------------------------------------------------
// Next level's parts.  Children of part s: left (2i+1, box hi = split, sum wl)
// and right (2i+2, box lo = split, sum S - wl), kept when non-empty
// (:586-588, :603-641).  Their split search starts on the next axis.
// ---------------------------------------------------------------------------
// Forwarding, filled by the move for each grandchild g (a part of the next
// real level) and each side of g's first candidate t3: the first pass of g's
// children, valid when g's split ends at t3.  Per (g, side): the count and
// weight of the points below the children's first candidate t4 (the midpoint
// of g's box along the children's axis, which g's split does not change) and
// the smallest key at or above it.  Per tile of g: the count below t3 (the
// tile's left count) and, per side, the count below t4 (the children's
// per-tile counts).
template <typename Acc>
struct Forward {
    uint32_t *cnt;      // [2 * part + side]
    Acc *w;
    uint32_t *mn;
    uint64_t *left;     // [tile]
    uint32_t *below2;   // [2 * tile + side]
};

// A part whose children's first pass was forwarded by the move that filled
// it: its split is its first candidate, so the forwarded sums below that
// candidate are the children's.
__device__ __forceinline__ bool forwarded(const Part &P) {
    if (P.status != kSplitAtT) return false;
    const int a = (int)P.axis;
    const float t = (sel3(a, P.lo[0], P.lo[1], P.lo[2]) + sel3(a, P.hi[0], P.hi[1], P.hi[2])) / 2.0f;
    return P.split == t;
}

// Split value of a node in the dense heap of the descent: NaN when every
// point went left, since !(c >= NaN) holds for every c.
__device__ inline float dense_split(const Part &P) {
    return P.status == kAllLeft ? __builtin_nanf("") : P.split;
}

__device__ inline uint32_t tiles_for(uint64_t count) { return (uint32_t)((count + kTile - 1) / kTile); }

__global__ void __launch_bounds__(kBlock)
k_children_count(const Part *__restrict__ parts, const uint32_t *n_parts, uint64_t *__restrict__ nchild,
                 uint64_t *__restrict__ ntiles) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_parts) return;
    const Part &P = parts[s];
    const bool inner = P.status != kDone && P.depth + 1 < P.levels;
    const uint64_t nl = inner ? P.cnt_left : 0, nr = inner ? P.count - P.cnt_left : 0;
    nchild[s] = (nl > 0) + (nr > 0);
    ntiles[s] = tiles_for(nl) + tiles_for(nr);
}

// Children counts, their exclusive scans and the next level's table sizes in
// one workgroup (k_children_count + two scans + k_set_sizes), for up to
// kPlanParts parts.
constexpr int kPlanThreads = 1024;
constexpr int kPlanItems = 8;
constexpr uint32_t kPlanParts = kPlanThreads * kPlanItems;
constexpr uint32_t kSelfPlanParts = 64;   // up to this many parts k_children_write scans by itself

__global__ void __launch_bounds__(kPlanThreads)
k_children_plan(const Part *__restrict__ parts, const uint32_t *n_parts, uint64_t *__restrict__ child_off,
                uint64_t *__restrict__ tile_off, uint64_t *parts_total, uint64_t *tiles_total,
                uint32_t *next_n_parts, uint32_t *next_n_tiles) {
    __shared__ uint64_t wc[kPlanThreads / 64], wt[kPlanThreads / 64];
    const uint32_t n = *n_parts;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid >> 6;
    uint64_t c[kPlanItems], t[kPlanItems], sc = 0, st = 0;
#pragma unroll
    for (int k = 0; k < kPlanItems; k++) {
        const uint32_t s = (uint32_t)tid * kPlanItems + k;
        c[k] = 0;
        t[k] = 0;
        if (s < n) {
            const Part &P = parts[s];
            const bool inner = P.status != kDone && P.depth + 1 < P.levels;
            const uint64_t nl = inner ? P.cnt_left : 0, nr = inner ? P.count - P.cnt_left : 0;
            c[k] = (nl > 0) + (nr > 0);
            t[k] = tiles_for(nl) + tiles_for(nr);
        }
        sc += c[k];
        st += t[k];
    }
    const uint64_t ic = wave_incl_add(sc), it = wave_incl_add(st);
    if (lane == 63) { wc[wid] = ic; wt[wid] = it; }
    __syncthreads();
    if (wid == 0) {
        const uint64_t vc = lane < kPlanThreads / 64 ? wc[lane] : 0, vt = lane < kPlanThreads / 64 ? wt[lane] : 0;
        const uint64_t jc = wave_incl_add(vc), jt = wave_incl_add(vt);
        if (lane < kPlanThreads / 64) { wc[lane] = jc - vc; wt[lane] = jt - vt; }
        if (lane == kPlanThreads / 64 - 1) {
            *parts_total = jc;
            *tiles_total = jt;
            *next_n_parts = (uint32_t)jc;
            *next_n_tiles = (uint32_t)jt;
        }
    }
    __syncthreads();
    uint64_t ec = wc[wid] + ic - sc, et = wt[wid] + it - st;
#pragma unroll
    for (int k = 0; k < kPlanItems; k++) {
        const uint32_t s = (uint32_t)tid * kPlanItems + k;
        if (s < n) {
            child_off[s] = ec;
            tile_off[s] = et;
        }
        ec += c[k];
        et += t[k];
    }
}

template <typename Acc>
__global__ void __launch_bounds__(kBlock)
k_children_write(Part *__restrict__ parts, const uint32_t *n_parts, int dim,
                 const uint64_t *__restrict__ child_off, const uint64_t *__restrict__ tile_off,
                 Part *__restrict__ next, Tile *__restrict__ next_tiles, Node *__restrict__ nodes,
                 uint64_t node_base, uint64_t next_node_base, float *__restrict__ heap_split,
                 Instance *__restrict__ insts, Forward<Acc> fsrc, uint32_t *__restrict__ bcnt,
                 Acc *__restrict__ bw, uint32_t *__restrict__ bmin, Forward<Acc> fclr, uint32_t *n_forwarded,
                 uint32_t *next_n_parts, uint32_t *next_n_tiles) {
    // fsrc (non-null): a forwarded part's children get their first-pass bins
    // from it.  fclr (non-null): the new parts' forwarding sums and their
    // tiles' counts are cleared for the move that will fill them.
    // child_off == nullptr (at most 64 parts): every workgroup scans the
    // parts' children and tiles itself, and workgroup 0 sets the next level's
    // table sizes (what k_children_plan does for up to kPlanParts parts).
    const uint32_t s = blockIdx.x;
    __shared__ uint64_t s_off[2];
    if (!child_off && threadIdx.x < 64) {
        const uint32_t n = *n_parts;
        const uint32_t r = threadIdx.x;
        uint64_t c = 0, t = 0;
        if (r < n) {
            const Part &Q = parts[r];
            const bool in = Q.status != kDone && Q.depth + 1 < Q.levels;
            const uint64_t nl = in ? Q.cnt_left : 0, nr = in ? Q.count - Q.cnt_left : 0;
            c = (nl > 0) + (nr > 0);
            t = tiles_for(nl) + tiles_for(nr);
        }
        const uint64_t ic = wave_incl_add(c), it = wave_incl_add(t);
        if (r == s) { s_off[0] = ic - c; s_off[1] = it - t; }
        if (s == 0 && r == 63) {
            *next_n_parts = (uint32_t)ic;
            *next_n_tiles = (uint32_t)it;
        }
    }
    if (!child_off) __syncthreads();
    if (s >= *n_parts) return;
    const Part P = parts[s];
    if (P.status == kDone) return;
    const bool fwd_in = fsrc.cnt != nullptr && forwarded(P);
    if (fwd_in && threadIdx.x == 0 && P.depth + 1 < P.levels) atomicAdd(&n_forwarded[xcd_copy()], 1u);
    // A part at its instance's last level has leaves, not parts (:589-601).
    const bool inner = P.depth + 1 < P.levels;
    const uint64_t nl = inner ? P.cnt_left : 0, nr = inner ? P.count - P.cnt_left : 0;
    const uint32_t a = P.axis;
    const uint32_t na = (a + 1) % dim;
    const Acc S = AccOps<Acc>::from_bits(P.sum);
    const Acc wl = AccOps<Acc>::from_bits(P.wl);
    uint64_t co = child_off ? child_off[s] : s_off[0];
    uint64_t to = child_off ? tile_off[s] : s_off[1];
    int32_t id0 = -1, id1 = -1;
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const uint64_t cnt = side == 0 ? nl : nr;
        if (cnt == 0) continue;
        const uint64_t b = side == 0 ? P.begin : P.begin + nl;
        const uint32_t nt = tiles_for(cnt);
        if (side == 0) id0 = (int32_t)co; else id1 = (int32_t)co;
        if (threadIdx.x == 0) {
            Part C;
            C.begin = b;
            C.count = cnt;
            C.heap_id = 2 * P.heap_id + 1 + side;
            C.sum = side == 0 ? AccOps<Acc>::to_bits(wl) : AccOps<Acc>::to_bits(AccOps<Acc>::sub(S, wl));
#pragma unroll
            for (int k = 0; k < kMaxDim; k++) {
                float lo = P.lo[k], hi = P.hi[k];
                if (k == (int)a) {
                    if (side == 0) hi = P.split; else lo = P.split;
                }
                C.lo[k] = lo;
                C.hi[k] = hi;
            }
            C.axis = na;
            C.blo = na == 0 ? C.lo[0] : (na == 1 ? C.lo[1] : C.lo[2]);
            C.bhi = na == 0 ? C.hi[0] : (na == 1 ? C.hi[1] : C.hi[2]);
            C.prev = ~0ull;
            C.status = kSearching;
            C.split = 0.0f;
            C.kstar = -1;
            C.wl = 0;
            C.cnt_left = 0;
            C.tile_begin = next_tiles ? (uint32_t)to : 0u;   // children searched in place own no tiles
            C.tile_count = next_tiles ? nt : 0u;
            C.child[0] = -1;
            C.child[1] = -1;
            C.inst = P.inst;
            C.depth = P.depth + 1;
            C.levels = P.levels;
            C.counts_ready = 0;
            C.tol = P.tol;
            next[co] = C;
            if (fwd_in) {   // bins of the child's first candidate (2 bins per part)
                const uint64_t f = 2 * (uint64_t)s + side;
                bcnt[2 * co] = fsrc.cnt[f];
                bw[2 * co] = fsrc.w[f];
                bmin[2 * co + 1] = fsrc.mn[f];
            }
            if (fclr.cnt) {
                fclr.cnt[2 * co] = 0;
                fclr.cnt[2 * co + 1] = 0;
                fclr.w[2 * co] = (Acc)0;
                fclr.w[2 * co + 1] = (Acc)0;
                fclr.mn[2 * co] = 0xffffffffu;
                fclr.mn[2 * co + 1] = 0xffffffffu;
            }
        }
        for (uint32_t k = threadIdx.x; next_tiles && k < nt; k += blockDim.x) {
            Tile U;
            U.begin = b + (uint64_t)k * kTile;
            uint64_t rem = cnt - (uint64_t)k * kTile;
            U.count = (uint32_t)(rem < (uint64_t)kTile ? rem : (uint64_t)kTile);
            U.part = (uint32_t)co;
            next_tiles[to + k] = U;
            if (fclr.left) {
                fclr.left[to + k] = 0;
                fclr.below2[2 * (to + k)] = 0;
                fclr.below2[2 * (to + k) + 1] = 0;
            }
        }
        co++;
        to += nt;
    }
    if (threadIdx.x == 0) {
        parts[s].child[0] = id0;
        parts[s].child[1] = id1;
        if (nodes) {                              // this split, for the tree descent
            Node nd;
            nd.split = P.split;
            nd.meta = P.axis | (P.status == kAllLeft ? kNodeAllLeft : 0u);
            nd.child[0] = id0 < 0 ? -1 : (int32_t)(next_node_base + (uint64_t)id0);
            nd.child[1] = id1 < 0 ? -1 : (int32_t)(next_node_base + (uint64_t)id1);
            nodes[node_base + s] = nd;
        }
        if (heap_split) heap_split[insts[P.inst].heap_base + P.heap_id] = dense_split(P);
        if (!inner) {                             // smallest leaf of the instance (:698-699)
            uint64_t v = P.cnt_left > 0 ? 2 * P.heap_id + 1 : 2 * P.heap_id + 2;
            atomicMin((unsigned long long *)&insts[P.inst].min_id, (unsigned long long)v);
        }
    }
}


// Root parts, one per instance (part index = instance index), and their
// tiles.  The box and the sum are set once the prologue has reduced them.
__global__ void __launch_bounds__(kBlock)
k_init_roots(const Instance *__restrict__ insts, uint32_t n_inst, Part *__restrict__ parts,
             Tile *__restrict__ tiles, uint32_t *ctr_words, uint32_t n_words, InstReduce *red) {
    // the call's counters (the prologue's flags among them) and the copies
    // of its reductions start empty (bounding boxes: min key = max, max key = 0)
    const uint32_t gt = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
    for (uint32_t k = gt; k < n_words; k += gridDim.x * gridDim.y * blockDim.x) ctr_words[k] = 0u;
    // block (q, y): instance q, a share y of its tiles
    const uint32_t q = blockIdx.x;
    if (q >= n_inst) return;
    if (blockIdx.y == 0 && threadIdx.x < kCopies) {
        InstReduce &R = red[(uint64_t)threadIdx.x * n_inst + q];
        for (int k = 0; k < kMaxDim; k++) {
            R.bmin[k] = ~0ull;
            R.bmax[k] = 0;
        }
        R.sum_u64 = 0;
        R.sum_f64 = 0.0;
    }
    const Instance I = insts[q];
    const uint32_t nt = tiles_for(I.count);
    if (threadIdx.x == 0 && blockIdx.y == 0) {
        Part R;
        memset(&R, 0, sizeof R);
        R.begin = I.begin;
        R.count = I.count;
        R.heap_id = 0;
        R.prev = ~0ull;
        R.status = (I.count == 0 || I.levels == 0) ? kDone : kSearching;
        R.kstar = -1;
        R.tile_begin = I.tile_begin;
        R.tile_count = nt;
        R.child[0] = -1;
        R.child[1] = -1;
        R.inst = q;
        R.depth = 0;
        R.levels = I.levels;
        R.tol = I.tol;
        parts[q] = R;
    }
    for (uint32_t k = blockIdx.y * blockDim.x + threadIdx.x; k < nt; k += gridDim.y * blockDim.x) {
        Tile U;
        U.begin = I.begin + (uint64_t)k * kTile;
        const uint64_t rem = I.count - (uint64_t)k * kTile;
        U.count = (uint32_t)(rem < (uint64_t)kTile ? rem : (uint64_t)kTile);
        U.part = q;
        tiles[I.tile_begin + k] = U;
    }
}

// Box (`as f32`, :604-605) and sum (:684) of every root part from the
// prologue's per-instance reductions; the id offset starts at "none".
__global__ void __launch_bounds__(kBlock)
k_roots_finalize(Instance *__restrict__ insts, uint32_t n_inst, Part *__restrict__ parts, int dim,
                 int real_mode, uint32_t *fwd_cnt, uint64_t *fwd_w, uint32_t *fwd_mn, uint32_t *n_parts0,
                 uint32_t *n_tiles0, uint32_t tiles0, const InstReduce *__restrict__ red) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q == 0) {   // level 0's table sizes: one root per instance, the level-0 tiles
        *n_parts0 = n_inst;
        *n_tiles0 = tiles0;
    }
    if (q >= n_inst) return;
    if (fwd_cnt) {   // the roots' forwarded sums start empty (any accumulator: zero bits)
        fwd_cnt[2 * q] = fwd_cnt[2 * q + 1] = 0;
        fwd_w[2 * q] = fwd_w[2 * q + 1] = 0;
        fwd_mn[2 * q] = fwd_mn[2 * q + 1] = 0xffffffffu;
    }
    Instance &I = insts[q];
    {   // the prologue's copies of the reductions, combined in copy order
        InstReduce r = red[q];
        for (int c = 1; c < kCopies; c++) {
            const InstReduce &x = red[(uint64_t)c * n_inst + q];
            for (int k = 0; k < kMaxDim; k++) {
                r.bmin[k] = x.bmin[k] < r.bmin[k] ? x.bmin[k] : r.bmin[k];
                r.bmax[k] = x.bmax[k] > r.bmax[k] ? x.bmax[k] : r.bmax[k];
            }
            r.sum_u64 += x.sum_u64;
            r.sum_f64 += x.sum_f64;
        }
        for (int k = 0; k < kMaxDim; k++) {
            I.bmin[k] = r.bmin[k];
            I.bmax[k] = r.bmax[k];
        }
        I.sum_u64 = r.sum_u64;
        I.sum_f64 = r.sum_f64;
    }
    Part &R = parts[q];
    R.sum = real_mode ? __builtin_bit_cast(uint64_t, I.sum_f64) : I.sum_u64;
    for (int k = 0; k < kMaxDim; k++) {
        R.lo[k] = k < dim ? (float)key_f64(I.bmin[k]) : 0.0f;
        R.hi[k] = k < dim ? (float)key_f64(I.bmax[k]) : 0.0f;
    }
    R.axis = 0;
    R.blo = R.lo[0];
    R.bhi = R.hi[0];
    // iter_count 0: every point in part 0 (offset 0)
    I.min_id = I.levels == 0 ? 0ull : ~0ull;
}

// ---------------------------------------------------------------------------
// Children's first split pass, fused into a partition.
// ---------------------------------------------------------------------------
// Fused first pass of the children: one candidate t2 (the midpoint of their
// box on the next axis, the same for both children).  Per child only the
// count and weight of {c < t2} and the smallest c >= t2 are needed by the
// decision at t2, so the bins stay in 8 registers (nb = 2 layout).
template <typename Acc>
struct ChildBins {
    uint32_t cnt0, cnt1, mn0, mn1;
    Acc w0, w1;
    __device__ __forceinline__ void reset() {
        cnt0 = cnt1 = 0;
        mn0 = mn1 = 0xffffffffu;
        w0 = w1 = (Acc)0;
    }
    __device__ __forceinline__ void add(bool right_child, bool below, Acc wv, uint32_t key) {
        const bool l0 = !right_child && below, l1 = right_child && below;
        cnt0 += l0 ? 1u : 0u;
        cnt1 += l1 ? 1u : 0u;
        w0 += l0 ? wv : (Acc)0;
        w1 += l1 ? wv : (Acc)0;
        const uint32_t k0 = (!right_child && !below) ? key : 0xffffffffu;
        const uint32_t k1 = (right_child && !below) ? key : 0xffffffffu;
        mn0 = k0 < mn0 ? k0 : mn0;
        mn1 = k1 < mn1 ? k1 : mn1;
    }
};

template <typename Acc>
struct ChildBinsShared {
    uint32_t cnt[kBlock / 64][2];
    Acc w[kBlock / 64][2];
    uint32_t mn[kBlock / 64][2];
};

template <typename Acc>
__device__ __forceinline__ void flush_child_bins(ChildBins<Acc> &r, ChildBinsShared<Acc> &sh, int32_t ch0,
                                                 int32_t ch1, uint32_t *__restrict__ gcnt, Acc *__restrict__ gw,
                                                 uint32_t *__restrict__ gmn) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const uint32_t c0 = wave_sum(r.cnt0), c1 = wave_sum(r.cnt1);
    const Acc x0 = wave_sum(r.w0), x1 = wave_sum(r.w1);
    const uint32_t m0 = wave_min(r.mn0), m1 = wave_min(r.mn1);
    if (lane == 0) {
        sh.cnt[wid][0] = c0; sh.cnt[wid][1] = c1;
        sh.w[wid][0] = x0; sh.w[wid][1] = x1;
        sh.mn[wid][0] = m0; sh.mn[wid][1] = m1;
    }
    __syncthreads();
    const int j = threadIdx.x;
    if (j < 2) {
        const int32_t ch = j == 0 ? ch0 : ch1;
        uint32_t c = 0, m = 0xffffffffu;
        Acc x = (Acc)0;
#pragma unroll
        for (int v = 0; v < kBlock / 64; v++) {
            c += sh.cnt[v][j];
            x += sh.w[v][j];
            m = sh.mn[v][j] < m ? sh.mn[v][j] : m;
        }
        if (ch >= 0) {
            const uint64_t g = (uint64_t)ch * 2;
            if (c) {
                atomicAdd(&gcnt[g], c);
                atomic_add_acc(&gw[g], x);
            }
            if (m != 0xffffffffu) atomicMin(&gmn[g + 1], m);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Partition of one level.  Every tile's left count is known before the
// partition: k_tile_left gathers it (from the fused counts of the previous
// partition when the part stopped at its first candidate, else by counting),
// a scan turns the counts into offsets, and every tile is then independent.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_tile_left(const Part *__restrict__ parts, const Tile *__restrict__ tiles, const uint32_t *n_tiles,
            const float *__restrict__ c0, const float *__restrict__ c1, const float *__restrict__ c2,
            uint32_t *__restrict__ below, const uint16_t *__restrict__ tile_bins,
            uint64_t *__restrict__ tile_left) {
    // one wavefront per tile
    const uint32_t nt = *n_tiles;
    const int lane = lane_id();
    const uint32_t waves = gridDim.x * (kBlock / 64);
    for (uint32_t t = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); t < nt; t += waves) {
        const Tile T = tiles[t];
        const Part &P = parts[T.part];
        uint64_t left;
        if (P.status == kDone || P.depth + 1 >= P.levels) {
            left = 0;                          // not partitioned
        } else if (P.status == kAllLeft) {
            left = T.count;
        } else if (P.counts_ready == kCountsBelow) {
            left = below[t];                   // counted by the previous partition
        } else if (P.counts_ready == kCountsBins) {
            // bins 0..kstar of this tile in the pass that decided the split
            const uint16_t *tb = tile_bins + (uint64_t)t * kMaxBins;
            uint32_t k = 0;
            for (int j = lane; j <= P.kstar; j += 64) k += tb[j];
            left = wave_sum(k);
        } else {
            const float *c = P.axis == 0 ? c0 : (P.axis == 1 ? c1 : c2);
            const float split = P.split;
            uint32_t k = 0;
            for (uint64_t i = T.begin + lane; i < T.begin + T.count; i += 64) k += !(c[i] >= split) ? 1u : 0u;
            left = wave_sum(k);
        }
        if (lane == 0) {
            tile_left[t] = left;
            below[t] = 0;                      // ready for the next partition
        }
    }
}

template <int D, typename WS, typename Acc>
__global__ void __launch_bounds__(kBlock)
k_partition_tiles(const Part *__restrict__ parts, const Part *__restrict__ next_parts,
                  const Tile *__restrict__ tiles, const uint32_t *n_tiles, const uint64_t *__restrict__ tile_prefix,
                  const float *__restrict__ i0, const float *__restrict__ i1, const float *__restrict__ i2,
                  const WS *__restrict__ iw, float *__restrict__ o0, float *__restrict__ o1,
                  float *__restrict__ o2, WS *__restrict__ ow, uint64_t n_total, uint32_t *__restrict__ bcnt,
                  Acc *__restrict__ bw, uint32_t *__restrict__ bmin, uint32_t *__restrict__ below) {
    __shared__ uint64_t stage[kTile];            // two 4-byte arrays, or one 8-byte array
    __shared__ uint32_t wtot[kItems][kBlock / 64];
    __shared__ uint32_t s_tile_left;
    __shared__ ChildBinsShared<Acc> sh;
    float *sf0 = reinterpret_cast<float *>(stage);
    float *sf1 = sf0 + kTile;
    WS *sw = reinterpret_cast<WS *>(stage);
    WS *sw1 = reinterpret_cast<WS *>(sf1);

    const int tid = threadIdx.x, lane = lane_id(), wid = tid >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t nt = *n_tiles;
    ChildBins<Acc> r;
    r.reset();
    uint32_t cur = 0xffffffffu;
    int32_t ch0 = -1, ch1 = -1;
    int a2 = 0;
    float t2 = 0.f;
    uint64_t cb0 = 0, cb1 = 0;        // first position of each child
    uint32_t ct0 = 0, ct1 = 0;        // first tile of each child

    uint32_t t_first, t_last;
    tile_range(nt, t_first, t_last);
    for (uint32_t t = t_first; t < t_last; t++) {
        const Tile T = tiles[t];
        const Part &P = parts[T.part];
        if (P.status == kDone || P.depth + 1 >= P.levels) continue;   // leaves: not moved
        if (T.part != cur) {
            if (cur != 0xffffffffu) { flush_child_bins(r, sh, ch0, ch1, bcnt, bw, bmin); r.reset(); }
            cur = T.part;
            ch0 = P.child[0];
            ch1 = P.child[1];
            a2 = (int)((P.axis + 1) % D);
            const float blo = sel3(a2, P.lo[0], P.lo[1], P.lo[2]);
            const float bhi = sel3(a2, P.hi[0], P.hi[1], P.hi[2]);
            t2 = (blo + bhi) / 2.0f;                     // the children's first candidate (:472)
            cb0 = P.begin;
            cb1 = P.begin + P.cnt_left;
            ct0 = ch0 >= 0 ? next_parts[ch0].tile_begin : 0;
            ct1 = ch1 >= 0 ? next_parts[ch1].tile_begin : 0;
        }
        const int ax = (int)P.axis;
        const bool all_left = P.status == kAllLeft;
        const float split = P.split;
        const uint64_t end = T.begin + T.count;

        float x0[kItems], x1[kItems], x2[kItems];
        WS wv[kItems];
        bool f[kItems];
        uint32_t pre[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            // clamped, unconditional loads: all of them issue back to back
            // (the values past the tile's end are never used)
            const uint64_t p = T.begin + (uint64_t)k * kBlock + tid;
            const uint64_t c = p < end ? p : end - 1;
            x0[k] = i0[c];
            x1[k] = i1[c];
            if constexpr (D == 3) x2[k] = i2[c];
            else x2[k] = 0.f;
            wv[k] = iw[c];
        }
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint64_t p = T.begin + (uint64_t)k * kBlock + tid;
            f[k] = p < end && (all_left || !(sel3(ax, x0[k], x1[k], x2[k]) >= split));
            const uint64_t bal = __ballot(f[k]);
            pre[k] = __popcll(bal & lt_mask);
            if (lane == 0) wtot[k][wid] = __popcll(bal);
        }
        __syncthreads();
        if (tid < 64) {   // exclusive scan of the kItems x waves counts, item-major
            static_assert(kItems * (kBlock / 64) <= 64, "one wavefront scans the counts");
            uint32_t *flat = &wtot[0][0];
            const uint32_t c = tid < kItems * (kBlock / 64) ? flat[tid] : 0u;
            const uint32_t inc = wave_incl_add(c);
            if (tid < kItems * (kBlock / 64)) flat[tid] = inc - c;
            if (tid == kItems * (kBlock / 64) - 1) s_tile_left = inc;
        }
        __syncthreads();
        const uint32_t tl = s_tile_left;
        const uint64_t prefix = tile_prefix[t] - tile_prefix[P.tile_begin];
        const uint64_t lbase = P.begin + prefix;
        const uint64_t rbase = P.begin + P.cnt_left + ((T.begin - P.begin) - prefix);
        const uint32_t cnt = T.count;
        uint32_t cp[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint32_t off = k * kBlock + tid;
            const uint32_t rl = wtot[k][wid] + pre[k];
            cp[k] = f[k] ? rl : tl + (off - rl);
        }
#define COUPE_STORE_RUN(EXPR)                                                  \
    _Pragma("unroll") for (int k = 0; k < kItems; k++) {                       \
        const uint32_t pos = k * kBlock + tid;                                 \
        if (pos < cnt) {                                                       \
            const uint64_t d = pos < tl ? lbase + pos : rbase + (pos - tl);    \
            if (d < n_total) { EXPR; }                                         \
        }                                                                      \
    }
#pragma unroll
        for (int k = 0; k < kItems; k++)
            if (k * kBlock + tid < (int)cnt) { sf0[cp[k]] = x0[k]; sf1[cp[k]] = x1[k]; }
        __syncthreads();
        COUPE_STORE_RUN(o0[d] = sf0[pos]; o1[d] = sf1[pos])
        __syncthreads();
        if constexpr (sizeof(WS) <= 4) {
#pragma unroll
            for (int k = 0; k < kItems; k++)
                if (k * kBlock + tid < (int)cnt) {
                    if constexpr (D == 3) sf0[cp[k]] = x2[k];
                    sw1[cp[k]] = wv[k];
                }
            __syncthreads();
            if constexpr (D == 3) {
                COUPE_STORE_RUN(o2[d] = sf0[pos]; ow[d] = sw1[pos])
            } else {
                COUPE_STORE_RUN(ow[d] = sw1[pos])
            }
            __syncthreads();
        } else {
            if constexpr (D == 3) {
#pragma unroll
                for (int k = 0; k < kItems; k++)
                    if (k * kBlock + tid < (int)cnt) sf0[cp[k]] = x2[k];
                __syncthreads();
                COUPE_STORE_RUN(o2[d] = sf0[pos])
                __syncthreads();
            }
#pragma unroll
            for (int k = 0; k < kItems; k++)
                if (k * kBlock + tid < (int)cnt) sw[cp[k]] = wv[k];
            __syncthreads();
            COUPE_STORE_RUN(ow[d] = sw[pos])
            __syncthreads();
        }
#undef COUPE_STORE_RUN
        // Children's first pass, and per child tile the points below their
        // first candidate: this tile's left run spans at most two tiles of
        // the left child, its right run two of the right child.
        const uint32_t jl = ct0 + (uint32_t)((lbase - cb0) / kTile);
        const uint32_t jr = ct1 + (uint32_t)((rbase - cb1) / kTile);
        uint32_t nl0 = 0, nl1 = 0, nr0 = 0, nr1 = 0;
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            if (k * kBlock + tid < (int)cnt) {
                const float xb = sel3(a2, x0[k], x1[k], x2[k]);
                const bool bel = !(xb >= t2);
                r.add(!f[k], bel, AccOps<Acc>::of(wv[k]), f32_key(xb));
                const uint64_t d = f[k] ? lbase + cp[k] : rbase + (cp[k] - tl);
                const uint32_t j = f[k] ? ct0 + (uint32_t)((d - cb0) / kTile) : ct1 + (uint32_t)((d - cb1) / kTile);
                const bool first = j == (f[k] ? jl : jr);
                nl0 += (bel && f[k] && first) ? 1u : 0u;
                nl1 += (bel && f[k] && !first) ? 1u : 0u;
                nr0 += (bel && !f[k] && first) ? 1u : 0u;
                nr1 += (bel && !f[k] && !first) ? 1u : 0u;
            }
        }
        nl0 = wave_sum(nl0);
        nl1 = wave_sum(nl1);
        nr0 = wave_sum(nr0);
        nr1 = wave_sum(nr1);
        if (lane == 0) {
            if (nl0) atomicAdd(&below[jl], nl0);
            if (nl1) atomicAdd(&below[jl + 1], nl1);
            if (nr0) atomicAdd(&below[jr], nr0);
            if (nr1) atomicAdd(&below[jr + 1], nr1);
        }
    }
    if (cur != 0xffffffffu) flush_child_bins(r, sh, ch0, ch1, bcnt, bw, bmin);
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

// Split records of the last level and the part id of each of its leaves
// (leaf heap index minus the smallest one, :589-601, :698-702).
__global__ void __launch_bounds__(kBlock)
k_last_nodes(const Part *__restrict__ parts, const uint32_t *n_parts, uint64_t node_base,
             const uint64_t *min_id, Node *__restrict__ nodes, uint64_t *__restrict__ leaf,
             float *__restrict__ heap_split) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_parts) return;
    const Part &P = parts[s];
    const uint64_t off = *min_id;
    Node nd;
    nd.split = P.split;
    nd.meta = P.axis | (P.status == kAllLeft ? kNodeAllLeft : 0u);
    nd.child[0] = -1;
    nd.child[1] = -1;
    nodes[node_base + s] = nd;
    leaf[2 * (uint64_t)s] = 2 * P.heap_id + 1 - off;
    leaf[2 * (uint64_t)s + 1] = 2 * P.heap_id + 2 - off;
    if (heap_split) heap_split[P.heap_id] = dense_split(P);
}

// Part ids in input order: every point walks the split tree from the root
// with its own coordinates (the prologue's SoA copy), taking the side the
// partitions gave it.  The top of the tree is cached in LDS.
constexpr int kNodeCache = 2048;

template <int D>
__global__ void __launch_bounds__(kBlock)
k_descend(const Node *__restrict__ nodes, uint64_t n_nodes, uint32_t levels, uint64_t last_base,
          const uint64_t *__restrict__ leaf, uint64_t n_leaf_parts, const float *__restrict__ c0,
          const float *__restrict__ c1, const float *__restrict__ c2, uint64_t n, uint64_t *__restrict__ out,
          uint32_t *error) {
    __shared__ Node cache[kNodeCache];
    const uint32_t cached = (uint32_t)(n_nodes < (uint64_t)kNodeCache ? n_nodes : (uint64_t)kNodeCache);
    for (uint32_t j = threadIdx.x; j < cached; j += kBlock) cache[j] = nodes[j];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const float x0 = c0[i], x1 = c1[i], x2 = D == 3 ? c2[i] : 0.f;
        uint64_t node = 0;
        uint64_t id = ~0ull;
        for (uint32_t l = 0; l < levels; l++) {
            const Node nd = node < cached ? cache[node] : nodes[node];
            const bool all_left = (nd.meta & kNodeAllLeft) != 0;
            const int side = (all_left || !(sel3((int)(nd.meta & 3u), x0, x1, x2) >= nd.split)) ? 0 : 1;
            if (l + 1 == levels) {
                const uint64_t s = node - last_base;
                if (node >= last_base && s < n_leaf_parts) id = leaf[2 * s + side];
                break;
            }
            const int32_t next = side ? nd.child[1] : nd.child[0];
            if (next < 0 || (uint64_t)next >= n_nodes) break;   // cannot happen for a consistent tree
            node = (uint64_t)next;
        }
        if (id == ~0ull) atomicOr(error, 2u);
        out[i] = id;
    }
}

// Dense variant for iter_count <= kDenseLevels: the split of node h (heap
// index) is heap_split[h], level l splits axis l mod D (:624, :636, :696),
// the children of h are 2h+1 and 2h+2, and the leaf's heap index minus the
// offset is the part id.  The top kDenseCached levels sit in LDS.
constexpr int kDenseLevels = 24;
constexpr int kDenseCached = 12;   // 4095 splits, 16 KB of LDS

// Six wavefronts per SIMD (72 VGPRs): more loads in flight than at five.
// NT: non-temporal loads of the coordinates and stores of the ids (chosen
// by the host for large calls, where the caches cannot hold either).
template <int D, bool NT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
k_descend_dense(const Tile *__restrict__ tiles, uint32_t n_tiles, const Instance *__restrict__ insts,
                const float *__restrict__ heap_split, const float *__restrict__ c0, const float *__restrict__ c1,
                const float *__restrict__ c2, uint64_t *__restrict__ out) {
    __shared__ float cache[(1u << kDenseCached) - 1];
    uint32_t cur = 0xffffffffu, levels = 0, cached_levels = 0;
    uint64_t hb = 0, off = 0;
    uint64_t src = 0, ibeg = 0;   // the instance's ids start at out[src]
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const Tile T = tiles[t];
        if (T.part != cur) {             // level-0 tiles: part index = instance index
            __syncthreads();
            cur = T.part;
            const Instance &I = insts[cur];
            levels = I.levels;
            hb = I.heap_base;
            off = I.min_id;
            src = I.src;
            ibeg = I.begin;
            cached_levels = levels < (uint32_t)kDenseCached ? levels : (uint32_t)kDenseCached;
            const uint32_t cached = (1u << cached_levels) - 1;
            for (uint32_t j = threadIdx.x; j < cached; j += kBlock) cache[j] = heap_split[hb + j];
            __syncthreads();
        }
        // kItems points per thread walk the heap side by side: their loads
        // are issued together and their lookup chains interleave.
        float x0[kItems], x1[kItems], x2[kItems];
        uint32_t h[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint32_t p = k * kBlock + threadIdx.x;
            const uint64_t c = T.begin + (p < T.count ? p : T.count - 1);   // clamped, unconditional loads
            if constexpr (NT) {
                x0[k] = __builtin_nontemporal_load(&c0[c]);
                x1[k] = __builtin_nontemporal_load(&c1[c]);
                x2[k] = D == 3 ? __builtin_nontemporal_load(&c2[c]) : 0.f;
            } else {
                x0[k] = c0[c];
                x1[k] = c1[c];
                x2[k] = D == 3 ? c2[c] : 0.f;
            }
            h[k] = 0;
        }
        // D levels at a time from the root: their axes are 0..D-1, so each
        // step compares a coordinate register directly
        uint32_t l = 0;
        for (; l + D <= cached_levels; l += D) {
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                h[k] = 2 * h[k] + (x0[k] >= cache[h[k]] ? 2u : 1u);
                h[k] = 2 * h[k] + (x1[k] >= cache[h[k]] ? 2u : 1u);
                if constexpr (D == 3) h[k] = 2 * h[k] + (x2[k] >= cache[h[k]] ? 2u : 1u);
            }
        }
        int a = 0;
        for (; l < cached_levels; l++) {
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                const float t = cache[h[k]];
                h[k] = 2 * h[k] + (sel3(a, x0[k], x1[k], x2[k]) >= t ? 2u : 1u);
            }
            a = a + 1 == D ? 0 : a + 1;
        }
        for (; l < levels; l++) {
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                const float t = heap_split[hb + h[k]];
                h[k] = 2 * h[k] + (sel3(a, x0[k], x1[k], x2[k]) >= t ? 2u : 1u);
            }
            a = a + 1 == D ? 0 : a + 1;
        }
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint32_t p = k * kBlock + threadIdx.x;
            if (p < T.count) {
                uint64_t *o = &out[src + (T.begin - ibeg) + p];
                if constexpr (NT) __builtin_nontemporal_store((uint64_t)h[k] - off, o);
                else *o = (uint64_t)h[k] - off;
            }
        }
    }
}

// Fills n 32-bit words (clears that a captured graph can replay, where a
// memset node is avoided).
__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t *out, uint64_t n, uint32_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = v;
}

// iter_count == 0: every point is in part 0.
__global__ void __launch_bounds__(kBlock) k_fill_u64(uint64_t *out, uint64_t n, uint64_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = v;
}

// Copies the 64-bit totals of the scans into the 32-bit table sizes.
__global__ void k_set_sizes(const uint64_t *parts_total, const uint64_t *tiles_total, uint32_t *n_parts,
                            uint32_t *n_tiles) {
    if (threadIdx.x == 0) {
        *n_parts = (uint32_t)*parts_total;
        *n_tiles = (uint32_t)*tiles_total;
    }
}

}  // namespace coupe_hip
