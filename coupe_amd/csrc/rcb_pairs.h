// rcb_pairs.h -- two levels per data move.
//
// The points of a part are moved once for two bisection levels: after level
// l decides a part p, its two children are left in place ("virtual": their
// points stay interleaved in p's range) and searched there; after level l+1
// decides them, every point moves straight to its grandchild (4 runs per
// tile).  Per pair of levels this reads 12 B/point for the children's first
// pass and moves the payload once, instead of moving it twice.
//
//   k_child_first   children's first pass over the parent's tiles: side of
//                   the parent's split, bins at the children's first
//                   candidate, per-tile left counts of the parent and per
//                   (tile, side) counts below that candidate
//   k_hist_virtual  later passes of the children over the parent's tiles,
//                   with per-(tile, side) bin counts
//   k_tile_count4   per tile, the points of each grandchild
//   k_partition4    the move parent -> grandchildren, LDS-staged runs, and
//                   the grandchildren's first pass
#pragma once

#include "rcb_kernels.h"

namespace coupe_hip {

// A part of level l whose children are searched in place and whose points
// move to grandchildren: neither done nor at its instance's last two levels.
__device__ __forceinline__ bool pair_active(const Part &P) {
    return P.status != kDone && P.depth + 2 < P.levels;
}
// A part of level l whose children are searched (in place) at all.
__device__ __forceinline__ bool children_searched(const Part &P) {
    return P.status != kDone && P.depth + 1 < P.levels;
}

// Block-wide sum of one value per thread (all threads get it).
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *sh4) {
    v = wave_sum(v);
    if (lane_id() == 0) sh4[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) s += sh4[i];
    __syncthreads();
    return s;
}

// Six wavefronts per SIMD (80 VGPRs, a few spilled): faster than five.
template <int D, typename WS, typename Acc>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
k_child_first(const Part *__restrict__ parts, const Tile *__restrict__ tiles, const uint32_t *n_tiles,
              const float *__restrict__ c0, const float *__restrict__ c1, const float *__restrict__ c2,
              const WS *__restrict__ w, uint32_t *__restrict__ bcnt, Acc *__restrict__ bw,
              uint32_t *__restrict__ bmin, uint32_t *__restrict__ below2, uint64_t *__restrict__ tile_left,
              uint32_t fwd_ok, unsigned long long *points_read) {
    __shared__ ChildBinsShared<Acc> sh;
    __shared__ uint32_t s4[3 * (kBlock / 64)];
    const int tid = threadIdx.x;
    const uint32_t nt = *n_tiles;
    ChildBins<Acc> r;
    r.reset();
    uint32_t cur = 0xffffffffu;
    int32_t ch0 = -1, ch1 = -1;
    const float *ca = c0, *cb = c1;
    float t2 = 0.f, split = 0.f;
    bool all_left = false;
    uint64_t read = 0;   // points of the tiles this workgroup reads
    uint32_t t_first, t_last;
    tile_range(nt, t_first, t_last);
    for (uint32_t t = t_first; t < t_last; t++) {
        const Tile T = tiles[t];
        const Part &P = parts[T.part];
        if (!children_searched(P) || (fwd_ok && forwarded(P))) {   // forwarded: nothing to read
            const uint32_t last = P.tile_begin + P.tile_count - 1;   // the part's tiles are contiguous
            if (P.tile_count > 0 && last > t) t = last;
            continue;
        }
        read += T.count;
        if (T.part != cur) {
            if (cur != 0xffffffffu) { flush_child_bins(r, sh, ch0, ch1, bcnt, bw, bmin); r.reset(); }
            cur = T.part;
            ch0 = P.child[0];
            ch1 = P.child[1];
            const int ax = (int)P.axis, a2 = (int)((P.axis + 1) % D);
            ca = ax == 0 ? c0 : (ax == 1 ? c1 : c2);
            cb = a2 == 0 ? c0 : (a2 == 1 ? c1 : c2);
            t2 = (sel3(a2, P.lo[0], P.lo[1], P.lo[2]) + sel3(a2, P.hi[0], P.hi[1], P.hi[2])) / 2.0f;
            split = P.split;
            all_left = P.status == kAllLeft;
        }
        // all loads of the tile first, then the points
        float xa[kItems], xb[kItems];
        WS wk[kItems];
        const uint64_t end = T.begin + T.count;
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint64_t i = T.begin + (uint64_t)k * kBlock + tid;
            const uint64_t c = i < end ? i : end - 1;   // clamped: loads issue back to back
            xa[k] = ca[c];
            xb[k] = cb[c];
            wk[k] = w[c];
        }
        uint32_t nleft = 0, b0 = 0, b1 = 0;   // wave-uniform counts from ballots
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const bool valid = T.begin + (uint64_t)k * kBlock + tid < end;
            const bool right = !all_left && (xa[k] >= split);
            const bool bel = !(xb[k] >= t2);
            if (valid) r.add(right, bel, AccOps<Acc>::of(wk[k]), f32_key(xb[k]));
            const uint64_t bv = __ballot(valid), br = __ballot(right), bb = __ballot(bel);
            nleft += __popcll(bv & ~br);
            b0 += __popcll(bv & ~br & bb);
            b1 += __popcll(bv & br & bb);
        }
        if (lane_id() == 0) {
            s4[(tid >> 6) * 3] = nleft;
            s4[(tid >> 6) * 3 + 1] = b0;
            s4[(tid >> 6) * 3 + 2] = b1;
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t a = 0, c = 0, e = 0;
#pragma unroll
            for (int v = 0; v < kBlock / 64; v++) { a += s4[3 * v]; c += s4[3 * v + 1]; e += s4[3 * v + 2]; }
            tile_left[t] = a;
            below2[2 * (uint64_t)t] = c;
            below2[2 * (uint64_t)t + 1] = e;
        }
        __syncthreads();
    }
    if (tid == 0 && read) atomicAdd(&points_read[xcd_copy()], (unsigned long long)read);
    if (cur != 0xffffffffu) flush_child_bins(r, sh, ch0, ch1, bcnt, bw, bmin);
}

// The roots' first pass (level 0) with their children's first pass
// forwarded: the root's 3 candidates along its axis in registers (as
// k_sweep), and, for a split at its first candidate e1 (the box midpoint),
// per side of e1 the count and weight below the children's first candidate
// (the midpoint of the box on the next axis) and the smallest key at or
// above it, and per tile the count below e1 and per side the count below the
// children's candidate.  Reads the two axes and the weight once.
// Six wavefronts per SIMD (80 VGPRs, a few spilled): faster than five.
template <int D, typename WS, typename Acc>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
k_root_first(const Part *__restrict__ parts, const Tile *__restrict__ tiles, const uint32_t *n_tiles,
             const float *__restrict__ c0, const float *__restrict__ c1, const float *__restrict__ c2,
             const WS *__restrict__ w, uint32_t *__restrict__ bcnt, Acc *__restrict__ bw,
             uint32_t *__restrict__ bmin, Forward<Acc> fw) {
    __shared__ RegBinsShared<kFirstBins, Acc> sh;
    __shared__ ChildBinsShared<Acc> shc;
    __shared__ uint32_t s4[3 * (kBlock / 64)];
    const int tid = threadIdx.x, lane = lane_id();
    const uint32_t nt = *n_tiles;
    RegBins<kFirstBins, Acc> r;
    r.reset();
    ChildBins<Acc> f;
    f.reset();
    uint32_t cur = 0xffffffffu;
    bool active = false;
    float e0 = 0.f, e1 = 0.f, e2 = 0.f, t2 = 0.f;
    const float *ca = c0, *cb = c1;
    // forwarded sums of a root: fw.* [2 root + side]
    auto flush_fwd = [&](uint32_t p) {
        const uint32_t x0 = wave_sum(f.cnt0), x1 = wave_sum(f.cnt1);
        const Acc y0 = wave_sum(f.w0), y1 = wave_sum(f.w1);
        const uint32_t m0 = wave_min(f.mn0), m1 = wave_min(f.mn1);
        const int wid = tid >> 6;
        if (lane == 0) {
            shc.cnt[wid][0] = x0; shc.cnt[wid][1] = x1;
            shc.w[wid][0] = y0; shc.w[wid][1] = y1;
            shc.mn[wid][0] = m0; shc.mn[wid][1] = m1;
        }
        __syncthreads();
        if (tid < 2) {
            uint32_t c = 0, m = 0xffffffffu;
            Acc x = (Acc)0;
#pragma unroll
            for (int v = 0; v < kBlock / 64; v++) {
                c += shc.cnt[v][tid];
                x += shc.w[v][tid];
                m = shc.mn[v][tid] < m ? shc.mn[v][tid] : m;
            }
            const uint64_t b = 2 * (uint64_t)p + tid;
            if (c) { atomicAdd(&fw.cnt[b], c); atomic_add_acc(&fw.w[b], x); }
            if (m != 0xffffffffu) atomicMin(&fw.mn[b], m);
        }
        __syncthreads();
    };
    uint32_t t_first, t_last;
    tile_range(nt, t_first, t_last);
    for (uint32_t t = t_first; t < t_last; t++) {
        const Tile T = tiles[t];
        if (T.part != cur) {
            if (active) {
                flush_bins<kFirstBins, kFirstBins>(r, sh, cur, -1, bcnt, bw, bmin);
                r.reset();
                flush_fwd(cur);
                f.reset();
            }
            cur = T.part;
            const Part &P = parts[cur];
            active = P.status == kSearching;
            e0 = candidate(P.blo, P.bhi, 0, kFirstBins);
            e1 = candidate(P.blo, P.bhi, 1, kFirstBins);
            e2 = candidate(P.blo, P.bhi, 2, kFirstBins);
            const int ax = (int)P.axis, a2 = (int)((P.axis + 1) % D);
            ca = ax == 0 ? c0 : (ax == 1 ? c1 : c2);
            cb = a2 == 0 ? c0 : (a2 == 1 ? c1 : c2);
            t2 = (sel3(a2, P.lo[0], P.lo[1], P.lo[2]) + sel3(a2, P.hi[0], P.hi[1], P.hi[2])) / 2.0f;
        }
        if (!active) continue;
        float xa[kItems], xb[kItems];
        WS wk[kItems];
        const uint32_t cnt = T.count;
        {
            const float *ba = ca + T.begin, *bb = cb + T.begin;
            const WS *bwt = w + T.begin;
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                const uint32_t p = (uint32_t)(k * kBlock + tid);
                const uint32_t c = p < cnt ? p : cnt - 1;   // clamped: loads issue back to back
                xa[k] = ba[c];
                xb[k] = bb[c];
                wk[k] = bwt[c];
            }
        }
        uint32_t nleft = 0, b0 = 0, b1 = 0;   // wave-uniform counts from ballots
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const bool valid = (uint32_t)(k * kBlock + tid) < cnt;
            const Acc wa = AccOps<Acc>::of(wk[k]);
            const bool right = xa[k] >= e1;
            const bool bel = !(xb[k] >= t2);
            if (valid) {
                r.add(bin4(e0, e1, e2, xa[k]), wa, f32_key(xa[k]));
                f.add(right, bel, wa, f32_key(xb[k]));
            }
            const uint64_t bv = __ballot(valid), br = __ballot(right), bb = __ballot(bel);
            nleft += __popcll(bv & ~br);
            b0 += __popcll(bv & ~br & bb);
            b1 += __popcll(bv & br & bb);
        }
        const int wid = tid >> 6;
        if (lane == 0) {
            s4[wid * 3] = nleft;
            s4[wid * 3 + 1] = b0;
            s4[wid * 3 + 2] = b1;
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t a = 0, c = 0, e = 0;
#pragma unroll
            for (int v = 0; v < kBlock / 64; v++) { a += s4[3 * v]; c += s4[3 * v + 1]; e += s4[3 * v + 2]; }
            fw.left[t] = a;
            fw.below2[2 * (uint64_t)t] = c;
            fw.below2[2 * (uint64_t)t + 1] = e;
        }
        __syncthreads();
    }
    if (active) {
        flush_bins<kFirstBins, kFirstBins>(r, sh, cur, -1, bcnt, bw, bmin);
        flush_fwd(cur);
    }
}

// Later passes of in-place children: LDS bins of both children of the tile's
// parent, per-(tile, side) bin counts for the move's offsets.
template <int D, typename WS, typename Acc>
__global__ void __launch_bounds__(kBlock)
k_hist_virtual(const Part *__restrict__ parts, const Tile *__restrict__ tiles, const uint32_t *n_tiles,
               const Part *__restrict__ children, const float *__restrict__ c0, const float *__restrict__ c1,
               const float *__restrict__ c2, const WS *__restrict__ w, int nb, uint32_t *__restrict__ bcnt,
               Acc *__restrict__ bw, uint32_t *__restrict__ bmin, const uint32_t *gate,
               uint16_t *__restrict__ tile_bins2, BinCopies<Acc> xc, uint32_t chunk) {
    if (gate && gate_count(gate) == 0) return;
    __shared__ float cand[2][kMaxBins];
    __shared__ uint32_t scnt[2][kMaxBins];
    __shared__ Acc sw[2][kMaxBins];
    __shared__ uint32_t smin[2][kMaxBins];
    const int tid = threadIdx.x;
    const uint32_t nt = *n_tiles;
    uint32_t cur = 0xffffffffu;
    bool act[2] = {false, false};
    int32_t ch[2] = {-1, -1};
    const float *ca = c0, *cb = c1;
    float split = 0.f;
    bool all_left = false;
    uint32_t before0 = 0, before1 = 0;
    float clo[2] = {0.f, 0.f}, cscale[2] = {0.f, 0.f};
    bool packed = false;   // both children's bins (the parent's points bound them)
    auto flush = [&](int s) {
        const BinsDest<Acc> d = bins_dest(xc, (uint32_t)ch[s], nb, bcnt, bw, bmin);
        flush_lds_bins(d.base, nb, scnt[s], sw[s], smin[s], d.cnt, d.w, d.mn, packed);
    };
    // chunks of `chunk` consecutive tiles go to the workgroups in turn
    for (uint32_t t0 = blockIdx.x * chunk; t0 < nt; t0 += gridDim.x * chunk)
    for (uint32_t t = t0; t < min(nt, t0 + chunk); t++) {
        const Tile T = tiles[t];
        const Part &P = parts[T.part];
        if (!children_searched(P)) continue;
        if (T.part != cur) {
            before0 = before1 = 0;
            __syncthreads();
            for (int s = 0; s < 2; s++)
                if (act[s]) flush(s);
            __syncthreads();
            cur = T.part;
            packed = packable<WS, Acc>() && pack_fits(P.count, gridDim.x);
            {
                const int ax = (int)P.axis, a2 = (int)((P.axis + 1) % D);
                ca = ax == 0 ? c0 : (ax == 1 ? c1 : c2);
                cb = a2 == 0 ? c0 : (a2 == 1 ? c1 : c2);
            }
            split = P.split;
            all_left = P.status == kAllLeft;
            for (int s = 0; s < 2; s++) {
                ch[s] = P.child[s];
                act[s] = ch[s] >= 0 && children[ch[s]].status == kSearching;
                if (act[s]) {
                    const Part &C = children[ch[s]];
                    for (int p = tid; p < nb - 1; p += kBlock) cand[s][p] = candidate(C.blo, C.bhi, p, nb);
                    for (int j = tid; j < nb; j += kBlock) { scnt[s][j] = 0; sw[s][j] = (Acc)0; smin[s][j] = 0xffffffffu; }
                    clo[s] = C.blo;
                    cscale[s] = (float)nb / (C.bhi - C.blo);
                }
            }
            __syncthreads();
        }
        if (!act[0] && !act[1]) continue;
        // every load of the tile issued before the binning (clamped, unconditional)
        const uint32_t cnt = T.count;
        float xa[kItems], xb[kItems];
        WS wk[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint32_t p = min((uint32_t)(k * kBlock + tid), cnt - 1u);
            xa[k] = ca[T.begin + p];
            xb[k] = cb[T.begin + p];
            wk[k] = w[T.begin + p];
        }
        EdgeBins<WS, Acc> edge0, edge1;
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            if ((uint32_t)(k * kBlock + tid) >= cnt) continue;
            const int s = (!all_left && (xa[k] >= split)) ? 1 : 0;
            if (!act[s]) continue;
            const float x = xb[k];
            const int b = bin_near(cand[s], nb, x, clo[s], cscale[s]);
            const bool taken = s ? edge1.take(b, nb, wk[k], x) : edge0.take(b, nb, wk[k], x);
            if (!taken) bin_add<WS, Acc>(packed, scnt[s], sw[s], smin[s], b, wk[k], x);
        }
        edge0.flush(packed, nb, scnt[0], sw[0], smin[0]);
        edge1.flush(packed, nb, scnt[1], sw[1], smin[1]);
        __syncthreads();
        if (tid < nb) {
            const uint32_t a0 = bin_count(packed, scnt[0], sw[0], tid), a1 = bin_count(packed, scnt[1], sw[1], tid);
            uint16_t *tb = tile_bins2 + (uint64_t)t * 2 * kMaxBins;
            if (act[0]) tb[tid] = (uint16_t)(a0 - before0);
            if (act[1]) tb[kMaxBins + tid] = (uint16_t)(a1 - before1);
            before0 = a0;
            before1 = a1;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int s = 0; s < 2; s++)
        if (act[s]) flush(s);
}

// Per tile of a parent, the points of each grandchild (q = 2 side + side of
// the child's split), from the counts the passes kept: the count below the
// child's first candidate, or the sum of the tile's bins up to the stopping
// candidate of a later pass (kstar + 1 u16 counts, read 8 at a time).  One
// thread per tile; a side whose count needs the points is done by the whole
// wavefront, one such tile after the other.
__device__ __forceinline__ uint32_t sum_tile_bins(const uint16_t *__restrict__ tb, int kstar) {
    const uint4 *v = reinterpret_cast<const uint4 *>(tb);   // kMaxBins u16 per (tile, side): 16-byte aligned
    const int nv = (kstar + 1 + 7) / 8;
    uint32_t k = 0;
    for (int i = 0; i < nv; i++) {
        const uint4 x = v[i];
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int h = 0; h < 8; h++) {
            const int j = i * 8 + h;
            const uint32_t c = (w[h >> 1] >> ((h & 1) * 16)) & 0xffffu;
            k += j <= kstar ? c : 0u;
        }
    }
    return k;
}

// The count of side s of a tile below the child's split, from the points
// (the whole wavefront).
__device__ uint32_t tile_count_points(const Part &P, const Part &C, int s, uint64_t tbegin, uint32_t tcount,
                                      const float *__restrict__ c0, const float *__restrict__ c1,
                                      const float *__restrict__ c2) {
    const int lane = lane_id();
    const int ax = (int)P.axis, a2 = (int)C.axis;
    const float *ca = ax == 0 ? c0 : (ax == 1 ? c1 : c2);
    const float *cb = a2 == 0 ? c0 : (a2 == 1 ? c1 : c2);
    const bool all_left = P.status == kAllLeft;
    uint32_t k = 0;
    for (uint64_t i = tbegin + lane; i < tbegin + tcount; i += 64) {
        const int si = (!all_left && (ca[i] >= P.split)) ? 1 : 0;
        k += (si == s && !(cb[i] >= C.split)) ? 1u : 0u;
    }
    return wave_sum(k);
}

template <int D>
__global__ void __launch_bounds__(kBlock)
k_tile_count4(const Part *__restrict__ parts, const Tile *__restrict__ tiles, const uint32_t *n_tiles,
              const Part *__restrict__ children, const float *__restrict__ c0, const float *__restrict__ c1,
              const float *__restrict__ c2, const uint64_t *__restrict__ tile_left,
              const uint32_t *__restrict__ below2, const uint16_t *__restrict__ tile_bins2, uint64_t *__restrict__ q0,
              uint64_t *__restrict__ q1, uint64_t *__restrict__ q2, uint64_t *__restrict__ q3,
              unsigned long long *__restrict__ runs) {
    // runs: per XCD, the destinations (non-empty slots) of the moved tiles,
    // then the moved tiles (the host's measure of how scattered a move is)
    const uint32_t nt = *n_tiles;
    const int lane = lane_id();
    uint32_t n_dests = 0, n_moved = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < nt; base += gridDim.x * kBlock) {
        const uint32_t t = base + threadIdx.x;
        uint64_t out[4] = {0, 0, 0, 0};
        uint64_t cnt[2] = {0, 0};
        bool slow0 = false, slow1 = false;
        Tile T{};
        if (t < nt) {
            T = tiles[t];
            const Part &P = parts[T.part];
            if (pair_active(P)) {
                cnt[0] = tile_left[t];
                cnt[1] = T.count - cnt[0];
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if (cnt[s] == 0 || P.child[s] < 0) continue;
                    const Part &C = children[P.child[s]];
                    uint64_t left;
                    if (C.status == kAllLeft) {
                        left = cnt[s];
                    } else if (C.counts_ready == kCountsBelow) {
                        left = below2[2 * (uint64_t)t + s];
                    } else if (C.counts_ready == kCountsBins) {
                        left = sum_tile_bins(tile_bins2 + ((uint64_t)t * 2 + s) * kMaxBins, C.kstar);
                    } else {
                        if (s == 0) slow0 = true; else slow1 = true;
                        continue;
                    }
                    out[2 * s] = left;
                    out[2 * s + 1] = cnt[s] - left;
                }
            }
        }
        uint64_t pending = __ballot(slow0 || slow1);
        while (pending) {   // wave-uniform loop over the lanes with a slow side
            const int src = __ffsll((unsigned long long)pending) - 1;
            pending &= pending - 1;
            const uint32_t tp = __shfl(T.part, src, 64), tc = __shfl(T.count, src, 64);
            const uint64_t tb = ((uint64_t)__shfl((uint32_t)(T.begin >> 32), src, 64) << 32) |
                                __shfl((uint32_t)T.begin, src, 64);
            const uint64_t c0s = ((uint64_t)__shfl((uint32_t)(cnt[0] >> 32), src, 64) << 32) |
                                 __shfl((uint32_t)cnt[0], src, 64);
            const bool sl0 = __shfl(slow0 ? 1 : 0, src, 64) != 0, sl1 = __shfl(slow1 ? 1 : 0, src, 64) != 0;
            const Part &P = parts[tp];
            if (sl0) {
                const uint32_t left = tile_count_points(P, children[P.child[0]], 0, tb, tc, c0, c1, c2);
                if (lane == src) { out[0] = left; out[1] = c0s - left; }
            }
            if (sl1) {
                const uint32_t left = tile_count_points(P, children[P.child[1]], 1, tb, tc, c0, c1, c2);
                if (lane == src) { out[2] = left; out[3] = (tc - c0s) - left; }
            }
        }
        if (t < nt) {
            q0[t] = out[0];
            q1[t] = out[1];
            q2[t] = out[2];
            q3[t] = out[3];
        }
        const uint32_t dests = (out[0] != 0) + (out[1] != 0) + (out[2] != 0) + (out[3] != 0);
        n_dests += dests;
        n_moved += dests ? 1u : 0u;
    }
    // one update of the counters per workgroup
    __shared__ uint32_t s_runs[2][kBlock / 64];
    const uint32_t rsum = wave_sum(n_dests), tsum = wave_sum(n_moved);
    if (lane == 0) {
        s_runs[0][threadIdx.x >> 6] = rsum;
        s_runs[1][threadIdx.x >> 6] = tsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0, m = 0;
        for (int k = 0; k < kBlock / 64; k++) {
            r += s_runs[0][k];
            m += s_runs[1][k];
        }
        if (m) {
            atomicAdd(&runs[xcd_copy()], (unsigned long long)r);
            atomicAdd(&runs[kCopies + xcd_copy()], (unsigned long long)m);
        }
    }
}

// The move of a pair of levels: every point of an active parent goes to its
// grandchild; stable within each grandchild.  p0..p3 are the exclusive scans
// of the per-tile counts of grandchild slots q = 0..3.
//
// Per tile: the points are classified (three ballots give each point's slot
// and rank), one wavefront scans the slot counts, every array is staged in
// LDS in slot order (run j: slot j's points, one contiguous range of its
// grandchild), and each wavefront stores a quarter of the staged tile, one
// run segment after the other.  That phase also sums, across the tiles of
// one parent, each grandchild's first pass (count, weight and smallest key at
// its first candidate t3) and, with FWD, the forwarding of its children's
// first pass (per side of t3: count, weight and smallest key at their first
// candidate t4; per destination tile: the count below t3 and per side the
// count below t4).
// skip_axis (0..2, or 3 for none): a coordinate array that no later level
// reads (3D, the last move: the children's axis is not split again), read for
// the classification but not written.  NT: the stores are non-temporal
// (measured faster where the arrays written are much larger than the
// last-level cache, slower where they fit in it).
template <int D, typename WS, typename Acc, bool FWD, bool NT>
__global__ void __launch_bounds__(kBlock)
k_partition4(const Part *__restrict__ parts, const Part *__restrict__ children, const Part *__restrict__ grand,
             const Tile *__restrict__ tiles, const uint32_t *n_tiles, const uint64_t *__restrict__ p0,
             const uint64_t *__restrict__ p1, const uint64_t *__restrict__ p2, const uint64_t *__restrict__ p3,
             const float *__restrict__ i0, const float *__restrict__ i1, const float *__restrict__ i2,
             const WS *__restrict__ iw, float *__restrict__ o0, float *__restrict__ o1, float *__restrict__ o2,
             WS *__restrict__ ow, uint64_t n_total, uint32_t *__restrict__ bcnt, Acc *__restrict__ bw,
             uint32_t *__restrict__ bmin, Forward<Acc> fw, uint32_t skip_axis) {
    static_assert(kBlock / 64 == 4, "one wavefront per grandchild slot");
    __shared__ float sfa[kTile], sfb[kTile], sfc[D == 3 ? kTile : 1];
    __shared__ WS sws[kTile];
    __shared__ uint32_t wtot[4][kItems * (kBlock / 64)];
    __shared__ uint32_t s_run[5];
    __shared__ uint64_t s_dst[4];
    __shared__ uint32_t s_off[4], s_tile[4];
    __shared__ uint64_t s_gbeg[4], s_gpj0[4];
    __shared__ uint32_t s_gtb[4];
    const int tid = threadIdx.x, lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: per-wave state in SGPRs
    const uint32_t nt = *n_tiles;
    uint32_t cur = 0xffffffffu;
    // the thresholds of the parent and of its two children; NaN where every
    // point goes left (kAllLeft, or no child), as !(x >= NaN) holds for all x
    const float kNone = __builtin_nanf("");
    float gsplit[2] = {kNone, kNone};
    int ax = 0, a2 = 0, a3 = 0, a4 = 0;
    float split = kNone;
    // lanes 0..3 serve grandchild slot `lane`: its scan array
    const int js = lane & 3;
    const uint64_t *pj = js == 0 ? p0 : (js == 1 ? p1 : (js == 2 ? p2 : p3));
    // the current parent's grandchild slots (LDS, threads 0..3 write slot
    // tid at a part change): the grandchild, its first candidate t3 and its
    // children's t4, and the sums of both first passes over the parent's
    // tiles, to which every wavefront adds those of the positions it stores;
    // threads 0..3 flush them after a barrier at the next part change
    __shared__ int32_t s_g[4];
    __shared__ float s_t3[4], s_t4[4];
    __shared__ uint32_t s_gcnt[4], s_gmn[4], s_fcn[2][4], s_fmn[2][4];
    __shared__ Acc s_gws[4], s_fws[2][4];
    auto reset_slot = [&](int j) {
        s_gcnt[j] = 0;
        s_gmn[j] = 0xffffffffu;
        s_gws[j] = (Acc)0;
        for (int h = 0; h < 2; h++) {
            s_fcn[h][j] = 0;
            s_fmn[h][j] = 0xffffffffu;
            s_fws[h][j] = (Acc)0;
        }
    };
    auto flush = [&]() {   // threads 0..3 (slot tid), after a workgroup barrier
        if (tid >= 4) return;
        const int32_t g = s_g[tid];
        if (g >= 0) {
            const uint64_t b = 2 * (uint64_t)g;
            if (s_gcnt[tid]) { atomicAdd(&bcnt[b], s_gcnt[tid]); atomic_add_acc(&bw[b], s_gws[tid]); }
            if (s_gmn[tid] != 0xffffffffu) atomicMin(&bmin[b + 1], s_gmn[tid]);
            if constexpr (FWD) {
                for (int h = 0; h < 2; h++) {
                    if (s_fcn[h][tid]) { atomicAdd(&fw.cnt[b + h], s_fcn[h][tid]); atomic_add_acc(&fw.w[b + h], s_fws[h][tid]); }
                    if (s_fmn[h][tid] != 0xffffffffu) atomicMin(&fw.mn[b + h], s_fmn[h][tid]);
                }
            }
        }
        reset_slot(tid);
    };
    if (tid < 4) {   // visible after the first tile's barriers
        s_g[tid] = -1;
        reset_slot(tid);
    }
    // this wavefront's sums for the slot of its last segment (registers):
    // added to the slot's LDS sums when its next segment is of another slot,
    // and before a part change
    int cur_j = -1;   // wave-uniform
    uint32_t gcnt = 0, fcn0 = 0, fcn1 = 0;
    Acc gws = (Acc)0, fws0 = (Acc)0, fws1 = (Acc)0;
    uint32_t gmn = 0xffffffffu, fmn0 = 0xffffffffu, fmn1 = 0xffffffffu;
    auto wave_flush = [&]() {
        if (cur_j < 0) return;
        const int j = cur_j;
        const Acc gx = wave_sum(gws);
        const uint32_t gm = wave_min(gmn);
        if (lane == 0) {
            if (gcnt) { atomicAdd(&s_gcnt[j], gcnt); atomic_add_acc(&s_gws[j], gx); }
            if (gm != 0xffffffffu) atomicMin(&s_gmn[j], gm);
        }
        if constexpr (FWD) {
            const Acc x0 = wave_sum(fws0), x1 = wave_sum(fws1);
            const uint32_t m0 = wave_min(fmn0), m1 = wave_min(fmn1);
            if (lane == 0) {
                if (fcn0) { atomicAdd(&s_fcn[0][j], fcn0); atomic_add_acc(&s_fws[0][j], x0); }
                if (fcn1) { atomicAdd(&s_fcn[1][j], fcn1); atomic_add_acc(&s_fws[1][j], x1); }
                if (m0 != 0xffffffffu) atomicMin(&s_fmn[0][j], m0);
                if (m1 != 0xffffffffu) atomicMin(&s_fmn[1][j], m1);
            }
        }
        gcnt = fcn0 = fcn1 = 0;
        gws = fws0 = fws1 = (Acc)0;
        gmn = fmn0 = fmn1 = 0xffffffffu;
        cur_j = -1;
    };

    uint32_t t_first, t_last;
    tile_range(nt, t_first, t_last);
    for (uint32_t t = t_first; t < t_last; t++) {
        const Tile T = tiles[t];
        const Part &P = parts[T.part];
        if (!pair_active(P)) {   // nothing of this part moves: skip its tiles
            const uint32_t last = P.tile_begin + P.tile_count - 1;
            if (P.tile_count > 0 && last > t) t = last;
            continue;
        }
        if (T.part != cur) {
            wave_flush();
            __syncthreads();   // every wavefront's stores and sums of the previous part are done
            if (cur != 0xffffffffu) flush();
            cur = T.part;
            ax = (int)P.axis;
            a2 = (int)((P.axis + 1) % D);
            a3 = (int)((P.axis + 2) % D);
            a4 = (int)((P.axis + 3) % D);
            split = P.status == kAllLeft ? kNone : P.split;
            int32_t g[4];
            for (int s = 0; s < 2; s++) {
                const int32_t ci = P.child[s];
                gsplit[s] = kNone;
                g[2 * s] = g[2 * s + 1] = -1;
                if (ci >= 0) {
                    const Part &C = children[ci];
                    gsplit[s] = C.status == kAllLeft ? kNone : C.split;
                    g[2 * s] = C.child[0];
                    g[2 * s + 1] = C.child[1];
                }
            }
            if (tid < 4) {
                const int32_t gj = tid == 0 ? g[0] : (tid == 1 ? g[1] : (tid == 2 ? g[2] : g[3]));
                s_gbeg[tid] = gj >= 0 ? grand[gj].begin : 0;
                s_gtb[tid] = gj >= 0 ? grand[gj].tile_begin : 0;
                s_gpj0[tid] = gj >= 0 ? pj[P.tile_begin] : 0;
                s_g[tid] = gj;
                if (gj >= 0) {
                    const Part &G = grand[gj];
                    s_t3[tid] = (G.blo + G.bhi) / 2.0f;   // the grandchild's first candidate (:472)
                    s_t4[tid] = (sel3(a4, G.lo[0], G.lo[1], G.lo[2]) + sel3(a4, G.hi[0], G.hi[1], G.hi[2])) / 2.0f;
                }
            }
        }
        // this tile's scan values: loaded by every lane without a branch, so
        // the load overlaps the points' and is consumed after the classification
        const uint64_t pv = pj[t];
        const uint32_t cnt = T.count;
        // the coordinates by role: rx on the parent's axis, ry on the
        // children's, rz on the grandchildren's (3D; in 2D it is rx's axis)
        float rx[kItems], ry[kItems], rz[kItems];
        WS wv[kItems];
        {
            // clamped, unconditional loads from the tile's base with 32-bit
            // offsets: all of them issue back to back (the values past the
            // tile's end are never used)
            const float *bx = (ax == 0 ? i0 : (ax == 1 ? i1 : i2)) + T.begin;
            const float *by = (a2 == 0 ? i0 : (a2 == 1 ? i1 : i2)) + T.begin;
            const float *bz = (a3 == 0 ? i0 : (a3 == 1 ? i1 : i2)) + T.begin;
            const WS *bwt = iw + T.begin;
            // the wave issuing its loads goes ahead of waves busy classifying
            // (keeps more reads in flight; measured 0.6% on the move)
            __builtin_amdgcn_s_setprio(3);
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                const uint32_t p = (uint32_t)(k * kBlock + tid);
                // min() keeps c <= p < kTile visible to the compiler: the
                // offsets are 32-bit, the loads take the base from SGPRs
                const uint32_t c = min(p, cnt - 1u);
                rx[k] = __builtin_nontemporal_load(&bx[c]);
                ry[k] = __builtin_nontemporal_load(&by[c]);
                if constexpr (D == 3) rz[k] = __builtin_nontemporal_load(&bz[c]);
                else rz[k] = 0.f;
                wv[k] = __builtin_nontemporal_load(&bwt[c]);
            }
        }
        __builtin_amdgcn_s_setprio(0);
        uint32_t pq[kItems];   // rank among the wavefront's points of the same slot | slot << 8
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const bool valid = (uint32_t)(k * kBlock + tid) < cnt;
            const bool s1 = rx[k] >= split;
            const bool s2 = ry[k] >= (s1 ? gsplit[1] : gsplit[0]);
            const uint32_t q = (s1 ? 2u : 0u) + (s2 ? 1u : 0u);
            // slot masks from three ballots (wave-uniform)
            const uint64_t bv = __ballot(valid), b1 = __ballot(s1), b2 = __ballot(s2);
            const uint64_t m0 = bv & ~b1 & ~b2, m1 = bv & ~b1 & b2, m2 = bv & b1 & ~b2, m3 = bv & b1 & b2;
            if (lane == 0) {
                const int e = k * (kBlock / 64) + wid;
                wtot[0][e] = __popcll(m0);
                wtot[1][e] = __popcll(m1);
                wtot[2][e] = __popcll(m2);
                wtot[3][e] = __popcll(m3);
            }
            const uint64_t mine = s1 ? (s2 ? m3 : m2) : (s2 ? m1 : m0);
            // rank: the set bits of `mine` below this lane
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mine >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)mine, 0u));
            pq[k] = rank | ((valid ? q : 4u) << 8);
            // computed here, not after the barrier: otherwise the compiler
            // keeps the four slot masks of every item in SGPRs until then
            asm volatile("" : "+v"(pq[k]));
        }
        __syncthreads();
        if (tid < 4) {   // after the barrier: another wavefront may still store the last tile's runs
            const uint64_t off = pv - s_gpj0[tid];   // this tile's first position in the grandchild
            s_dst[tid] = s_gbeg[tid] + off;
            s_off[tid] = (uint32_t)(off % kTile);
            s_tile[tid] = s_gtb[tid] + (uint32_t)(off / kTile);
        }
        if (tid < 64) {   // exclusive scans of the counts of each grandchild slot, item-major
            constexpr int kE = kItems * (kBlock / 64);
            static_assert(kE <= 64, "one wavefront scans the counts");
            uint32_t run = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t c = tid < kE ? wtot[j][tid] : 0u;
                const uint32_t inc = wave_incl_add(c);
                if (tid < kE) wtot[j][tid] = run + inc - c;
                run += __shfl(inc, kE - 1, 64);
                if (tid == 0) s_run[j + 1] = run;
            }
            if (tid == 0) s_run[0] = 0;
        }
        __syncthreads();
        uint32_t cp[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const uint32_t q = (pq[k] >> 8) < 4 ? (pq[k] >> 8) : 0u;
            cp[k] = wtot[q][k * (kBlock / 64) + wid] + (pq[k] & 0xffu);   // includes the run start of q
        }
        // every array staged in slot order (3D: a3 = rz, a4 = the parent's
        // axis rx, a2 = ry; 2D: a3 = rx, a4 = ry); then each wavefront stores
        // a quarter of the staged tile, run segment by run segment: the runs'
        // lengths follow the data (a tile whose points all go to one
        // grandchild is one run), the quarters do not.  The run starts and
        // destinations stay in LDS until the next tile's first barrier, which
        // every wavefront reaches after its stores.
#pragma unroll
        for (int k = 0; k < kItems; k++)
            if (k * kBlock + tid < (int)cnt) {
                if constexpr (D == 3) {
                    sfa[cp[k]] = rz[k];
                    sfb[cp[k]] = rx[k];
                    sfc[cp[k]] = ry[k];
                } else {
                    sfa[cp[k]] = rx[k];
                    sfb[cp[k]] = ry[k];
                }
                sws[cp[k]] = wv[k];
            }
        __syncthreads();
        {
            float *oa2 = a2 == 0 ? o0 : (a2 == 1 ? o1 : o2);
            float *oa3 = a3 == 0 ? o0 : (a3 == 1 ? o1 : o2);
            float *oa4 = a4 == 0 ? o0 : (a4 == 1 ? o1 : o2);
            const bool keep2 = (uint32_t)a2 != skip_axis;   // wave-uniform
            const uint32_t total = s_run[4];
            const uint32_t quarter = ((total + 255u) / 256u) * 64u;   // whole 64-position chunks
            const uint32_t p0 = min((uint32_t)wid * quarter, total), p1 = min(p0 + quarter, total);
            for (int j = 0; j < 4; j++) {   // wave-uniform: the runs this quarter meets
                const uint32_t rbj = s_run[j];
                const uint32_t sa = max(p0, rbj), sb = min(p1, s_run[j + 1]);
                if (sa >= sb) continue;
                const uint32_t roff = sa - rbj;   // the segment's first position in run j
                const uint64_t dv = s_dst[j] + roff;
                // wave-uniform: into SGPRs once, for the stores' bases
                const uint64_t db = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(dv >> 32)) << 32) |
                                    __builtin_amdgcn_readfirstlane((uint32_t)dv);
                uint32_t len = sb - sa;
                len = db + len <= n_total ? len : (db < n_total ? (uint32_t)(n_total - db) : 0u);   // never past the arrays
                // positions of the segment from si on are in the destination
                // tile after the one of run j's first position
                const int32_t si = (int32_t)(kTile - s_off[j]) - (int32_t)roff;
                const uint32_t run_tile = s_tile[j];
                const float t3w = s_t3[j], t4w = s_t4[j];
                if (j != cur_j) {   // the sums go on in registers while the slot stays
                    wave_flush();
                    cur_j = j;
                }
                uint32_t c_left = 0, c_b0 = 0, c_b1 = 0;
                // FWD: the same counts over the positions in the first destination
                // tile (positions below si), taken where the segment crosses it
                uint32_t f_left = 0, f_b0 = 0, f_b1 = 0;
                bool crossed = si <= 0;
                // the segment's destination as uniform bases with 32-bit offsets
                // (a segment lies in one run: i < kTile + 128, so the mask below
                // changes nothing and shows the compiler the offsets' range)
                float *pa2 = oa2 + db, *pa3 = oa3 + db, *pa4 = oa4 + db;
                WS *pw = ow + db;
                // two positions per lane per iteration, every lane runs every
                // iteration: validity is in uniform lane masks, and the ballots
                // take the comparisons alone
                for (uint32_t base = 0; base < len; base += 128) {
                    float xa[2], xb[2], xc[2];
                    WS wq[2];
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        const uint32_t i = base + (uint32_t)u * 64 + (uint32_t)lane;
                        const uint32_t c = sa + (i < len ? i : 0u);
                        xa[u] = sfa[c];
                        xb[u] = sfb[c];
                        if constexpr (D == 3) xc[u] = sfc[c];
                        wq[u] = sws[c];
                    }
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        const uint32_t pos = base + (uint32_t)u * 64;   // wave-uniform
                        const bool vl = pos + (uint32_t)lane < len;
                        const uint64_t vm = __builtin_amdgcn_ballot_w64(vl);   // lanes holding a position of the segment
                        const uint32_t i = (pos + (uint32_t)lane) & (4 * kTile - 1);
                        if (vl) {
                            if constexpr (NT) {
                                if constexpr (D == 3)
                                    if (keep2) __builtin_nontemporal_store(xc[u], &pa2[i]);
                                __builtin_nontemporal_store(xa[u], &pa3[i]);
                                __builtin_nontemporal_store(xb[u], &pa4[i]);
                                __builtin_nontemporal_store(wq[u], &pw[i]);
                            } else {
                                if constexpr (D == 3)
                                    if (keep2) pa2[i] = xc[u];
                                pa3[i] = xa[u];
                                pa4[i] = xb[u];
                                pw[i] = wq[u];
                            }
                        }
                        const bool lt3 = !(xa[u] >= t3w);
                        const uint64_t bb = __builtin_amdgcn_ballot_w64(lt3) & vm;
                        const bool bel3 = vl && lt3;
                        const Acc wa = AccOps<Acc>::of(wq[u]);
                        // the grandchild's first pass
                        c_left += __popcll(bb);
                        gws += bel3 ? wa : (Acc)0;
                        const uint32_t ka = (vl && !lt3) ? f32_key(xa[u]) : 0xffffffffu;
                        gmn = ka < gmn ? ka : gmn;
                        if constexpr (FWD) {
                            const bool lt4 = !(xb[u] >= t4w);
                            const uint64_t b4 = __builtin_amdgcn_ballot_w64(lt4) & vm;
                            if (!crossed && (uint32_t)si < pos + 64) {
                                // the chunk that holds the destination tile's end (once
                                // per segment, uniform): the counts of the positions before it
                                const uint64_t lo = (1ull << ((uint32_t)si - pos)) - 1ull;   // si - pos in [0, 63]
                                f_left = c_left - __popcll(bb) + __popcll(bb & lo);
                                f_b0 = c_b0 + __popcll(bb & b4 & lo);
                                f_b1 = c_b1 + __popcll(~bb & b4 & lo);
                                crossed = true;
                            }
                            c_b0 += __popcll(bb & b4);
                            c_b1 += __popcll(~bb & b4);
                            const uint32_t key = f32_key(xb[u]);
                            const bool below4 = vl && lt4;
                            fws0 += (below4 && lt3) ? wa : (Acc)0;
                            fws1 += (below4 && !lt3) ? wa : (Acc)0;
                            const bool above4 = vl && !lt4;
                            const uint32_t k0 = (above4 && lt3) ? key : 0xffffffffu;
                            const uint32_t k1 = (above4 && !lt3) ? key : 0xffffffffu;
                            fmn0 = k0 < fmn0 ? k0 : fmn0;
                            fmn1 = k1 < fmn1 ? k1 : fmn1;
                        }
                    }
                }
                gcnt += c_left;
                if constexpr (FWD) {
                    fcn0 += c_b0;
                    fcn1 += c_b1;
                    if (!crossed) { f_left = c_left; f_b0 = c_b0; f_b1 = c_b1; }
                    const uint32_t h_left = c_left - f_left, h_b0 = c_b0 - f_b0, h_b1 = c_b1 - f_b1;
                    if (lane == 0) {
                        if (len > 0) {   // per destination tile
                            const uint64_t dt = run_tile;
                            if (f_left) atomicAdd((unsigned long long *)&fw.left[dt], (unsigned long long)f_left);
                            if (f_b0) atomicAdd(&fw.below2[2 * dt], f_b0);
                            if (f_b1) atomicAdd(&fw.below2[2 * dt + 1], f_b1);
                            if (h_left) atomicAdd((unsigned long long *)&fw.left[dt + 1], (unsigned long long)h_left);
                            if (h_b0) atomicAdd(&fw.below2[2 * dt + 2], h_b0);
                            if (h_b1) atomicAdd(&fw.below2[2 * dt + 3], h_b1);
                        }
                    }
                }
            }
        }
    }
    wave_flush();
    __syncthreads();   // every wavefront's stores and sums of the last part are done
    if (cur != 0xffffffffu) flush();
}

}  // namespace coupe_hip
