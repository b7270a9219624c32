/*
 * grid_oracle.c -- CPU restatement of coupe's RCB on cartesian grids.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker of coupe_hip_grid_rcb); the
 * product library never links or calls it.
 *
 *   Grid::rcb (2D, 3D)      coupe/src/cartesian/mod.rs:127-180
 *   Grid::index_of / position_of   mod.rs:58-124
 *   SubGrid::axis / split_at       mod.rs:183-218
 *   weighted_median         coupe/src/cartesian/rcb.rs:52-99 (TOLERANCE 0.01, :44)
 *   recurse_2d / recurse_3d rcb.rs:101-281
 *   IterationResult::part_of rcb.rs:21-42
 *
 * Weights are int32 (kind 0, sums wrap mod 2^32 as Rust's release
 * arithmetic), int64 (kind 1) or double (kind 2).  Sums follow the
 * reference's order: axis sums run over the cross-section in the order of
 * recurse_2d / recurse_3d, the weighted median's chunk sums are sequential
 * within a chunk and its prefix runs over the chunks in order.  The total
 * weight of the grid (a rayon sum in the reference) is taken in index
 * order.  `threads` plays rayon::current_num_threads() in the median's
 * chunk size; with 1 the reference's search can loop forever, here it stops
 * when a round changes nothing.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef union { int64_t i; double f; } gval;

static gval g_add(int kind, gval a, gval b) {
    gval r;
    if (kind == 2) r.f = a.f + b.f;
    else if (kind == 0) r.i = (int32_t)(uint32_t)((uint32_t)a.i + (uint32_t)b.i);
    else r.i = (int64_t)((uint64_t)a.i + (uint64_t)b.i);
    return r;
}
static gval g_sub(int kind, gval a, gval b) {
    gval r;
    if (kind == 2) r.f = a.f - b.f;
    else if (kind == 0) r.i = (int32_t)(uint32_t)((uint32_t)a.i - (uint32_t)b.i);
    else r.i = (int64_t)((uint64_t)a.i - (uint64_t)b.i);
    return r;
}
static int g_lt(int kind, gval a, gval b) { return kind == 2 ? a.f < b.f : a.i < b.i; }
static double g_f64(int kind, gval a) { return kind == 2 ? a.f : (double)a.i; }
static gval g_zero(void) { gval r; r.i = 0; return r; }
/* f64 -> W with Rust's `as`: truncation, saturation, NaN -> 0 */
static gval g_from_f64(int kind, double x) {
    gval r;
    if (kind == 2) { r.f = x; return r; }
    double lo = kind == 0 ? -2147483648.0 : -9223372036854775808.0;
    double hi = kind == 0 ? 2147483647.0 : 9223372036854775807.0;
    if (x != x) r.i = 0;
    else if (x <= lo) r.i = kind == 0 ? INT32_MIN : INT64_MIN;
    else if (x >= hi) r.i = kind == 0 ? INT32_MAX : INT64_MAX;
    else r.i = (int64_t)x;
    return r;
}

typedef struct {
    int kind;
    size_t dim;
    size_t size[3];
    const void *w;
    int threads;
    /* split tree, heap-indexed from 1 (children 2h, 2h+1): position or -1 (whole) */
    int64_t *pos;
    size_t n_nodes;
} gctx;

static gval load(const gctx *c, size_t i) {
    gval r;
    if (c->kind == 0) r.i = ((const int32_t *)c->w)[i];
    else if (c->kind == 1) r.i = ((const int64_t *)c->w)[i];
    else r.f = ((const double *)c->w)[i];
    return r;
}

static size_t index_of(const gctx *c, const size_t *p) {
    if (c->dim == 2) return p[0] + c->size[0] * p[1];
    return p[0] + c->size[0] * (p[1] + c->size[1] * p[2]);
}

/* weighted_median (rcb.rs:52-99) */
static void weighted_median(int kind, const gval *w, size_t len, gval total, int threads, size_t *out_pos,
                            gval *out_left) {
    const double ideal = g_f64(kind, total) / 2.0;
    const gval minp = g_from_f64(kind, ideal * (1.0 - 0.01));
    const gval maxp = g_from_f64(kind, ideal * (1.0 + 0.01));
    size_t min = 0, max = len;
    gval left = g_zero();
    for (;;) {
        size_t cs = (max - min) / (size_t)threads;
        if (cs < 1) cs = 1;
        const size_t min0 = min, max0 = max;
        const gval left0 = left;
        gval prefix = g_zero();
        for (size_t start = min0; start < max0; start += cs) {
            size_t end = start + cs < max0 ? start + cs : max0;
            gval chunk = g_zero();
            for (size_t i = start; i < end; i++) chunk = g_add(kind, chunk, w[i]);
            const gval pcw = g_add(kind, left0, prefix);
            prefix = g_add(kind, prefix, chunk);
            if (g_lt(kind, pcw, minp)) {
                min = start;
                left = pcw;
            } else if (g_lt(kind, maxp, pcw)) {
                max = start;
                break;
            } else {
                *out_pos = start;
                *out_left = pcw;
                return;
            }
        }
        if (min + 1 >= max || (min == min0 && max == max0)) {
            *out_pos = min;
            *out_left = left;
            return;
        }
    }
}

/* recurse_2d / recurse_3d (rcb.rs:101-281) */
static int recurse(gctx *c, const size_t *off, const size_t *sz, gval total, size_t iters, size_t coord,
                   size_t heap) {
    if (sz[coord] == 0 || iters == 0) return 0;
    const size_t m = sz[coord];
    gval *aw = malloc((m ? m : 1) * sizeof(gval));
    if (!aw) return 1;
    const size_t d = c->dim;
    /* cross-section order: 2D coord 0 sums over y, coord 1 over x;
       3D coord 0: y outer z inner; coord 1: z outer x inner; coord 2: x outer y inner */
    const size_t o1 = d == 2 ? 1 - coord : (coord == 0 ? 1 : (coord == 1 ? 2 : 0));
    const size_t o2 = d == 2 ? coord : (coord == 0 ? 2 : (coord == 1 ? 0 : 1));
    for (size_t a = 0; a < m; a++) {
        gval s = g_zero();
        size_t p[3] = {0, 0, 0};
        p[coord] = off[coord] + a;
        if (d == 2) {
            for (size_t b = 0; b < sz[o1]; b++) {
                p[o1] = off[o1] + b;
                s = g_add(c->kind, s, load(c, index_of(c, p)));
            }
        } else {
            for (size_t b = 0; b < sz[o1]; b++)
                for (size_t e = 0; e < sz[o2]; e++) {
                    p[o1] = off[o1] + b;
                    p[o2] = off[o2] + e;
                    s = g_add(c->kind, s, load(c, index_of(c, p)));
                }
        }
        aw[a] = s;
    }
    size_t pos;
    gval left;
    weighted_median(c->kind, aw, m, total, c->threads, &pos, &left);
    free(aw);
    const size_t split = pos + off[coord];
    const gval right = g_sub(c->kind, total, left);
    if (heap < c->n_nodes) c->pos[heap] = (int64_t)split;
    size_t loff[3], lsz[3], hoff[3], hsz[3];
    memcpy(loff, off, sizeof loff);
    memcpy(lsz, sz, sizeof lsz);
    memcpy(hoff, off, sizeof hoff);
    memcpy(hsz, sz, sizeof hsz);
    lsz[coord] = split - off[coord];
    hsz[coord] -= split - off[coord];
    hoff[coord] = split;
    const size_t next = (coord + 1) % d;
    if (recurse(c, loff, lsz, left, iters - 1, next, 2 * heap)) return 1;
    return recurse(c, hoff, hsz, right, iters - 1, next, 2 * heap + 1);
}

/* Grid::rcb: ids of every cell, in memory order.  Returns 0, 1 (allocation)
 * or 3 (dimension).  `splits` (optional, 2^iter_count entries) receives
 * the heap of split positions (-1 where a subgrid was whole). */
int oracle_grid_rcb(uint64_t *part, size_t dim, const size_t *size, int kind, const void *weights,
                    size_t iter_count, int threads, int64_t *splits) {
    if (dim != 2 && dim != 3) return 3;
    gctx c;
    memset(&c, 0, sizeof c);
    c.kind = kind;
    c.dim = dim;
    for (size_t k = 0; k < dim; k++) c.size[k] = size[k];
    c.w = weights;
    c.threads = threads < 1 ? 1 : threads;
    size_t n = 1;
    for (size_t k = 0; k < dim; k++) n *= size[k];
    c.n_nodes = iter_count < 40 ? ((size_t)1 << iter_count) : 0;
    c.pos = malloc((c.n_nodes ? c.n_nodes : 1) * sizeof(int64_t));
    if (!c.pos) return 1;
    for (size_t h = 0; h < c.n_nodes; h++) c.pos[h] = -1;
    gval total = g_zero();
    for (size_t i = 0; i < n; i++) total = g_add(kind, total, load(&c, i));
    size_t off[3] = {0, 0, 0}, sz[3] = {c.size[0], c.size[1], dim == 3 ? c.size[2] : 1};
    if (recurse(&c, off, sz, total, iter_count, 1, 1)) {
        free(c.pos);
        return 1;
    }
    /* IterationResult::part_of(pos, 1) (rcb.rs:21-42) */
    for (size_t i = 0; i < n; i++) {
        size_t p[3];
        p[0] = i % c.size[0];
        p[1] = dim == 2 ? i / c.size[0] : (i / c.size[0]) % c.size[1];
        p[2] = dim == 3 ? i / c.size[0] / c.size[1] : 0;
        size_t h = 1, coord = 1;
        uint64_t id = 0;
        while (h < c.n_nodes && c.pos[h] >= 0) {
            if ((int64_t)p[coord] < c.pos[h]) {
                id = 2 * id;
                h = 2 * h;
            } else {
                id = 2 * id + 1;
                h = 2 * h + 1;
            }
            coord = (coord + 1) % dim;
        }
        part[i] = id;
    }
    if (splits)
        for (size_t h = 0; h < c.n_nodes; h++) splits[h] = c.pos[h];
    free(c.pos);
    return 0;
}

/* weighted_median alone, for the reference's property test (rcb.rs:300-314). */
void oracle_weighted_median(int kind, const void *w, size_t len, int threads, size_t *pos, void *left) {
    gval *a = malloc((len ? len : 1) * sizeof(gval));
    gctx c;
    memset(&c, 0, sizeof c);
    c.kind = kind;
    c.w = w;
    gval total = g_zero();
    for (size_t i = 0; i < len; i++) {
        a[i] = load(&c, i);
        total = g_add(kind, total, a[i]);
    }
    gval l;
    weighted_median(kind, a, len, total, threads < 1 ? 1 : threads, pos, &l);
    memcpy(left, &l, 8);
    free(a);
}
