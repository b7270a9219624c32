/*
 * rcb_oracle.c -- CPU restatement of coupe's Recursive Coordinate Bisection.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * implementation in coupe_amd/csrc and the CPU baseline timed beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * Every function cites the reference lines it restates
 * (paths relative to the coupe repository root):
 *
 *   rcb()            coupe/src/algorithms/recursive_bisection.rs:644-705
 *   rcb_recurse()    coupe/src/algorithms/recursive_bisection.rs:575-642
 *   par_rcb_split()  coupe/src/algorithms/recursive_bisection.rs:456-573
 *   reorder_split_scalar()  coupe/src/algorithms/recursive_bisection.rs:122-181
 *   BoundingBox::from_points  coupe/src/geometry.rs:33-78
 *   RcbWeight arithmetic      coupe/src/algorithms/recursive_bisection.rs:707-722
 *   coupe_rcb weight types    coupe-ffi/src/data.rs:249-255 (int, int64, double)
 *
 * Deterministic choices where the reference is not deterministic:
 *  - nearest point: the reference keeps the index of the smallest f32
 *    distance fl(c - t) over c >= t, with ties broken by rayon's chunking
 *    (recursive_bisection.rs:496-512).  Here the pivot is the smallest
 *    coordinate c >= t ("canonical pivot"); it is one of the reference's
 *    possible outcomes.  When another, distinct coordinate has the same
 *    rounded distance the oracle counts an "ambiguous pivot" event so a test
 *    can tell that case apart.
 *  - weight sums for non-integer f64 weights follow the current element order
 *    (the reference's rayon reduction order is unspecified).
 *
 * Parity status of each function (how it is pinned) is listed in the module
 * header of oracle/oracle.py.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, SSE float math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Weight kinds, numbered as coupe_type (coupe-ffi/include/coupe.h:86-93). */
enum { OR_W_INT = 0, OR_W_INT64 = 1, OR_W_DOUBLE = 2 };

/* Error codes, numbered as enum coupe_err (coupe-ffi/include/coupe.h:16-57). */
enum { OR_OK = 0, OR_ALLOC = 1, OR_BAD_DIMENSION = 3, OR_LEN_MISMATCH = 6 };

/* One bisection decision, recorded for debugging and for the tests. */
typedef struct {
    uint64_t heap_id;    /* iter_id of the node being split */
    uint64_t count;      /* points in the node */
    uint64_t count_left; /* points sent left */
    double weight_left;  /* weight_left as f64 */
    float split_pos;     /* split_pos (f32) */
    uint32_t steps;      /* candidates evaluated */
    uint32_t level;      /* depth of the node (0 = root) */
    uint32_t all_left;   /* 1 when the "no point right of t twice" exit fired */
} oracle_node;

typedef struct {
    uint64_t ambiguous_pivots; /* distinct c with equal fl(c - t) at a stop */
    uint64_t candidates;       /* total split candidates evaluated */
    uint64_t nodes;            /* splits performed */
    oracle_node *trace;        /* optional, caller-owned */
    uint64_t trace_cap;
    uint64_t trace_len;
} oracle_stats;

/* ---- W arithmetic (RcbWeight, recursive_bisection.rs:707-722) -------------
 * A W value is held in a union; integer kinds use two's complement wrapping
 * arithmetic, as Rust release builds do for i32/i64 `+` and `-`.           */
typedef union { int64_t i; double f; } wval;

static inline wval w_add(int kind, wval a, wval b) {
    wval r;
    if (kind == OR_W_DOUBLE) r.f = a.f + b.f;
    else if (kind == OR_W_INT) r.i = (int32_t)(uint32_t)((uint32_t)a.i + (uint32_t)b.i);
    else r.i = (int64_t)((uint64_t)a.i + (uint64_t)b.i);
    return r;
}
static inline wval w_sub(int kind, wval a, wval b) {
    wval r;
    if (kind == OR_W_DOUBLE) r.f = a.f - b.f;
    else if (kind == OR_W_INT) r.i = (int32_t)(uint32_t)((uint32_t)a.i - (uint32_t)b.i);
    else r.i = (int64_t)((uint64_t)a.i - (uint64_t)b.i);
    return r;
}
static inline int w_lt(int kind, wval a, wval b) {
    return kind == OR_W_DOUBLE ? (a.f < b.f) : (a.i < b.i);
}
/* ToPrimitive::to_f64 (num-traits 0.2.19): exact for |x| <= 2^53. */
static inline double w_f64(int kind, wval a) {
    return kind == OR_W_DOUBLE ? a.f : (double)a.i;
}
static inline wval w_zero(void) { wval r; r.i = 0; return r; }

/* ---- working set (Items, recursive_bisection.rs:105-109) ------------------ */
typedef struct {
    int dim;
    int kind;
    double tol;
    float *c[3];     /* SoA f32 coordinates, permuted in place */
    wval *w;         /* weights, permuted in place */
    uint32_t *idx;   /* original index: stands for the &AtomicUsize pointers */
    uint64_t *part;  /* output ids, indexed by original index */
    oracle_stats *st;
    int par;         /* 1 = fork-join variant */
    int nchunks;     /* chunks per parallel pass */
} items;

static inline void swap_item(items *it, size_t a, size_t b) {
    for (int k = 0; k < it->dim; k++) {
        float t = it->c[k][a]; it->c[k][a] = it->c[k][b]; it->c[k][b] = t;
    }
    wval tw = it->w[a]; it->w[a] = it->w[b]; it->w[b] = tw;
    uint32_t ti = it->idx[a]; it->idx[a] = it->idx[b]; it->idx[b] = ti;
}

/* reorder_split_scalar, recursive_bisection.rs:122-181.
 * Moves the element `pivot` to the front, runs the Hoare-style loop on the
 * rest with `c < pivot` going left and `pivot <= c` going right, then swaps
 * the pivot into place.  Returns the size of the left part.               */
static size_t reorder_split(items *it, size_t begin, size_t end, size_t pivot, int coord) {
    swap_item(it, begin, pivot);
    const float pv = it->c[coord][begin];
    const float *cs = it->c[coord];
    size_t l = begin + 1, r = end;
    for (;;) {
        while (l < r && cs[l] < pv) l++;
        while (l < r && pv <= cs[r - 1]) r--;
        if (r <= l) break;
        r--;
        swap_item(it, l, r);
        l++;
    }
    /* `l` counts from begin+1 in the sub-slice; the reference swaps the
     * pivot (slot 0) with slot l of the sub-slice, i.e. absolute l - 1 + 1. */
    size_t left = l - (begin + 1);
    swap_item(it, begin, begin + left);
    return left;
}

typedef struct {
    uint64_t count;
    wval wl;
    int have;      /* some c >= t */
    float m;       /* min c >= t */
    size_t m_idx;
} pass_out;

/* One candidate pass of par_rcb_split (recursive_bisection.rs:478-520):
 * count(c < t), sum of weights (c < t), and the smallest c >= t.
 * The side test is the reference's `point - split_target < 0.0`.        */
static pass_out split_pass(const items *it, size_t begin, size_t end, int coord, float t) {
    pass_out o;
    o.count = 0; o.wl = w_zero(); o.have = 0; o.m = INFINITY; o.m_idx = begin;
    const float *cs = it->c[coord];
    if (it->par && end - begin >= 65536) {
        /* Fork-join pass: fixed chunks run as tasks and are combined in
         * order, like rayon's fold/reduce (min chunk 4096, :480).          */
        enum { MAXCH = 256 };
        pass_out partial[MAXCH];
        size_t len = end - begin;
        size_t nch = (size_t)it->nchunks;
        if (nch > MAXCH) nch = MAXCH;
        if (nch > len / 4096) nch = len / 4096;
        if (nch < 1) nch = 1;
        for (size_t ch = 0; ch < nch; ch++) {
#ifdef _OPENMP
#pragma omp task default(shared) firstprivate(ch)
#endif
            {
                size_t b = begin + len * ch / nch, e = begin + len * (ch + 1) / nch;
                pass_out p;
                p.count = 0; p.wl = w_zero(); p.have = 0; p.m = INFINITY; p.m_idx = b;
                for (size_t i = b; i < e; i++) {
                    float d = cs[i] - t;
                    if (d < 0.0f) { p.count++; p.wl = w_add(it->kind, p.wl, it->w[i]); }
                    else if (!p.have || cs[i] < p.m) { p.m = cs[i]; p.m_idx = i; p.have = 1; }
                }
                partial[ch] = p;
            }
        }
#ifdef _OPENMP
#pragma omp taskwait
#endif
        for (size_t ch = 0; ch < nch; ch++) {
            pass_out *p = &partial[ch];
            o.count += p->count;
            o.wl = w_add(it->kind, o.wl, p->wl);
            if (p->have && (!o.have || p->m < o.m)) { o.m = p->m; o.m_idx = p->m_idx; o.have = 1; }
        }
        return o;
    }
    for (size_t i = begin; i < end; i++) {
        float d = cs[i] - t;
        if (d < 0.0f) {
            o.count++;
            o.wl = w_add(it->kind, o.wl, it->w[i]);
        } else if (!o.have || cs[i] < o.m) {
            o.m = cs[i]; o.m_idx = i; o.have = 1;
        }
    }
    return o;
}

typedef struct {
    size_t left;    /* size of the left part */
    wval wl;        /* weight_left */
    float split;    /* split_pos */
} split_res;

/* par_rcb_split, recursive_bisection.rs:456-573. */
static split_res rcb_split(items *it, size_t begin, size_t end, int coord,
                           float min, float max, wval sum, oracle_node *node) {
    uint64_t prev = UINT64_MAX;               /* :470 */
    uint32_t steps = 0;
    for (;;) {
        float t = (min + max) / 2.0f;         /* :472 */
        steps++;
        pass_out o = split_pass(it, begin, end, coord, t);
        if (!o.have) {                        /* :522-545 */
            if (prev == o.count) {
                split_res r = { end - begin, sum, max };
                node->steps = steps; node->all_left = 1;
                return r;
            }
            max = t;
            prev = o.count;
            continue;
        }
        float nd = o.m - t;                   /* nearest_distance */
        double ideal = w_f64(it->kind, sum) / 2.0;               /* :547-551 */
        double imb = fabs((w_f64(it->kind, o.wl) - ideal) / ideal);
        if (o.count == prev || max <= t + nd || imb <= it->tol) { /* :552-563 */
            /* Count distinct coordinates whose rounded distance ties the
             * nearest one: the reference may pick any of them as pivot. */
            const float *cs = it->c[coord];
            for (size_t i = begin; i < end; i++) {
                if (cs[i] > o.m && cs[i] - t == nd) { it->st->ambiguous_pivots++; break; }
            }
            size_t left = reorder_split(it, begin, end, o.m_idx, coord);
            split_res r = { left, o.wl, t };
            node->steps = steps; node->all_left = 0;
            return r;
        }
        prev = o.count;                       /* :564 */
        wval wr = w_sub(it->kind, sum, o.wl); /* :566-571 */
        if (w_lt(it->kind, o.wl, wr)) min = t; else max = t;
    }
}

static void record(items *it, const oracle_node *n) {
    oracle_stats *st = it->st;
#ifdef _OPENMP
#pragma omp critical(oracle_record)
#endif
    {
        st->nodes++;
        st->candidates += n->steps;
        if (st->trace && st->trace_len < st->trace_cap) st->trace[st->trace_len++] = *n;
    }
}

/* rcb_recurse, recursive_bisection.rs:575-642.  The box is kept in f64 like
 * BoundingBox and cast to f32 where the reference casts (:604-605, :613-616). */
static void rcb_recurse(items *it, size_t begin, size_t end, size_t iter_count,
                        uint64_t iter_id, int coord, wval sum,
                        const double *bmin, const double *bmax, uint32_t level) {
    if (begin == end) return;                 /* :586-588 */
    if (iter_count == 0) {                    /* :589-601 */
        for (size_t i = begin; i < end; i++) it->part[it->idx[i]] = iter_id;
        return;
    }
    float mn = (float)bmin[coord];
    float mx = (float)bmax[coord];
    oracle_node node;
    memset(&node, 0, sizeof node);
    node.heap_id = iter_id; node.count = end - begin; node.level = level;
    split_res s = rcb_split(it, begin, end, coord, mn, mx, sum, &node);
    node.count_left = s.left; node.weight_left = w_f64(it->kind, s.wl); node.split_pos = s.split;
    record(it, &node);

    double lmin[3], lmax[3], rmin[3], rmax[3];
    for (int k = 0; k < it->dim; k++) { lmin[k] = rmin[k] = bmin[k]; lmax[k] = rmax[k] = bmax[k]; }
    lmax[coord] = (double)s.split;
    rmin[coord] = (double)s.split;
    int next = (coord + 1) % it->dim;
    wval wr = w_sub(it->kind, sum, s.wl);
    /* Rust usize arithmetic; wraps like release builds for huge iter_count. */
    uint64_t idl = 2 * iter_id + 1, idr = 2 * iter_id + 2;
    size_t mid = begin + s.left;
    if (it->par && end - begin >= 32768) {
#ifdef _OPENMP
#pragma omp task default(shared) firstprivate(idl, next, mid, level, iter_count)
#endif
        rcb_recurse(it, begin, mid, iter_count - 1, idl, next, s.wl, lmin, lmax, level + 1);
        rcb_recurse(it, mid, end, iter_count - 1, idr, next, wr, rmin, rmax, level + 1);
#ifdef _OPENMP
#pragma omp taskwait
#endif
    } else {
        rcb_recurse(it, begin, mid, iter_count - 1, idl, next, s.wl, lmin, lmax, level + 1);
        rcb_recurse(it, mid, end, iter_count - 1, idr, next, wr, rmin, rmax, level + 1);
    }
}

static wval load_w(int kind, const void *w, size_t i) {
    wval r;
    if (kind == OR_W_INT) r.i = ((const int32_t *)w)[i];
    else if (kind == OR_W_INT64) r.i = ((const int64_t *)w)[i];
    else r.f = ((const double *)w)[i];
    return r;
}

/* rcb(), recursive_bisection.rs:644-705, behind the coupe_rcb checks of
 * coupe-ffi/src/lib.rs:272-304 (length mismatch first, then dimension).
 * `weights_const` != 0 means the weight data set repeats weights[0]
 * (coupe_data_constant, coupe-ffi/src/data.rs:19-56).                      */
static int oracle_run(uint64_t *part, size_t n_part, size_t dim,
                      const double *points, size_t n_points,
                      int kind, const void *weights, size_t n_weights, int weights_const,
                      size_t iter_count, double tol, oracle_stats *st, int par, int nthreads) {
    if (n_points != n_weights) return OR_LEN_MISMATCH;   /* ffi lib.rs:285-288 */
    if (dim != 2 && dim != 3) return OR_BAD_DIMENSION;   /* ffi lib.rs:297-301 */
    if (n_weights != n_part || n_points != n_part) return OR_LEN_MISMATCH; /* :661-672 */
    oracle_stats local;
    memset(&local, 0, sizeof local);
    if (!st) st = &local;
    size_t n = n_part;
    items it;
    memset(&it, 0, sizeof it);
    it.dim = (int)dim; it.kind = kind; it.tol = tol; it.part = part; it.st = st; it.par = par;
    it.nchunks = 1;
    for (size_t k = 0; k < dim; k++) {
        it.c[k] = (float *)malloc((n ? n : 1) * sizeof(float));
        if (!it.c[k]) return OR_ALLOC;
    }
    it.w = (wval *)malloc((n ? n : 1) * sizeof(wval));
    it.idx = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!it.w || !it.idx) return OR_ALLOC;
#ifdef _OPENMP
    if (par && nthreads > 0) omp_set_num_threads(nthreads);
    if (par) it.nchunks = 4 * omp_get_max_threads();
#endif
    /* :674-680 -- SoA f32 copies (`as f32`, round to nearest even) and Vec<W>. */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (par)
#endif
    for (size_t i = 0; i < n; i++) {
        for (size_t k = 0; k < dim; k++) it.c[k][i] = (float)points[i * dim + k];
        it.w[i] = load_w(kind, weights, weights_const ? 0 : i);
        it.idx[i] = (uint32_t)i;
    }
    /* :684 -- sum of the weights in W.  Integer kinds wrap, so any order
     * gives the same bits; f64 is summed in index order.                   */
    wval sum = w_zero();
    if (par && kind != OR_W_DOUBLE) {
        uint64_t acc = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : acc) schedule(static)
#endif
        for (size_t i = 0; i < n; i++) acc += (uint64_t)it.w[i].i;
        sum.i = kind == OR_W_INT ? (int64_t)(int32_t)(uint32_t)acc : (int64_t)acc;
    } else {
        for (size_t i = 0; i < n; i++) sum = w_add(kind, sum, it.w[i]);
    }
    /* geometry.rs:33-78 -- None for an empty input, then Ok(()) (:685-688). */
    int rc = OR_OK;
    if (n > 0) {
        double bmin[3], bmax[3];
        for (size_t k = 0; k < 3; k++) { bmin[k] = DBL_MAX; bmax[k] = -DBL_MAX; }
#ifdef _OPENMP
#pragma omp parallel for reduction(min : bmin[:3]) reduction(max : bmax[:3]) schedule(static) if (par)
#endif
        for (size_t i = 0; i < n; i++) {
            for (size_t k = 0; k < dim; k++) {
                double v = points[i * dim + k];
                if (v < bmin[k]) bmin[k] = v;
                if (bmax[k] < v) bmax[k] = v;
            }
        }
        if (par) {
#ifdef _OPENMP
#pragma omp parallel
#pragma omp single
#endif
            rcb_recurse(&it, 0, n, iter_count, 0, 0, sum, bmin, bmax, 0);
        } else {
            rcb_recurse(&it, 0, n, iter_count, 0, 0, sum, bmin, bmax, 0);
        }
        /* :698-702 -- ids start at zero. */
        uint64_t off = UINT64_MAX;
        for (size_t i = 0; i < n; i++) if (part[i] < off) off = part[i];
        for (size_t i = 0; i < n; i++) part[i] -= off;
    }
    for (size_t k = 0; k < dim; k++) free(it.c[k]);
    free(it.w);
    free(it.idx);
    return rc;
}

/* Sequential canonical oracle (parity). */
int oracle_rcb(uint64_t *part, size_t n_part, size_t dim, const double *points, size_t n_points,
               int kind, const void *weights, size_t n_weights, int weights_const,
               size_t iter_count, double tol, oracle_stats *st) {
    return oracle_run(part, n_part, dim, points, n_points, kind, weights, n_weights, weights_const,
                      iter_count, tol, st, 0, 0);
}

/* Fork-join variant: OpenMP tasks where the reference uses rayon::join
 * (:618-641) and parallel passes where it uses par_iter fold/reduce
 * (:478-520); the reorder stays serial per part as in the reference.
 * Integer weights give the same ids as oracle_rcb.                        */
int oracle_rcb_par(uint64_t *part, size_t n_part, size_t dim, const double *points, size_t n_points,
                   int kind, const void *weights, size_t n_weights, int weights_const,
                   size_t iter_count, double tol, oracle_stats *st, int nthreads) {
    return oracle_run(part, n_part, dim, points, n_points, kind, weights, n_weights, weights_const,
                      iter_count, tol, st, 1, nthreads);
}

/* BoundingBox::from_points, geometry.rs:33-78: returns 0 for an empty input
 * (None), else 1 with the per-axis f64 min/max.                          */
int oracle_bbox(const double *points, size_t n, size_t dim, double *pmin, double *pmax) {
    if (n == 0) return 0;
    for (size_t k = 0; k < dim; k++) { pmin[k] = DBL_MAX; pmax[k] = -DBL_MAX; }
    for (size_t i = 0; i < n; i++)
        for (size_t k = 0; k < dim; k++) {
            double v = points[i * dim + k];
            if (v < pmin[k]) pmin[k] = v;
            if (pmax[k] < v) pmax[k] = v;
        }
    return 1;
}

/* One split of a single node, exposed for the property test that restates
 * test_par_rcb_split (recursive_bisection.rs:1041-1075): f32 coordinates,
 * u32 weights (carried as int64), tolerance, min/max given by the caller.
 * Outputs the permuted coordinates/weights and the split result.          */
int oracle_split_1d(float *coords, int64_t *weights, size_t n, double tol, float min, float max,
                    uint64_t *count_left, int64_t *weight_left, float *split_pos) {
    items it;
    oracle_stats st;
    memset(&it, 0, sizeof it);
    memset(&st, 0, sizeof st);
    it.dim = 1; it.kind = OR_W_INT64; it.tol = tol; it.st = &st;
    it.c[0] = coords;
    it.w = (wval *)malloc((n ? n : 1) * sizeof(wval));
    it.idx = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!it.w || !it.idx) return OR_ALLOC;
    wval sum = w_zero();
    for (size_t i = 0; i < n; i++) {
        it.w[i].i = weights[i]; it.idx[i] = (uint32_t)i;
        sum = w_add(OR_W_INT64, sum, it.w[i]);
    }
    oracle_node node;
    memset(&node, 0, sizeof node);
    split_res r = rcb_split(&it, 0, n, 0, min, max, sum, &node);
    for (size_t i = 0; i < n; i++) weights[i] = it.w[i].i;
    *count_left = r.left; *weight_left = r.wl.i; *split_pos = r.split;
    free(it.w);
    free(it.idx);
    return OR_OK;
}

/* reorder_split_scalar alone (restating test_reorder_split_scalar,
 * recursive_bisection.rs:952-983).  Returns the size of the left part.    */
size_t oracle_reorder_1d(float *coords, size_t n, size_t pivot) {
    items it;
    oracle_stats st;
    memset(&it, 0, sizeof it);
    memset(&st, 0, sizeof st);
    it.dim = 1; it.kind = OR_W_INT; it.st = &st;
    it.c[0] = coords;
    it.w = (wval *)calloc(n ? n : 1, sizeof(wval));
    it.idx = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    size_t left = reorder_split(&it, 0, n, pivot, 0);
    free(it.w);
    free(it.idx);
    return left;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
