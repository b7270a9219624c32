/*
 * workloads.c -- synthetic RCB inputs for benchmarks and tests.
 *
 * PRNG: SplitMix64.  Output k (k = 0, 1, ...) of the stream seeded with s is
 * mix(s + (k + 1) * 0x9E3779B97F4A7C15), so every point consumes a fixed
 * number of draws at a fixed offset and generation runs in parallel while
 * producing exactly the sequential stream.  u = (x >> 11) * 2^-53 in [0, 1).
 * A normal draw uses two uniforms (Box-Muller, cosine branch):
 * sqrt(-2 ln(1 - u1)) * cos(2 pi u2).  round() is C's (half away from 0).
 *
 * Workloads (points are AoS f64, weights f64):
 *   uniform      n points uniform in [0,1)^D, weight 1 or 1 + floor(wmax u)
 *   gauss_mix    8 isotropic Gaussians (centres uniform in [0,1)^3, sigma s),
 *                weight 1 + floor(100 u)
 *   jgrid        cell centroids of an m^3 grid, index i + m (j + m k),
 *                each jittered by uniform +-0.25 cell, weight
 *                round(exp(N(0,1)) * 10) + 1; or, unjittered with unit weights
 *   box          n points uniform in an axis-aligned box, weight 1 + floor(wmax u)
 *   batch_header instance sizes / dimensions / boxes / iteration counts
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#define GOLDEN 0x9E3779B97F4A7C15ull

static inline uint64_t sm64(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * GOLDEN;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t seed, uint64_t k) { return (double)(sm64(seed, k) >> 11) * 0x1.0p-53; }
static inline double normal(uint64_t seed, uint64_t k) {
    double u1 = u01(seed, k), u2 = u01(seed, k + 1);
    return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}

uint64_t wl_splitmix64(uint64_t seed, uint64_t k) { return sm64(seed, k); }
double wl_uniform01(uint64_t seed, uint64_t k) { return u01(seed, k); }

/* uniform: D+1 draws per point (the last one for the weight). */
void wl_uniform(double *pts, double *w, uint64_t n, int dim, uint64_t seed, int wmax) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint64_t b = (uint64_t)i * (dim + 1);
        for (int k = 0; k < dim; k++) pts[i * dim + k] = u01(seed, b + k);
        w[i] = wmax > 0 ? 1.0 + floor(wmax * u01(seed, b + dim)) : 1.0;
    }
}

/* gauss_mix: 24 draws for the centres, then 8 per point: component,
 * three normals, weight. */
void wl_gauss_mix(double *pts, double *w, uint64_t n, uint64_t seed, double sigma) {
    double ctr[8][3];
    for (int c = 0; c < 8; c++)
        for (int k = 0; k < 3; k++) ctr[c][k] = u01(seed, (uint64_t)(c * 3 + k));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint64_t b = 24 + (uint64_t)i * 8;
        int c = (int)floor(8.0 * u01(seed, b));
        if (c > 7) c = 7;
        for (int k = 0; k < 3; k++) pts[i * 3 + k] = ctr[c][k] + sigma * normal(seed, b + 1 + 2 * k);
        w[i] = 1.0 + floor(100.0 * u01(seed, b + 7));
    }
}

/* jgrid: 5 draws per cell (3 jitters, one normal). */
void wl_jgrid(double *pts, double *w, uint64_t m, uint64_t seed, int jitter) {
    const uint64_t n = m * m * m;
    const double inv = 1.0 / (double)m;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < (int64_t)n; idx++) {
        uint64_t i = (uint64_t)idx % m, j = ((uint64_t)idx / m) % m, k = (uint64_t)idx / (m * m);
        uint64_t b = (uint64_t)idx * 5;
        uint64_t c[3] = {i, j, k};
        for (int a = 0; a < 3; a++) {
            double off = jitter ? (u01(seed, b + a) - 0.5) / 2.0 : 0.0;
            pts[idx * 3 + a] = ((double)c[a] + 0.5 + off) * inv;
        }
        w[idx] = jitter ? round(exp(normal(seed, b + 3)) * 10.0) + 1.0 : 1.0;
    }
}

/* box: D+1 draws per point; coordinates lo + ext * u. */
void wl_box(double *pts, double *w, uint64_t n, int dim, const double *lo, const double *ext, uint64_t seed,
            int wmax) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint64_t b = (uint64_t)i * (dim + 1);
        for (int k = 0; k < dim; k++) pts[i * dim + k] = lo[k] + ext[k] * u01(seed, b + k);
        w[i] = 1.0 + floor(wmax * u01(seed, b + dim));
    }
}

/* batch_header: 9 draws per instance:
 *   n = round(10^(lg_lo + (lg_hi - lg_lo) u0)), D = u1 < 1/2 ? 2 : 3,
 *   lo_k = 200 u_{2+k} - 100, ext_k = 10^(4 u_{5+k} - 2), L = l_lo + floor((l_hi - l_lo + 1) u8).
 * dim2_only / dim3_only force the dimension when non-zero. */
void wl_batch_header(uint64_t count, uint64_t seed, double lg_lo, double lg_hi, int l_lo, int l_hi, int force_dim,
                     uint64_t *n, int *dim, double *lo, double *ext, int *iters) {
    for (uint64_t q = 0; q < count; q++) {
        uint64_t b = q * 9;
        n[q] = (uint64_t)llround(pow(10.0, lg_lo + (lg_hi - lg_lo) * u01(seed, b)));
        int d = u01(seed, b + 1) < 0.5 ? 2 : 3;
        dim[q] = force_dim ? force_dim : d;
        for (int k = 0; k < 3; k++) {
            lo[q * 3 + k] = 200.0 * u01(seed, b + 2 + k) - 100.0;
            ext[q * 3 + k] = pow(10.0, 4.0 * u01(seed, b + 5 + k) - 2.0);
        }
        int l = l_lo + (int)floor((double)(l_hi - l_lo + 1) * u01(seed, b + 8));
        if (l > l_hi) l = l_hi;
        iters[q] = l;
    }
}

/* Per-instance seed of batch instance q (its points use their own stream). */
uint64_t wl_instance_seed(uint64_t seed, uint64_t q) { return sm64(seed ^ 0xA5A5A5A5A5A5A5A5ull, q); }
