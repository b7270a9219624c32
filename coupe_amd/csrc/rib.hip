// rib.hip -- Recursive Inertial Bisection on device (coupe::Rib,
// coupe/src/algorithms/recursive_bisection.rs:829-930).
//
// Rib = RCB after a change of basis whose first vector is the inertia axis
// of the points (coupe/src/geometry.rs:134-160, :256-319):
//   1. centroid c = (sum of the points) / n                 -- device pass 1
//   2. inertia matrix M = sum (p - c)(p - c)^T              -- device pass 2
//   3. inertia axis v = eigenvector of M's largest eigenvalue      -- host
//   4. H = I - 2 w w^T / (w^T w), w = v + sign(v0) |v| e0 (reflection)
//   5. obb_to_aabb = H^-1; every point is mapped p -> obb_to_aabb p inside
//      the RCB prologue (no mapped copy of the points is written).
// Both reductions run on a fixed grid with fixed-order block and final
// reductions, so the rotation is the same on every run and every device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>

#include "rcb_engine.h"
#include "rcb_device.h"

namespace coupe_hip {

namespace {

constexpr int kMomBlocks = 1024;   // fixed grid: deterministic reduction order
constexpr int kMomTerms = 6;       // centroid: D sums; inertia: D(D+1)/2 products

void ck(hipError_t e) {
    if (e != hipSuccess) throw HipFailure{e};
}

// Block sum of `kMomTerms` per-thread values, in a fixed order: wavefront
// butterflies, then the wavefront results in order.
__device__ void block_sum_terms(double (&v)[kMomTerms], double *out) {
    __shared__ double sh[kBlock / 64][kMomTerms];
#pragma unroll
    for (int k = 0; k < kMomTerms; k++) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    }
    const int wid = threadIdx.x >> 6;
    if (lane_id() == 0)
#pragma unroll
        for (int k = 0; k < kMomTerms; k++) sh[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x < kMomTerms) {
        double s = sh[0][threadIdx.x];
        for (int j = 1; j < kBlock / 64; j++) s += sh[j][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// Pass 1 (STAGE 0): per-block sums of the coordinates.  Pass 2 (STAGE 1):
// per-block sums of the products (p - c)_a (p - c)_b, a <= b, in the order
// (0,0) (0,1) [(0,2)] (1,1) [(1,2) (2,2)].
template <int D, int STAGE>
__global__ void __launch_bounds__(kBlock)
k_moments(const double *__restrict__ pts, uint64_t n, const double *__restrict__ centroid,
          double *__restrict__ partial) {
    double v[kMomTerms] = {0, 0, 0, 0, 0, 0};
    double c[D];
#pragma unroll
    for (int a = 0; a < D; a++) c[a] = STAGE ? centroid[a] : 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        double p[D];
#pragma unroll
        for (int a = 0; a < D; a++) p[a] = pts[i * D + a];
        if constexpr (STAGE == 0) {
#pragma unroll
            for (int a = 0; a < D; a++) v[a] += p[a];
        } else {
            double o[D];
#pragma unroll
            for (int a = 0; a < D; a++) o[a] = p[a] - c[a];
            int t = 0;
#pragma unroll
            for (int a = 0; a < D; a++)
#pragma unroll
                for (int b = a; b < D; b++) v[t++] += o[a] * o[b];
        }
    }
    block_sum_terms(v, partial + (uint64_t)blockIdx.x * kMomTerms);
}

// Sum of the block partials in a fixed order (one workgroup).  STAGE 0
// divides by n (the centroid, geometry.rs:257-258).
template <int STAGE>
__global__ void __launch_bounds__(kBlock)
k_moments_final(const double *__restrict__ partial, uint32_t n_blocks, uint64_t n, double *__restrict__ out) {
    double v[kMomTerms] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < n_blocks; b += kBlock)
#pragma unroll
        for (int k = 0; k < kMomTerms; k++) v[k] += partial[(uint64_t)b * kMomTerms + k];
    __shared__ double res[kMomTerms];
    block_sum_terms(v, res);
    __syncthreads();
    if (threadIdx.x < kMomTerms) out[threadIdx.x] = STAGE == 0 ? res[threadIdx.x] / (double)n : res[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Host linear algebra on D x D (D <= 3) row-major matrices.
// ---------------------------------------------------------------------------

// Eigen-decomposition of a symmetric matrix by cyclic Jacobi rotations
// (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.5.2-8.5.3).
// `vec` holds the eigenvectors as columns.
void sym_eigen(int D, const double *m, double *val, double *vec) {
    double a[3][3], v[3][3];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            a[i][j] = m[i * D + j];
            v[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) {
                tot += a[i][j] * a[i][j];
                if (i != j) off += a[i][j] * a[i][j];
            }
        if (off == 0.0 || off <= 1e-36 * tot) break;
        for (int p = 0; p < D - 1; p++)
            for (int q = p + 1; q < D; q++) {
                if (a[p][q] == 0.0) continue;
                const double tau = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < D; k++) {   // A <- A J
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < D; k++) {   // A <- J^T A
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                a[p][q] = a[q][p] = 0.0;
                for (int k = 0; k < D; k++) {   // V <- V J
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < D; i++) {
        val[i] = a[i][i];
        for (int j = 0; j < D; j++) vec[i * D + j] = v[i][j];
    }
}

// approx's `Ulps::default().eq` on one f64 (epsilon f64::EPSILON, 4 ulps).
bool ulps_eq(double a, double b) {
    if (std::fabs(a - b) <= 2.220446049250313e-16) return true;
    if (std::signbit(a) != std::signbit(b)) return false;
    int64_t ia, ib;
    memcpy(&ia, &a, 8);
    memcpy(&ib, &b, 8);
    const int64_t d = ia > ib ? ia - ib : ib - ia;
    return d <= 4;
}

double norm(int D, const double *x) {
    double s = 0.0;
    for (int i = 0; i < D; i++) s += x[i] * x[i];
    return std::sqrt(s);
}

// householder_reflection (geometry.rs:290-305).
void householder(int D, const double *e, double *h) {
    const double nrm = norm(D, e);
    bool is_e0 = true;
    for (int i = 0; i < D; i++) is_e0 = is_e0 && ulps_eq(e[i] / nrm, i == 0 ? 1.0 : 0.0);
    if (is_e0) {
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) h[i * D + j] = i == j ? 1.0 : 0.0;
        return;
    }
    const double sign = e[0] > 0.0 ? -1.0 : 1.0;
    double w[3];
    for (int i = 0; i < D; i++) w[i] = e[i] + (i == 0 ? sign * nrm : sign * 0.0 * nrm);
    double ww = 0.0;
    for (int i = 0; i < D; i++) ww += w[i] * w[i];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) h[i * D + j] = (i == j ? 1.0 : 0.0) - (2.0 * w[i]) * w[j] / ww;
}

// Inverse by the adjugate (closed form for 2x2 and 3x3); false if singular.
bool inverse(int D, const double *m, double *r) {
    if (D == 2) {
        const double det = m[0] * m[3] - m[2] * m[1];
        if (det == 0.0) return false;
        r[0] = m[3] / det;
        r[1] = -m[1] / det;
        r[2] = -m[2] / det;
        r[3] = m[0] / det;
        return true;
    }
    const double m11 = m[0], m12 = m[1], m13 = m[2], m21 = m[3], m22 = m[4], m23 = m[5], m31 = m[6], m32 = m[7],
                 m33 = m[8];
    const double c11 = m22 * m33 - m32 * m23;
    const double c21 = m21 * m33 - m31 * m23;
    const double c31 = m21 * m32 - m31 * m22;
    const double det = m11 * c11 - m12 * c21 + m13 * c31;
    if (det == 0.0) return false;
    r[0] = c11 / det;
    r[1] = (m13 * m32 - m33 * m12) / det;
    r[2] = (m12 * m23 - m22 * m13) / det;
    r[3] = -c21 / det;
    r[4] = (m11 * m33 - m31 * m13) / det;
    r[5] = (m13 * m21 - m23 * m11) / det;
    r[6] = c31 / det;
    r[7] = (m12 * m31 - m32 * m11) / det;
    r[8] = (m11 * m22 - m21 * m12) / det;
    return true;
}

template <int D>
void moments(Context &cx, hipStream_t st, const double *pts, uint64_t n, double *h_centroid, double *h_inertia) {
    cx.rib_moments.ensure((kMomBlocks + 2) * kMomTerms * sizeof(double));
    double *partial = (double *)cx.rib_moments.ptr;
    double *d_c = partial + kMomBlocks * kMomTerms;
    double *d_m = d_c + kMomTerms;
    const unsigned g = (unsigned)std::min<uint64_t>(kMomBlocks, (n + kBlock - 1) / kBlock);
    const double bytes = (double)n * 8.0 * D;
    // one record per reading pass; the one-workgroup final sums go with it
    int r = cx.prof.begin(st, COUPE_HIP_K_MOMENTS, bytes);
    k_moments<D, 0><<<g, kBlock, 0, st>>>(pts, n, nullptr, partial);
    k_moments_final<0><<<1, kBlock, 0, st>>>(partial, g, n, d_c);
    ck(hipGetLastError());
    cx.prof.end(st, r);
    r = cx.prof.begin(st, COUPE_HIP_K_MOMENTS, bytes);
    k_moments<D, 1><<<g, kBlock, 0, st>>>(pts, n, d_c, partial);
    k_moments_final<1><<<1, kBlock, 0, st>>>(partial, g, n, d_m);
    ck(hipGetLastError());
    cx.prof.end(st, r);
    double h[2 * kMomTerms];
    ck(hipMemcpyAsync(h, d_c, sizeof h, hipMemcpyDeviceToHost, st));
    ck(hipStreamSynchronize(st));
    for (int a = 0; a < D; a++) h_centroid[a] = h[a];
    int t = 0;
    for (int a = 0; a < D; a++)
        for (int b = a; b < D; b++) {
            h_inertia[a * D + b] = h_inertia[b * D + a] = h[kMomTerms + t];
            t++;
        }
}

}  // namespace

// coupe::Rib on device-resident inputs: the rotation, then the RCB of the
// rotated points with the rotation fused into the prologue.
enum coupe_err rib_device(Context &cx, hipStream_t st, uint64_t *d_out, uint64_t n, size_t dim,
                          const double *d_points, int wkind, const void *d_weights, size_t iter_count, double tol) {
    if (dim != 2 && dim != 3) return COUPE_ERR_BAD_DIMENSION;
    if (wkind != COUPE_INT && wkind != COUPE_INT64 && wkind != COUPE_DOUBLE) return COUPE_ERR_BAD_TYPE;
    memset(&cx.rib, 0, sizeof cx.rib);
    cx.rib.dimension = dim;
    if (n == 0) return COUPE_ERR_OK;   // OrientedBoundingBox::from_points is None (:846-849)
    const int D = (int)dim;
    try {
        ck(hipSetDevice(cx.device));
        cx.prof.reset();
        double cen[3] = {0, 0, 0}, mat[9] = {0}, val[3], vec[9], h[9], inv[9];
        if (D == 2) moments<2>(cx, st, d_points, n, cen, mat);
        else moments<3>(cx, st, d_points, n, cen, mat);
        sym_eigen(D, mat, val, vec);
        // inertia_vector (geometry.rs:269-288): the eigenvector of the
        // largest eigenvalue, the first one among equals.
        int best = 0;
        for (int i = 1; i < D; i++)
            if (val[i] > val[best]) best = i;
        double axis[3];
        for (int i = 0; i < D; i++) axis[i] = vec[i * D + best];
        householder(D, axis, h);
        if (!inverse(D, h, inv)) return COUPE_ERR_CRASH;   // try_inverse().unwrap() panics
        for (int i = 0; i < D; i++) {
            cx.rib.centroid[i] = cen[i];
            cx.rib.axis[i] = axis[i];
            for (int j = 0; j < D; j++) {
                cx.rib.inertia[i * D + j] = mat[i * D + j];
                cx.rib.aabb_to_obb[i * D + j] = h[i * D + j];
                cx.rib.obb_to_aabb[i * D + j] = inv[i * D + j];
            }
        }
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        return map_hip_error(f.err);
    } catch (const std::bad_alloc &) {
        return COUPE_ERR_ALLOC;
    }
    Rotation rot;
    memset(&rot, 0, sizeof rot);
    for (int i = 0; i < D * D; i++) rot.m[i] = cx.rib.obb_to_aabb[i];
    return rcb_batch(cx, st, d_out, dim, d_points, wkind, d_weights, &n, &iter_count, &tol, 1, &rot);
}

}  // namespace coupe_hip
