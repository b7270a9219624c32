/*
 * rib_oracle.c -- CPU restatement of coupe's Recursive Inertial Bisection.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker of coupe_hip_rib_device and
 * coupe_rib); the product library never links or calls it.
 *
 *   rib()                     coupe/src/algorithms/recursive_bisection.rs:829-851
 *   OrientedBoundingBox::from_points  coupe/src/geometry.rs:217-231
 *   inertia_matrix()          coupe/src/geometry.rs:256-267
 *   inertia_vector()          coupe/src/geometry.rs:269-288
 *   householder_reflection()  coupe/src/geometry.rs:290-305
 *   Matrix::try_inverse       nalgebra 0.34.1 (Cargo.lock), closed-form
 *                             adjugate / determinant for 2x2 and 3x3
 *   Matrix * Vector           nalgebra 0.34.1, columns accumulated left to right
 *
 * The eigen-decomposition of the reference is nalgebra's symmetric_eigen
 * (an iterative QR method).  The oracle uses the closed-form eigenvalues of
 * symmetric 2x2 / 3x3 matrices (trigonometric solution of the characteristic
 * cubic, O. K. Smith, CACM 4(4):168, 1961) and takes the eigenvector from
 * cross products of the rows of M - lambda I: an independent method, so a
 * mistake in either this or the product's Jacobi solver shows as a
 * disagreement.  Its last bits differ from nalgebra's; ids of Rib are
 * therefore pinned given the rotation (see oracle/oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int oracle_rcb(uint64_t *part, size_t n_part, size_t dim, const double *points, size_t n_points,
               int kind, const void *weights, size_t n_weights, int weights_const,
               size_t iter_count, double tol, void *st);

/* inertia_matrix: centroid = (sum p) / n, M = sum (p - c)(p - c)^T, both
 * summed in input order. */
void oracle_inertia(const double *p, size_t n, size_t dim, double *centroid, double *m) {
    double c[3] = {0, 0, 0};
    for (size_t i = 0; i < n; i++)
        for (size_t a = 0; a < dim; a++) c[a] += p[i * dim + a];
    for (size_t a = 0; a < dim; a++) c[a] = c[a] / (double)n;
    double s[9] = {0};
    for (size_t i = 0; i < n; i++) {
        double o[3];
        for (size_t a = 0; a < dim; a++) o[a] = p[i * dim + a] - c[a];
        for (size_t a = 0; a < dim; a++)
            for (size_t b = 0; b < dim; b++) s[a * dim + b] += o[a] * o[b];
    }
    memcpy(centroid, c, dim * sizeof(double));
    memcpy(m, s, dim * dim * sizeof(double));
}

static void cross(const double *a, const double *b, double *r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot(const double *a, const double *b, size_t d) {
    double s = 0.0;
    for (size_t i = 0; i < d; i++) s += a[i] * b[i];
    return s;
}

/* Unit eigenvector of the largest eigenvalue of a symmetric matrix; for a
 * diagonal matrix, the first axis among equal largest entries (nalgebra
 * leaves a diagonal matrix unrotated and inertia_vector's sort keeps the
 * first of equal eigenvalues).  Returns the eigenvalue. */
double oracle_inertia_vector(const double *m, size_t dim, double *v) {
    if (dim == 2) {
        const double a = m[0], b = m[1], d = m[3];
        if (b == 0.0) {
            const int k = a >= d ? 0 : 1;
            v[0] = k == 0;
            v[1] = k == 1;
            return k == 0 ? a : d;
        }
        const double h = 0.5 * (a - d);
        const double lam = 0.5 * (a + d) + sqrt(h * h + b * b);
        double u1[2] = {lam - d, b}, u2[2] = {b, lam - a};
        double *u = dot(u1, u1, 2) >= dot(u2, u2, 2) ? u1 : u2;
        const double nn = sqrt(dot(u, u, 2));
        v[0] = u[0] / nn;
        v[1] = u[1] / nn;
        return lam;
    }
    const double p1 = m[1] * m[1] + m[2] * m[2] + m[5] * m[5];
    if (p1 == 0.0) {
        int k = 0;
        for (int i = 1; i < 3; i++)
            if (m[i * 4] > m[k * 4]) k = i;
        for (int i = 0; i < 3; i++) v[i] = i == k;
        return m[k * 4];
    }
    const double q = (m[0] + m[4] + m[8]) / 3.0;
    const double p2 = (m[0] - q) * (m[0] - q) + (m[4] - q) * (m[4] - q) + (m[8] - q) * (m[8] - q) + 2.0 * p1;
    const double p = sqrt(p2 / 6.0);
    double B[9];
    for (int i = 0; i < 9; i++) B[i] = (m[i] - (i % 4 == 0 ? q : 0.0)) / p;
    double r = (B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                B[2] * (B[3] * B[7] - B[4] * B[6])) / 2.0;
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    const double phi = acos(r) / 3.0;
    const double lam = q + 2.0 * p * cos(phi);
    double rows[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) rows[i][j] = m[i * 3 + j] - (i == j ? lam : 0.0);
    double best[3] = {0, 0, 0}, bn = -1.0;
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) {
            double c[3];
            cross(rows[i], rows[j], c);
            const double cn = dot(c, c, 3);
            if (cn > bn) {
                bn = cn;
                memcpy(best, c, sizeof best);
            }
        }
    const double nn = sqrt(bn);
    for (int i = 0; i < 3; i++) v[i] = best[i] / nn;
    return lam;
}

/* approx 0.5.1 `Ulps::default().eq` for one f64: |a-b| <= f64::EPSILON, or
 * same sign and at most 4 units in the last place apart. */
static int ulps_eq(double a, double b) {
    if (fabs(a - b) <= 2.220446049250313e-16) return 1;
    if (signbit(a) != signbit(b)) return 0;
    int64_t ia, ib;
    memcpy(&ia, &a, 8);
    memcpy(&ib, &b, 8);
    return (ia > ib ? ia - ib : ib - ia) <= 4;
}

/* householder_reflection (geometry.rs:290-305). */
void oracle_householder(const double *e, size_t dim, double *h) {
    double nrm = sqrt(dot(e, e, dim));
    int same = 1;
    for (size_t i = 0; i < dim; i++) same &= ulps_eq(e[i] / nrm, i == 0 ? 1.0 : 0.0);
    if (same) {
        for (size_t i = 0; i < dim; i++)
            for (size_t j = 0; j < dim; j++) h[i * dim + j] = i == j;
        return;
    }
    const double sign = e[0] > 0.0 ? -1.0 : 1.0;
    double w[3];
    for (size_t i = 0; i < dim; i++) w[i] = e[i] + (sign * (i == 0 ? 1.0 : 0.0)) * nrm;
    const double ww = dot(w, w, dim);
    for (size_t i = 0; i < dim; i++)
        for (size_t j = 0; j < dim; j++) h[i * dim + j] = (i == j ? 1.0 : 0.0) - ((2.0 * w[i]) * w[j]) / ww;
}

/* Matrix::try_inverse for 2x2 / 3x3: adjugate over determinant, expanded
 * along the first row; returns 0 when the determinant is zero. */
int oracle_inverse(const double *m, size_t dim, double *r) {
    if (dim == 2) {
        const double det = m[0] * m[3] - m[2] * m[1];
        if (det == 0.0) return 0;
        r[0] = m[3] / det;
        r[1] = -m[1] / det;
        r[2] = -m[2] / det;
        r[3] = m[0] / det;
        return 1;
    }
    const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], k = m[8];
    const double A = e * k - h * f, B = d * k - g * f, C = d * h - g * e;
    const double det = a * A - b * B + c * C;
    if (det == 0.0) return 0;
    r[0] = A / det;
    r[1] = (c * h - k * b) / det;
    r[2] = (b * f - e * c) / det;
    r[3] = -B / det;
    r[4] = (a * k - g * c) / det;
    r[5] = (c * d - f * a) / det;
    r[6] = C / det;
    r[7] = (b * g - h * a) / det;
    r[8] = (a * e - d * b) / det;
    return 1;
}

/* OrientedBoundingBox::from_points up to the map: fills the centroid,
 * inertia matrix, axis, aabb_to_obb (H) and obb_to_aabb (H^-1).  Returns 0
 * for an empty input, -1 when H is singular (try_inverse().unwrap()). */
int oracle_rib_frame(const double *p, size_t n, size_t dim, double *centroid, double *inertia, double *axis,
                     double *h, double *inv) {
    if (n == 0) return 0;
    oracle_inertia(p, n, dim, centroid, inertia);
    oracle_inertia_vector(inertia, dim, axis);
    oracle_householder(axis, dim, h);
    return oracle_inverse(h, dim, inv) ? 1 : -1;
}

/* obb_to_aabb * p for every point: sum over columns left to right. */
void oracle_map_points(const double *p, size_t n, size_t dim, const double *m, double *out) {
    for (size_t i = 0; i < n; i++)
        for (size_t a = 0; a < dim; a++) {
            double s = m[a * dim] * p[i * dim];
            for (size_t b = 1; b < dim; b++) s = s + m[a * dim + b] * p[i * dim + b];
            out[i * dim + a] = s;
        }
}

/* rib(): RCB of the mapped points.  `obb_to_aabb` (dim x dim, row-major)
 * replaces the oracle's own rotation when not NULL. */
int oracle_rib(uint64_t *part, size_t n_part, size_t dim, const double *points, size_t n_points, int kind,
               const void *weights, size_t n_weights, int weights_const, size_t iter_count, double tol, void *st,
               const double *obb_to_aabb) {
    if (n_points == 0) return 0;   /* OrientedBoundingBox::from_points is None */
    if (dim != 2 && dim != 3) return 3;
    double cen[3], mat[9], axis[3], h[9], inv[9];
    const double *m = obb_to_aabb;
    if (!m) {
        if (oracle_rib_frame(points, n_points, dim, cen, mat, axis, h, inv) < 0) return 2;
        m = inv;
    }
    double *mapped = malloc(n_points * dim * sizeof(double));
    if (!mapped) return 1;
    oracle_map_points(points, n_points, dim, m, mapped);
    int rc = oracle_rcb(part, n_part, dim, mapped, n_points, kind, weights, n_weights, weights_const, iter_count,
                        tol, st);
    free(mapped);
    return rc;
}
