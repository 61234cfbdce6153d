// linalg.h -- small dense symmetric eigen-decomposition and Gram
// pseudo-inverse, shared by the host set-up (listener and ghost-cell
// weights) and the device (fresh-cell interpolation of moving objects).
#pragma once

#include <cfloat>
#include <cmath>

#include "geom.h"

namespace fdtd {

// ---- small dense linear algebra (symmetric Jacobi eigen-decomposition) ----
// A (n x n symmetric, row-major) -> eigenvalues lam, eigenvectors V (columns)
FDTD_HD inline void sym_eigen(int n, const double *A_in, double *lam, double *V)
{
    double A[64];
    for (int i = 0; i < n * n; ++i) A[i] = A_in[i];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; ++i) {
            diag += A[i * n + i] * A[i * n + i];
            for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
        }
        if (off <= 1e-34 * diag || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0) continue;
                const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) lam[i] = A[i * n + i];
}

// pseudo-inverse of a symmetric PSD Gram matrix G = B^T B with the rank rule
// sigma > sigma_max * rank_scale * eps (sigma = sqrt(lambda)); returns cond.
FDTD_HD inline double gram_pinv(int n, const double *G, double rank_scale, double *Gp)
{
    double lam[8], V[64];
    sym_eigen(n, G, lam, V);
    double lmax = 0, lmin = INFINITY;
    for (int i = 0; i < n; ++i) { lmax = fmax(lmax, lam[i]); lmin = fmin(lmin, lam[i]); }
    const double smax = sqrt(fmax(lmax, 0.0));
    double thr = smax * rank_scale * DBL_EPSILON;
    thr = fmax(thr * thr, DBL_MIN);
    for (int i = 0; i < n * n; ++i) Gp[i] = 0;
    for (int k = 0; k < n; ++k) {
        if (!(lam[k] > thr)) continue;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Gp[i * n + j] += V[i * n + k] * V[j * n + k] / lam[k];
    }
    return lmin > 0 ? sqrt(lmax / lmin) : INFINITY;
}

// Order-1 MLS weights at p over m <= 8 points (MLSModeInterpolator::lookup
// with Fleishman weights w = exp(-d^2 / (2 r^2))): the weighted design rows
// [w, w dx/h, w dy/h, w dz/h] with d = point - p, the Gram pseudo-inverse
// with the rank rule of gram_pinv, and c0 of the min-norm solution as
// sum_i beta_i val_i.
FDTD_HD inline void mls_weights(int m, const double *pts, const double *p, double h, double radius, double *beta)
{
    double A[32], w[8], G[16], Gp[16];
    const double r2 = radius * radius;
    for (int i = 0; i < m; ++i) {
        const double dx = pts[3 * i] - p[0], dy = pts[3 * i + 1] - p[1], dz = pts[3 * i + 2] - p[2];
        w[i] = exp(-(dx * dx + dy * dy + dz * dz) / (2. * r2));
        A[4 * i + 0] = w[i]; A[4 * i + 1] = w[i] * dx / h; A[4 * i + 2] = w[i] * dy / h; A[4 * i + 3] = w[i] * dz / h;
    }
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            double s = 0;
            for (int i = 0; i < m; ++i) s += A[4 * i + a] * A[4 * i + b];
            G[4 * a + b] = s;
        }
    // row 0 of G^+: by Cholesky when every pivot keeps more than 1e-10 of its
    // diagonal (G is then well inside full rank and G^+ = G^-1), else from
    // the eigen-decomposition with the rank rule
    double z[4], L[16];
    bool chol = m >= 4;
    for (int j = 0; j < 4 && chol; ++j) {
        double d = G[5 * j];
        for (int q = 0; q < j; ++q) d -= L[4 * j + q] * L[4 * j + q];
        if (!(d > 1e-10 * G[5 * j])) { chol = false; break; }
        L[5 * j] = sqrt(d);
        for (int i = j + 1; i < 4; ++i) {
            double v = G[4 * i + j];
            for (int q = 0; q < j; ++q) v -= L[4 * i + q] * L[4 * j + q];
            L[4 * i + j] = v / L[5 * j];
        }
    }
    if (chol) {
        double y[4];   // L y = e0, then L^T z = y
        for (int i = 0; i < 4; ++i) {
            double v = i == 0 ? 1.0 : 0.0;
            for (int q = 0; q < i; ++q) v -= L[4 * i + q] * y[q];
            y[i] = v / L[5 * i];
        }
        for (int i = 3; i >= 0; --i) {
            double v = y[i];
            for (int q = i + 1; q < 4; ++q) v -= L[4 * q + i] * z[q];
            z[i] = v / L[5 * i];
        }
    } else {
        gram_pinv(4, G, (double)(m < 4 ? m : 4), Gp);
        for (int b = 0; b < 4; ++b) z[b] = Gp[b];
    }
    for (int i = 0; i < m; ++i) {
        double s = 0;
        for (int b = 0; b < 4; ++b) s += z[b] * A[4 * i + b];
        beta[i] = s * w[i];
    }
}

}  // namespace fdtd
