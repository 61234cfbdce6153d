// linalg.h -- small dense symmetric eigen-decomposition and Gram
// pseudo-inverse, shared by the host set-up (listener and ghost-cell
// weights) and the device (fresh-cell interpolation of moving objects).
#pragma once

#include <cfloat>
#include <cmath>

#include "geom.h"

namespace fdtd {

// ---- small dense linear algebra (symmetric Jacobi eigen-decomposition) ----
// The sizes are compile-time constants so that, in device code, every array
// index is static after unrolling and the arrays live in registers rather
// than in scratch memory.
// A (N x N symmetric, row-major) -> eigenvalues lam, eigenvectors V (columns)
template <int N>
FDTD_HD inline void sym_eigen(const double *A_in, double *lam, double *V)
{
    double A[N * N];
#pragma unroll
    for (int i = 0; i < N * N; ++i) A[i] = A_in[i];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) V[i * N + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, diag = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            diag += A[i * N + i] * A[i * N + i];
#pragma unroll
            for (int j = i + 1; j < N; ++j) off += A[i * N + j] * A[i * N + j];
        }
        if (off <= 1e-34 * diag || off == 0) break;
#pragma unroll
        for (int p = 0; p < N - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < N; ++q) {
                const double apq = A[p * N + q];
                if (apq == 0) continue;
                const double theta = (A[q * N + q] - A[p * N + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double akp = A[k * N + p], akq = A[k * N + q];
                    A[k * N + p] = c * akp - s * akq;
                    A[k * N + q] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double apk = A[p * N + k], aqk = A[q * N + k];
                    A[p * N + k] = c * apk - s * aqk;
                    A[q * N + k] = s * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double vkp = V[k * N + p], vkq = V[k * N + q];
                    V[k * N + p] = c * vkp - s * vkq;
                    V[k * N + q] = s * vkp + c * vkq;
                }
            }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) lam[i] = A[i * N + i];
}

// pseudo-inverse of a symmetric PSD Gram matrix G = B^T B (N x N) with the
// rank rule sigma > sigma_max * rank_scale * eps (sigma = sqrt(lambda));
// returns cond.
template <int N>
FDTD_HD inline double gram_pinv(const double *G, double rank_scale, double *Gp)
{
    double lam[N], V[N * N];
    sym_eigen<N>(G, lam, V);
    double lmax = 0, lmin = INFINITY;
#pragma unroll
    for (int i = 0; i < N; ++i) { lmax = fmax(lmax, lam[i]); lmin = fmin(lmin, lam[i]); }
    const double smax = sqrt(fmax(lmax, 0.0));
    double thr = smax * rank_scale * DBL_EPSILON;
    thr = fmax(thr * thr, DBL_MIN);
#pragma unroll
    for (int i = 0; i < N * N; ++i) Gp[i] = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (!(lam[k] > thr)) continue;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) Gp[i * N + j] += V[i * N + k] * V[j * N + k] / lam[k];
    }
    return lmin > 0 ? sqrt(lmax / lmin) : INFINITY;
}

// Order-1 MLS interpolation at p (MLSModeInterpolator::lookup with
// Fleishman weights w = exp(-d^2 / (2 r^2))) of two value sets at once, as
// running sums: each point adds its weighted design row A_i = [w, w dx/h,
// w dy/h, w dz/h] (d = point - p) to the Gram matrix G = sum A_i A_i^T and
// to the moments b = sum w_i A_i v_i of each value set.  With z the first
// row of G^+ (Cholesky when every pivot keeps more than 1e-10 of its
// diagonal, so G is well inside full rank and G^+ = G^-1; else the
// eigen-decomposition with the rank rule of gram_pinv), the interpolated
// value c0 of the min-norm solution is sum_i beta_i v_i with
// beta_i = w_i (z . A_i), that is z . b.  Nothing per point outlives its add,
// so on the device the sums stay in registers.
struct MlsSums {
    double G[16] = {0}, bl[4] = {0}, bc[4] = {0};
    int m = 0;

    FDTD_HD void add(const double *pt, const double *p, double h, double radius, double vl, double vc)
    {
        const double dx = pt[0] - p[0], dy = pt[1] - p[1], dz = pt[2] - p[2];
        const double w = exp(-(dx * dx + dy * dy + dz * dz) / (2. * radius * radius));
        const double A[4] = {w, w * dx / h, w * dy / h, w * dz / h};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int b = 0; b < 4; ++b) G[4 * a + b] += A[a] * A[b];
            bl[a] += w * A[a] * vl;
            bc[a] += w * A[a] * vc;
        }
        ++m;
    }

    FDTD_HD void solve(double *vl, double *vc) const
    {
        double z[4], L[16];
        bool chol = m >= 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!chol) break;
            double d = G[5 * j];
#pragma unroll
            for (int q = 0; q < j; ++q) d -= L[4 * j + q] * L[4 * j + q];
            if (!(d > 1e-10 * G[5 * j])) { chol = false; break; }
            L[5 * j] = sqrt(d);
#pragma unroll
            for (int i = j + 1; i < 4; ++i) {
                double v = G[4 * i + j];
#pragma unroll
                for (int q = 0; q < j; ++q) v -= L[4 * i + q] * L[4 * j + q];
                L[4 * i + j] = v / L[5 * j];
            }
        }
        if (chol) {
            double y[4];   // L y = e0, then L^T z = y
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double v = i == 0 ? 1.0 : 0.0;
#pragma unroll
                for (int q = 0; q < i; ++q) v -= L[4 * i + q] * y[q];
                y[i] = v / L[5 * i];
            }
#pragma unroll
            for (int i = 3; i >= 0; --i) {
                double v = y[i];
#pragma unroll
                for (int q = i + 1; q < 4; ++q) v -= L[4 * q + i] * z[q];
                z[i] = v / L[5 * i];
            }
        } else {
            double Gp[16];
            gram_pinv<4>(G, (double)(m < 4 ? m : 4), Gp);
#pragma unroll
            for (int b = 0; b < 4; ++b) z[b] = Gp[b];
        }
        double sl = 0.0, sc = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) { sl += z[b] * bl[b]; sc += z[b] * bc[b]; }
        *vl = sl;
        *vc = sc;
    }
};

}  // namespace fdtd
