/*
 * fdtd_oracle.c -- FP64 CPU restatement of the reference's collocated
 * time-stepping path.  TEST INFRASTRUCTURE ONLY (the checker, never the
 * product); see fdtd_oracle.h for the list of reference lines it follows.
 *
 * Parity status of each part is recorded beside the function; every part is
 * pinned by a test in tests/test_oracle_*.py (closed forms, invariants, the
 * reference's own known-answer inputs) unless it says "parity unpinned".
 */
#define _GNU_SOURCE
#include "fdtd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[512];
static void set_err(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); }
const char *oracle_error(void) { return g_err; }

/* Wavesolver_ConstantsAndTypes.h:72-98, config.h:41-46 */
#define SMALL_PML_BUFFER 1e-12        /* InsidePML detection buffer, MAC_Grid.cpp:4031 */
#define GAUSSIAN_CHECK_BOUND 2.0
#define CONDN_LIM 25.0
#define REF_EPS 1e-10                 /* EPS in config.h */
#define MLS_RADIUS 0.018              /* MLSConstantsAndTypes.h:8 */

/* ======================================================================== */
/* small vector helpers                                                      */
/* ======================================================================== */
static inline double dot3(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }

/* ======================================================================== */
/* dense linear algebra: one-sided Jacobi SVD (Hestenes), pseudo-inverse     */
/* solve with Eigen 3.3 JacobiSVD's rank threshold                           */
/* ======================================================================== */
/* A is m x n row-major, m >= n.  On return U (m x n, row-major, columns are
 * left singular vectors scaled out), sv[n] (unsorted), V (n x n row-major). */
static void jacobi_svd(int m, int n, const double *A, double *U, double *sv, double *V)
{
    memcpy(U, A, sizeof(double) * m * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i*n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < m; ++i) {
                    const double up = U[i*n + p], uq = U[i*n + q];
                    alpha += up * up; beta += uq * uq; gamma += up * uq;
                }
                if (gamma == 0.0) continue;
                const double conv = fabs(gamma) / sqrt(alpha * beta);
                if (!(conv > 1e-300)) continue;
                if (conv > off) off = conv;
                if (conv < 1e-17) continue;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int i = 0; i < m; ++i) {
                    const double up = U[i*n + p], uq = U[i*n + q];
                    U[i*n + p] = c * up - s * uq;
                    U[i*n + q] = s * up + c * uq;
                }
                for (int i = 0; i < n; ++i) {
                    const double vp = V[i*n + p], vq = V[i*n + q];
                    V[i*n + p] = c * vp - s * vq;
                    V[i*n + q] = s * vp + c * vq;
                }
            }
        if (off < 1e-16) break;
    }
    for (int j = 0; j < n; ++j) {
        double s2 = 0;
        for (int i = 0; i < m; ++i) s2 += U[i*n + j] * U[i*n + j];
        sv[j] = sqrt(s2);
        if (sv[j] > 0)
            for (int i = 0; i < m; ++i) U[i*n + j] /= sv[j];
    }
}

int oracle_svd_values(int m, int n, const double *A, double *sv)
{
    if (m < n || n <= 0 || n > 16 || m > 256) return -1;
    double U[256 * 16], V[16 * 16], s[16];
    jacobi_svd(m, n, A, U, s, V);
    /* sort decreasing */
    for (int i = 0; i < n; ++i) sv[i] = s[i];
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (sv[j] > sv[i]) { double t = sv[i]; sv[i] = sv[j]; sv[j] = t; }
    return 0;
}

/* x = pinv(A) b with Eigen's threshold: rank counts sv > max(sv) * min(m,n) * eps
 * (Eigen/src/SVD/SVDBase.h threshold(), JacobiSVD::_solve_impl). */
static int lsq_solve_internal(int m, int n, const double *A, const double *b, double *x, double *cond_out)
{
    double U[256 * 16], V[16 * 16], s[16];
    jacobi_svd(m, n, A, U, s, V);
    double smax = 0, smin = INFINITY;
    for (int j = 0; j < n; ++j) { if (s[j] > smax) smax = s[j]; if (s[j] < smin) smin = s[j]; }
    if (cond_out) *cond_out = smax / smin;
    double thr = smax * (double)(m < n ? m : n) * DBL_EPSILON;
    if (thr < DBL_MIN) thr = DBL_MIN;
    int rank = 0;
    double tmp[16];
    for (int j = 0; j < n; ++j) {
        if (s[j] > thr) {
            double acc = 0;
            for (int i = 0; i < m; ++i) acc += U[i*n + j] * b[i];
            tmp[j] = acc / s[j];
            ++rank;
        } else tmp[j] = 0.0;
    }
    for (int i = 0; i < n; ++i) {
        double acc = 0;
        for (int j = 0; j < n; ++j) acc += V[i*n + j] * tmp[j];
        x[i] = acc;
    }
    return rank;
}

int oracle_lsq_solve(int m, int n, const double *A, const double *b, double *x)
{
    if (m < n || n <= 0 || n > 16 || m > 256) return -1;
    return lsq_solve_internal(m, n, A, b, x, NULL);
}

/* ======================================================================== */
/* MLS lookup, order 1, Fleishman weights                                    */
/* MLSModeInterpolator.hpp:282-429 (lookup), 588-597 (distanceWeight),       */
/* 712-738 (initializeWeights); PolyGen<3> rows [1, dx, dy, dz] with         */
/* d = point - p.  Solve: JacobiSVD(A) min-norm, output c(0).               */
/* ======================================================================== */
double oracle_mls_lookup(const double *p, int npts, const double *pts, const double *vals)
{
    double A[256 * 4], b[256], c[4];
    if (npts > 256) npts = 256;
    const double r2 = MLS_RADIUS * MLS_RADIUS;
    for (int i = 0; i < npts; ++i) {
        const double dx = pts[3*i+0] - p[0], dy = pts[3*i+1] - p[1], dz = pts[3*i+2] - p[2];
        const double d2 = (p[0]-pts[3*i+0])*(p[0]-pts[3*i+0]) + (p[1]-pts[3*i+1])*(p[1]-pts[3*i+1])
                        + (p[2]-pts[3*i+2])*(p[2]-pts[3*i+2]);
        const double w = exp(-(d2) / (2. * r2));
        A[i*4 + 0] = 1.0 * w; A[i*4 + 1] = dx * w; A[i*4 + 2] = dy * w; A[i*4 + 3] = dz * w;
        b[i] = vals[i] * w;
    }
    lsq_solve_internal(npts, 4, A, b, c, NULL);
    return c[0];
}

/* ======================================================================== */
/* analytic geometry (replaces FDTD_RigidObject's mesh + ADF + kd-tree)      */
/* ======================================================================== */
static void to_object(const fdtd_object *o, const double *x, double *y)
{
    const double d[3] = {x[0]-o->center[0], x[1]-o->center[1], x[2]-o->center[2]};
    /* y = R^T d */
    for (int i = 0; i < 3; ++i) y[i] = o->rotation[0*3+i]*d[0] + o->rotation[1*3+i]*d[1] + o->rotation[2*3+i]*d[2];
}
static void to_world_vec(const fdtd_object *o, const double *y, double *x)
{
    for (int i = 0; i < 3; ++i) x[i] = o->rotation[i*3+0]*y[0] + o->rotation[i*3+1]*y[1] + o->rotation[i*3+2]*y[2];
}

/* Eberly, "Distance from a Point to an Ellipse, an Ellipsoid, or a
 * Hyperellipsoid" (2013): bisection on the largest root of
 * F(s) = sum (r_i z_i / (s + r_i))^2 - 1, z_i = y_i / e_i, r_i = (e_i/e_min)^2,
 * first-octant point, axes sorted decreasing. */
static double robust_len2(double a, double b)
{
    const double m = fmax(fabs(a), fabs(b));
    if (m == 0) return 0;
    const double x = a / m, y = b / m;
    return m * sqrt(x*x + y*y);
}
static double robust_len3(double a, double b, double c)
{
    const double m = fmax(fabs(a), fmax(fabs(b), fabs(c)));
    if (m == 0) return 0;
    const double x = a / m, y = b / m, z = c / m;
    return m * sqrt(x*x + y*y + z*z);
}
static double get_root2(double r0, double z0, double z1, double g)
{
    const double n0 = r0 * z0;
    double s0 = z1 - 1.0, s1 = (g < 0 ? 0.0 : robust_len2(n0, z1) - 1.0), s = 0;
    for (int it = 0; it < 2200; ++it) {
        s = 0.5 * (s0 + s1);
        if (s == s0 || s == s1) break;
        const double a = n0 / (s + r0), b = z1 / (s + 1.0);
        g = a*a + b*b - 1.0;
        if (g > 0) s0 = s; else if (g < 0) s1 = s; else break;
    }
    return s;
}
static double get_root3(double r0, double r1, double z0, double z1, double z2, double g)
{
    const double n0 = r0 * z0, n1 = r1 * z1;
    double s0 = z2 - 1.0, s1 = (g < 0 ? 0.0 : robust_len3(n0, n1, z2) - 1.0), s = 0;
    for (int it = 0; it < 2200; ++it) {
        s = 0.5 * (s0 + s1);
        if (s == s0 || s == s1) break;
        const double a = n0 / (s + r0), b = n1 / (s + r1), c = z2 / (s + 1.0);
        g = a*a + b*b + c*c - 1.0;
        if (g > 0) s0 = s; else if (g < 0) s1 = s; else break;
    }
    return s;
}
/* 2D: e0 >= e1 > 0, y0, y1 >= 0 */
static void ellipse_closest(double e0, double e1, double y0, double y1, double *x0, double *x1)
{
    if (y1 > 0) {
        if (y0 > 0) {
            const double z0 = y0 / e0, z1 = y1 / e1, g = z0*z0 + z1*z1 - 1.0;
            if (g != 0) {
                const double r0 = (e0 / e1) * (e0 / e1);
                const double sb = get_root2(r0, z0, z1, g);
                *x0 = r0 * y0 / (sb + r0); *x1 = y1 / (sb + 1.0);
            } else { *x0 = y0; *x1 = y1; }
        } else { *x0 = 0; *x1 = e1; }
    } else {
        const double numer0 = e0 * y0, denom0 = e0*e0 - e1*e1;
        if (numer0 < denom0) {
            const double xde0 = numer0 / denom0;
            *x0 = e0 * xde0; *x1 = e1 * sqrt(fmax(0.0, 1.0 - xde0*xde0));
        } else { *x0 = e0; *x1 = 0; }
    }
}
/* 3D: e0 >= e1 >= e2 > 0, y >= 0 */
static void ellipsoid_closest_sorted(const double *e, const double *y, double *x)
{
    if (y[2] > 0) {
        if (y[1] > 0) {
            if (y[0] > 0) {
                const double z0 = y[0]/e[0], z1 = y[1]/e[1], z2 = y[2]/e[2];
                const double g = z0*z0 + z1*z1 + z2*z2 - 1.0;
                if (g != 0) {
                    const double r0 = (e[0]/e[2])*(e[0]/e[2]), r1 = (e[1]/e[2])*(e[1]/e[2]);
                    const double sb = get_root3(r0, r1, z0, z1, z2, g);
                    x[0] = r0 * y[0] / (sb + r0); x[1] = r1 * y[1] / (sb + r1); x[2] = y[2] / (sb + 1.0);
                } else { x[0] = y[0]; x[1] = y[1]; x[2] = y[2]; }
            } else { x[0] = 0; ellipse_closest(e[1], e[2], y[1], y[2], &x[1], &x[2]); }
        } else {
            if (y[0] > 0) { x[1] = 0; ellipse_closest(e[0], e[2], y[0], y[2], &x[0], &x[2]); }
            else { x[0] = 0; x[1] = 0; x[2] = e[2]; }
        }
    } else {
        const double d0 = e[0]*e[0] - e[2]*e[2], d1 = e[1]*e[1] - e[2]*e[2];
        const double n0 = e[0]*y[0], n1 = e[1]*y[1];
        int done = 0;
        if (n0 < d0 && n1 < d1) {
            const double a = n0 / d0, b = n1 / d1, disc = 1.0 - a*a - b*b;
            if (disc > 0) { x[0] = e[0]*a; x[1] = e[1]*b; x[2] = e[2]*sqrt(disc); done = 1; }
        }
        if (!done) { x[2] = 0; ellipse_closest(e[0], e[1], y[0], y[1], &x[0], &x[1]); }
    }
}

/* Closest surface point xb, outward unit normal n at xb, signed distance
 * (negative inside); the semantics of FDTD_RigidObject::DistanceToMesh and
 * NormalToMesh (FDTD_RigidObject.cpp:256-327) for analytic shapes. */
double oracle_object_closest(const fdtd_object *obj, const double *x, double *xb, double *n)
{
    if (obj->type == FDTD_OBJ_SPHERE) {
        const double R = obj->semi_axes[0];
        double d[3] = {x[0]-obj->center[0], x[1]-obj->center[1], x[2]-obj->center[2]};
        double r = norm3(d);
        double u[3];
        if (r > 0) { u[0] = d[0]/r; u[1] = d[1]/r; u[2] = d[2]/r; }
        else { u[0] = 1; u[1] = 0; u[2] = 0; }
        for (int i = 0; i < 3; ++i) { xb[i] = obj->center[i] + R*u[i]; n[i] = u[i]; }
        return r - R;
    }
    double y[3];
    to_object(obj, x, y);
    /* reflect into the first octant and sort the axes decreasing */
    double sgn[3], ya[3];
    for (int i = 0; i < 3; ++i) { sgn[i] = (y[i] < 0) ? -1.0 : 1.0; ya[i] = fabs(y[i]); }
    int perm[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (obj->semi_axes[perm[j]] > obj->semi_axes[perm[i]]) { int t = perm[i]; perm[i] = perm[j]; perm[j] = t; }
    double es[3], ys[3], xs[3];
    for (int i = 0; i < 3; ++i) { es[i] = obj->semi_axes[perm[i]]; ys[i] = ya[perm[i]]; }
    ellipsoid_closest_sorted(es, ys, xs);
    double xo[3];
    for (int i = 0; i < 3; ++i) xo[perm[i]] = xs[i] * sgn[perm[i]];
    const double f = (y[0]/obj->semi_axes[0])*(y[0]/obj->semi_axes[0]) + (y[1]/obj->semi_axes[1])*(y[1]/obj->semi_axes[1])
                   + (y[2]/obj->semi_axes[2])*(y[2]/obj->semi_axes[2]) - 1.0;
    const double dd[3] = {xo[0]-y[0], xo[1]-y[1], xo[2]-y[2]};
    double dist = norm3(dd);
    if (f < 0) dist = -dist;
    double no[3] = {xo[0]/(obj->semi_axes[0]*obj->semi_axes[0]), xo[1]/(obj->semi_axes[1]*obj->semi_axes[1]),
                    xo[2]/(obj->semi_axes[2]*obj->semi_axes[2])};
    const double nn = norm3(no);
    for (int i = 0; i < 3; ++i) no[i] /= nn;
    double xw[3];
    to_world_vec(obj, xo, xw);
    for (int i = 0; i < 3; ++i) xb[i] = xw[i] + obj->center[i];
    to_world_vec(obj, no, n);
    return dist;
}

/* inside test: distance < DISTANCE_TOLERANCE (= 0), FDTD_Objects.cpp:35-49 */
static int object_occupies(const fdtd_object *obj, const double *x)
{
    if (obj->type == FDTD_OBJ_SPHERE) {
        const double d[3] = {x[0]-obj->center[0], x[1]-obj->center[1], x[2]-obj->center[2]};
        return sqrt(dot3(d, d)) - obj->semi_axes[0] < 0.0;
    }
    double y[3];
    to_object(obj, x, y);
    const double f = (y[0]/obj->semi_axes[0])*(y[0]/obj->semi_axes[0]) + (y[1]/obj->semi_axes[1])*(y[1]/obj->semi_axes[1])
                   + (y[2]/obj->semi_axes[2])*(y[2]/obj->semi_axes[2]) - 1.0;
    return f < 0.0;
}

static double object_distance(const fdtd_object *obj, const double *x)
{
    double xb[3], n[3];
    return oracle_object_closest(obj, x, xb, n);
}

/* FDTD_RigidObject::ReflectAgainstBoundary (FDTD_RigidObject.cpp:429-536):
 * inside: erected = xb - x, image = xb + erected; outside: erected = x - xb,
 * image = x + erected; on the surface: erected = surface normal, image = x. */
/* oracle_reflect plus the outward surface normal at the boundary point */
static double reflect_normal(const fdtd_object *obj, const double *x, double *image, double *xb, double *erected, double *n)
{
    const double dist = oracle_object_closest(obj, x, xb, n);
    const int inside = dist < 0;
    const double diff[3] = {xb[0]-x[0], xb[1]-x[1], xb[2]-x[2]};
    if (norm3(diff) > REF_EPS) {
        if (inside) for (int i = 0; i < 3; ++i) { erected[i] = xb[i] - x[i]; image[i] = xb[i] + erected[i]; }
        else        for (int i = 0; i < 3; ++i) { erected[i] = -xb[i] + x[i]; image[i] = x[i] + erected[i]; }
    } else {
        for (int i = 0; i < 3; ++i) { erected[i] = n[i]; image[i] = x[i]; }
    }
    const double en = norm3(erected);
    for (int i = 0; i < 3; ++i) erected[i] /= en;
    return dist;
}

double oracle_reflect(const fdtd_object *obj, const double *x, double *image, double *xb, double *erected)
{
    double n[3];
    return reflect_normal(obj, x, image, xb, erected, n);
}

/* tight world AABB of an object (stands in for GetUnionBBox) */
static void object_aabb(const fdtd_object *o, double *mn, double *mx)
{
    for (int i = 0; i < 3; ++i) {
        double ext;
        if (o->type == FDTD_OBJ_SPHERE) ext = o->semi_axes[0];
        else {
            ext = 0;
            for (int j = 0; j < 3; ++j) ext += (o->rotation[i*3+j]*o->semi_axes[j]) * (o->rotation[i*3+j]*o->semi_axes[j]);
            ext = sqrt(ext);
        }
        mn[i] = o->center[i] - ext; mx[i] = o->center[i] + ext;
    }
}

/* ======================================================================== */
/* Vandermonde fill (MAC_Grid.cpp:4444-4510) and the fully-coupled weights   */
/* ======================================================================== */
static void fill_regular(const double *np, double *row)
{
    const double x = np[0], y = np[1], z = np[2];
    row[0] = x*y*z; row[1] = x*y; row[2] = x*z; row[3] = y*z;
    row[4] = x; row[5] = y; row[6] = z; row[7] = 1;
}
static void fill_boundary(const double *np, const double *nr, double *row)
{
    const double x = np[0], y = np[1], z = np[2], nx = nr[0], ny = nr[1], nz = nr[2];
    row[0] = nx*y*z + ny*x*z + nz*x*y; row[1] = nx*y + ny*x; row[2] = nx*z + nz*x;
    row[3] = ny*z + nz*y; row[4] = nx; row[5] = ny; row[6] = nz; row[7] = 0;
}
static void scaled(const double *p, const double *origin, double h, double *np)
{
    for (int i = 0; i < 3; ++i) { np[i] = p[i]; np[i] -= origin[i]; np[i] /= h; np[i] += 0.5; }
}

/* V rows = sample rows; solve V^T w = v (JacobiSVD of V^T), return cond. */
double oracle_vandermonde_weights(const double *rows, int s_i, const double *bp, const double *nrm,
                                  const double *image, const double *origin, double h, double *w)
{
    double V[64], VT[64], v[8], np[3];
    for (int r = 0; r < 8; ++r) {
        if (r == s_i) { scaled(bp, origin, h, np); fill_boundary(np, nrm, &V[r*8]); }
        else          { scaled(&rows[3*r], origin, h, np); fill_regular(np, &V[r*8]); }
    }
    scaled(image, origin, h, np);
    fill_regular(np, v);
    for (int i = 0; i < 8; ++i) for (int j = 0; j < 8; ++j) VT[i*8 + j] = V[j*8 + i];
    double cond;
    lsq_solve_internal(8, 8, VT, v, w, &cond);
    return cond;
}

/* ======================================================================== */
/* PML profile (MAC_Grid.cpp:4027-4102) and coefficients (MAC_Grid.h:528-571)*/
/* ======================================================================== */
static void pressure_bbox(const fdtd_settings *s, double *bmin, double *bmax)
{
    for (int d = 0; d < 3; ++d) {
        bmin[d] = s->min_bound[d];
        bmax[d] = bmin[d];
        bmax[d] += s->cell_size * (double)s->divisions[d];   /* ScalarField.cpp:43-46 */
    }
}

static int inside_pml(const fdtd_settings *s, const double *x, double W)
{
    double bmin[3], bmax[3];
    pressure_bbox(s, bmin, bmax);
    const double pw = W + SMALL_PML_BUFFER;
    for (int d = 0; d < 3; ++d) {
        const double hMin = x[d] - bmin[d], hMax = bmax[d] - x[d];
        switch (s->boundary_condition_preset) {
        case 0: if (hMin <= pw || hMax <= pw) return 1; break;
        case 1: if (hMin <= pw || (d == 1 && hMax <= pw)) return 1; break;
        case 2: if (d == 2 && hMax <= pw) return 1; break;
        default: break;
        }
    }
    return 0;
}

static double pml_sigma(const fdtd_settings *s, const double *x, double W, int d)
{
    double bmin[3], bmax[3];
    pressure_bbox(s, bmin, bmax);
    const double S = s->pml_strength;
    const double hMin = x[d] - bmin[d], hMax = bmax[d] - x[d];
    const double a2 = W * W, distMin = W - hMin, distMax = W - hMax;
    switch (s->boundary_condition_preset) {
    case 0:
        if (hMin <= W) return S * pow(distMin, 2) / a2;
        else if (hMax <= W) return S * pow(distMax, 2) / a2;
        return 0.0;
    case 1:
        if (hMin <= W) return S * pow(distMin, 2) / a2;
        else if (d == 1 && hMax <= W) return S * pow(distMax, 2) / a2;
        return 0.0;
    case 2:
        if (d != 2 || hMin <= W) return 0.0;
        else if (d == 2 && hMax <= W) return S * pow(distMax, 2) / a2;
        return 0.0;
    default: return 0.0;
    }
}

double oracle_pml_sigma(const fdtd_settings *s, const double *x, int dim)
{
    return pml_sigma(s, x, s->pml_width * s->cell_size, dim);
}
int oracle_inside_pml(const fdtd_settings *s, const double *x)
{
    return inside_pml(s, x, s->pml_width * s->cell_size);
}

static inline double f_vel_update(double sg, double dt) { double c = 1.0/dt - sg/2.0; c /= 1.0/dt + sg/2.0; return c; }
static inline double f_grad(double sg, double dt, double h, double rho) { double c = 1.0/dt + sg/2.0; c *= rho*h; return -1.0/c; }
static inline double f_dir(double sg, double dt) { return 1.0/dt + sg/2.0; }
static inline double f_pupd(double sg, double dt, double D) { return (1.0/dt - sg/2.0) / D; }
static inline double f_div(double rho, double c, double D) { return -1.0*rho*c*c / D; }

/* ======================================================================== */
/* solver state                                                               */
/* ======================================================================== */
typedef struct { int index; double upd[3], div[3]; int vl[3], vr[3]; } pml_pcell;
typedef struct { int index, dim; double upd, grad; int pr, pl; } pml_vcell;

typedef struct {
    int owner, neighbour, topology, object;
    int type;          /* 0: object ghost, 1: plane-constraint ghost (MakeGhostCell's gctype) */
    uint64_t key;
    double position[3];
    double distance;
    double pressure;
} ghost_t;

struct oracle {
    fdtd_scene sc;
    fdtd_object *objs;
    fdtd_impulse_source *isrc;
    fdtd_acc_noise_source *ans;   /* acceleration-noise sources, impacts copied */
    fdtd_kinematics *kin;         /* keyframe animations, frames copied (as in the file) */
    fdtd_pressure_source *psrc;
    double *listeners;
    double src_bmin[64][3], src_bmax[64][3], src_pos[64][3];

    int n[3];
    long long N, Nv[3];
    double h, W, dt, c, rho;
    double bmin[3], bmax[3];

    double *pc[3], *lap[2], *pdir[3], *v[3];
    unsigned char *bulk, *ghostc, *pmlc, *cavity, *history, *classified;
    int pc_ind, lap_ind;

    pml_pcell *pp; long long npp;
    pml_vcell *pv; long long npv;

    ghost_t *gh; int ngh, capgh;
    /* the last assembled ghost system (CSR), kept for inspection */
    int sys_n, sys_nnz; int *sys_rp, *sys_ci; double *sys_vv, *sys_rhs;
    /* open-addressing hash: key -> ghost index */
    uint64_t *hkeys; int *hvals; int hcap;
    /* owner -> children range (ghost indices sorted by owner then key) */
    int *child_order;
    int reclass;
    int n_coupled;

    double t;
    long long step;
    int evals_off;   /* time-parallel chunk past its stop_boundary_acc_time */
    fdtd_object *objs0;  /* poses at t = 0 of moving objects                */
    fdtd_speaker_source *spk;  /* speaker sources with copied samples       */
    fdtd_plane_constraint *cons; /* plane constraints                       */
    long long *solid; long long nsolid;  /* solid cells of the last classification */
    int moving;          /* some object has a velocity or angular velocity  */
    int fresh_error;     /* a fresh cell lay inside an object               */
    /* staggered (MAC) scheme: pFull; pdir[3] and v[3] are the full-grid fields */
    int scheme;
    double *pfull;
};

static inline long long pidx(const oracle *o, int i, int j, int k) { return ((long long)i * o->n[1] + j) * o->n[2] + k; }
static inline long long vidx(const oracle *o, int d, int i, int j, int k)
{
    const int ny = o->n[1] + (d == 1), nz = o->n[2] + (d == 2);
    return ((long long)i * ny + j) * nz + k;
}
static inline void pidx3(const oracle *o, long long c, int *ijk)
{
    ijk[0] = (int)(c / ((long long)o->n[1] * o->n[2]));
    c -= (long long)ijk[0] * o->n[1] * o->n[2];
    ijk[1] = (int)(c / o->n[2]);
    ijk[2] = (int)(c - (long long)ijk[1] * o->n[2]);
}
/* ScalarField::cellPosition (ScalarField.cpp:126-136) */
static inline void ppos(const oracle *o, const int *ijk, double *x)
{
    for (int d = 0; d < 3; ++d) { x[d] = o->bmin[d]; x[d] += (0.5 + (double)ijk[d]) * o->h; }
}
/* velocity field d: min = pmin - h/2 e_d (MAC_Grid.cpp:103-113) */
static inline void vpos(const oracle *o, int d, const int *ijk, double *x)
{
    for (int e = 0; e < 3; ++e) {
        double mn = o->bmin[e] - (e == d ? o->h / 2.0 : 0.0);
        x[e] = mn; x[e] += (0.5 + (double)ijk[e]) * o->h;
    }
}

/* GhostCell::MakeKey, Szudzik pairing (MAC_Grid.h:108-114) */
static inline uint64_t szudzik(long long a_, long long b_)
{
    const uint64_t a = (uint64_t)(a_ < b_ ? a_ : b_), b = (uint64_t)(a_ < b_ ? b_ : a_);
    return (a >= b ? a*a + a + b : a + b*b);
}

static int hash_find(const oracle *o, uint64_t key)
{
    if (!o->hcap) return -1;
    uint64_t hsh = key * 0x9E3779B97F4A7C15ull;
    int pos = (int)(hsh >> 40) & (o->hcap - 1);
    while (o->hvals[pos] >= 0) {
        if (o->hkeys[pos] == key) return o->hvals[pos];
        pos = (pos + 1) & (o->hcap - 1);
    }
    return -1;
}
static void hash_build(oracle *o)
{
    int cap = 16;
    while (cap < 2 * o->ngh + 16) cap <<= 1;
    if (cap != o->hcap) {
        free(o->hkeys); free(o->hvals);
        o->hkeys = (uint64_t *)malloc(sizeof(uint64_t) * cap);
        o->hvals = (int *)malloc(sizeof(int) * cap);
        o->hcap = cap;
    }
    for (int i = 0; i < cap; ++i) o->hvals[i] = -1;
    for (int g = 0; g < o->ngh; ++g) {
        uint64_t hsh = o->gh[g].key * 0x9E3779B97F4A7C15ull;
        int pos = (int)(hsh >> 40) & (cap - 1);
        while (o->hvals[pos] >= 0) pos = (pos + 1) & (cap - 1);
        o->hkeys[pos] = o->gh[g].key; o->hvals[pos] = g;
    }
}

static double ghost_pressure(const oracle *o, long long a, long long b)
{
    const int g = hash_find(o, szudzik(a, b));
    if (g < 0) { fprintf(stderr, "oracle: missing ghost key\n"); abort(); }
    return o->gh[g].pressure;
}

/* ======================================================================== */
/* MAC_Grid::classifyCells (MAC_Grid.cpp:2397-2693): PML lists.              */
/* The analytic-object scenes go through FDTD_Objects, never _boundaryFields,*/
/* so every cell is bulk here and the interfacial lists stay empty.          */
/* ======================================================================== */
static void classify_static(oracle *o)
{
    const fdtd_settings *s = &o->sc.settings;
    const double dt = o->dt, dx = o->h, c = o->c, rho = o->rho;
    o->npp = 0; o->npv = 0;
    for (long long ci = 0; ci < o->N; ++ci) {
        int ijk[3]; double x[3];
        pidx3(o, ci, ijk); ppos(o, ijk, x);
        o->bulk[ci] = 1;
        o->pmlc[ci] = 0;
        /* MAC_Grid::classifyCells (MAC_Grid.cpp:2457-2473): without use_mesh
           every cell is bulk and no PML cell is made */
        if (s->use_mesh && inside_pml(s, x, o->W)) {
            pml_pcell *pc = &o->pp[o->npp++];
            pc->index = (int)ci;
            for (int d = 0; d < 3; ++d) {
                const double sg = pml_sigma(s, x, o->W, d);
                const double D = f_dir(sg, dt);
                pc->upd[d] = f_pupd(sg, dt, D);
                pc->div[d] = f_div(rho, c, D);
                int t[3] = {ijk[0], ijk[1], ijk[2]};
                pc->vl[d] = (int)vidx(o, d, t[0], t[1], t[2]);
                t[d] += 1;
                pc->vr[d] = (int)vidx(o, d, t[0], t[1], t[2]);
            }
            o->pmlc[ci] = 1;
        }
    }
    for (int d = 0; d < 3; ++d) {
        const int vn[3] = {o->n[0] + (d == 0), o->n[1] + (d == 1), o->n[2] + (d == 2)};
        for (int i = 0; i < vn[0]; ++i)
        for (int j = 0; j < vn[1]; ++j)
        for (int k = 0; k < vn[2]; ++k) {
            int ijk[3] = {i, j, k};
            if (ijk[d] == 0 || ijk[d] == vn[d] - 1) continue;  /* domain walls stay 0 */
            double x[3];
            vpos(o, d, ijk, x);
            if (!inside_pml(s, x, o->W)) continue;
            const double sg = pml_sigma(s, x, o->W, d);
            pml_vcell *pv = &o->pv[o->npv++];
            pv->index = (int)vidx(o, d, i, j, k);
            pv->dim = d;
            pv->upd = f_vel_update(sg, dt);
            pv->grad = f_grad(sg, dt, dx, rho);
            pv->pr = (int)pidx(o, ijk[0], ijk[1], ijk[2]);
            ijk[d] -= 1;
            pv->pl = (int)pidx(o, ijk[0], ijk[1], ijk[2]);
            if (!o->pmlc[pv->pr] && !o->pmlc[pv->pl]) {
                fprintf(stderr, "oracle: pml face with no pml neighbour\n"); abort();
            }
        }
    }
}

/* ScalarField::GetIterationBox (ScalarField.cpp:288-314) */
static void iteration_box(const oracle *o, const double *in_min, const double *in_max, int *start, int *range)
{
    double mn[3] = {in_min[0], in_min[1], in_min[2]}, mx[3] = {in_max[0], in_max[1], in_max[2]};
    int inside_min = 1, inside_max = 1;
    for (int d = 0; d < 3; ++d) {
        if (!(in_min[d] >= o->bmin[d] && in_min[d] <= o->bmax[d])) inside_min = 0;
        if (!(in_max[d] >= o->bmin[d] && in_max[d] <= o->bmax[d])) inside_max = 0;
    }
    if (!inside_min) for (int d = 0; d < 3; ++d) mn[d] = fmax(o->bmin[d], mn[d]);
    if (!inside_max) for (int d = 0; d < 3; ++d) mx[d] = fmin(o->bmax[d], mx[d]);
    for (int d = 0; d < 3; ++d) {
        start[d] = (int)floor((mn[d] - o->bmin[d]) / o->h);
        const int end = (int)ceil((mx[d] - o->bmin[d]) / o->h);
        range[d] = end - start[d] > 1 ? end - start[d] : 1;
    }
}

static int occupy_by_object(const oracle *o, const double *x)
{
    for (int i = 0; i < o->sc.n_objects; ++i)
        if (object_occupies(&o->objs[i], x)) return i;
    return -1;
}

/* FDTD_Objects::LowestObjectDistance (FDTD_Objects.cpp:240-257) */
static void lowest_object_distance(const oracle *o, const double *x, double *dist, int *id)
{
    *dist = DBL_MAX; *id = 0;
    for (int i = 0; i < o->sc.n_objects; ++i) {
        const double d = object_distance(&o->objs[i], x);
        if (fabs(d) < fabs(*dist)) { *dist = d; *id = i; }
    }
}

static int cmp_ghost(const void *a, const void *b)
{
    const ghost_t *x = (const ghost_t *)a, *y = (const ghost_t *)b;
    if (x->owner != y->owner) return x->owner < y->owner ? -1 : 1;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return 0;
}

/* MAC_Grid::classifyCells_FAST (MAC_Grid.cpp:3455-3891) for rigid objects:
 * bulk = the centre outside every object (ONLY_SAMPLE_CENTER, the compiled
 * default, 3558-3564), or all 7 samples (centre, +-h/4 per axis) outside
 * every object (3566-3577); solid cells with a bulk 6-neighbour become ghost
 * cells, one ghost entry per (solid owner, bulk neighbour) pair via
 * MakeGhostCell (4310-4340). */
/* FDTD_Objects::OccupyByConstraint (FDTD_Objects.cpp:54-71) with
 * FDTD_PlaneConstraint::Inside (FDTD_PlaneConstraint.h:38-42) */
static int occupy_by_constraint(const oracle *o, const double *x)
{
    for (int i = 0; i < o->sc.n_constraints; ++i) {
        const fdtd_plane_constraint *c = &o->cons[i];
        if ((x[c->direction] - c->height) * (double)c->sign < 0.0) return 1;
    }
    return 0;
}

static void classify_fast(oracle *o)
{
    const int nobj = o->sc.n_objects;
    /* classifyCellsDynamic_FAST returns at once without use_mesh (MAC_Grid.cpp:3014) */
    if (!o->sc.settings.use_mesh) { memcpy(o->history, o->bulk, (size_t)o->N); return; }
    memcpy(o->history, o->bulk, (size_t)o->N);
    memset(o->bulk, 1, (size_t)o->N);
    memset(o->cavity, 0, (size_t)o->N);
    memset(o->ghostc, 0, (size_t)o->N);
    memset(o->classified, 0, (size_t)o->N);
    o->ngh = 0;
    o->nsolid = 0;
    const int ncons = o->sc.n_constraints;
    if (nobj == 0 && ncons == 0) { hash_build(o); return; }
    const double q = 0.25 * o->h;
    long long nsolid = 0;
    long long *solid = NULL; long long capsolid = 0;
    unsigned char *stype = NULL;
    /* iteration boxes (MAC_Grid.cpp:3511-3550): each object's box with the
       volume behind every constraint cut away, then each constraint's solid
       half-space */
    for (int box = 0; box < nobj + ncons; ++box) {
        double mn[3], mx[3]; int st[3], rg[3];
        if (box < nobj) {
            object_aabb(&o->objs[box], mn, mx);
            for (int c = 0; c < ncons; ++c) {
                const fdtd_plane_constraint *pc = &o->cons[c];
                if (pc->sign > 0) mn[pc->direction] = fmax(mn[pc->direction], pc->height);
                else              mx[pc->direction] = fmin(mx[pc->direction], pc->height);
            }
        } else {
            const fdtd_plane_constraint *pc = &o->cons[box - nobj];
            for (int d = 0; d < 3; ++d) { mn[d] = -DBL_MAX; mx[d] = DBL_MAX; }
            if (pc->sign > 0) mx[pc->direction] = pc->height;
            else              mn[pc->direction] = pc->height;
        }
        iteration_box(o, mn, mx, st, rg);
        for (int i = st[0]; i < st[0] + rg[0]; ++i)
        for (int j = st[1]; j < st[1] + rg[1]; ++j)
        for (int k = st[2]; k < st[2] + rg[2]; ++k) {
            if (i < 0 || j < 0 || k < 0 || i >= o->n[0] || j >= o->n[1] || k >= o->n[2]) continue;
            const long long ci = pidx(o, i, j, k);
            if (o->classified[ci]) continue;
            int ijk[3] = {i, j, k}; double x[3];
            ppos(o, ijk, x);
            /* isFluid (MAC_Grid.cpp:3566-3577): the centre outside every
               constraint, all seven samples outside every object; a solid
               cell is a constraint cell (type 1) when a constraint holds its
               centre */
            const int in_con = occupy_by_constraint(o, x);
            int fluid = !in_con && occupy_by_object(o, x) < 0;
            for (int d = 0; d < 3 && fluid && o->sc.settings.sampling == FDTD_SAMPLE_SEVEN; ++d)
                for (int sgn = -1; sgn <= 1 && fluid; sgn += 2) {
                    double y[3] = {x[0], x[1], x[2]};
                    y[d] += (double)sgn * q;
                    if (occupy_by_object(o, y) >= 0) fluid = 0;
                }
            if (!fluid) {
                o->bulk[ci] = 0;
                if (nsolid == capsolid) {
                    capsolid = capsolid ? 2*capsolid : 1024;
                    solid = (long long *)realloc(solid, sizeof(long long)*capsolid);
                    stype = (unsigned char *)realloc(stype, (size_t)capsolid);
                }
                stype[nsolid] = (unsigned char)in_con;
                solid[nsolid++] = ci;
            }
            o->classified[ci] = 1;
        }
    }
    /* ghost entries, MakeGhostCell + LowestObjectDistance */
    for (long long s = 0; s < nsolid; ++s) {
        const long long ci = solid[s];
        o->pc[0][ci] = 0.0; o->pc[1][ci] = 0.0; o->pc[2][ci] = 0.0;
        int ijk[3]; pidx3(o, ci, ijk);
        int isg = 0;
        for (int d = 0; d < 3; ++d)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                int t[3] = {ijk[0], ijk[1], ijk[2]};
                t[d] += sgn;
                if (t[d] < 0 || t[d] >= o->n[d]) continue;
                const long long nb = pidx(o, t[0], t[1], t[2]);
                if (!o->bulk[nb]) continue;
                isg = 1;
                if (o->ngh == o->capgh) { o->capgh = o->capgh ? 2*o->capgh : 1024; o->gh = (ghost_t *)realloc(o->gh, sizeof(ghost_t)*o->capgh); }
                ghost_t *g = &o->gh[o->ngh++];
                g->owner = (int)ci; g->neighbour = (int)nb; g->topology = sgn * (d + 1);
                g->type = stype[s];
                g->pressure = 0.0;
                double x[3]; ppos(o, ijk, x);
                lowest_object_distance(o, x, &g->distance, &g->object);
                g->position[0] = x[0]; g->position[1] = x[1]; g->position[2] = x[2];
                /* MakeGhostCell (MAC_Grid.cpp:4322-4331): h/4 towards the
                   neighbour without ONLY_SAMPLE_CENTER; with it (and
                   STAGGER_ACC_SAMPLE) h/2 for rasterized ghosts, none for
                   fully-coupled ones */
                double off = 0.25;
                if (o->sc.settings.sampling != FDTD_SAMPLE_SEVEN)
                    off = o->sc.settings.boundary_handling == FDTD_BOUNDARY_RASTERIZE ? 0.5 : 0.0;
                g->position[d] += (g->topology > 0 ? 1.0 : -1.0) * off * o->h;
                g->key = szudzik(ci, nb);
            }
        o->ghostc[ci] = (unsigned char)isg;
    }
    free(o->solid);
    o->solid = solid;
    o->nsolid = nsolid;
    free(stype);
    qsort(o->gh, o->ngh, sizeof(ghost_t), cmp_ghost);
    hash_build(o);
}

static void update_evals(oracle *o);
static void animate_objects(oracle *o, double t);

/* HarmonicVibrationalSource::Evaluate (HarmonicVibrationalSource.cpp:5-28):
 * v_bar omega sin(omega t + phase), decayed when a decay rate is set; zero
 * outside the source's range (the reference's range is unbounded unless
 * SetRange is called; here t1 <= t0 means unbounded) */
static double harmonic_an(const fdtd_impulse_source *s, double t)
{
    if (s->t1 > s->t0 && (t < s->t0 || t > s->t1)) return 0.0;
    const double v_bar = s->amplitude;
    const double scale = v_bar * s->omega;
    const double val = sin(s->omega * t + s->phase) * scale;
    if (s->decay_rate < 0) return val;
    return exp(-s->decay_rate * (t - s->decay_center)) * val;
}

/* ChirpVibrationalSource::Evaluate (ChirpVibrationalSource.cpp:5-19): zero
 * before _time[0]; alpha runs to 0.5 at _time[1] and is 1 after; the phase
 * jumps once alpha reaches 0.52; unit amplitude */
static double chirp_an(const fdtd_impulse_source *s, double t)
{
    const double time0 = s->t0, time1 = s->t1, omega0 = s->omega, omega1 = s->omega1;
    if (t < time0) return 0.0;
    const double alpha = (t < time1 ? 0.5 * (t - time0) / (time1 - time0) : 1.0);
    const double phase = (alpha < 0.52 ? 0 : (time1) * 0.5 * (omega1 - omega0));
    const double omega = (1.0 - alpha) * omega0 + alpha * omega1;
    const double amp = 1.0;
    return sin(omega * t + phase) * amp;
}

/* AccelerationNoiseVibrationalSource::Evaluate(vertexID, time)
 * (AccelerationNoiseVibrationalSource.cpp:63-90) over the impacts
 * ImpulseSeriesObject::GetImpulseWithinSupport returns (t_0 <= t < t_0 + tau,
 * ImpulseSeriesObject.cpp:88-101); ComputeS is the half sine (:19-29);
 * ObjectToWorldVector applies the object's rotation */
static void acc_noise(const fdtd_acc_noise_source *an, const fdtd_object *ob, double t, double *out)
{
    double acc[3] = {0, 0, 0};
    for (int q = 0; q < an->n_impacts; ++q) {
        const fdtd_impact *im = &an->impacts[q];
        if (!(t >= im->timestamp && t < im->timestamp + im->support)) continue;
        if (im->support < 1e-12) continue;   /* SMALL_NUM */
        const double *J = im->impulse;
        double r[3];
        for (int d = 0; d < 3; ++d) r[d] = im->position[d] - an->center_of_mass[d];
        double S = 0.0;
        if (t <= im->timestamp + im->support && t >= im->timestamp) S = sin(M_PI * (t - im->timestamp) / im->support);
        /* translational acceleration */
        for (int d = 0; d < 3; ++d) acc[d] += J[d] * (M_PI * S / (2.0 * im->support * an->mass));
        /* rotational acceleration: alpha = I^-1 (r x J) pi S / (2 tau), added as alpha x r */
        const double rxJ[3] = {r[1] * J[2] - r[2] * J[1], r[2] * J[0] - r[0] * J[2], r[0] * J[1] - r[1] * J[0]};
        double alpha[3];
        for (int d = 0; d < 3; ++d) {
            alpha[d] = an->inv_inertia[3 * d] * rxJ[0] + an->inv_inertia[3 * d + 1] * rxJ[1] + an->inv_inertia[3 * d + 2] * rxJ[2];
            alpha[d] *= (M_PI * S) / (2.0 * im->support);
        }
        acc[0] += alpha[1] * r[2] - alpha[2] * r[1];
        acc[1] += alpha[2] * r[0] - alpha[0] * r[2];
        acc[2] += alpha[0] * r[1] - alpha[1] * r[0];
    }
    for (int d = 0; d < 3; ++d)
        out[d] += ob->rotation[3 * d] * acc[0] + ob->rotation[3 * d + 1] * acc[1] + ob->rotation[3 * d + 2] * acc[2];
}

/* FDTD_RigidObject::EvaluateBoundaryAcceleration (FDTD_RigidObject.cpp:
 * 331-368): the sum over the object's vibrational sources of their
 * acceleration vectors at a surface point with outward normal nrm -- a_n nrm
 * for the Gaussian-impulse and harmonic sources, the object's one vector for
 * acceleration noise; zero once the evaluations are disabled */
static void object_accel(const oracle *o, int obj, double t, const double *nrm, double *out)
{
    out[0] = out[1] = out[2] = 0.0;
    if (o->evals_off) return;
    double a = 0.0;
    for (int i = 0; i < o->sc.n_impulse_sources; ++i) {
        const fdtd_impulse_source *s = &o->isrc[i];
        if (s->object != obj) continue;
        if (s->kind == FDTD_VIB_HARMONIC) {
            a += harmonic_an(s, t);
        } else if (s->kind == FDTD_VIB_CHIRP) {
            a += chirp_an(s, t);
        } else {
            const double u = t - s->onset;
            a += s->amplitude * exp(-(u * u) / (2.0 * s->width * s->width));
        }
    }
    for (int d = 0; d < 3; ++d) out[d] = a * nrm[d];
    for (int i = 0; i < o->sc.n_acc_noise_sources; ++i)
        if (o->ans[i].object == obj) acc_noise(&o->ans[i], &o->objs[obj], t, out);
}

void oracle_object_accel(const oracle *o, int obj, double t, const double *nrm, double *out)
{
    object_accel(o, obj, t, nrm, out);
}

/* ScalarField::enclosingNeighbours (ScalarField.cpp:261-286): sorted unique */
static int enclosing(const oracle *o, const double *p, long long *out)
{
    double diff[3];
    for (int d = 0; d < 3; ++d) { diff[d] = p[d] - o->bmin[d]; diff[d] -= o->h / 2.0; }
    int a[3], b[3];
    for (int d = 0; d < 3; ++d) {
        int v0 = (int)(diff[d] / o->h);
        if (v0 < 0) v0 = 0;
        if (v0 > o->n[d] - 1) v0 = o->n[d] - 1;
        a[d] = v0;
        b[d] = diff[d] < 0.0 ? 0 : (v0 + 1 < o->n[d] - 1 ? v0 + 1 : o->n[d] - 1);
    }
    long long tmp[8]; int m = 0;
    for (int x = 0; x < 2; ++x) for (int y = 0; y < 2; ++y) for (int z = 0; z < 2; ++z)
        tmp[m++] = pidx(o, x ? b[0] : a[0], y ? b[1] : a[1], z ? b[2] : a[2]);
    /* sort + unique */
    for (int i = 0; i < 8; ++i) for (int j = i + 1; j < 8; ++j) if (tmp[j] < tmp[i]) { long long t = tmp[i]; tmp[i] = tmp[j]; tmp[j] = t; }
    int u = 0;
    for (int i = 0; i < 8; ++i) if (u == 0 || tmp[i] != out[u-1]) out[u++] = tmp[i];
    return u;
}

static int find_children(const oracle *o, int owner, int *first)
{
    /* gh sorted by owner: binary search */
    int lo = 0, hi = o->ngh;
    while (lo < hi) { int mid = (lo + hi) / 2; if (o->gh[mid].owner < owner) lo = mid + 1; else hi = mid; }
    *first = lo;
    int cnt = 0;
    while (lo + cnt < o->ngh && o->gh[lo + cnt].owner == owner) ++cnt;
    return cnt;
}

/* ======================================================================== */
/* MAC_Grid::PML_pressureUpdateGhostCells (MAC_Grid.cpp:1131-1441)          */
/* ======================================================================== */
typedef struct { int row, col; double val; } trip_t;

static void sparse_matvec(int n, const int *rp, const int *ci, const double *vv, const double *x, double *y)
{
    for (int r = 0; r < n; ++r) {
        double acc = 0;
        for (int q = rp[r]; q < rp[r+1]; ++q) acc += vv[q] * x[ci[q]];
        y[r] = acc;
    }
}

/* BiCGSTAB with a diagonal (Jacobi) preconditioner, Eigen's defaults:
 * tolerance eps, at most 2n iterations (Eigen/src/IterativeLinearSolvers). */
static void bicgstab(int n, const int *rp, const int *ci, const double *vv, const double *b, double *x)
{
    double *dinv = (double *)malloc(sizeof(double) * n);
    for (int r = 0; r < n; ++r) {
        double d = 0;
        for (int q = rp[r]; q < rp[r+1]; ++q) if (ci[q] == r) d += vv[q];
        dinv[r] = d != 0 ? 1.0 / d : 1.0;
    }
    double *r = (double *)calloc(n, sizeof(double)), *r0 = (double *)calloc(n, sizeof(double));
    double *p = (double *)calloc(n, sizeof(double)), *vq = (double *)calloc(n, sizeof(double));
    double *y = (double *)calloc(n, sizeof(double)), *z = (double *)calloc(n, sizeof(double));
    double *s = (double *)calloc(n, sizeof(double)), *t = (double *)calloc(n, sizeof(double));
    double bn = 0;
    for (int i = 0; i < n; ++i) { x[i] = dinv[i] * b[i]; bn += b[i] * b[i]; }
    if (bn == 0) { for (int i = 0; i < n; ++i) x[i] = 0; goto done; }
    sparse_matvec(n, rp, ci, vv, x, r);
    for (int i = 0; i < n; ++i) { r[i] = b[i] - r[i]; r0[i] = r[i]; }
    {
        const double tol2 = DBL_EPSILON * DBL_EPSILON * bn;
        double rho = 1, alpha = 1, w = 1;
        for (int it = 0; it < 2 * n + 10; ++it) {
            double rn = 0;
            for (int i = 0; i < n; ++i) rn += r[i] * r[i];
            if (rn <= tol2) break;
            double rho_old = rho;
            rho = 0;
            for (int i = 0; i < n; ++i) rho += r0[i] * r[i];
            if (fabs(rho) < 1e-300) {
                /* restart */
                for (int i = 0; i < n; ++i) r0[i] = r[i];
                rho = rn; rho_old = 1; alpha = 1; w = 1;
                for (int i = 0; i < n; ++i) { vq[i] = 0; p[i] = 0; }
            }
            const double beta = (rho / rho_old) * (alpha / w);
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - w * vq[i]);
            for (int i = 0; i < n; ++i) y[i] = dinv[i] * p[i];
            sparse_matvec(n, rp, ci, vv, y, vq);
            double r0v = 0;
            for (int i = 0; i < n; ++i) r0v += r0[i] * vq[i];
            alpha = rho / r0v;
            for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * vq[i];
            for (int i = 0; i < n; ++i) z[i] = dinv[i] * s[i];
            sparse_matvec(n, rp, ci, vv, z, t);
            double tt = 0, ts = 0;
            for (int i = 0; i < n; ++i) { tt += t[i] * t[i]; ts += t[i] * s[i]; }
            w = tt > 0 ? ts / tt : 0;
            for (int i = 0; i < n; ++i) { x[i] += alpha * y[i] + w * z[i]; r[i] = s[i] - w * t[i]; }
        }
    }
done:
    free(dinv); free(r); free(r0); free(p); free(vq); free(y); free(z); free(s); free(t);
}

static int cmp_trip(const void *a, const void *b)
{
    const trip_t *x = (const trip_t *)a, *y = (const trip_t *)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    return x->col < y->col ? -1 : (x->col > y->col);
}

static void ghost_update(oracle *o, const double *p)
{
    const int G = o->ngh;
    if (G == 0) return;
    const double h = o->h, rho = o->rho, t = o->t;
    const int coupled = o->sc.settings.boundary_handling == FDTD_BOUNDARY_FULLY_COUPLED;
    double *rhs = (double *)malloc(sizeof(double) * G);
    trip_t *trips = (trip_t *)malloc(sizeof(trip_t) * (size_t)G * 9);
    int *ntr = (int *)calloc(G, sizeof(int));
    int n_coupled = 0;
#pragma omp parallel for schedule(static) reduction(+:n_coupled)
    for (int g = 0; g < G; ++g) {
        ghost_t *gc = &o->gh[g];
        trip_t *tr = &trips[(size_t)g * 9];
        const double p_neig = p[gc->neighbour];
        if (gc->type == 1) {
            /* a plane constraint reflects perfectly (MAC_Grid.cpp:1198-1204) */
            tr[0].row = g; tr[0].col = g; tr[0].val = 1.0; ntr[g] = 1;
            rhs[g] = p_neig;
            continue;
        }
        const fdtd_object *obj = &o->objs[gc->object];
        double image[3], bp[3], en[3], nrm_b[3];
        const double distance = reflect_normal(obj, gc->position, image, bp, en, nrm_b);
        double grad_p[3];
        object_accel(o, gc->object, t, nrm_b, grad_p);
        for (int i = 0; i < 3; ++i) grad_p[i] *= (-rho);
        double cell_n[3] = {0, 0, 0};
        cell_n[abs(gc->topology) - 1] = gc->topology > 0 ? 1.0 : -1.0;
        const double grad_p_nr = dot3(grad_p, cell_n);
        const double grad_p_ne = dot3(grad_p, en);
        /* ONLY_SAMPLE_CENTER: pg = p_neig - h grad_p_nr; else 0.75 h (MAC_Grid.cpp:1277-1281) */
        const double pg = o->sc.settings.sampling == FDTD_SAMPLE_SEVEN ? p_neig - 0.75 * h * grad_p_nr : p_neig - h * grad_p_nr;
        double weights = (object_distance(obj, gc->position) < 0.0 ? -2.0 * fabs(distance) : -fabs(distance));
        double weighted = grad_p_ne * weights;
        long long nbrs[8];
        const int nn = enclosing(o, image, nbrs);
        int use_ii = (nn == 8) && coupled;
        if (!use_ii) {
            tr[0].row = g; tr[0].col = g; tr[0].val = 1.0; ntr[g] = 1;
            rhs[g] = pg;
            continue;
        }
        double origin[3] = {0, 0, 0};
        for (int q = 0; q < 8; ++q) { int ijk[3]; double x[3]; pidx3(o, nbrs[q], ijk); ppos(o, ijk, x); for (int d = 0; d < 3; ++d) origin[d] += x[d]; }
        for (int d = 0; d < 3; ++d) origin[d] /= 8.0;
        double rows[24], r[8];
        int s_i = -1, known[8], unk_g[8];
        for (int q = 0; q < 8; ++q) { known[q] = 0; unk_g[q] = -1; r[q] = 0; }
        for (int q = 0; q < 8; ++q) {
            const long long ni = nbrs[q];
            if (ni == gc->owner) {
                rows[3*q] = bp[0]; rows[3*q+1] = bp[1]; rows[3*q+2] = bp[2];
                r[q] = grad_p_ne * h;
                s_i = q;
            } else if (!o->ghostc[ni]) {
                int ijk[3]; pidx3(o, ni, ijk); ppos(o, ijk, &rows[3*q]);
                r[q] = p[ni];
                known[q] = 1;
                if (!o->bulk[ni]) { /* a solid non-ghost cell cannot be here */ }
            } else {
                int first; const int cnt = find_children(o, (int)ni, &first);
                if (cnt == 0) { fprintf(stderr, "oracle: ghost without children\n"); abort(); }
                double best = DBL_MAX; int bi = first;
                for (int cidx = first; cidx < first + cnt; ++cidx) {
                    const double *gp = o->gh[cidx].position;
                    const double d2 = (gp[0]-image[0])*(gp[0]-image[0]) + (gp[1]-image[1])*(gp[1]-image[1]) + (gp[2]-image[2])*(gp[2]-image[2]);
                    if (d2 < best) { best = d2; bi = cidx; }
                }
                rows[3*q] = o->gh[bi].position[0]; rows[3*q+1] = o->gh[bi].position[1]; rows[3*q+2] = o->gh[bi].position[2];
                unk_g[q] = bi;
            }
        }
        double w[8];
        const double cond = oracle_vandermonde_weights(rows, s_i, bp, en, image, origin, h, w);
        if (cond > CONDN_LIM && s_i != -1) {
            for (int q = 0; q < 8; ++q) { w[q] = 0; r[q] = 0; known[q] = 0; unk_g[q] = -1; }
            weighted = pg;
        } else {
            ++n_coupled;
        }
        double rr = -weighted;
        int m = 0;
        tr[m].row = g; tr[m].col = g; tr[m].val = -1.0; ++m;
        for (int q = 0; q < 8; ++q) if (known[q]) rr -= w[q] * r[q];
        if (s_i != -1) rr -= w[s_i] * r[s_i];
        for (int q = 0; q < 8; ++q) if (unk_g[q] >= 0) { tr[m].row = g; tr[m].col = unk_g[q]; tr[m].val = w[q]; ++m; }
        ntr[g] = m;
        rhs[g] = rr;
    }
    o->n_coupled = n_coupled;
    /* assemble CSR (duplicates summed, as Eigen::setFromTriplets does) */
    int nt = 0;
    for (int g = 0; g < G; ++g) nt += ntr[g];
    trip_t *all = (trip_t *)malloc(sizeof(trip_t) * (nt ? nt : 1));
    nt = 0;
    for (int g = 0; g < G; ++g) for (int q = 0; q < ntr[g]; ++q) all[nt++] = trips[(size_t)g * 9 + q];
    qsort(all, nt, sizeof(trip_t), cmp_trip);
    int *rp = (int *)calloc(G + 1, sizeof(int)), *ci = (int *)malloc(sizeof(int) * (nt ? nt : 1));
    double *vv = (double *)malloc(sizeof(double) * (nt ? nt : 1));
    int m = 0;
    for (int q = 0; q < nt; ++q) {
        if (m > 0 && all[q].row == all[q-1].row && all[q].col == ci[m-1]) { vv[m-1] += all[q].val; continue; }
        ci[m] = all[q].col; vv[m] = all[q].val; rp[all[q].row + 1] = m + 1; ++m;
    }
    for (int r = 1; r <= G; ++r) if (rp[r] < rp[r-1]) rp[r] = rp[r-1];
    double *sol = (double *)malloc(sizeof(double) * G);
    bicgstab(G, rp, ci, vv, rhs, sol);
    for (int g = 0; g < G; ++g) o->gh[g].pressure = sol[g];
    /* the assembled system stays readable (oracle_ghost_system) until the next update */
    free(o->sys_rp); free(o->sys_ci); free(o->sys_vv); free(o->sys_rhs);
    o->sys_n = G; o->sys_nnz = m;
    o->sys_rp = rp; o->sys_ci = ci; o->sys_vv = vv; o->sys_rhs = rhs;
    free(sol); free(all); free(trips); free(ntr);
}

int oracle_ghost_count(const oracle *o) { return o->ngh; }

int oracle_ghost_update_at(oracle *o, const double *p, double t, double *val, double *pos, long long *nbr)
{
    const double keep = o->t;
    o->t = t;
    ghost_update(o, p);
    o->t = keep;
    for (int g = 0; g < o->ngh; ++g) {
        if (val) val[g] = o->gh[g].pressure;
        if (pos) for (int d = 0; d < 3; ++d) pos[3 * g + d] = o->gh[g].position[d];
        if (nbr) nbr[g] = o->gh[g].neighbour;
    }
    return o->ngh;
}

int oracle_ghost_system(const oracle *o, int *n, int *nnz, int *rp, int *ci, double *vv, double *rhs)
{
    if (n) *n = o->sys_n;
    if (nnz) *nnz = o->sys_nnz;
    if (!o->sys_rp) return -1;
    if (rp) memcpy(rp, o->sys_rp, sizeof(int) * (o->sys_n + 1));
    if (ci) memcpy(ci, o->sys_ci, sizeof(int) * o->sys_nnz);
    if (vv) memcpy(vv, o->sys_vv, sizeof(double) * o->sys_nnz);
    if (rhs) memcpy(rhs, o->sys_rhs, sizeof(double) * o->sys_n);
    return 0;
}

/* ======================================================================== */
/* MAC_Grid::PML_velocityUpdateCollocated (MAC_Grid.cpp:461-479)             */
/* ======================================================================== */
static void pml_velocity(oracle *o, const double *pFull)
{
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < o->npv; ++q) {
        const pml_vcell *c = &o->pv[q];
        double *v = o->v[c->dim];
        v[c->index] *= c->upd;
        v[c->index] += c->grad * pFull[c->pr];
        v[c->index] -= c->grad * pFull[c->pl];
    }
}

/* ======================================================================== */
/* MAC_Grid::pressureFieldLaplacianGhostCell (MAC_Grid.cpp:187-299)          */
/* ======================================================================== */
static void laplacian_ghost(oracle *o, const double *val, double *lap)
{
    const double scale = 1.0 / pow(o->h, 2);
    memset(lap, 0, sizeof(double) * o->N);
#pragma omp parallel for schedule(static)
    for (long long ci = 0; ci < o->N; ++ci) {
        if (!o->bulk[ci]) continue;
        int ijk[3]; pidx3(o, ci, ijk);
        int surrounded = 0;
        double acc = 0.0;
        for (int d = 0; d < 3; ++d) {
            int pos[3] = {ijk[0], ijk[1], ijk[2]}, neg[3] = {ijk[0], ijk[1], ijk[2]};
            if (ijk[d] != o->n[d] - 1) pos[d] += 1;
            if (ijk[d] != 0) neg[d] -= 1;
            const long long ip = pidx(o, pos[0], pos[1], pos[2]), in = pidx(o, neg[0], neg[1], neg[2]);
            const int gp = o->ghostc[ip], gn = o->ghostc[in];
            if (o->sc.settings.sampling != FDTD_SAMPLE_SEVEN) {
                /* ONLY_SAMPLE_CENTER (MAC_Grid.cpp:247-266): ghost values
                   with the plain weights */
                if (gp && gn) {
                    acc += (ghost_pressure(o, ci, ip) + ghost_pressure(o, ci, in) - 2.0 * val[ci]);
                    ++surrounded;
                } else if (gp && !gn) {
                    acc += (ghost_pressure(o, ci, ip) + val[in] - 2.0 * val[ci]);
                } else if (!gp && gn) {
                    acc += (val[ip] + ghost_pressure(o, ci, in) - 2.0 * val[ci]);
                } else {
                    acc += (val[ip] + val[in] - 2.0 * val[ci]);
                }
            } else if (gp && gn) {
                acc += (ghost_pressure(o, ci, ip) + ghost_pressure(o, ci, in) - 2.0 * val[ci]) / (9.0 / 16.0);
                ++surrounded;
            } else if (gp && !gn) {
                acc += (ghost_pressure(o, ci, ip) + 0.75 * val[in] - (7.0 / 4.0) * val[ci]) / (21.0 / 32.0);
            } else if (!gp && gn) {
                acc += (0.75 * val[ip] + ghost_pressure(o, ci, in) - (7.0 / 4.0) * val[ci]) / (21.0 / 32.0);
            } else {
                acc += (val[ip] + val[in] - 2.0 * val[ci]);
            }
        }
        o->cavity[ci] = (unsigned char)(surrounded > 1);
        lap[ci] = acc * scale;
    }
}

/* GaussianPressureSource::Evaluate (GaussianPressureSource.cpp:6-30) */
static double pressure_sources(const oracle *o, const double *x, double t)
{
    double acc = 0;
    if (o->evals_off) return 0.0;   /* FDTD_Objects::EvaluatePressureSources, _disableAllEvals */
    for (int i = 0; i < o->sc.n_pressure_sources; ++i) {
        const fdtd_pressure_source *s = &o->psrc[i];
        const double *mn = o->src_bmin[i], *mx = o->src_bmax[i];
        if (!(x[0] < mx[0] && x[0] > mn[0] && x[1] < mx[1] && x[1] > mn[1] && x[2] < mx[2] && x[2] > mn[2])) continue;
        const double *sp = o->src_pos[i];
        double value = (x[0]-sp[0])*(x[0]-sp[0]) + (x[1]-sp[1])*(x[1]-sp[1]) + (x[2]-sp[2])*(x[2]-sp[2]);
        value = -value / (2.0 * s->width_space * s->width_space);
        value = exp(value);
        value *= exp(-pow((t - s->offset_time) / s->width_time, 2) / 2.0);
        value *= 1. / 3.6836 * s->normalize_constant;
        acc += value;
    }
    /* SpeakerPressureSource::Evaluate (SpeakerPressureSource.cpp:91-113): the
       bounding box is never set for this source type, so every cell counts;
       WavReader::SampleRate is the sample period */
    for (int i = 0; i < o->sc.n_speaker_sources; ++i) {
        const fdtd_speaker_source *sp = &o->spk[i];
        const double *q = sp->position;
        double value = pow(x[0] - q[0], 2.0) + pow(x[1] - q[1], 2.0) + pow(x[2] - q[2], 2.0);
        const double r = sqrt(value);
        value = -value / (2.0 * pow(sp->width_space, 2.0));
        value = exp(value);
        const double delay = t - r / o->c;
        const int idx = (int)((delay - sp->start_time) / (1.0 / sp->sample_rate));
        if (idx >= 0 && idx < sp->n_samples) value *= sp->samples[idx];
        else value = 0.0;
        acc += value;
    }
    return acc;
}

/* ======================================================================== */
/* The Engquist-Majda absorbing boundary of the layer-1 cells that no PML    */
/* covers (MAC_Grid.cpp:521-919; ENGQUIST_ORDER 2, MAC_Grid.cpp:33): order 2 */
/* on faces, order 1 on edges and corners.  The extrapolated pressures of    */
/* the exterior (layer-0) neighbours are written into pCurr, where the next  */
/* step reads them back as pLast; edges and corners also zero the diagonal   */
/* exterior cell.  Returns p^{n+1} of the cell.  No interface blending (one  */
/* simulation unit) and no ghost cells next to layer 1 (the caller checks).  */
/* ======================================================================== */
static long long abc_at(const oracle *o, const int *ijk, int da, int a, int db, int b, int dc, int c)
{
    int q[3] = {ijk[0], ijk[1], ijk[2]};
    q[da] += a; q[db] += b; q[dc] += c;
    return pidx(o, q[0], q[1], q[2]);
}

static double engquist(const oracle *o, const int *ijk, const int *nml, const double *pLast, double *pCurr)
{
    const double lam = o->c * o->dt / o->h, lam2 = lam * lam;
    const long long ci = pidx(o, ijk[0], ijk[1], ijk[2]);
    const double pc = pCurr[ci], pl = pLast[ci];
    const int btype = abs(nml[0]) + abs(nml[1]) + abs(nml[2]);
    if (btype == 1) {
        int di = 0;
        while (!nml[di]) ++di;
        const int dj = (di + 1) % 3, dk = (di + 2) % 3, s = nml[di];
        const int jm = ijk[dj] > 0 ? -1 : 0, jp = ijk[dj] < o->n[dj] - 1 ? 1 : 0;
        const int km = ijk[dk] > 0 ? -1 : 0, kp = ijk[dk] < o->n[dk] - 1 ? 1 : 0;
        const double pl_out = pLast[abc_at(o, ijk, di, s, dj, 0, dk, 0)];
        const double p_in = pCurr[abc_at(o, ijk, di, -s, dj, 0, dk, 0)];
        const double pl_in = pLast[abc_at(o, ijk, di, -s, dj, 0, dk, 0)];
        const double tang = pCurr[abc_at(o, ijk, di, 0, dj, jp, dk, 0)] + pCurr[abc_at(o, ijk, di, 0, dj, jm, dk, 0)] +
                            pCurr[abc_at(o, ijk, di, 0, dj, 0, dk, kp)] + pCurr[abc_at(o, ijk, di, 0, dj, 0, dk, km)];
        double p1 = (2. / lam2 + 4. / lam - 6. - 4. * lam) * pc - (1. / lam2 + 2. / lam) * pl + pl_out - pl_in + 2. * p_in +
                    (lam + 1.) * tang;
        p1 *= lam2 / (1. + 2. * lam);
        const double p_out = pl_out + p_in - pl_in - 2. / lam * (p1 + pl - 2. * pc) + lam * (tang - 4. * pc);
        pCurr[abc_at(o, ijk, di, s, dj, 0, dk, 0)] = p_out;
        return 2.0 * pc - pl + lam2 * (p_in + tang + p_out - 6. * pc);
    }
    if (btype == 2) {
        int di = 0;
        while (!nml[di]) ++di;
        int dj = di + 1;
        while (!nml[dj]) ++dj;
        const int dk = (dj != di + 1) ? (di + 1) % 3 : (di + 2) % 3;
        const int si = nml[di], sj = nml[dj];
        const int km = ijk[dk] > 0 ? -1 : 0, kp = ijk[dk] < o->n[dk] - 1 ? 1 : 0;
        const double pi_in = pCurr[abc_at(o, ijk, di, -si, dj, 0, dk, 0)];
        const double pj_in = pCurr[abc_at(o, ijk, di, 0, dj, -sj, dk, 0)];
        const double pk = pCurr[abc_at(o, ijk, di, 0, dj, 0, dk, kp)] + pCurr[abc_at(o, ijk, di, 0, dj, 0, dk, km)];
        double p1 = (2. / lam2 - 6.) * pc + (2.0 * lam - 1.0) / lam2 * pl + 2. * pi_in + 2. * pj_in + pk;
        p1 *= lam2 / (1. + 2. * lam);
        const double pi_out = (pl - p1) / lam + pi_in, pj_out = (pl - p1) / lam + pj_in;
        pCurr[abc_at(o, ijk, di, si, dj, 0, dk, 0)] = pi_out;
        pCurr[abc_at(o, ijk, di, 0, dj, sj, dk, 0)] = pj_out;
        pCurr[abc_at(o, ijk, di, si, dj, sj, dk, 0)] = 0.0;
        return 2.0 * pc - pl + lam2 * (pi_in + pj_in + pk + pi_out + pj_out - 6. * pc);
    }
    /* corner */
    const double pi_in = pCurr[abc_at(o, ijk, 0, -nml[0], 1, 0, 2, 0)];
    const double pj_in = pCurr[abc_at(o, ijk, 0, 0, 1, -nml[1], 2, 0)];
    const double pk_in = pCurr[abc_at(o, ijk, 0, 0, 1, 0, 2, -nml[2])];
    double p1 = (2. / lam2 - 6.) * pc + (3.0 * lam - 1.0) / lam2 * pl + 2. * pi_in + 2. * pj_in + 2. * pk_in;
    p1 *= lam2 / (1. + 3. * lam);
    const double pi_out = (pl - p1) / lam + pi_in, pj_out = (pl - p1) / lam + pj_in, pk_out = (pl - p1) / lam + pk_in;
    pCurr[abc_at(o, ijk, 0, nml[0], 1, 0, 2, 0)] = pi_out;
    pCurr[abc_at(o, ijk, 0, 0, 1, nml[1], 2, 0)] = pj_out;
    pCurr[abc_at(o, ijk, 0, 0, 1, 0, 2, nml[2])] = pk_out;
    pCurr[abc_at(o, ijk, 0, nml[0], 1, nml[1], 2, nml[2])] = 0.0;
    return 2.0 * pc - pl + lam2 * (pi_in + pj_in + pk_in + pi_out + pj_out + pk_out - 6. * pc);
}

/* ======================================================================== */
/* MAC_Grid::PML_pressureUpdateCollocated (MAC_Grid.cpp:481-946), PML slab   */
/* plus the non-PML bulk cells, the Engquist boundary at layer 1 included    */
/* ======================================================================== */
static void pressure_update(oracle *o, double *pLast, double *pCurr, double *pNext, const double *lapCurr, const double *lapLast)
{
    const fdtd_settings *s = &o->sc.settings;
    const double dt = o->dt;
    const double one_over_dx = 1.0 / o->h;
    const double c2_k2 = pow(o->c * dt, 2);
    const double kcalp = o->c * s->air_viscosity_alpha * dt;
    const int visc = s->air_viscosity_alpha > 0.0;
    for (long long q = 0; q < o->npp; ++q) {
        const pml_pcell *c = &o->pp[q];
        const int ci = c->index;
        for (int d = 0; d < 3; ++d) {
            o->pdir[d][ci] *= c->upd[d];
            o->pdir[d][ci] += c->div[d] * o->v[d][c->vr[d]] * one_over_dx;
            o->pdir[d][ci] -= c->div[d] * o->v[d][c->vl[d]] * one_over_dx;
        }
        pNext[ci] = o->pdir[0][ci] + o->pdir[1][ci] + o->pdir[2][ci];
    }
    const int has_src = o->sc.n_pressure_sources > 0 || o->sc.n_speaker_sources > 0;
    const double tsrc = o->t + 0.5 * dt;
#pragma omp parallel for schedule(static)
    for (long long ci = 0; ci < o->N; ++ci) {
        if (!o->bulk[ci] || o->pmlc[ci]) continue;
        int ijk[3]; pidx3(o, ci, ijk);
        if (ijk[0] == 0 || ijk[1] == 0 || ijk[2] == 0 || ijk[0] == o->n[0]-1 || ijk[1] == o->n[1]-1 || ijk[2] == o->n[2]-1)
            continue;  /* layer 0: skipped (MAC_Grid.cpp:524-526) */
        int nml[3];   /* ScalarField::boundaryCellNormal(cell, 1) (ScalarField.h:202-213) */
        for (int d = 0; d < 3; ++d) nml[d] = ijk[d] == 1 ? -1 : ijk[d] == o->n[d] - 2 ? 1 : 0;
        if (nml[0] || nml[1] || nml[2])
            pNext[ci] = engquist(o, ijk, nml, pLast, pCurr);
        else if (o->cavity[ci] && o->history[ci])
            pNext[ci] = pLast[ci] = pCurr[ci];
        else if (visc && o->history[ci])
            pNext[ci] = 2.0 * pCurr[ci] - pLast[ci] - lapLast[ci] * kcalp + lapCurr[ci] * (c2_k2 + kcalp);
        else
            pNext[ci] = 2.0 * pCurr[ci] - pLast[ci] + lapCurr[ci] * c2_k2;
        if (has_src) {
            double x[3]; ppos(o, ijk, x);
            pNext[ci] += pressure_sources(o, x, tsrc) * dt;
        }
    }
}

/* ======================================================================== */
/* PML_WaveSolver::FetchScalarData (PML_WaveSolver.cpp:210-317)              */
/* ======================================================================== */
static void fetch_listeners(const oracle *o, const double *scalar, double *out)
{
    const int L = o->sc.n_listeners;
#pragma omp parallel for schedule(static)
    for (int li = 0; li < L; ++li) {
        const double *p = &o->listeners[3*li];
        long long nb[8];
        const int nn = enclosing(o, p, nb);
        double pts[3*256], vals[256];
        int np = 0, kept = 0;
        long long keptidx[8];
        for (int q = 0; q < nn; ++q) {
            const long long c = nb[q];
            if (isnan(scalar[c])) continue;
            keptidx[kept++] = c;
            if (!o->ghostc[c]) {
                int ijk[3]; pidx3(o, c, ijk); ppos(o, ijk, &pts[3*np]);
                vals[np++] = scalar[c];
            } else {
                int first; const int cnt = find_children(o, (int)c, &first);
                for (int g = first; g < first + cnt && np < 256; ++g) {
                    pts[3*np] = o->gh[g].position[0]; pts[3*np+1] = o->gh[g].position[1]; pts[3*np+2] = o->gh[g].position[2];
                    vals[np++] = o->gh[g].pressure;
                }
            }
        }
        if (kept < 1) { out[li] = NAN; continue; }
        if (kept < 4) {
            long long best = -1; double bd = DBL_MAX;
            for (int q = 0; q < kept; ++q) {
                int ijk[3]; double x[3]; pidx3(o, keptidx[q], ijk); ppos(o, ijk, x);
                const double d2 = (x[0]-p[0])*(x[0]-p[0]) + (x[1]-p[1])*(x[1]-p[1]) + (x[2]-p[2])*(x[2]-p[2]);
                if (d2 < bd) { bd = d2; best = keptidx[q]; }
            }
            out[li] = scalar[best];
        } else {
            out[li] = oracle_mls_lookup(p, np, pts, vals);
        }
    }
}

/* ======================================================================== */
/* staggered MAC leapfrog (PML_WaveSolver::stepLeapfrog, 710-773), free      */
/* field: every cell and every velocity face is bulk                         */
/* ======================================================================== */

/* GaussianPressureSource::Evaluate with USE_COLLOCATED undefined
 * (GaussianPressureSource.cpp:6-30): the time factor is the pulse integrated
 * from t = 0, erf((t - t0)/(sqrt2 wt)) - erf(-t0/(sqrt2 wt)), and the
 * amplitude is the normalisation constant itself. */
static double pressure_sources_mac(const oracle *o, const double *x, double t)
{
    if (o->evals_off) return 0.0;
    const double sqrt_2 = sqrt(2.0);   /* Wavesolver_ConstantsAndTypes.h:75 */
    double acc = 0;
    for (int i = 0; i < o->sc.n_pressure_sources; ++i) {
        const fdtd_pressure_source *s = &o->psrc[i];
        const double *mn = o->src_bmin[i], *mx = o->src_bmax[i];
        if (!(x[0] < mx[0] && x[0] > mn[0] && x[1] < mx[1] && x[1] > mn[1] && x[2] < mx[2] && x[2] > mn[2])) continue;
        const double *sp = o->src_pos[i];
        double value = (x[0]-sp[0])*(x[0]-sp[0]) + (x[1]-sp[1])*(x[1]-sp[1]) + (x[2]-sp[2])*(x[2]-sp[2]);
        value = -value / (2.0 * s->width_space * s->width_space);
        value = exp(value);
        value *= erf((t - s->offset_time) / (sqrt_2 * s->width_time)) - erf(-s->offset_time / (sqrt_2 * s->width_time));
        value *= s->normalize_constant;
        acc += value;
    }
    return acc;
}

/* MAC_Grid::PML_velocityUpdate (MAC_Grid.cpp:301-370), bulk faces: zero on
 * the two walls of each axis; in the PML (sigma_d > 1e-12 at the face)
 * v = f1 v + f2 (p_R - p_L), elsewhere v += -dt/(h rho) (p_R - p_L). */
static void mac_velocity(oracle *o)
{
    const fdtd_settings *s = &o->sc.settings;
    const double *p = o->pfull;
    const double ndt = -o->dt / (o->h * o->rho);
    for (int d = 0; d < 3; ++d) {
        const int nv[3] = {o->n[0] + (d == 0), o->n[1] + (d == 1), o->n[2] + (d == 2)};
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nv[0]; ++i)
            for (int j = 0; j < nv[1]; ++j)
                for (int k = 0; k < nv[2]; ++k) {
                    int ijk[3] = {i, j, k};
                    double *v = &o->v[d][vidx(o, d, i, j, k)];
                    if (ijk[d] == 0 || ijk[d] == nv[d] - 1) { *v = 0.0; continue; }
                    double x[3];
                    vpos(o, d, ijk, x);
                    const double sg = pml_sigma(s, x, o->W, d);
                    const long long right = pidx(o, i, j, k);
                    ijk[d] -= 1;
                    const long long left = pidx(o, ijk[0], ijk[1], ijk[2]);
                    if (sg > 1e-12) {
                        const double upd = f_vel_update(sg, o->dt), grad = f_grad(sg, o->dt, o->h, o->rho);
                        *v *= upd;
                        *v += grad * p[right];
                        *v -= grad * p[left];
                    } else {
                        *v += ndt * p[right];
                        *v -= ndt * p[left];
                    }
                }
    }
}

/* MAC_Grid::PML_pressureUpdate (MAC_Grid.cpp:948-1020) for each axis, then
 * MAC_Grid::UpdatePMLPressure (1102-1129): pFull = pDir0 + pDir1 + pDir2.
 * The pressure source is added to pDir_d of every axis d whose sigma_d is
 * zero at the cell. */
static void mac_pressure(oracle *o)
{
    const fdtd_settings *s = &o->sc.settings;
    const double nrc = -o->rho * o->c * o->c * o->dt;
    const double hinv = 1.0 / o->h;
    const int has_src = o->sc.n_pressure_sources > 0 || o->sc.n_speaker_sources > 0;
    for (int d = 0; d < 3; ++d) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < o->n[0]; ++i)
            for (int j = 0; j < o->n[1]; ++j)
                for (int k = 0; k < o->n[2]; ++k) {
                    int ijk[3] = {i, j, k};
                    double x[3];
                    ppos(o, ijk, x);
                    const double sg = pml_sigma(s, x, o->W, d);
                    double *pd = &o->pdir[d][pidx(o, i, j, k)];
                    const long long left = vidx(o, d, i, j, k);
                    ijk[d] += 1;
                    const long long right = vidx(o, d, ijk[0], ijk[1], ijk[2]);
                    if (sg > 1e-12) {
                        const double D = f_dir(sg, o->dt), upd = f_pupd(sg, o->dt, D), div = f_div(o->rho, o->c, D);
                        *pd *= upd;
                        *pd += div * o->v[d][right] / o->h;
                        *pd -= div * o->v[d][left] / o->h;
                    } else {
                        *pd += nrc * o->v[d][right] * hinv;
                        *pd -= nrc * o->v[d][left] * hinv;
                        if (has_src) *pd += pressure_sources_mac(o, x, o->t + 0.5 * o->dt) * o->dt;
                    }
                }
    }
    #pragma omp parallel for schedule(static)
    for (long long c = 0; c < o->N; ++c) o->pfull[c] = o->pdir[0][c] + o->pdir[1][c] + o->pdir[2][c];
}

/* ======================================================================== */
/* construction (PML_WaveSolver::Reinitialize_PML_WaveSolver, 102-146)       */
/* ======================================================================== */
oracle *oracle_create(const fdtd_scene *scene, int reclassify_each_step)
{
    const fdtd_settings *s = &scene->settings;
    if (s->divisions[0] < 4 || s->divisions[1] < 4 || s->divisions[2] < 4) { set_err("grid too small"); return NULL; }
    if (!s->use_mesh && scene->n_objects > 0) {
        set_err("use_mesh=0 leaves objects unclassified; the restatement takes free-field scenes there");
        return NULL;
    }
    if (scene->n_pressure_sources > 64) { set_err("too many pressure sources"); return NULL; }
    oracle *o = (oracle *)calloc(1, sizeof(oracle));
    o->sc = *scene;
    o->objs = (fdtd_object *)malloc(sizeof(fdtd_object) * (scene->n_objects + 1));
    memcpy(o->objs, scene->objects, sizeof(fdtd_object) * scene->n_objects);
    o->isrc = (fdtd_impulse_source *)malloc(sizeof(fdtd_impulse_source) * (scene->n_impulse_sources + 1));
    memcpy(o->isrc, scene->impulse_sources, sizeof(fdtd_impulse_source) * scene->n_impulse_sources);
    o->ans = (fdtd_acc_noise_source *)calloc((size_t)scene->n_acc_noise_sources + 1, sizeof(fdtd_acc_noise_source));
    for (int i = 0; i < scene->n_acc_noise_sources; ++i) {
        o->ans[i] = scene->acc_noise_sources[i];
        fdtd_impact *copy = (fdtd_impact *)malloc(sizeof(fdtd_impact) * ((size_t)o->ans[i].n_impacts + 1));
        memcpy(copy, scene->acc_noise_sources[i].impacts, sizeof(fdtd_impact) * (size_t)o->ans[i].n_impacts);
        o->ans[i].impacts = copy;
    }
    o->sc.acc_noise_sources = o->ans;
    o->psrc = (fdtd_pressure_source *)malloc(sizeof(fdtd_pressure_source) * (scene->n_pressure_sources + 1));
    memcpy(o->psrc, scene->pressure_sources, sizeof(fdtd_pressure_source) * scene->n_pressure_sources);
    o->listeners = (double *)malloc(sizeof(double) * 3 * (scene->n_listeners + 1));
    memcpy(o->listeners, scene->listeners, sizeof(double) * 3 * scene->n_listeners);
    o->sc.objects = o->objs; o->sc.impulse_sources = o->isrc; o->sc.pressure_sources = o->psrc; o->sc.listeners = o->listeners;
    o->spk = (fdtd_speaker_source *)calloc((size_t)scene->n_speaker_sources + 1, sizeof(fdtd_speaker_source));
    for (int i = 0; i < scene->n_speaker_sources; ++i) {
        o->spk[i] = scene->speaker_sources[i];
        double *copy = (double *)malloc(sizeof(double) * ((size_t)o->spk[i].n_samples + 1));
        memcpy(copy, scene->speaker_sources[i].samples, sizeof(double) * (size_t)o->spk[i].n_samples);
        o->spk[i].samples = copy;
    }
    o->sc.speaker_sources = o->spk;
    if (scene->n_constraints < 0 || (scene->n_constraints > 0 && !scene->constraints)) { set_err("bad constraint list"); oracle_free(o); return NULL; }
    for (int i = 0; i < scene->n_constraints; ++i)
        if (scene->constraints[i].direction < 0 || scene->constraints[i].direction > 2 || (scene->constraints[i].sign != 1 && scene->constraints[i].sign != -1)) {
            set_err("constraint direction must be 0, 1 or 2 and sign +-1"); oracle_free(o); return NULL;
        }
    if (scene->n_constraints > 0 && (!s->use_mesh || s->scheme != FDTD_SCHEME_COLLOCATED)) {
        set_err("plane constraints are classified with use_mesh = 1 in the collocated scheme"); oracle_free(o); return NULL;
    }
    o->cons = (fdtd_plane_constraint *)malloc(sizeof(fdtd_plane_constraint) * ((size_t)scene->n_constraints + 1));
    if (scene->n_constraints) memcpy(o->cons, scene->constraints, sizeof(fdtd_plane_constraint) * (size_t)scene->n_constraints);
    o->sc.constraints = o->cons;
    o->objs0 = (fdtd_object *)malloc(sizeof(fdtd_object) * (scene->n_objects + 1));
    memcpy(o->objs0, scene->objects, sizeof(fdtd_object) * scene->n_objects);
    for (int i = 0; i < scene->n_objects; ++i)
        for (int d = 0; d < 3; ++d)
            o->moving |= scene->objects[i].velocity[d] != 0.0 || scene->objects[i].angular_velocity[d] != 0.0;
    o->kin = (fdtd_kinematics *)calloc((size_t)scene->n_kinematics + 1, sizeof(fdtd_kinematics));
    for (int i = 0; i < scene->n_kinematics; ++i) {
        o->kin[i] = scene->kinematics[i];
        double *copy = (double *)malloc(sizeof(double) * 7 * ((size_t)o->kin[i].n_frames + 1));
        memcpy(copy, scene->kinematics[i].frames, sizeof(double) * 7 * (size_t)o->kin[i].n_frames);
        o->kin[i].frames = copy;
        o->moving = 1;
    }
    o->sc.kinematics = o->kin;
    if (o->moving) animate_objects(o, s->start_time);
    for (int i = 0; i < scene->n_pressure_sources; ++i) {
        /* GaussianPressureSource ctor: bbox = pos -/+ ws*2, position = bbox centroid */
        for (int d = 0; d < 3; ++d) {
            o->src_bmin[i][d] = o->psrc[i].position[d] - o->psrc[i].width_space * GAUSSIAN_CHECK_BOUND;
            o->src_bmax[i][d] = o->psrc[i].position[d] + o->psrc[i].width_space * GAUSSIAN_CHECK_BOUND;
            o->src_pos[i][d] = (o->src_bmax[i][d] + o->src_bmin[i][d]) / 2.0;
        }
    }
    for (int d = 0; d < 3; ++d) o->n[d] = s->divisions[d];
    o->N = (long long)o->n[0] * o->n[1] * o->n[2];
    for (int d = 0; d < 3; ++d) o->Nv[d] = (long long)(o->n[0] + (d==0)) * (o->n[1] + (d==1)) * (o->n[2] + (d==2));
    o->h = s->cell_size; o->dt = s->time_step; o->c = s->sound_speed; o->rho = s->density;
    o->W = s->pml_width * s->cell_size;   /* MAC_Grid::setPMLBoundaryWidth */
    pressure_bbox(s, o->bmin, o->bmax);
    for (int r = 0; r < 3; ++r) o->pc[r] = (double *)calloc(o->N, sizeof(double));
    for (int r = 0; r < 2; ++r) o->lap[r] = (double *)calloc(o->N, sizeof(double));
    for (int d = 0; d < 3; ++d) o->pdir[d] = (double *)calloc(o->N, sizeof(double));
    for (int d = 0; d < 3; ++d) o->v[d] = (double *)calloc(o->Nv[d], sizeof(double));
    o->bulk = (unsigned char *)malloc(o->N); o->ghostc = (unsigned char *)calloc(o->N, 1);
    o->pmlc = (unsigned char *)calloc(o->N, 1); o->cavity = (unsigned char *)calloc(o->N, 1);
    o->history = (unsigned char *)malloc(o->N); o->classified = (unsigned char *)calloc(o->N, 1);
    o->pp = (pml_pcell *)malloc(sizeof(pml_pcell) * o->N);
    o->pv = (pml_vcell *)malloc(sizeof(pml_vcell) * (o->Nv[0] + o->Nv[1] + o->Nv[2]));
    o->pc_ind = 1; o->lap_ind = 1;
    o->t = s->start_time;
    o->evals_off = 0;
    update_evals(o);
    o->reclass = reclassify_each_step || o->moving;
    o->scheme = s->scheme;
    if (o->scheme == FDTD_SCHEME_STAGGERED) {
        if (scene->n_objects > 0) {
            set_err("the staggered scheme is restated for free-field scenes only");
            oracle_free(o);
            return NULL;
        }
        o->pfull = (double *)calloc(o->N, sizeof(double));
    }
    classify_static(o);
    memset(o->history, 1, (size_t)o->N);
    classify_fast(o);
    memset(o->history, 1, (size_t)o->N);   /* ResetCellHistory(true) */
    return o;
}

void oracle_free(oracle *o)
{
    if (!o) return;
    for (int r = 0; r < 3; ++r) free(o->pc[r]);
    for (int r = 0; r < 2; ++r) free(o->lap[r]);
    for (int d = 0; d < 3; ++d) { free(o->pdir[d]); free(o->v[d]); }
    free(o->bulk); free(o->ghostc); free(o->pmlc); free(o->cavity); free(o->history); free(o->classified);
    free(o->pp); free(o->pv); free(o->gh); free(o->hkeys); free(o->hvals);
    for (int i = 0; o->spk && i < o->sc.n_speaker_sources; ++i) free((void *)o->spk[i].samples);
    free(o->spk);
    for (int i = 0; o->ans && i < o->sc.n_acc_noise_sources; ++i) free((void *)o->ans[i].impacts);
    for (int i = 0; o->kin && i < o->sc.n_kinematics; ++i) free((void *)o->kin[i].frames);
    free(o->kin);
    free(o->ans);
    free(o->objs); free(o->objs0); free(o->isrc); free(o->psrc); free(o->listeners);
    free(o->pfull);
    free(o->sys_rp); free(o->sys_ci); free(o->sys_vv); free(o->sys_rhs);
    free(o->cons); free(o->solid);
    free(o);
}

/* Rigid motion with constant velocity v and angular velocity w about the
 * centre, from the pose at t = 0: c(t) = c0 + v t, R(t) = Rot(w/|w|, |w| t) R0
 * (Rodrigues' formula); stands in for FDTD_Objects::AnimateObjects with the
 * animator's GetRigidObjectTransform (FDTD_Objects.cpp:565-582). */
/* Quaternion<double> (w, v) of src/linearalgebra/Quaternion.hpp */
typedef struct { double w, x, y, z; } quat;

static quat quat_mul(quat a, quat b)   /* operator* (Quaternion.hpp:82-89) */
{
    quat r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    return r;
}

static void quat_norm(quat *q)   /* normalize (Quaternion.hpp:186-195, EPS 1e-12) */
{
    double len = q->w * q->w + q->x * q->x + q->y * q->y + q->z * q->z;
    if (fabs(len - 1.0) > 1e-12) {
        len = 1.0 / sqrt(len);
        q->w *= len; q->x *= len; q->y *= len; q->z *= len;
    }
}

static quat quat_slerp(quat a, double r, quat b)   /* slerp (Quaternion.hpp:365-420) */
{
    quat v0 = a, v1 = b;
    quat_norm(&v0);
    quat_norm(&v1);
    double dot = v0.w * v1.w + v0.x * v1.x + v0.y * v1.y + v0.z * v1.z;
    const double DOT_THRESHOLD = 0.9995;
    if (fabs(dot) > DOT_THRESHOLD) {
        quat ret = {v0.w + (v1.w - v0.w) * r, v0.x + (v1.x - v0.x) * r, v0.y + (v1.y - v0.y) * r, v0.z + (v1.z - v0.z) * r};
        quat_norm(&ret);
        return ret;
    }
    if (dot < 0.0) {
        v1.w = -v1.w; v1.x = -v1.x; v1.y = -v1.y; v1.z = -v1.z;
        dot = -dot;
    }
    if (dot < -1.0) dot = -1.0;
    if (dot > 1.0) dot = 1.0;
    const double theta_0 = acos(dot), theta = theta_0 * r;
    quat v2 = {v1.w - v0.w * dot, v1.x - v0.x * dot, v1.y - v0.y * dot, v1.z - v0.z * dot};
    quat_norm(&v2);
    quat ret = {v0.w * cos(theta) + v2.w * sin(theta), v0.x * cos(theta) + v2.x * sin(theta),
                v0.y * cos(theta) + v2.y * sin(theta), v0.z * cos(theta) + v2.z * sin(theta)};
    return ret;
}

/* frame f of a kinematics file as ReadObjectKinData keeps it
 * (FDTD_RigidObject_Animator.cpp:98-118): w clamped to [-1, 1], then the
 * Blender z-up rotation turned by -90 degrees about x, q * fromAxisRotD(x, -90) */
static void kin_frame(const fdtd_kinematics *k, int f, double *c, quat *q)
{
    const double *r = k->frames + 7 * (size_t)f;
    for (int d = 0; d < 3; ++d) c[d] = r[d];
    double w = r[3];
    w = fmax(-1.0, fmin(1.0, w));
    const quat file = {w, r[4], r[5], r[6]};
    const double rad = -90.0 * M_PI / 180.0;
    const quat corr = {cos(rad / 2), sin(rad / 2), 0.0, 0.0};   /* fromAxisRotD (Quaternion.hpp:202-208) */
    *q = quat_mul(file, corr);
}

/* pose of a keyframed object at time t: GetRigidObjectTransform
 * (FDTD_RigidObject_Animator.cpp:137-160, frames step apart from time 0),
 * the normalize of FDTD_Objects::AnimateObjects (FDTD_Objects.cpp:563-579)
 * and SetRigidBodyTransform (FDTD_RigidObject.cpp:550-566: angle 2 acos(w)
 * about the normalized vector part, Eigen's angle-axis matrix; w clamped to
 * [-1, 1] for acos) */
static void kin_pose(const fdtd_kinematics *k, double t, fdtd_object *b)
{
    const int n = k->n_frames;
    const double pos = t / k->step_size;
    int idx_0 = pos > 0.0 ? (int)pos : 0;
    if (idx_0 > n - 1) idx_0 = n - 1;
    const int idx_1 = idx_0 + 1 < n - 1 ? idx_0 + 1 : n - 1;
    const double t_0 = idx_0 * k->step_size, t_1 = idx_1 * k->step_size;
    const double alpha = (idx_0 == idx_1 ? 1.0 : (t - t_0) / (t_1 - t_0));
    double c0[3], c1[3];
    quat q0, q1;
    kin_frame(k, idx_0, c0, &q0);
    kin_frame(k, idx_1, c1, &q1);
    for (int d = 0; d < 3; ++d) b->center[d] = c0[d] * (1.0 - alpha) + c1[d] * alpha;
    quat q = quat_slerp(q0, alpha, q1);
    quat_norm(&q);
    const double rad = 2.0 * acos(fmax(-1.0, fmin(1.0, q.w)));
    double ax[3] = {0.0, 1.0, 0.0};
    if (!(rad < 1e-12)) {
        const double l = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
        ax[0] = q.x / l; ax[1] = q.y / l; ax[2] = q.z / l;
    }
    const double cs = cos(rad), sn = sin(rad);
    /* Eigen AngleAxis::toRotationMatrix: cos I + sin [a]x + (1 - cos) a a^T */
    const double cross[9] = {0, -ax[2], ax[1], ax[2], 0, -ax[0], -ax[1], ax[0], 0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            b->rotation[r * 3 + c] = (r == c ? cs : 0.0) + sn * cross[r * 3 + c] + (1.0 - cs) * ax[r] * ax[c];
}

static void animate_objects(oracle *o, double t)
{
    for (int i = 0; i < o->sc.n_objects; ++i) {
        const fdtd_object *a = &o->objs0[i];
        fdtd_object *b = &o->objs[i];
        int keyed = 0;
        for (int q = 0; q < o->sc.n_kinematics; ++q)
            if (o->kin[q].object == i) { kin_pose(&o->kin[q], t, b); keyed = 1; }
        if (keyed) continue;
        for (int d = 0; d < 3; ++d) b->center[d] = a->center[d] + a->velocity[d] * t;
        const double *w = a->angular_velocity;
        const double wn = sqrt(w[0]*w[0] + w[1]*w[1] + w[2]*w[2]);
        if (wn == 0.0) { memcpy(b->rotation, a->rotation, sizeof(b->rotation)); continue; }
        const double k[3] = {w[0]/wn, w[1]/wn, w[2]/wn}, th = wn * t, sn = sin(th), cs = 1.0 - cos(th);
        const double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
        double Rt[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double k2 = 0;
                for (int m = 0; m < 3; ++m) k2 += K[r*3+m] * K[m*3+c];
                Rt[r*3+c] = (r == c ? 1.0 : 0.0) + sn * K[r*3+c] + cs * k2;
            }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double v = 0;
                for (int m = 0; m < 3; ++m) v += Rt[r*3+m] * a->rotation[m*3+c];
                b->rotation[r*3+c] = v;
            }
    }
}

/* pose of a keyframed object at time t (test hook of kin_pose) */
void oracle_keyframe_pose(const fdtd_kinematics *k, double t, double *center, double *rotation)
{
    fdtd_object b;
    memset(&b, 0, sizeof b);
    kin_pose(k, t, &b);
    memcpy(center, b.center, sizeof b.center);
    memcpy(rotation, b.rotation, sizeof b.rotation);
}

/* pose of a moving object at time t (test hook of animate_objects) */
void oracle_object_pose(const fdtd_object *obj, double t, fdtd_object *out)
{
    oracle tmp;
    memset(&tmp, 0, sizeof tmp);
    tmp.sc.n_objects = 1;
    tmp.objs0 = (fdtd_object *)obj;
    *out = *obj;
    tmp.objs = out;
    animate_objects(&tmp, t);
}

/* MAC_Grid::InterpolateFreshPressureCell (MAC_Grid.cpp:1993-2254), its MLS
 * branch: a bulk cell without valid history (solid in the previous step)
 * takes the order-1 MLS value at its image point across the nearest object's
 * boundary, over the enclosing cells with valid history, plus rho a_n |d|
 * (d the cell's distance to the boundary, a_n the boundary acceleration along
 * the erected normal); fewer than four valid cells: the one nearest the image
 * point, else the first bulk axis neighbour (+x, -x, +y, ...), else zero.
 * Cells are visited in index order, as in the reference. */
static void fresh_fill(oracle *o, double *p)
{
    if (o->sc.n_objects == 0) return;
    for (long long ci = 0; ci < o->N; ++ci) {
        if (o->history[ci] || !o->bulk[ci]) continue;
        int ijk[3]; double x[3];
        pidx3(o, ci, ijk); ppos(o, ijk, x);
        double dist; int id;
        lowest_object_distance(o, x, &dist, &id);
        if (dist < 0.0) { o->fresh_error = 1; continue; }   /* DISTANCE_TOLERANCE = 0: the reference throws */
        double image[3], bp[3], en[3], nb[3];
        const double travelled = reflect_normal(&o->objs[id], x, image, bp, en, nb);
        long long cells[8];
        const int ne = enclosing(o, image, cells);
        long long valid[8]; int nv = 0;
        double pts[24], vals[8];
        for (int q = 0; q < ne; ++q) {
            if (!o->history[cells[q]]) continue;
            int cq[3]; pidx3(o, cells[q], cq); ppos(o, cq, &pts[3*nv]);
            vals[nv] = p[cells[q]];
            valid[nv++] = cells[q];
        }
        if (nv < 4) {
            if (nv > 0) {
                double bd = DBL_MAX; long long best = -1;
                for (int q = 0; q < nv; ++q) {
                    const double *y = &pts[3*q];
                    const double d2 = (image[0]-y[0])*(image[0]-y[0]) + (image[1]-y[1])*(image[1]-y[1]) + (image[2]-y[2])*(image[2]-y[2]);
                    if (d2 < bd) { bd = d2; best = valid[q]; }
                }
                p[ci] = p[best];
            } else {
                long long best = -1;
                for (int d = 0; d < 3 && best < 0; ++d)
                    for (int sg = 1; sg >= -1 && best < 0; sg -= 2) {
                        int t3[3] = {ijk[0], ijk[1], ijk[2]};
                        t3[d] += sg;
                        const long long nbc = pidx(o, t3[0], t3[1], t3[2]);
                        if (o->bulk[nbc]) best = nbc;
                    }
                p[ci] = best >= 0 ? p[best] : 0.0;
            }
        } else {
            const double mls = oracle_mls_lookup(image, nv, pts, vals);
            double acc[3];
            object_accel(o, id, o->t, nb, acc);
            const double an = acc[0]*en[0] + acc[1]*en[1] + acc[2]*en[2];
            p[ci] = mls + o->rho * an * fabs(travelled);
        }
    }
}

/* SimWorld::UpdateObjectState (SimWorld.cpp:468-479), after the clock moved
 * on: a time-parallel chunk switches every source off for good once the
 * clock reaches its stop_boundary_acc_time */
static void update_evals(oracle *o)
{
    const fdtd_settings *s = &o->sc.settings;
    if (s->time_parallel && o->t >= s->stop_boundary_acc_time) o->evals_off = 1;
}

/* PML_WaveSolver::stepCollocated (PML_WaveSolver.cpp:775-826) */
int oracle_step(oracle *o, int nsteps, double *rows)
{
    if (o->scheme == FDTD_SCHEME_STAGGERED) {
        for (int st = 0; st < nsteps; ++st) {
            mac_velocity(o);
            mac_pressure(o);
            o->t += o->dt;
            o->step += 1;
            update_evals(o);
            if (rows) fetch_listeners(o, o->pfull, rows + (size_t)st * o->sc.n_listeners);
        }
        return 0;
    }
    for (int st = 0; st < nsteps; ++st) {
        double *pLast = o->pc[(o->pc_ind + 2) % 3];
        double *pCurr = o->pc[o->pc_ind];
        double *pNext = o->pc[(o->pc_ind + 1) % 3];
        double *lapLast = o->lap[(o->lap_ind + 1) % 2];
        double *lapCurr = o->lap[o->lap_ind];
        if (o->reclass) {
            classify_fast(o);
            fresh_fill(o, pLast);
            fresh_fill(o, pCurr);
            if (o->fresh_error) { set_err("a fresh cell lies inside an object"); return -1; }
        } else {
            memcpy(o->history, o->bulk, (size_t)o->N);
            memset(o->cavity, 0, (size_t)o->N);
            /* the per-step classification zeroes every solid cell's pressure
               (MAC_Grid.cpp:3784-3786); only solid PML cells (behind a
               constraint) are written by the update in between */
            if (o->sc.n_constraints > 0)
                for (long long q = 0; q < o->nsolid; ++q) {
                    const long long ci = o->solid[q];
                    o->pc[0][ci] = 0.0; o->pc[1][ci] = 0.0; o->pc[2][ci] = 0.0;
                }
        }
        ghost_update(o, pCurr);
        pml_velocity(o, pCurr);
        laplacian_ghost(o, pCurr, lapCurr);
        pressure_update(o, pLast, pCurr, pNext, lapCurr, lapLast);
        o->pc_ind = (o->pc_ind + 1) % 3;
        o->lap_ind = (o->lap_ind + 1) % 2;
        o->t += o->dt;
        o->step += 1;
        update_evals(o);
        if (rows) fetch_listeners(o, o->pc[o->pc_ind], rows + (size_t)st * o->sc.n_listeners);
        if (o->moving) animate_objects(o, o->t);   /* SimWorld::UpdateObjectState after the step */
    }
    return 0;
}

void oracle_get_pressure(const oracle *o, double *out)
{
    memcpy(out, o->scheme == FDTD_SCHEME_STAGGERED ? o->pfull : o->pc[o->pc_ind], sizeof(double) * o->N);
}

void oracle_set_state(oracle *o, const double *p_last, const double *p_curr)
{
    memcpy(o->pc[(o->pc_ind + 2) % 3], p_last, sizeof(double) * o->N);
    memcpy(o->pc[o->pc_ind], p_curr, sizeof(double) * o->N);
}

void oracle_get_state(const oracle *o, double *p_last, double *p_curr)
{
    if (p_last) memcpy(p_last, o->pc[(o->pc_ind + 2) % 3], sizeof(double) * o->N);
    if (p_curr) memcpy(p_curr, o->pc[o->pc_ind], sizeof(double) * o->N);
}
void oracle_get_pml_state(const oracle *o, double *pdir, double *v)
{
    if (pdir) for (int d = 0; d < 3; ++d) memcpy(pdir + d * o->N, o->pdir[d], sizeof(double) * o->N);
    if (v) {
        long long off = 0;
        for (int d = 0; d < 3; ++d) { memcpy(v + off, o->v[d], sizeof(double) * o->Nv[d]); off += o->Nv[d]; }
    }
}

void oracle_counts(const oracle *o, long long *c)
{
    long long solid = 0, bulk = 0, cav = 0;
    for (long long i = 0; i < o->N; ++i) { solid += !o->bulk[i]; bulk += o->bulk[i]; cav += o->cavity[i]; }
    c[0] = o->N; c[1] = o->npp; c[2] = o->npv; c[3] = o->ngh; c[4] = solid; c[5] = bulk; c[6] = cav;
    c[7] = o->n_coupled; c[8] = o->step;
}

double oracle_time(const oracle *o) { return o->t; }
