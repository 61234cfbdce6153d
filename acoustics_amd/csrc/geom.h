// geom.h -- analytic rigid objects behind the FDTD_RigidObject query
// semantics (DistanceToMesh / NormalToMesh / closest point,
// src/wavesolver/FDTD_RigidObject.cpp:256-327, 429-536), shared by the host
// set-up and the device kernels of moving objects.
//
// Ellipsoid projection: the closest point of y (object frame, semi-axes a)
// is x_i = a_i^2 y_i / (t + a_i^2), where t is the largest root of
// F(t) = sum_i (a_i y_i / (t + a_i^2))^2 - 1 (Eberly 2013, "Distance from a
// Point to an Ellipse, an Ellipsoid, or a Hyperellipsoid").  The root is
// found with a safeguarded Newton iteration (started from the Newton step at
// t = 0, steps kept inside a shrinking bracket, bisection when a step leaves
// it); points on principal planes follow the same case analysis as the paper.
#pragma once

#include <cfloat>
#include <cmath>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define FDTD_HD __host__ __device__
#else
#define FDTD_HD
#endif

namespace fdtd {

struct Geom {
    int type;     // FDTD_OBJ_SPHERE (0) | FDTD_OBJ_ELLIPSOID (1)
    double c[3];
    double a[3];
    double R[9];  // world-from-object, row-major
};

FDTD_HD inline void geom_obj_frame(const Geom &g, const double *x, double *y)
{
    const double d[3] = {x[0] - g.c[0], x[1] - g.c[1], x[2] - g.c[2]};
    for (int i = 0; i < 3; ++i) y[i] = g.R[0 + i] * d[0] + g.R[3 + i] * d[1] + g.R[6 + i] * d[2];
}

FDTD_HD inline void geom_world_vec(const Geom &g, const double *y, double *x)
{
    for (int i = 0; i < 3; ++i) x[i] = g.R[3 * i + 0] * y[0] + g.R[3 * i + 1] * y[1] + g.R[3 * i + 2] * y[2];
}

// F(t) and F'(t), one division per axis (FP64 division is a long dependent
// instruction sequence on the GPU)
FDTD_HD inline double geom_root_f(int m, const double *e, const double *y, double t, double *dF)
{
    double f = -1.0, d = 0.0;
    for (int i = 0; i < m; ++i) {
        const double r = 1.0 / (t + e[i] * e[i]);
        const double q = e[i] * y[i] * r;
        f += q * q;
        d += -2.0 * q * q * r;
    }
    if (dF) *dF = d;
    return f;
}

// Largest root of F(t) = sum_{i<m} (e_i y_i / (t + e_i^2))^2 - 1 with all
// y_i > 0, e sorted decreasing; the root lies in (-e_{m-1}^2, inf).
FDTD_HD inline double geom_largest_root(int m, const double *e, const double *y)
{
    const double emin2 = e[m - 1] * e[m - 1];
    double lo = -emin2, hi;
    // upper bracket: F(t) < 0 for t >= |e.z| (each term <= (e_i z_i / t)^2)
    double nz = 0;
    for (int i = 0; i < m; ++i) nz += (e[i] * y[i]) * (e[i] * y[i]);
    hi = fmax(sqrt(nz), 0.0) + 1e-300;
    if (geom_root_f(m, e, y, hi, nullptr) > 0) hi *= 2;
    // inside points (F(0) < 0): root in (lo, 0]; outside: in [0, hi)
    double d0;
    const double f0 = geom_root_f(m, e, y, 0.0, &d0);
    if (f0 == 0.0) return 0.0;
    if (f0 < 0) hi = 0.0; else lo = 0.0;
    // F is convex and decreasing, so the Newton step from 0 lands at or left
    // of the root, from where the iteration climbs to it without overshoot;
    // for points near the surface it starts next to the root
    double t = -f0 / d0;
    if (!(t > lo && t < hi)) {
        t = 0.5 * (lo + hi);
        if (lo == -emin2) t = hi - 0.5 * (hi - lo);  // stay away from the pole
    }
    for (int it = 0; it < 400; ++it) {
        double d;
        const double f = geom_root_f(m, e, y, t, &d);
        if (f == 0.0) return t;
        if (f > 0) lo = t; else hi = t;
        double tn = t - f / d;
        // converged: the Newton correction is within two ulps of the scale
        // (tested before the bracket check, which would otherwise turn a
        // converged step that rounds onto the bracket end into bisection)
        if (fabs(tn - t) <= 2.0 * DBL_EPSILON * fmax(fabs(t), emin2)) return t;
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (fabs(tn - t) <= 1e-16 * fmax(fabs(t), emin2)) return tn;
        t = tn;
    }
    return t;
}

// 2D: e0 >= e1 > 0, y >= 0
FDTD_HD inline void geom_ellipse_project(double e0, double e1, double y0, double y1, double *x0, double *x1)
{
    if (y1 > 0) {
        if (y0 > 0) {
            const double e[2] = {e0, e1}, yy[2] = {y0, y1};
            const double t = geom_largest_root(2, e, yy);
            *x0 = e0 * e0 * y0 / (t + e0 * e0);
            *x1 = e1 * e1 * y1 / (t + e1 * e1);
        } else { *x0 = 0; *x1 = e1; }
    } else {
        const double num = e0 * y0, den = e0 * e0 - e1 * e1;
        if (num < den) {
            const double r = num / den;
            *x0 = e0 * r; *x1 = e1 * sqrt(fmax(0.0, 1.0 - r * r));
        } else { *x0 = e0; *x1 = 0; }
    }
}

// 3D: e sorted decreasing, y >= 0
FDTD_HD inline void geom_ellipsoid_project(const double *e, const double *y, double *x)
{
    if (y[2] > 0) {
        if (y[1] > 0) {
            if (y[0] > 0) {
                const double t = geom_largest_root(3, e, y);
                for (int i = 0; i < 3; ++i) x[i] = e[i] * e[i] * y[i] / (t + e[i] * e[i]);
            } else { x[0] = 0; geom_ellipse_project(e[1], e[2], y[1], y[2], &x[1], &x[2]); }
        } else {
            if (y[0] > 0) { x[1] = 0; geom_ellipse_project(e[0], e[2], y[0], y[2], &x[0], &x[2]); }
            else { x[0] = 0; x[1] = 0; x[2] = e[2]; }
        }
    } else {
        const double d0 = e[0] * e[0] - e[2] * e[2], d1 = e[1] * e[1] - e[2] * e[2];
        const double n0 = e[0] * y[0], n1 = e[1] * y[1];
        if (n0 < d0 && n1 < d1) {
            const double r0 = n0 / d0, r1 = n1 / d1, disc = 1.0 - r0 * r0 - r1 * r1;
            if (disc > 0) { x[0] = e[0] * r0; x[1] = e[1] * r1; x[2] = e[2] * sqrt(disc); return; }
        }
        x[2] = 0;
        geom_ellipse_project(e[0], e[1], y[0], y[1], &x[0], &x[1]);
    }
}

// signed distance (negative inside), closest surface point, outward unit normal
FDTD_HD inline double geom_closest(const Geom &g, const double *x, double *xb, double *n)
{
    if (g.type == 0) {
        const double d[3] = {x[0] - g.c[0], x[1] - g.c[1], x[2] - g.c[2]};
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double u[3] = {1, 0, 0};
        if (r > 0) for (int i = 0; i < 3; ++i) u[i] = d[i] / r;
        for (int i = 0; i < 3; ++i) { xb[i] = g.c[i] + g.a[0] * u[i]; n[i] = u[i]; }
        return r - g.a[0];
    }
    double y[3];
    geom_obj_frame(g, x, y);
    // axes by decreasing semi-axis (a stable three-element sort)
    int perm[3] = {0, 1, 2};
    for (int i = 1; i < 3; ++i)
        for (int j = i; j > 0 && g.a[perm[j]] > g.a[perm[j - 1]]; --j) { const int t = perm[j]; perm[j] = perm[j - 1]; perm[j - 1] = t; }
    double e[3], ys[3], xs[3];
    for (int i = 0; i < 3; ++i) { e[i] = g.a[perm[i]]; ys[i] = fabs(y[perm[i]]); }
    geom_ellipsoid_project(e, ys, xs);
    double xo[3];
    for (int i = 0; i < 3; ++i) xo[perm[i]] = (y[perm[i]] < 0 ? -xs[i] : xs[i]);
    double f = -1.0;
    for (int i = 0; i < 3; ++i) f += (y[i] / g.a[i]) * (y[i] / g.a[i]);
    double dist = sqrt((xo[0] - y[0]) * (xo[0] - y[0]) + (xo[1] - y[1]) * (xo[1] - y[1]) + (xo[2] - y[2]) * (xo[2] - y[2]));
    if (f < 0) dist = -dist;
    double no[3];
    for (int i = 0; i < 3; ++i) no[i] = xo[i] / (g.a[i] * g.a[i]);
    const double nn = sqrt(no[0] * no[0] + no[1] * no[1] + no[2] * no[2]);
    for (int i = 0; i < 3; ++i) no[i] /= nn;
    double xw[3];
    geom_world_vec(g, xo, xw);
    for (int i = 0; i < 3; ++i) xb[i] = xw[i] + g.c[i];
    geom_world_vec(g, no, n);
    return dist;
}

// strict inside test (distance < 0)
FDTD_HD inline bool geom_inside(const Geom &g, const double *x)
{
    if (g.type == 0) {
        const double d[3] = {x[0] - g.c[0], x[1] - g.c[1], x[2] - g.c[2]};
        return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - g.a[0] < 0.0;
    }
    double y[3];
    geom_obj_frame(g, x, y);
    double f = -1.0;
    for (int i = 0; i < 3; ++i) f += (y[i] / g.a[i]) * (y[i] / g.a[i]);
    return f < 0.0;
}

// Rigid motion with constant velocity v and angular velocity w about the
// centre, from the pose g0 at t = 0: c(t) = c0 + v t and
// R(t) = Rot(w / |w|, |w| t) R0 by Rodrigues' formula.
FDTD_HD inline Geom geom_at(const Geom &g0, const double *v, const double *w, double t)
{
    Geom g = g0;
    for (int d = 0; d < 3; ++d) g.c[d] = g0.c[d] + v[d] * t;
    const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (wn == 0.0) return g;
    const double k[3] = {w[0] / wn, w[1] / wn, w[2] / wn}, th = wn * t, s = sin(th), cm = 1.0 - cos(th);
    const double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
    double Q[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double k2 = 0;
            for (int m = 0; m < 3; ++m) k2 += K[r * 3 + m] * K[m * 3 + c];
            Q[r * 3 + c] = (r == c ? 1.0 : 0.0) + s * K[r * 3 + c] + cm * k2;
        }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double x = 0;
            for (int m = 0; m < 3; ++m) x += Q[r * 3 + m] * g0.R[m * 3 + c];
            g.R[r * 3 + c] = x;
        }
    return g;
}

// Quaternion (w, x, y, z) scaled to unit length unless its squared length
// is within 1e-12 of one (Quaternion::normalize, Quaternion.hpp:186-195).
FDTD_HD inline void quat_normalize(double *q)
{
    const double len = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (fabs(len - 1.0) > 1e-12) {
        const double s = 1.0 / sqrt(len);
        for (int i = 0; i < 4; ++i) q[i] *= s;
    }
}

// Quaternion::slerp (Quaternion.hpp:365-420, its active branch): a normalized
// lerp when the two are within acos(0.9995) (either sign), else the great-arc
// blend along the shorter way.
FDTD_HD inline void quat_slerp(const double *a, const double *b, double r, double *out)
{
    double v0[4] = {a[0], a[1], a[2], a[3]}, v1[4] = {b[0], b[1], b[2], b[3]};
    quat_normalize(v0);
    quat_normalize(v1);
    double dot = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2] + v0[3] * v1[3];
    if (fabs(dot) > 0.9995) {
        for (int i = 0; i < 4; ++i) out[i] = v0[i] + (v1[i] - v0[i]) * r;
        quat_normalize(out);
        return;
    }
    if (dot < 0.0) {
        for (int i = 0; i < 4; ++i) v1[i] = -v1[i];
        dot = -dot;
    }
    dot = fmin(fmax(dot, -1.0), 1.0);
    const double th = acos(dot) * r;
    double v2[4];
    for (int i = 0; i < 4; ++i) v2[i] = v1[i] - v0[i] * dot;
    quat_normalize(v2);
    const double c = cos(th), s = sin(th);
    for (int i = 0; i < 4; ++i) out[i] = v0[i] * c + v2[i] * s;
}

// Pose at time t of an object animated by keyframes (8 doubles per frame:
// centre x y z, rotation w x y z, unused), frame f at time f * step:
// FDTD_RigidObject_Blender_Animator::GetRigidObjectTransform
// (FDTD_RigidObject_Animator.cpp:137-160), FDTD_Objects::AnimateObjects
// (FDTD_Objects.cpp:563-579: the blended quaternion normalized) and
// FDTD_RigidObject::SetRigidBodyTransform (FDTD_RigidObject.cpp:550-566: the
// rotation of angle 2 acos(w) about the normalized vector part, the centre
// moved to the blended one).  w is clamped to [-1, 1] before acos.
FDTD_HD inline Geom geom_keyframe(const Geom &shape, const double *frames, int n, double step, double t)
{
    const double pos = t / step;
    int f0 = pos > 0.0 ? (int)pos : 0;   // std::max<int>(0, pos) truncates toward zero
    f0 = f0 < n - 1 ? f0 : n - 1;
    const int f1 = f0 + 1 < n - 1 ? f0 + 1 : n - 1;
    const double t0 = f0 * step, t1 = f1 * step;
    const double alpha = f0 == f1 ? 1.0 : (t - t0) / (t1 - t0);
    const double *A = frames + 8 * f0, *B = frames + 8 * f1;
    Geom g = shape;
    for (int d = 0; d < 3; ++d) g.c[d] = A[d] * (1.0 - alpha) + B[d] * alpha;
    double q[4];
    quat_slerp(A + 3, B + 3, alpha, q);
    quat_normalize(q);
    const double w = fmin(fmax(q[0], -1.0), 1.0);
    const double th = 2.0 * acos(w);
    double k[3] = {0.0, 1.0, 0.0};
    if (!(th < 1e-12)) {
        const double vn = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        if (vn > 0.0)
            for (int d = 0; d < 3; ++d) k[d] = q[1 + d] / vn;
    }
    const double s = sin(th), cm = 1.0 - cos(th);
    const double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double k2 = 0;
            for (int m = 0; m < 3; ++m) k2 += K[r * 3 + m] * K[m * 3 + c];
            g.R[r * 3 + c] = (r == c ? 1.0 : 0.0) + s * K[r * 3 + c] + cm * k2;
        }
    return g;
}

}  // namespace fdtd
