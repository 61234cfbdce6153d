// kernels.hip -- CDNA4 (gfx950) kernels of the collocated FDTD step.
//
// step_kernel advances every tile of the work list by one step
// (PML_WaveSolver::stepCollocated, src/wavesolver/PML_WaveSolver.cpp:775-826):
//   * TILE_INTERIOR: p^{n+1} = 2 p^n - p^{n-1} + (c dt/h)^2 L(p^n) + S dt
//     (MAC_Grid.cpp:924-944), FP32.  Threads own a (j,k) column and march the
//     tile's planes along i, keeping p^n(i-1), p^n(i), p^n(i+1) in registers;
//     k is the fastest axis so a wavefront streams 64 consecutive floats.
//   * TILE_PML: staggered face velocity v <- f1 v + f2 (p_R - p_L) on the
//     cell's faces, then pDir_d <- f3 pDir_d + f5/h (v_R - v_L) and
//     p^{n+1} = sum_d pDir_d (MAC_Grid.cpp:461-508), fused per thread.  The
//     PML state lives in compact per-slab arrays (PmlSlab).  Face velocities
//     are double-buffered (vin -> vout), so a thread may recompute the face it
//     shares with a neighbour without racing the neighbour's write.
//   * TILE_SPECIAL: one thread per bulk cell next to a rigid object: the
//     non-uniform ghost-cell Laplacian (MAC_Grid.cpp:187-299) and the
//     one-cell-cavity freeze, in FP64; ghost values are FP64.
// No MFMA anywhere: a 7-point stencil is not a contraction.
#include "kernels.h"

#include <cfloat>
#include <climits>

#include "linalg.h"

namespace fdtd {

// Pointers read out of a SceneDev are generic to the compiler, which then
// emits flat memory instructions; those count against both the vector-memory
// and the LDS counters, so every wait becomes a full wait.  Every scene array
// lives in device global memory: casting to the global address space gives
// global_load/global_store with vmcnt-only waits, and read-only tables cast to
// the constant address space are fetched with scalar loads when the index is
// uniform across the wavefront.
#if defined(__HIP_DEVICE_COMPILE__)
#define AS_GLOBAL __attribute__((address_space(1)))
#define AS_CONST __attribute__((address_space(4)))
#else  // the host pass only parses device code
#define AS_GLOBAL
#define AS_CONST
#endif
template <class T>
__device__ __forceinline__ AS_GLOBAL T *gmem(T *p) { return (AS_GLOBAL T *)p; }
template <class T>
__device__ __forceinline__ const AS_GLOBAL T *gmem(const T *p) { return (const AS_GLOBAL T *)p; }
template <class T>
__device__ __forceinline__ const AS_CONST T *cmem(const T *p) { return (const AS_CONST T *)p; }

// a_s(obj, t): the sources whose acceleration vector is a_n times the normal
// (Gaussian impulse; HarmonicVibrationalSource::Evaluate,
// HarmonicVibrationalSource.cpp:5-28; the chirp)
__device__ __forceinline__ double impulse_accel(const ImpDev *isrc, int n, int obj, double t)
{
    double a = 0.0;
    for (int q = 0; q < n; ++q) {
        const ImpDev s = isrc[q];
        if (s.obj != obj) continue;
        if (s.kind == FDTD_VIB_HARMONIC) {
            if (s.t1 > s.t0 && (t < s.t0 || t > s.t1)) continue;
            const double v = sin(s.omega * t + s.phase) * (s.A * s.omega);
            a += s.decay < 0.0 ? v : exp(-s.decay * (t - s.dcenter)) * v;
        } else if (s.kind == FDTD_VIB_CHIRP) {
            // ChirpVibrationalSource::Evaluate (ChirpVibrationalSource.cpp:5-19)
            if (t < s.t0) continue;
            const double al = t < s.t1 ? 0.5 * (t - s.t0) / (s.t1 - s.t0) : 1.0;
            const double ph = al < 0.52 ? 0.0 : s.t1 * 0.5 * (s.omega1 - s.omega);
            const double om = (1.0 - al) * s.omega + al * s.omega1;
            a += sin(om * t + ph);
        } else {
            const double u = t - s.ton;
            a += s.A * exp(-(u * u) / (2.0 * s.w * s.w));
        }
    }
    return a;
}

// a_v(obj, t)[d]: acceleration noise, the half-sine contact of every impact
// whose support holds t (AccelerationNoiseVibrationalSource::Evaluate,
// AccelerationNoiseVibrationalSource.cpp:63-90, and ImpulseSeriesObject::
// GetImpulseWithinSupport: t0 <= t < t0 + tau)
__device__ __forceinline__ double noise_accel(const ImpactDev *imp, int n, int obj, int d, double t)
{
    double a = 0.0;
    for (int q = 0; q < n; ++q) {
        const ImpactDev r = imp[q];
        if (r.obj != obj || t < r.t0 || !(t < r.t0 + r.tau)) continue;
        a += sin(3.141592653589793 * (t - r.t0) / r.tau) * r.v[d];
    }
    return a;
}

__device__ __forceinline__ float source_amp(const PSrcDev &ps, double tn, double dt, int scheme)
{
    // S(x, t_n + dt/2) dt without the spatial factor (GaussianPressureSource.cpp:6-30):
    // the pulse itself (collocated) or its time integral (staggered)
    const double t = tn + 0.5 * dt;
    if (scheme == FDTD_SCHEME_STAGGERED) {
        const double s2 = 1.4142135623730951 * ps.wt;
        return (float)((erf((t - ps.t0) / s2) - erf(-ps.t0 / s2)) * ps.amp * dt);
    }
    const double u = (t - ps.t0) / ps.wt;
    return (float)(exp(-(u * u) / 2.0) * ps.amp * dt);
}

// Speaker sources at cell (i, j, k) in step `step`, times dt
// (SpeakerPressureSource::Evaluate, SpeakerPressureSource.cpp:91-113, at
// t_n + dt/2): exp(-r^2 / (2 w^2)) samples[(int)((t - r/c - t0) / period)],
// zero past either end of the samples.
__device__ inline double speaker_terms(const SceneDev &S, long long step, int i, int j, int k)
{
    if (S.n_spk == 0 || (S.src_steps >= 0 && step >= S.src_steps)) return 0.0;
    const double t = gmem(S.t_tab)[amp_row(step)] + 0.5 * S.dt;
    const int ijk[3] = {i, j, k};
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = S.bmin[d] + (0.5 + (double)ijk[d]) * S.h;
    double v = 0.0;
    for (int s = 0; s < S.n_spk; ++s) {
        const double *q = S.spk + 8 * s;
        const double r2 = (x[0] - q[0]) * (x[0] - q[0]) + (x[1] - q[1]) * (x[1] - q[1]) + (x[2] - q[2]) * (x[2] - q[2]);
        const double g = exp(-r2 / (2.0 * q[3] * q[3]));
        const double delay = t - sqrt(r2) / S.c;
        const int idx = (int)((delay - q[4]) / q[5]);
        if (idx >= 0 && idx < (int)q[7]) v += g * S.spk_samples[(int)q[6] + idx];
    }
    return v * S.dt;
}

// The speaker term of cell (i, j, k) for the step being run, from the field
// speaker_kernel wrote over the speakers' box before the step kernels (zero
// outside the box, where every speaker is below exp(-32) of its peak).
__device__ __forceinline__ float speaker_at(const SceneDev &S, int i, int j, int k)
{
    const int a = i - S.sp_lo[0], b = j - S.sp_lo[1], c = k - S.sp_lo[2];
    if (a < 0 || b < 0 || c < 0 || a >= S.sp_n[0] || b >= S.sp_n[1] || c >= S.sp_n[2]) return 0.f;
    return gmem(S.spk_field)[(a * S.sp_n[1] + b) * S.sp_n[2] + c];
}

// ---------------------------------------------------------------------------
// fully-coupled ghost values (MAC_Grid.cpp:1131-1441): the coupled system's
// static solution operator applied to p^n and the accelerations of step n,
// x_g = sum_q M_gq p_q + sum_o m_go a_n(o), FP64, in row order.  Computed where
// it is used (the object-adjacent cell next to the ghost, and listener terms),
// four independent index loads and gathers at a time.
// -f h grad_p . cell_n of a rasterised ghost (f = 1 with centre sampling,
// 0.75 with seven samples), grad_p = -rho (a_s n_b + a_v) (MAC_Grid.cpp:
// 1256-1281): k_raster a_s, plus k_vec a_v along the ghost's axis in a
// scene with acceleration noise
__device__ __forceinline__ double raster_accel_term(const SceneDev &S, const GhostDev &gd, long long step)
{
    const AS_GLOBAL double *acc = gmem(S.acc_tab) + amp_row(step) * S.n_ch + gd.obj * S.n_cpo;
    double v = gd.k_raster * acc[0];
    if (S.n_cpo > 1) v = fma(gd.k_vec, acc[1 + gd.vaxis], v);
    return v;
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ double coupled_ghost(const SceneDev &S, int g, const AS_GLOBAL float *__restrict__ pcur,
                                                long long step)
{
    const AS_GLOBAL int32_t *idx = gmem(S.op_idx);
    const AS_GLOBAL double *w = gmem(S.op_w);
    int q = gmem(S.op_ptr)[g];
    const int q1 = gmem(S.op_ptr)[g + 1];
    double x = 0.0;
    for (; q + 4 <= q1; q += 4) {
        const int i0 = idx[q], i1 = idx[q + 1], i2 = idx[q + 2], i3 = idx[q + 3];
        const double w0 = w[q], w1 = w[q + 1], w2 = w[q + 2], w3 = w[q + 3];
        const float v0 = pcur[i0], v1 = pcur[i1], v2 = pcur[i2], v3 = pcur[i3];
        x = fma(w0, (double)v0, x);
        x = fma(w1, (double)v1, x);
        x = fma(w2, (double)v2, x);
        x = fma(w3, (double)v3, x);
    }
    for (; q < q1; ++q) x = fma(w[q], (double)pcur[idx[q]], x);
    const AS_GLOBAL double *acc = gmem(S.acc_tab) + amp_row(step) * S.n_ch;
    const int c1 = gmem(S.oc_ptr)[g + 1];
    for (int c = gmem(S.oc_ptr)[g]; c < c1; ++c) x = fma(gmem(S.oc_w)[c], acc[gmem(S.oc_ch)[c]], x);
    return x;
}

// ---------------------------------------------------------------------------
// listener terms (PML_WaveSolver::FetchScalarData, MLSModeInterpolator::lookup
// with precomputed weights): the thread that produces a term's sample writes
// beta * sample into the scene's series ring; the terms of a listener are
// summed when the ring is read out.  A tile lists its terms in [tb, te).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void emit_cell_terms(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step,
                                                int idx, float pn)
{
    for (int q = t.tb; q < t.te; ++q) {
        const TermRef r = cmem(terms)[q];
        if (r.kind == 0 && r.cell == idx) gmem(S.terms)[(size_t)(step % S.cap) * S.T + r.slot] = r.beta * (double)pn;
    }
}

__device__ __forceinline__ void emit_ghost_terms(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step)
{
    const AS_GLOBAL float *__restrict__ pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    for (int q = t.tb + (int)threadIdx.x; q < t.te; q += (int)blockDim.x) {
        const TermRef r = cmem(terms)[q];
        if (r.kind == 0) continue;
        double v;
        if (S.raster) {
            const GhostDev gd = gmem(S.ghosts)[r.cell];
            v = (double)pcur[gd.nbr] + raster_accel_term(S, gd, step);
        } else {
            v = coupled_ghost(S, r.cell, pcur, step);
        }
        gmem(S.terms)[(size_t)(step % S.cap) * S.T + r.slot] = r.beta * v;
    }
}

// ---------------------------------------------------------------------------
// interior leapfrog
// ---------------------------------------------------------------------------
// The leapfrog update in increment form, p^{n+1} = p^n + ((p^n - p^{n-1}) +
// lam2 sum_q (p_q - p^n)).  A smooth field changes little between steps and
// between neighbours, so every difference here is exact in FP32 (Sterbenz:
// operands within a factor of two) and the only rounding at the scale of p
// itself is the last addition.  Forming 2 p^n - p^{n-1} and the six-neighbour
// sum instead rounds three times at that scale per step, and the rounding
// errors, fed back as velocity, add up along every propagation path: 1.3e-4
// relative L2 at the listeners after 4000 steps at 256^3, against 1e-4 allowed.
__device__ __forceinline__ float leapfrog1(float c, float prev, float xm, float xp, float ym, float yp, float zm, float zp,
                                           float lam2)
{
    const float lap = ((xm - c) + (xp - c)) + ((ym - c) + (yp - c)) + ((zm - c) + (zp - c));
    return c + fmaf(lam2, lap, c - prev);
}

__device__ __forceinline__ void interior_tile(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step, bool skip_check)
{
    const int tk = t.tk_log2;
    const int tx = threadIdx.x & ((1 << tk) - 1), ty = threadIdx.x >> tk;
    const int k = t.k0 + tx, j = t.j0 + ty;
    if (k >= t.k1 || j >= t.j1) return;
    const int sx = S.sx, sy = S.sy;
    const AS_GLOBAL float *__restrict__ pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    // p^{n+1} overwrites p^{n-1}: each thread reads and writes only its own cells there
    const AS_GLOBAL float *pprev = gmem(S.field + (size_t)slot_next(step) * S.npad);
    AS_GLOBAL float *pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    const float lam2 = S.lam2f;

    int idx = t.i0 * sx + j * sy + k;
    if (S.lapv) {
        // air viscosity (MAC_Grid.cpp:926-930): with L = (sum of the six
        // neighbours - 6 p) / h^2, p^{n+1} = 2 p^n - p^{n-1} - c alpha dt L^{n-1}
        // + ((c dt)^2 + c alpha dt) L^n; L^n replaces L^{n-1} in place
        AS_GLOBAL float *lapv = gmem(S.lapv);
        for (int i = t.i0; i < t.i1; ++i, idx += sx) {
            const float c = pcur[idx];
            const float lc = (((pcur[idx - sx] - c) + (pcur[idx + sx] - c)) + ((pcur[idx - sy] - c) + (pcur[idx + sy] - c)) +
                              ((pcur[idx - 1] - c) + (pcur[idx + 1] - c))) * S.inv_h2f;
            const float ll = lapv[idx];
            float pn = c + (((c - pprev[idx]) - S.kcalpf * ll) + S.c2k2_kcalpf * lc);
            if (t.src)
                for (int s = 0; s < S.n_psrc; ++s) {
                    const PSrcDev ps = cmem(S.psrc)[s];
                    const AS_CONST float *g = cmem(S.gtab);
                    pn += gmem(S.src_tab)[amp_row(step) * S.n_psrc + s] * g[ps.goff[0] + i] * g[ps.goff[1] + j] * g[ps.goff[2] + k];
                }
            if (t.src && S.n_spk) pn += speaker_at(S, i, j, k);
            if (!skip_check || !gmem(S.skip)[idx]) {
                pnext[idx] = pn;
                lapv[idx] = lc;
                if (t.te > t.tb) emit_cell_terms(S, t, terms, step, idx, pn);
            }
        }
        return;
    }
    if (!t.src && !skip_check && t.te == t.tb && t.i1 - t.i0 == 2) {
        // two planes with every load issued before any arithmetic
        const float a0 = pcur[idx - sx], a1 = pcur[idx], a2 = pcur[idx + sx], a3 = pcur[idx + 2 * sx];
        const float ym1 = pcur[idx - sy], yp1 = pcur[idx + sy], zm1 = pcur[idx - 1], zp1 = pcur[idx + 1];
        const float ym2 = pcur[idx + sx - sy], yp2 = pcur[idx + sx + sy], zm2 = pcur[idx + sx - 1], zp2 = pcur[idx + sx + 1];
        const float pv1 = pprev[idx], pv2 = pprev[idx + sx];
        pnext[idx] = leapfrog1(a1, pv1, a0, a2, ym1, yp1, zm1, zp1, lam2);
        pnext[idx + sx] = leapfrog1(a2, pv2, a1, a3, ym2, yp2, zm2, zp2, lam2);
        return;
    }
    float pm = pcur[idx - sx];
    float pc = pcur[idx];
    if (!t.src && !skip_check && t.te == t.tb) {
        // the streaming case: 12 bytes of HBM traffic per cell
        for (int i = t.i0; i < t.i1; ++i, idx += sx) {
            const float pp = pcur[idx + sx];
            const float ym = pcur[idx - sy], yp = pcur[idx + sy];
            const float zm = pcur[idx - 1], zp = pcur[idx + 1];
            const float prev = pprev[idx];
            pnext[idx] = leapfrog1(pc, prev, pm, pp, ym, yp, zm, zp, lam2);
            pm = pc;
            pc = pp;
        }
        return;
    }
    // tiles with pressure sources and/or object cells to leave alone
    for (int i = t.i0; i < t.i1; ++i, idx += sx) {
        const float pp = pcur[idx + sx];
        const float ym = pcur[idx - sy], yp = pcur[idx + sy];
        const float zm = pcur[idx - 1], zp = pcur[idx + 1];
        const float prev = pprev[idx];
        float pn = leapfrog1(pc, prev, pm, pp, ym, yp, zm, zp, lam2);
        if (t.src)
            for (int s = 0; s < S.n_psrc; ++s) {
                const PSrcDev ps = cmem(S.psrc)[s];
                const AS_CONST float *g = cmem(S.gtab);
                pn += gmem(S.src_tab)[amp_row(step) * S.n_psrc + s] * g[ps.goff[0] + i] * g[ps.goff[1] + j] * g[ps.goff[2] + k];
            }
        if (t.src && S.n_spk) pn += speaker_at(S, i, j, k);
        if (!skip_check || !gmem(S.skip)[idx]) {
            pnext[idx] = pn;
            if (t.te > t.tb) emit_cell_terms(S, t, terms, step, idx, pn);
        }
        pm = pc;
        pc = pp;
    }
}

// Stores components [lo, n) of v at p (16-byte aligned; 0 <= lo < n <= 4).
__device__ __forceinline__ void store4(AS_GLOBAL float *p, float4 v, int n, int lo = 0)
{
    if (n == 4 && lo == 0) {
        *(AS_GLOBAL float4 *)p = v;
    } else {
        if (lo == 0) p[0] = v.x;
        if (n > 1 && lo <= 1) p[1] = v.y;
        if (n > 2 && lo <= 2) p[2] = v.z;
        if (n > 3) p[3] = v.w;
    }
}

// Leapfrog update of four cells along k: centre c, previous level prev, x
// neighbours xm/xp, y neighbours ym/yp, and the k neighbours beyond the four
// (zl below c.x, zr above c.w).
__device__ __forceinline__ float4 leapfrog4(float4 c, float4 prev, float4 xm, float4 xp, float4 ym, float4 yp, float zl,
                                            float zr, float lam2)
{
    float4 o;
    o.x = leapfrog1(c.x, prev.x, xm.x, xp.x, ym.x, yp.x, zl, c.y, lam2);
    o.y = leapfrog1(c.y, prev.y, xm.y, xp.y, ym.y, yp.y, c.x, c.z, lam2);
    o.z = leapfrog1(c.z, prev.z, xm.z, xp.z, ym.z, yp.z, c.y, c.w, lam2);
    o.w = leapfrog1(c.w, prev.w, xm.w, xp.w, ym.w, yp.w, c.z, zr, lam2);
    return o;
}

// Whole rows, four floats per lane: wavefront w of the workgroup owns rows
// j0 + w R .. j0 + w R + R - 1 (R = rpw rows of nk4 lanes each, R nk4 <= 64,
// so short rows still fill the wavefront), lane l of a row the cells
// k0 + 4l .. k0 + 4l + 3; one load moves up to 1 KiB.  The k neighbours come
// from the adjacent lanes; the two row ends read theirs from memory.  Plain
// cells only (no source, no object cell); k0 and the row stride are
// multiples of four.
__device__ __forceinline__ void interior_vec_tile(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step)
{
    // the last lane of a row may hold a partial group: its cells past k1
    // (PML or row padding) are read, used as k neighbours, and not written;
    // likewise the first lane's first t.lead cells when the row's own cells
    // start past the float4 boundary k0
    const int nk4 = (t.k1 - t.k0 + 3) >> 2, R = t.rpw, rem = (t.k1 - t.k0) & 3;
    const int j0w = t.j0 + (int)(threadIdx.x >> 6) * R;   // first row of this wavefront
    if (j0w >= t.j1) return;                                 // (whole wavefront)
    const int l64 = threadIdx.x & 63, row = R == 1 ? 0 : l64 / nk4;
    const int lr = l64 - row * nk4;               // lane within its row
    const int jr = j0w + row;
    // lanes past the wavefront's rows or past a row's last group, and rows
    // past the tile, load a valid group and store nothing
    const bool act = row < R && lr < nk4 && jr < t.j1;
    const int j = act ? jr : t.j0;
    const int lane = act ? lr : 0;
    const int k = t.k0 + 4 * lane;
    const int sx = S.sx, sy = S.sy;
    const AS_GLOBAL float *__restrict__ pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    AS_GLOBAL float *pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    const float lam2 = S.lam2f;
    int idx = t.i0 * sx + j * sy + k;
    if (t.i1 - t.i0 == 2 && t.te == t.tb) {
        // two planes with every load issued before any arithmetic (twice the
        // bytes in flight per wave of the marching form below)
        const float4 a0 = *(const AS_GLOBAL float4 *)(pcur + idx - sx);
        const float4 a1 = *(const AS_GLOBAL float4 *)(pcur + idx);
        const float4 a2 = *(const AS_GLOBAL float4 *)(pcur + idx + sx);
        const float4 a3 = *(const AS_GLOBAL float4 *)(pcur + idx + 2 * sx);
        const float4 ym1 = *(const AS_GLOBAL float4 *)(pcur + idx - sy);
        const float4 yp1 = *(const AS_GLOBAL float4 *)(pcur + idx + sy);
        const float4 ym2 = *(const AS_GLOBAL float4 *)(pcur + idx + sx - sy);
        const float4 yp2 = *(const AS_GLOBAL float4 *)(pcur + idx + sx + sy);
        const float4 pv1 = *(const AS_GLOBAL float4 *)(pnext + idx);
        const float4 pv2 = *(const AS_GLOBAL float4 *)(pnext + idx + sx);
        float zl1 = __shfl_up(a1.w, 1), zr1 = __shfl_down(a1.x, 1);
        float zl2 = __shfl_up(a2.w, 1), zr2 = __shfl_down(a2.x, 1);
        if (lane == 0) { zl1 = pcur[idx - 1]; zl2 = pcur[idx + sx - 1]; }
        if (lane == nk4 - 1) { zr1 = pcur[idx + 4]; zr2 = pcur[idx + sx + 4]; }
        const float4 o1 = leapfrog4(a1, pv1, a0, a2, ym1, yp1, zl1, zr1, lam2);
        const float4 o2 = leapfrog4(a2, pv2, a1, a3, ym2, yp2, zl2, zr2, lam2);
        if (rem == 0 && t.lead == 0) {   // whole groups only (uniform over the tile)
            if (act) {
                *(AS_GLOBAL float4 *)(pnext + idx) = o1;
                *(AS_GLOBAL float4 *)(pnext + idx + sx) = o2;
            }
        } else if (act) {
            const int n = (lane == nk4 - 1 && rem) ? rem : 4, lo = lane == 0 ? t.lead : 0;
            store4(pnext + idx, o1, n, lo);
            store4(pnext + idx + sx, o2, n, lo);
        }
        return;
    }
    float4 pm = *(const AS_GLOBAL float4 *)(pcur + idx - sx);
    float4 pc = *(const AS_GLOBAL float4 *)(pcur + idx);
    for (int i = t.i0; i < t.i1; ++i, idx += sx) {
        const float4 pp = *(const AS_GLOBAL float4 *)(pcur + idx + sx);
        const float4 ym = *(const AS_GLOBAL float4 *)(pcur + idx - sy);
        const float4 yp = *(const AS_GLOBAL float4 *)(pcur + idx + sy);
        const float4 prev = *(const AS_GLOBAL float4 *)(pnext + idx);
        float zl = __shfl_up(pc.w, 1), zr = __shfl_down(pc.x, 1);
        if (lane == 0) zl = pcur[idx - 1];
        if (lane == nk4 - 1) zr = pcur[idx + 4];
        const float4 o = leapfrog4(pc, prev, pm, pp, ym, yp, zl, zr, lam2);
        if (act) {
            const int n = (lane == nk4 - 1 && rem) ? rem : 4, lo = lane == 0 ? t.lead : 0;
            store4(pnext + idx, o, n, lo);
            if (t.te > t.tb) {
                if (lo == 0) emit_cell_terms(S, t, terms, step, idx, o.x);
                if (n > 1 && lo <= 1) emit_cell_terms(S, t, terms, step, idx + 1, o.y);
                if (n > 2 && lo <= 2) emit_cell_terms(S, t, terms, step, idx + 2, o.z);
                if (n > 3) emit_cell_terms(S, t, terms, step, idx + 3, o.w);
            }
        }
        pm = pc;
        pc = pp;
    }
}

// ---------------------------------------------------------------------------
// PML slab: fused staggered velocity + split pressure update
// (MAC_Grid::PML_velocityUpdateCollocated, MAC_Grid.cpp:383-459, and the PML
// branch of PML_pressureUpdateCollocated, 496-508), with the merged
// transverse state of PmlSlab for cells damped along the slab axis only.
// ---------------------------------------------------------------------------
// Face velocity update on one axis for a face whose old value is v, between
// pressures pl (left) and pr (right); inactive faces stay at zero.  The fma
// is explicit: a face is computed by the two cells that share it and both
// must round it the same way.
__device__ __forceinline__ float face_update(bool active, float4 c, float v, float pl, float pr)
{
    return active ? fmaf(c.y, pr - pl, c.x * v) : 0.f;
}

// Split pressure update pDir <- f3 pDir + f5/h (v_R - v_L).
__device__ __forceinline__ float split_update(float4 c, float q, float dv)
{
    return fmaf(c.w, dv, c.z * q);
}

// One transverse axis of the merged update: f5h (f2_R (p_R - p) - f2_L (p - p_L)),
// cL the cell's table entry (its low face and itself), cR the high face's.
__device__ __forceinline__ float transverse_term(float4 cL, float4 cR, float pl, float pc, float pr)
{
    return cL.w * fmaf(cR.y, pr - pc, -(cL.y * (pc - pl)));
}

// Loads and stores at a 32-bit byte offset from a wavefront-uniform base, so
// the compiler can use the scalar-base form of global memory instructions
// (one VGPR per address instead of a 64-bit pair).  Offsets stay below 4 GiB:
// the host caps a pressure slot at 2^30 cells.
template <class T>
__device__ __forceinline__ T ldb(const AS_GLOBAL T *base, uint32_t off)
{
    return *(const AS_GLOBAL T *)((const AS_GLOBAL char *)base + off);
}
template <class T>
__device__ __forceinline__ void stb(AS_GLOBAL T *base, uint32_t off, T v)
{
    *(AS_GLOBAL T *)((AS_GLOBAL char *)base + off) = v;
}

// A = the slab's damped axis.  Threads run along z (lanes) and y (rows) and
// march along x, carrying the x face between planes.  A cell is merged when
// the slab merges and both transverse axes are plain at it; in y and z slabs
// x is plain over the whole slab, so merging is constant along a column.
template <int A>
__device__ __forceinline__ void pml_tile_axis(const SceneDev &S, const TileDesc &t, const PmlSlab &sl, const TermRef *terms,
                                              long long step)
{
    const int tk = t.tk_log2;
    const int tx = threadIdx.x & ((1 << tk) - 1), ty = threadIdx.x >> tk;
    const int k = t.k0 + tx, j = t.j0 + ty;
    if (k >= t.k1 || j >= t.j1) return;
    const int nx = S.n[0], ny = S.n[1], nz = S.n[2];
    const AS_GLOBAL float *__restrict__ pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    AS_GLOBAL float *__restrict__ pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    // compact storage of this slab: cells (and x faces) have plane stride
    // e1*e2; y faces (e1+1)*e2; z faces e1*(e2+1)
    const int e0 = sl.e[0], e1 = sl.e[1], e2 = sl.e[2];
    const int li0 = t.i0 - sl.lo[0], lj = j - sl.lo[1], lk = k - sl.lo[2];
    const int vb = (int)(step & 1);
    AS_GLOBAL float *__restrict__ pd0 = gmem(S.pd[0] + sl.cell_base);
    AS_GLOBAL float *__restrict__ pd1 = gmem(S.pd[1] + sl.cell_base);
    AS_GLOBAL float *__restrict__ pd2 = gmem(S.pd[2] + sl.cell_base);
    const AS_GLOBAL float *__restrict__ vi0 = gmem(S.vel[vb][0] + sl.v_base[0]);
    const AS_GLOBAL float *__restrict__ vi1 = gmem(S.vel[vb][1] + sl.v_base[1]);
    const AS_GLOBAL float *__restrict__ vi2 = gmem(S.vel[vb][2] + sl.v_base[2]);
    AS_GLOBAL float *__restrict__ vo0 = gmem(S.vel[vb ^ 1][0] + sl.v_base[0]);
    AS_GLOBAL float *__restrict__ vo1 = gmem(S.vel[vb ^ 1][1] + sl.v_base[1]);
    AS_GLOBAL float *__restrict__ vo2 = gmem(S.vel[vb ^ 1][2] + sl.v_base[2]);
    const AS_CONST float4 *__restrict__ cx = cmem(S.coef[0]);
    const AS_CONST int *__restrict__ fx = cmem(S.flags[0]);

    // y and z coefficients and face predicates are constant along the column
    const AS_CONST int *fy = cmem(S.flags[1]), *fz = cmem(S.flags[2]);
    const AS_CONST float4 *cy = cmem(S.coef[1]), *cz = cmem(S.coef[2]);
    const int fyL = fy[j], fzL = fz[k];
    const bool has_yR = j + 1 < ny, has_zR = k + 1 < nz;
    const int fyR = has_yR ? fy[j + 1] : 0, fzR = has_zR ? fz[k + 1] : 0;
    const float4 cyL = cy[j], czL = cz[k];
    const float4 cyR = cy[min(j + 1, ny)], czR = cz[min(k + 1, nz)];
    const bool cj = fyL & 1, ck = fzL & 1;
    const bool ox = cj || ck;
    // behind a plane constraint (flag 16 on any axis) a cell is solid: its
    // pressure stays zero, as the reference's per-step classification resets
    // it, while its faces and split state update as usual
    const bool solid_col = (fyL | fzL) & 16;
    // a face is updated when it lies in the PML or its cell is in the PML on
    // another axis (ci, the x flag of the plane, joins per plane); wall faces stay zero
    const bool yl_on = j >= 1 && ((fyL & 2) || ck), yl_ci = j >= 1;
    const bool yr_on = has_yR && ((fyR & 2) || ck), yr_ci = has_yR;
    const bool zl_on = k >= 1 && ((fzL & 2) || cj), zl_ci = k >= 1;
    const bool zr_on = has_zR && ((fzR & 2) || cj), zr_ci = has_zR;
    // the high y / z face of this column is stored here when it is the slab's last layer
    const bool own_yR = lj == e1 - 1, own_zR = lk == e2 - 1;

    // merged or split; a split axis keeps its pDir and faces.  The low face of
    // a merged cell is still updated when the cell below it in the slab is
    // split on that axis (that cell reads it as its high face).
    const bool mg = sl.merge != 0;
    const bool py = A == 1 || (fyL & 8), pz = A == 2 || (fzL & 8);
    const bool m = mg && py && pz;
    const bool sx = A == 0 || !m, sy = A == 1 || !m, sz = A == 2 || !m;
    const int fyM = lj >= 1 ? fy[j - 1] : 0, fzM = lk >= 1 ? fz[k - 1] : 0;
    const bool need_yL = sy || (lj >= 1 && !(mg && (A == 1 || (fyM & 8)) && pz));
    const bool need_zL = sz || (lk >= 1 && !(mg && py && (A == 2 || (fzM & 8))));
    // merged state: the increment d in the lower transverse axis's array; q
    // itself is p - pDir_A
    AS_GLOBAL float *__restrict__ darr = A == 0 ? pd1 : pd0;

    // byte offsets
    const uint32_t sx4 = 4u * (uint32_t)S.sx, sy4 = 4u * (uint32_t)S.sy;
    const uint32_t cs4 = 4u * (uint32_t)(e1 * e2), ys4 = 4u * (uint32_t)((e1 + 1) * e2), zs4 = 4u * (uint32_t)(e1 * (e2 + 1));
    const uint32_t e24 = 4u * (uint32_t)e2;
    const uint32_t ym4 = j > 0 ? sy4 : 0u, yp4 = j < ny - 1 ? sy4 : 0u;
    const uint32_t zm4 = k > 0 ? 4u : 0u, zp4 = k < nz - 1 ? 4u : 0u;
    uint32_t po = 4u * (uint32_t)(t.i0 * S.sx + j * S.sy + k);        // pressure cell
    uint32_t co = 4u * (uint32_t)((li0 * e1 + lj) * e2 + lk);         // PML cell and x face
    uint32_t yo = 4u * (uint32_t)((li0 * (e1 + 1) + lj) * e2 + lk);   // y face
    uint32_t zo = 4u * (uint32_t)((li0 * e1 + lj) * (e2 + 1) + lk);   // z face

    float pm = ldb(pcur, po - (t.i0 > 0 ? sx4 : 0u));
    float pc = ldb(pcur, po);
    // the x face on the low side of the current cell (updated value), carried
    // from one plane to the next: cell i's high face is cell i+1's low face
    float vxL = 0.f;
    if (sx) vxL = face_update(t.i0 >= 1 && ((fx[t.i0] & 2) || ox), cx[t.i0], ldb(vi0, co), pm, pc);
    const int ilast = sl.lo[0] + e0 - 1;
    for (int i = t.i0; i < t.i1; ++i, po += sx4, co += cs4, yo += ys4, zo += zs4) {
        const bool last = i + 1 >= nx;
        const float pp = ldb(pcur, po + (last ? 0u : sx4));
        const float ym = ldb(pcur, po - ym4), yp = ldb(pcur, po + yp4);
        const float zm = ldb(pcur, po - zm4), zp = ldb(pcur, po + zp4);
        float vxr_old = 0.f, vyl_old = 0.f, vyr_old = 0.f, vzl_old = 0.f, vzr_old = 0.f;
        float q0 = 0.f, q1 = 0.f, q2 = 0.f, qd = 0.f;
        if (sx) { vxr_old = ldb(vi0, co + cs4); q0 = ldb(pd0, co); }
        if (need_yL) vyl_old = ldb(vi1, yo);
        if (sy) { vyr_old = ldb(vi1, yo + e24); q1 = ldb(pd1, co); }
        if (need_zL) vzl_old = ldb(vi2, zo);
        if (sz) { vzr_old = ldb(vi2, zo + 4u); q2 = ldb(pd2, co); }
        if (m) qd = ldb(darr, co);
        const int fxC = fx[i], fxR = last ? 0 : fx[i + 1];
        const float4 cxC = cx[i], cxR = cx[min(i + 1, nx)];
        const bool ci = fxC & 1;
        float n0 = 0.f, n1 = 0.f, n2 = 0.f;
        // x: high face (new), low face carried
        if (sx) {
            const float vxR = last ? 0.f : face_update((fxR & 2) || ox, cxR, vxr_old, pc, pp);
            stb(vo0, co, vxL);
            if (i == ilast) stb(vo0, co + cs4, vxR);
            n0 = split_update(cxC, q0, vxR - vxL);
            stb(pd0, co, n0);
            vxL = vxR;
        }
        // y
        float vyL = 0.f;
        if (need_yL) {
            vyL = face_update(yl_on || (ci && yl_ci), cyL, vyl_old, ym, pc);
            stb(vo1, yo, vyL);
        }
        if (sy) {
            const float vyR = face_update(yr_on || (ci && yr_ci), cyR, vyr_old, pc, yp);
            if (own_yR) stb(vo1, yo + e24, vyR);
            n1 = split_update(cyL, q1, vyR - vyL);
            stb(pd1, co, n1);
        }
        // z
        float vzL = 0.f;
        if (need_zL) {
            vzL = face_update(zl_on || (ci && zl_ci), czL, vzl_old, zm, pc);
            stb(vo2, zo, vzL);
        }
        if (sz) {
            const float vzR = face_update(zr_on || (ci && zr_ci), czR, vzr_old, pc, zp);
            if (own_zR) stb(vo2, zo + 4u, vzR);
            n2 = split_update(czL, q2, vzR - vzL);
            stb(pd2, co, n2);
        }
        float pn;
        if (m) {
            float lap = 0.f;
            if constexpr (A != 0) lap += transverse_term(cxC, cxR, pm, pc, pp);
            if constexpr (A != 1) lap += transverse_term(cyL, cyR, ym, pc, yp);
            if constexpr (A != 2) lap += transverse_term(czL, czR, zm, pc, zp);
            // p^{n+1} = p^n + (pDir_A^{n+1} - pDir_A^n) + d^{n+1}: only
            // increments are rounded before the last addition
            const float dn = qd + lap;
            const float dA = (A == 0 ? n0 : A == 1 ? n1 : n2) - (A == 0 ? q0 : A == 1 ? q1 : q2);
            stb(darr, co, dn);
            pn = pc + (dA + dn);
        } else {
            pn = n0 + n1 + n2;
        }
        if (solid_col || (fxC & 16)) pn = 0.f;
        stb(pnext, po, pn);
        if (t.te > t.tb) emit_cell_terms(S, t, terms, step, (int)(po >> 2), pn);
        pm = pc;
        pc = pp;
    }
}

__device__ __forceinline__ void pml_tile(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step)
{
    const PmlSlab sl = cmem(S.slabs)[t.slab];
    if (sl.axis == 0) pml_tile_axis<0>(S, t, sl, terms, step);
    else if (sl.axis == 1) pml_tile_axis<1>(S, t, sl, terms, step);
    else pml_tile_axis<2>(S, t, sl, terms, step);
}

// ---------------------------------------------------------------------------
// staggered (MAC) leapfrog, PML_WaveSolver::stepLeapfrog (PML_WaveSolver.cpp:
// 710-773): face velocities v <- f1 v + f2 (p_R - p_L) with zero wall faces
// (MAC_Grid.cpp:301-370), split pressures pDir_d <- f3 pDir_d + f5/h
// (v_R - v_L) plus the source on every axis whose sigma is zero at the cell
// (948-1020), p = sum_d pDir_d (1102-1129).  Outside the PML f1 = f3 = 1 and
// the split pressures are not stored: p^{n+1} = p^n + sum_d increments.  Face
// velocities are full-grid and double-buffered; a thread writes the low
// faces of its cells and recomputes the high ones from the old values.
// ---------------------------------------------------------------------------
// One staggered face update, v <- f1 v + f2 (p_R - p_L), with its rounding
// fixed by an explicit fma: a face is computed by two threads (as the high
// face of one cell and the low face of the next) and must come out the same.
__device__ __forceinline__ float mac_face(float4 c, float v, float pl, float pr)
{
    return fmaf(c.y, pr - pl, c.x * v);
}

template <bool PML>
__device__ __forceinline__ void mac_tile(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step)
{
    const int tk = t.tk_log2;
    const int tx = threadIdx.x & ((1 << tk) - 1), ty = threadIdx.x >> tk;
    const int k = t.k0 + tx, j = t.j0 + ty;
    if (k >= t.k1 || j >= t.j1) return;
    const int nx = S.n[0], ny = S.n[1], nz = S.n[2];
    const int sx = S.sx, sy = S.sy;
    const AS_GLOBAL float *__restrict__ pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    AS_GLOBAL float *__restrict__ pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    const int vb = (int)(step & 1);
    const AS_GLOBAL float *__restrict__ vi0 = gmem(S.vel[vb][0]);
    const AS_GLOBAL float *__restrict__ vi1 = gmem(S.vel[vb][1]);
    const AS_GLOBAL float *__restrict__ vi2 = gmem(S.vel[vb][2]);
    AS_GLOBAL float *__restrict__ vo0 = gmem(S.vel[vb ^ 1][0]);
    AS_GLOBAL float *__restrict__ vo1 = gmem(S.vel[vb ^ 1][1]);
    AS_GLOBAL float *__restrict__ vo2 = gmem(S.vel[vb ^ 1][2]);
    const AS_CONST float4 *__restrict__ cx = cmem(S.coef[0]);
    const AS_CONST int *__restrict__ fx = cmem(S.flags[0]);
    const AS_CONST float4 *cy = cmem(S.coef[1]), *cz = cmem(S.coef[2]);
    const AS_CONST int *fy = cmem(S.flags[1]), *fz = cmem(S.flags[2]);
    const float4 cyL = cy[j], cyR = cy[min(j + 1, ny)], czL = cz[k], czR = cz[min(k + 1, nz)];
    const bool has_yL = j > 0, has_yR = j + 1 < ny, has_zL = k > 0, has_zR = k + 1 < nz;
    const bool src_y = fy[j] & 4, src_z = fz[k] & 4;   // sigma_d == 0 at the cell
    const int ym_off = has_yL ? sy : 0, yp_off = has_yR ? sy : 0;
    const int zm_off = has_zL ? 1 : 0, zp_off = has_zR ? 1 : 0;

    AS_GLOBAL float *pd0 = nullptr, *pd1 = nullptr, *pd2 = nullptr;
    int c = 0, cs = 0;
    if constexpr (PML) {
        const PmlSlab sl = cmem(S.slabs)[t.slab];
        pd0 = gmem(S.pd[0] + sl.cell_base);
        pd1 = gmem(S.pd[1] + sl.cell_base);
        pd2 = gmem(S.pd[2] + sl.cell_base);
        cs = sl.e[1] * sl.e[2];
        c = ((t.i0 - sl.lo[0]) * sl.e[1] + (j - sl.lo[1])) * sl.e[2] + (k - sl.lo[2]);
    }
    int idx = t.i0 * sx + j * sy + k;
    float pm = pcur[idx - (t.i0 > 0 ? sx : 0)];
    float pc = pcur[idx];
    float vxL = t.i0 > 0 ? mac_face(cx[t.i0], vi0[idx], pm, pc) : 0.f;
    for (int i = t.i0; i < t.i1; ++i, idx += sx, c += cs) {
        const bool has_xR = i + 1 < nx;
        const float pp = pcur[idx + (has_xR ? sx : 0)];
        const float ym = pcur[idx - ym_off], yp = pcur[idx + yp_off];
        const float zm = pcur[idx - zm_off], zp = pcur[idx + zp_off];
        const float4 cxC = cx[i], cxR = cx[min(i + 1, nx)];
        const float vxR = has_xR ? mac_face(cxR, vi0[idx + sx], pc, pp) : 0.f;
        const float vyL = has_yL ? mac_face(cyL, vi1[idx], ym, pc) : 0.f;
        const float vyR = has_yR ? mac_face(cyR, vi1[idx + sy], pc, yp) : 0.f;
        const float vzL = has_zL ? mac_face(czL, vi2[idx], zm, pc) : 0.f;
        const float vzR = has_zR ? mac_face(czR, vi2[idx + 1], pc, zp) : 0.f;
        vo0[idx] = vxL;
        vo1[idx] = vyL;
        vo2[idx] = vzL;
        // products and sums rounded one by one (no contraction), as in the
        // float4 row tiles, so every tile plan gives the same bits
        const float ix = __fmul_rn(cxC.w, vxR - vxL), iy = __fmul_rn(cyL.w, vyR - vyL), iz = __fmul_rn(czL.w, vzR - vzL);
        float src = 0.f;
        if (t.src)
            for (int s = 0; s < S.n_psrc; ++s) {
                const PSrcDev ps = cmem(S.psrc)[s];
                const AS_CONST float *g = cmem(S.gtab);
                src += gmem(S.src_tab)[amp_row(step) * S.n_psrc + s] * g[ps.goff[0] + i] * g[ps.goff[1] + j] * g[ps.goff[2] + k];
            }
        float pn;
        if constexpr (PML) {
            const bool src_x = fx[i] & 4;
            const float n0 = fmaf(cxC.z, pd0[c], ix) + (src_x ? src : 0.f);
            const float n1 = fmaf(cyL.z, pd1[c], iy) + (src_y ? src : 0.f);
            const float n2 = fmaf(czL.z, pd2[c], iz) + (src_z ? src : 0.f);
            pd0[c] = n0;
            pd1[c] = n1;
            pd2[c] = n2;
            pn = n0 + n1 + n2;
        } else {
            pn = __fadd_rn(pc, fmaf(3.f, src, __fadd_rn(__fadd_rn(ix, iy), iz)));
        }
        pnext[idx] = pn;
        if (t.te > t.tb) emit_cell_terms(S, t, terms, step, idx, pn);
        vxL = vxR;
        pm = pc;
        pc = pp;
    }
}

// Staggered scheme, cells outside the PML, whole rows with four cells per
// lane (the lane and row layout of interior_vec_tile).  No face there is a
// wall (walls lie in the PML); each lane updates the five z faces k .. k + 4 of its
// group (the fifth is the next lane's first, recomputed the same way), its
// cells' low x and y faces (the high ones are recomputed by the neighbours
// that own them) and its four pressures, marching the tile's planes with the
// x face carried, and writes the listener terms of its cells.  Arithmetic as
// mac_tile<false>, cell by cell.  The host
// builds these rows only where each axis's coefficient table holds a single
// value over the tile's faces and cells (outside the PML: f1 = 1, sigma = 0),
// so one entry per axis is held instead of one per face.
__device__ __forceinline__ void mac_vec_tile(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step)
{
    const int nk4 = (t.k1 - t.k0 + 3) >> 2, R = t.rpw, rem = (t.k1 - t.k0) & 3;
    const int j0w = t.j0 + (int)(threadIdx.x >> 6) * R;
    if (j0w >= t.j1) return;
    const int l64 = threadIdx.x & 63, row = R == 1 ? 0 : l64 / nk4;
    const int lr = l64 - row * nk4, jr = j0w + row;
    const bool act = row < R && lr < nk4 && jr < t.j1;
    const int j = act ? jr : t.j0;
    const int lane = act ? lr : 0;
    const int k = t.k0 + 4 * lane;
    const int sx = S.sx, sy = S.sy;
    const AS_GLOBAL float *__restrict__ pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    AS_GLOBAL float *__restrict__ pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    const int vb = (int)(step & 1);
    const AS_GLOBAL float *__restrict__ vi0 = gmem(S.vel[vb][0]);
    const AS_GLOBAL float *__restrict__ vi1 = gmem(S.vel[vb][1]);
    const AS_GLOBAL float *__restrict__ vi2 = gmem(S.vel[vb][2]);
    AS_GLOBAL float *__restrict__ vo0 = gmem(S.vel[vb ^ 1][0]);
    AS_GLOBAL float *__restrict__ vo1 = gmem(S.vel[vb ^ 1][1]);
    AS_GLOBAL float *__restrict__ vo2 = gmem(S.vel[vb ^ 1][2]);
    const AS_CONST float4 *__restrict__ cx = cmem(S.coef[0]);
    const AS_CONST float4 *cy = cmem(S.coef[1]), *cz = cmem(S.coef[2]);
    const float4 cxv = cx[t.i0], cyv = cy[j], czv = cz[t.k0 + t.lead];
    int idx = t.i0 * sx + j * sy + k;
    float4 pm = *(const AS_GLOBAL float4 *)(pcur + idx - sx);
    float4 pc = *(const AS_GLOBAL float4 *)(pcur + idx);
    float4 vxL;
    {
        const float4 v = *(const AS_GLOBAL float4 *)(vi0 + idx);
        vxL = make_float4(mac_face(cxv, v.x, pm.x, pc.x), mac_face(cxv, v.y, pm.y, pc.y), mac_face(cxv, v.z, pm.z, pc.z),
                          mac_face(cxv, v.w, pm.w, pc.w));
    }
    for (int i = t.i0; i < t.i1; ++i, idx += sx) {
        const float4 pp = *(const AS_GLOBAL float4 *)(pcur + idx + sx);
        const float4 ym = *(const AS_GLOBAL float4 *)(pcur + idx - sy);
        const float4 yp = *(const AS_GLOBAL float4 *)(pcur + idx + sy);
        const float4 v0R = *(const AS_GLOBAL float4 *)(vi0 + idx + sx);
        const float4 v1L = *(const AS_GLOBAL float4 *)(vi1 + idx);
        const float4 v1R = *(const AS_GLOBAL float4 *)(vi1 + idx + sy);
        const float4 v2 = *(const AS_GLOBAL float4 *)(vi2 + idx);
        float zl = __shfl_up(pc.w, 1), zr = __shfl_down(pc.x, 1), v2n = __shfl_down(v2.x, 1);
        if (lane == 0) zl = pcur[idx - 1];
        if (lane == nk4 - 1) { zr = pcur[idx + 4]; v2n = vi2[idx + 4]; }
        const float4 vxR = make_float4(mac_face(cxv, v0R.x, pc.x, pp.x), mac_face(cxv, v0R.y, pc.y, pp.y),
                                       mac_face(cxv, v0R.z, pc.z, pp.z), mac_face(cxv, v0R.w, pc.w, pp.w));
        const float4 vyL = make_float4(mac_face(cyv, v1L.x, ym.x, pc.x), mac_face(cyv, v1L.y, ym.y, pc.y),
                                       mac_face(cyv, v1L.z, ym.z, pc.z), mac_face(cyv, v1L.w, ym.w, pc.w));
        const float4 vyR = make_float4(mac_face(cyv, v1R.x, pc.x, yp.x), mac_face(cyv, v1R.y, pc.y, yp.y),
                                       mac_face(cyv, v1R.z, pc.z, yp.z), mac_face(cyv, v1R.w, pc.w, yp.w));
        const float f0 = mac_face(czv, v2.x, zl, pc.x), f1 = mac_face(czv, v2.y, pc.x, pc.y);
        const float f2 = mac_face(czv, v2.z, pc.y, pc.z), f3 = mac_face(czv, v2.w, pc.z, pc.w);
        const float f4 = mac_face(czv, v2n, pc.w, zr);
        float4 pn;
        pn.x = __fadd_rn(pc.x, fmaf(3.f, 0.f, __fadd_rn(__fadd_rn(__fmul_rn(cxv.w, vxR.x - vxL.x), __fmul_rn(cyv.w, vyR.x - vyL.x)),
                                                      __fmul_rn(czv.w, f1 - f0))));
        pn.y = __fadd_rn(pc.y, fmaf(3.f, 0.f, __fadd_rn(__fadd_rn(__fmul_rn(cxv.w, vxR.y - vxL.y), __fmul_rn(cyv.w, vyR.y - vyL.y)),
                                                      __fmul_rn(czv.w, f2 - f1))));
        pn.z = __fadd_rn(pc.z, fmaf(3.f, 0.f, __fadd_rn(__fadd_rn(__fmul_rn(cxv.w, vxR.z - vxL.z), __fmul_rn(cyv.w, vyR.z - vyL.z)),
                                                      __fmul_rn(czv.w, f3 - f2))));
        pn.w = __fadd_rn(pc.w, fmaf(3.f, 0.f, __fadd_rn(__fadd_rn(__fmul_rn(cxv.w, vxR.w - vxL.w), __fmul_rn(cyv.w, vyR.w - vyL.w)),
                                                      __fmul_rn(czv.w, f4 - f3))));
        if (act) {
            const int n = (lane == nk4 - 1 && rem) ? rem : 4, lo = lane == 0 ? t.lead : 0;
            store4(vo0 + idx, vxL, n, lo);
            store4(vo1 + idx, vyL, n, lo);
            store4(vo2 + idx, make_float4(f0, f1, f2, f3), n, lo);
            store4(pnext + idx, pn, n, lo);
            if (t.te > t.tb) {
                if (lo == 0) emit_cell_terms(S, t, terms, step, idx, pn.x);
                if (n > 1 && lo <= 1) emit_cell_terms(S, t, terms, step, idx + 1, pn.y);
                if (n > 2 && lo <= 2) emit_cell_terms(S, t, terms, step, idx + 2, pn.z);
                if (n > 3) emit_cell_terms(S, t, terms, step, idx + 3, pn.w);
            }
        }
        vxL = vxR;
        pm = pc;
        pc = pp;
    }
}

// ---------------------------------------------------------------------------
// bulk cells next to rigid objects (FP64)
// ---------------------------------------------------------------------------
// One axis of the ghost-aware Laplacian (MAC_Grid::
// pressureFieldLaplacianGhostCell, MAC_Grid.cpp:187-299) at a cell with
// pressure c, neighbours nm / np and ghost values gm / gp where the
// neighbour is solid: the plain 3-point stencil with ghost values (centre
// sampling, 247-266), or the non-uniform one for ghosts 3h/4 away (seven
// samples, 268-286).
__device__ __forceinline__ double ghost_lap_axis(int seven, bool ghm, bool ghp, double gm, double gp, double nm, double np,
                                                 double c)
{
    const double m = ghm ? gm : nm, p = ghp ? gp : np;
    if (!seven || (!ghm && !ghp)) return p + m - 2.0 * c;
    if (ghm && ghp) return (gp + gm - 2.0 * c) / (9.0 / 16.0);
    if (ghp) return (gp + 0.75 * nm - (7.0 / 4.0) * c) / (21.0 / 32.0);
    return (0.75 * np + gm - (7.0 / 4.0) * c) / (21.0 / 32.0);
}

__device__ __forceinline__ void special_tile(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step)
{
    const int r = t.i0 + (int)threadIdx.x;
    if (r >= t.i1) return;
    const SpecialRec rec = gmem(S.recs)[r];
    const int sx = S.sx, sy = S.sy;
    const AS_GLOBAL float *__restrict__ pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    const AS_GLOBAL float *pprev = gmem(S.field + (size_t)slot_next(step) * S.npad);
    AS_GLOBAL float *pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    const int idx = rec.cell;
    const double c = pcur[idx];
    const double nb[6] = {pcur[idx - sx], pcur[idx + sx], pcur[idx - sy], pcur[idx + sy], pcur[idx - 1], pcur[idx + 1]};
    double g[6];
    for (int q = 0; q < 6; ++q) {
        g[q] = 0.0;
        const int gid = rec.gid[q];
        if (gid >= 0) {
            if (S.raster) {
                const GhostDev gd = gmem(S.ghosts)[gid];
                g[q] = c + raster_accel_term(S, gd, step);
            } else {
                g[q] = coupled_ghost(S, gid, pcur, step);
            }
        }
    }
    double pn;
    if (rec.flags & 1) {
        pn = c;  // one-cell cavity: p^{n+1} = p^n (MAC_Grid.cpp:924-925)
    } else {
        double lap = 0.0;
        for (int d = 0; d < 3; ++d)
            lap += ghost_lap_axis(S.seven, rec.gid[2 * d] >= 0, rec.gid[2 * d + 1] >= 0, g[2 * d], g[2 * d + 1], nb[2 * d],
                                  nb[2 * d + 1], c);
        if (S.lapv) {
            // air viscosity: the ghost-aware Laplacian L^n = lap / h^2 (FP64)
            // and the previous step's L^{n-1} of this cell
            const double lc = lap * S.inv_h2;
            pn = 2.0 * c - (double)pprev[idx] - S.kcalp * (double)gmem(S.lapv)[idx] + (S.c2k2 + S.kcalp) * lc;
            gmem(S.lapv)[idx] = (float)lc;
        } else {
            pn = 2.0 * c - (double)pprev[idx] + lap * S.lam2;
        }
    }
    if (S.n_psrc > 0) {
        const int i = idx / sx, j = (idx - i * sx) / sy, k = idx - i * sx - j * sy;
        for (int s = 0; s < S.n_psrc; ++s) {
            const PSrcDev ps = cmem(S.psrc)[s];
            const AS_CONST float *g3 = cmem(S.gtab);
            pn += (double)(gmem(S.src_tab)[amp_row(step) * S.n_psrc + s] * g3[ps.goff[0] + i] * g3[ps.goff[1] + j] * g3[ps.goff[2] + k]);
        }
    }
    if (S.n_spk) {
        const int i = idx / sx, j = (idx - i * sx) / sy, k = idx - i * sx - j * sy;
        pn += (double)speaker_at(S, i, j, k);
    }
    pnext[idx] = (float)pn;
    if (t.te > t.tb) emit_cell_terms(S, t, terms, step, idx, (float)pn);
}

// ---------------------------------------------------------------------------
// Engquist-Majda absorbing boundary on the layer-1 cells of a grid without a
// PML (MAC_Grid.cpp:521-919): order 2 on the faces, order 1 on edges and
// corners, FP64 arithmetic on FP32 fields.  The update extrapolates the
// pressures of the exterior (layer-0) neighbours into p^n, where the next
// step reads them back as p^{n-1}; edges and corners also zero the diagonal
// exterior cell.  Layer 0 itself is never updated.
//
// A face cell also reads p^{n-1} of its inner neighbour, which the interior
// update overwrites in place during the same step.  The face cell keeps that
// value itself: after reading p^{n-1} at its exterior cell from the slot
// p^{n+1} goes to, it stores the inner neighbour's p^n there, and reads it
// back in the next step from the same cell, then the current slot.  Only the
// face cell's own thread touches its exterior cell, in either slot.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void abc_tile(const SceneDev &S, const TileDesc &t, const TermRef *terms, long long step)
{
    const int nk = t.k1 - t.k0, nj = t.j1 - t.j0;
    const int q = (int)threadIdx.x, r = q / nk;
    const int k = t.k0 + (q - r * nk), j = t.j0 + r % nj, i = t.i0 + r / nj;
    if (i >= t.i1) return;
    const int ijk[3] = {i, j, k};
    const int st[3] = {S.sx, S.sy, 1};
    int nml[3];
    for (int d = 0; d < 3; ++d) nml[d] = ijk[d] == 1 ? -1 : (ijk[d] == S.n[d] - 2 ? 1 : 0);
    const int btype = abs(nml[0]) + abs(nml[1]) + abs(nml[2]);
    if (btype == 0) return;
    AS_GLOBAL float *pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    AS_GLOBAL float *pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    const int idx = i * S.sx + j * S.sy + k;
    const double lam = S.lam, lam2 = lam * lam;
    const double pc = pcur[idx], pl = pnext[idx];
    double pn;
    if (btype == 1) {
        const int di = nml[0] ? 0 : (nml[1] ? 1 : 2);
        const int dj = di == 2 ? 0 : di + 1, dk = di == 0 ? 2 : di - 1;
        const int out = nml[di] * st[di];
        const int jp = ijk[dj] < S.n[dj] - 1 ? st[dj] : 0, jm = ijk[dj] > 0 ? st[dj] : 0;
        const int kp = ijk[dk] < S.n[dk] - 1 ? st[dk] : 0, km = ijk[dk] > 0 ? st[dk] : 0;
        const double pl_out = pnext[idx + out];   // p^{n-1} of the exterior cell
        const float pin_f = pcur[idx - out];
        const double pl_in = pcur[idx + out];     // the kept p^{n-1} of the inner neighbour
        const double p_in = pin_f;
        const double tang = ((double)pcur[idx + jp] + (double)pcur[idx - jm]) + ((double)pcur[idx + kp] + (double)pcur[idx - km]);
        double p1 = (2. / lam2 + 4. / lam - 6. - 4. * lam) * pc - (1. / lam2 + 2. / lam) * pl + pl_out - pl_in + 2. * p_in +
                    (lam + 1.) * tang;
        p1 *= lam2 / (1. + 2. * lam);
        const double p_out = pl_out + p_in - pl_in - 2. / lam * (p1 + pl - 2. * pc) + lam * (tang - 4. * pc);
        pcur[idx + out] = (float)p_out;
        pnext[idx + out] = pin_f;
        pn = 2. * pc - pl + lam2 * (p_in + tang + p_out - 6. * pc);
    } else if (btype == 2) {
        const int di = nml[0] ? 0 : 1;
        const int dj = (di == 0 && nml[1]) ? 1 : 2;
        const int dk = 3 - di - dj;
        const int oi = nml[di] * st[di], oj = nml[dj] * st[dj];
        const int kp = ijk[dk] < S.n[dk] - 1 ? st[dk] : 0, km = ijk[dk] > 0 ? st[dk] : 0;
        const double pi_in = pcur[idx - oi], pj_in = pcur[idx - oj];
        const double pk = (double)pcur[idx + kp] + (double)pcur[idx - km];
        double p1 = (2. / lam2 - 6.) * pc + (2. * lam - 1.) / lam2 * pl + 2. * pi_in + 2. * pj_in + pk;
        p1 *= lam2 / (1. + 2. * lam);
        const double pi_out = (pl - p1) / lam + pi_in, pj_out = (pl - p1) / lam + pj_in;
        pcur[idx + oi] = (float)pi_out;
        pcur[idx + oj] = (float)pj_out;
        pcur[idx + oi + oj] = 0.f;
        pn = 2. * pc - pl + lam2 * (pi_in + pj_in + pk + pi_out + pj_out - 6. * pc);
    } else {
        const int oi = nml[0] * st[0], oj = nml[1] * st[1], ok = nml[2] * st[2];
        const double pi_in = pcur[idx - oi], pj_in = pcur[idx - oj], pk_in = pcur[idx - ok];
        double p1 = (2. / lam2 - 6.) * pc + (3. * lam - 1.) / lam2 * pl + 2. * pi_in + 2. * pj_in + 2. * pk_in;
        p1 *= lam2 / (1. + 3. * lam);
        const double e = (pl - p1) / lam;
        const double pi_out = e + pi_in, pj_out = e + pj_in, pk_out = e + pk_in;
        pcur[idx + oi] = (float)pi_out;
        pcur[idx + oj] = (float)pj_out;
        pcur[idx + ok] = (float)pk_out;
        pcur[idx + oi + oj + ok] = 0.f;
        pn = 2. * pc - pl + lam2 * (pi_in + pj_in + pk_in + pi_out + pj_out + pk_out - 6. * pc);
    }
    if (t.src)
        for (int s = 0; s < S.n_psrc; ++s) {
            const PSrcDev ps = cmem(S.psrc)[s];
            const AS_CONST float *g = cmem(S.gtab);
            pn += (double)(gmem(S.src_tab)[amp_row(step) * S.n_psrc + s] * g[ps.goff[0] + i] * g[ps.goff[1] + j] * g[ps.goff[2] + k]);
        }
    if (t.src && S.n_spk) pn += (double)speaker_at(S, i, j, k);
    pnext[idx] = (float)pn;
    if (t.te > t.tb) emit_cell_terms(S, t, terms, step, idx, (float)pn);
}

// ---------------------------------------------------------------------------
// moving objects (FP64): per-step reclassification of the cells the objects
// can reach (MAC_Grid::classifyCells_FAST), fresh cells
// (MAC_Grid::InterpolateFreshPressureCell, MAC_Grid.cpp:1993-2254) and the
// ghost-aware update of the fluid cells among them
// ---------------------------------------------------------------------------
__device__ inline void mv_centre(const SceneDev &S, int i, int j, int k, double *x)
{
    const int ijk[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) x[d] = S.bmin[d] + (0.5 + (double)ijk[d]) * S.h;
}

// fluid: the centre, and by the seven-sample rule also the six points a
// quarter cell away along the axes, lie outside every object
__device__ inline bool mv_fluid_at(const SceneDev &S, const Geom *pose, const double *x)
{
    const double q4 = 0.25 * S.h;
    const int nsmp = S.seven ? 7 : 1;
    for (int smp = 0; smp < nsmp; ++smp) {
        double y[3] = {x[0], x[1], x[2]};
        if (smp > 0) y[(smp - 1) >> 1] += ((smp - 1) & 1) ? q4 : -q4;
        for (int o = 0; o < S.n_obj; ++o)
            if (geom_inside(pose[o], y)) return false;
    }
    return true;
}

__device__ inline int mv_uidx(const SceneDev &S, int i, int j, int k)
{
    const int a = i - S.u_lo[0], b = j - S.u_lo[1], c = k - S.u_lo[2];
    if (a < 0 || b < 0 || c < 0 || a >= S.u_n[0] || b >= S.u_n[1] || c >= S.u_n[2]) return -1;
    return (a * S.u_n[1] + b) * S.u_n[2] + c;
}

// LowestObjectDistance (FDTD_Objects.cpp:240-257): the object nearest to x by |distance|
__device__ inline int mv_nearest(const SceneDev &S, const Geom *pose, const double *x, double *dist)
{
    if (S.n_obj == 1) { *dist = 0.0; return 0; }   // the only object (the distance is not used then)
    double best = DBL_MAX;
    int id = 0;
    for (int o = 0; o < S.n_obj; ++o) {
        double xb[3], n[3];
        const double d = geom_closest(pose[o], x, xb, n);
        if (fabs(d) < fabs(best)) { best = d; id = o; }
    }
    *dist = best;
    return id;
}

// this step's classification of cell q (cells outside the box are fluid):
// the same function of the poses the box's own threads evaluate for their
// cells, so a thread recomputes it for the neighbours it reads
__device__ inline bool mv_fluid_now(const SceneDev &S, const Geom *pose, const int *q)
{
    if (mv_uidx(S, q[0], q[1], q[2]) < 0) return true;
    double x[3];
    mv_centre(S, q[0], q[1], q[2], x);
    return mv_fluid_at(S, pose, x);
}

// fresh values of cell (i, j, k) in both pressure slots (p^{s-1} in plast,
// p^s in pcur; one geometry and one set of MLS weights for both); prev_f is
// the box's classification of the previous step, acc[o] object o's normal
// acceleration a_s at this step
__device__ void mv_fresh(const SceneDev &S, const Geom *pose, const double *acc, const float *plast, const float *pcur,
                         const uint8_t *prev_f, int i, int j, int k, double *vl, double *vc)
{
    double x[3];
    mv_centre(S, i, j, k, x);
    double dist;
    const int id = mv_nearest(S, pose, x, &dist);
    // ReflectAgainstBoundary (FDTD_RigidObject.cpp:429-536) for a point outside
    double xb[3], nb[3], en[3], image[3];
    const double travelled = geom_closest(pose[id], x, xb, nb);
    if (travelled < 0.0) { *vl = *vc = __builtin_nan(""); return; }   // the reference throws: a fresh cell inside an object
    const double dd = sqrt((xb[0] - x[0]) * (xb[0] - x[0]) + (xb[1] - x[1]) * (xb[1] - x[1]) + (xb[2] - x[2]) * (xb[2] - x[2]));
    if (dd > 1e-10) for (int e = 0; e < 3; ++e) { en[e] = x[e] - xb[e]; image[e] = x[e] + en[e]; }
    else for (int e = 0; e < 3; ++e) { en[e] = nb[e]; image[e] = x[e]; }
    const double enn = sqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
    for (int e = 0; e < 3; ++e) en[e] /= enn;
    // enclosing cells of the image point (ScalarField::enclosingNeighbours), in
    // flat-index order without repeats, those with valid history; each adds
    // its row to the MLS sums and competes for the nearest-neighbour fallback
    int lo3[3], hi3[3];
    for (int d = 0; d < 3; ++d) {
        double diff = image[d] - S.bmin[d];
        diff -= S.h / 2.0;
        int v0 = (int)(diff / S.h);
        v0 = min(max(v0, 0), S.n[d] - 1);
        lo3[d] = v0;
        hi3[d] = diff < 0.0 ? 0 : min(v0 + 1, S.n[d] - 1);
    }
    MlsSums mls;
    double bd = DBL_MAX, nl = 0.0, nc = 0.0;
    // every corner's history flag and values are loaded first, in one round
    // trip (the clamped corners are always inside the grid)
    bool hist[8];
    float vl8[8], vc8[8];
#pragma unroll
    for (int sl = 0; sl < 8; ++sl) {
        const int q[3] = {(sl >> 2) ? hi3[0] : lo3[0], ((sl >> 1) & 1) ? hi3[1] : lo3[1], (sl & 1) ? hi3[2] : lo3[2]};
        const int u = mv_uidx(S, q[0], q[1], q[2]);
        const int qi = q[0] * S.sx + q[1] * S.sy + q[2];
        hist[sl] = u < 0 || prev_f[u];
        vl8[sl] = plast[qi];
        vc8[sl] = pcur[qi];
    }
#pragma unroll
    for (int sl = 0; sl < 8; ++sl) {
        const int a = sl >> 2, b = (sl >> 1) & 1, c = sl & 1;
        const int q[3] = {a ? hi3[0] : lo3[0], b ? hi3[1] : lo3[1], c ? hi3[2] : lo3[2]};
        if ((a && hi3[0] == lo3[0]) || (b && hi3[1] == lo3[1]) || (c && hi3[2] == lo3[2])) continue;
        if (!hist[sl]) continue;   // no valid history
        double y[3];
        mv_centre(S, q[0], q[1], q[2], y);
        // a cell that turned solid this step was zeroed
        const bool fluid = mv_fluid_now(S, pose, q);
        const double l = fluid ? (double)vl8[sl] : 0.0, cv = fluid ? (double)vc8[sl] : 0.0;
        const double d2 = (image[0] - y[0]) * (image[0] - y[0]) + (image[1] - y[1]) * (image[1] - y[1]) + (image[2] - y[2]) * (image[2] - y[2]);
        if (d2 < bd) { bd = d2; nl = l; nc = cv; }
        mls.add(y, image, S.h, 0.018, l, cv);
    }
    if (mls.m >= 4) {
        // MLS at the image point plus rho a_n |d|, a_n along the erected normal
        double ml, mc;
        mls.solve(&ml, &mc);
        const double add = S.rho * acc[id] * (nb[0] * en[0] + nb[1] * en[1] + nb[2] * en[2]) * fabs(travelled);
        *vl = ml + add;
        *vc = mc + add;
        return;
    }
    if (mls.m > 0) { *vl = nl; *vc = nc; return; }
    // no valid enclosing cell: the first fluid axis neighbour (+x, -x, +y, ...);
    // a neighbour that is itself fresh this step counts with the zero it held
    const int ijk[3] = {i, j, k};
    for (int d = 0; d < 3; ++d)
        for (int sg = 1; sg >= -1; sg -= 2) {
            int q[3] = {ijk[0], ijk[1], ijk[2]};
            q[d] += sg;
            const int u = mv_uidx(S, q[0], q[1], q[2]);
            if (u >= 0 && !mv_fluid_now(S, pose, q)) continue;
            if (u >= 0 && !prev_f[u]) { *vl = *vc = 0.0; return; }
            const int qi = q[0] * S.sx + q[1] * S.sy + q[2];
            *vl = plast[qi];
            *vc = pcur[qi];
            return;
        }
    *vl = *vc = 0.0;
}

// Moving objects, before the step kernels of step s, in one launch over the
// box of each moving scene (six threads per cell along x, scenes along y).
// Each workgroup takes the objects' poses at t_s from the pose ring into LDS.
// Thread (q, c), q = 0..5 the directions -x, +x, -y, +y, -z, +z:
// * q = 0 writes cell c's classification (the seven-sample or centre rule)
//   into the row of this step's parity, zeroes the pressures (both slots, as
//   classifyCells_FAST zeroes every pressure buffer) of a cell that turned
//   solid and fills a fresh cell (p^{s-1}, then p^s);
// * every q of a fluid cell writes the rasterized ghost term of its solid
//   neighbour in direction q.
// Classifications of other cells are recomputed from the poses, not read, so
// no thread depends on another's writes: fresh values come from cells with
// valid history that are still fluid, and cells that turned solid count as
// zero without being read.  The step kernels after this launch read the row.
__global__ __launch_bounds__(256) void motion_kernel(const SceneDev *__restrict__ scenes, const MotionDesc *__restrict__ list, int k)
{
    const SceneDev &S = scenes[list[blockIdx.y].scene];
    const long long step = (long long)S.clock[CLK_STEP] + k;
    if (step >= (long long)S.clock[CLK_TARGET]) return;
    const int ncell = S.u_n[0] * S.u_n[1] * S.u_n[2];
    if ((int)blockIdx.x * (int)blockDim.x >= 6 * ncell) return;
    // the poses and the objects' normal accelerations a_s at t_s, and this
    // thread's previous classification, all loaded before the barrier so no
    // dependent load waits at the end of the chain
    __shared__ Geom pose[MAX_MOVING_OBJ];
    __shared__ double acc[MAX_MOVING_OBJ];
    for (int o = threadIdx.x; o < S.n_obj; o += blockDim.x) {
        pose[o] = S.pose[(size_t)amp_row(step) * S.n_obj + o];
        acc[o] = gmem(S.acc_tab)[amp_row(step) * S.n_ch + o * S.n_cpo];
    }
    const int tid = (int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x;
    const int q = tid / ncell, c = tid - q * ncell;
    uint8_t *now_row = S.ucls + (size_t)(step & 1) * ncell;
    const uint8_t *prev_f = S.ucls + (size_t)((step + 1) & 1) * ncell;
    const bool prev = tid < 6 * ncell ? prev_f[c] != 0 : false;
    __syncthreads();
    if (tid >= 6 * ncell) return;
    const int kk = c % S.u_n[2], r = c / S.u_n[2];
    const int j = S.u_lo[1] + r % S.u_n[1], i = S.u_lo[0] + r / S.u_n[1], kz = S.u_lo[2] + kk;
    double x[3];
    mv_centre(S, i, j, kz, x);
    const bool now = mv_fluid_at(S, pose, x);
    if (q == 0) {
        now_row[c] = now ? 1 : 0;
        if (now != prev) {
            const int idx = i * S.sx + j * S.sy + kz;
            float *pcur = S.field + (size_t)slot_cur(step) * S.npad;
            float *plast = S.field + (size_t)slot_next(step) * S.npad;
            if (!now) {
                pcur[idx] = 0.f;
                plast[idx] = 0.f;
            } else {
                double fl, fc;
                mv_fresh(S, pose, acc, plast, pcur, prev_f, i, j, kz, &fl, &fc);
                plast[idx] = (float)fl;
                pcur[idx] = (float)fc;
            }
        }
    }
    // rasterized ghost term of a fluid cell's solid neighbour in direction q:
    // f h rho a_n (n_b . n_cell), the ghost on the face (f = 1) or a quarter
    // cell from the solid cell's centre (seven samples, f = 0.75) towards
    // this cell, its object the one nearest the solid cell's centre, n_b the
    // outward normal at the ghost's closest boundary point
    if (!now) return;
    const int d = q >> 1, sg = (q & 1) ? 1 : -1;
    int m[3] = {i, j, kz};
    m[d] += sg;
    if (mv_uidx(S, m[0], m[1], m[2]) < 0 || mv_fluid_now(S, pose, m)) return;
    double xo[3], gp[3], dist, xb[3], nb[3];
    mv_centre(S, m[0], m[1], m[2], xo);
    const int id = mv_nearest(S, pose, xo, &dist);
    for (int e = 0; e < 3; ++e) gp[e] = xo[e];
    gp[d] -= sg * (S.seven ? 0.25 : 0.5) * S.h;
    geom_closest(pose[id], gp, xb, nb);
    S.ughost[(size_t)q * ncell + c] = (S.seven ? 0.75 : 1.0) * S.h * S.rho * acc[id] * nb[d] * -(double)sg;
}

// A box of a moving scene's cells (one thread per cell): solid cells are
// left alone; a fluid cell takes the ghost-aware Laplacian (MAC_Grid.cpp:
// 187-299) with rasterized ghost values p + (the ghost term motion_kernel
// wrote for that direction) for its solid neighbours, and a cell with ghosts
// on both sides along two axes keeps p^n when it was fluid in the previous
// step.
__device__ void moving_tile(const SceneDev &S, const TileDesc &t, long long step)
{
    const int nk = t.k1 - t.k0, nj = t.j1 - t.j0;
    const int q = (int)threadIdx.x, r = q / nk;
    const int k = t.k0 + (q - r * nk), j = t.j0 + r % nj, i = t.i0 + r / nj;
    if (i >= t.i1) return;
    const int ncell = S.u_n[0] * S.u_n[1] * S.u_n[2];
    const uint8_t *now_f = S.ucls + (size_t)(step & 1) * ncell;
    const uint8_t *prev_f = S.ucls + (size_t)((step + 1) & 1) * ncell;
    const int u = mv_uidx(S, i, j, k);
    if (u < 0 || !now_f[u]) return;
    const AS_GLOBAL float *pcur = gmem(S.field + (size_t)slot_cur(step) * S.npad);
    AS_GLOBAL float *pnext = gmem(S.field + (size_t)slot_next(step) * S.npad);
    const int st[3] = {S.sx, S.sy, 1};
    const int idx = i * S.sx + j * S.sy + k;
    const double c = pcur[idx];
    const int ijk[3] = {i, j, k};
    double lap = 0.0;
    int surrounded = 0;
    for (int d = 0; d < 3; ++d) {
        double g[2], nbv[2];
        bool solid[2];
        for (int side = 0; side < 2; ++side) {
            const int sg = side ? 1 : -1;
            int m[3] = {ijk[0], ijk[1], ijk[2]};
            m[d] += sg;
            const int um = mv_uidx(S, m[0], m[1], m[2]);
            solid[side] = um >= 0 && !now_f[um];
            nbv[side] = pcur[idx + sg * st[d]];
            g[side] = 0.0;
            if (solid[side]) g[side] = c + S.ughost[(size_t)(2 * d + side) * ncell + u];
        }
        // side 0 is the negative neighbour, side 1 the positive one
        lap += ghost_lap_axis(S.seven, solid[0], solid[1], g[0], g[1], nbv[0], nbv[1], c);
        if (solid[1] && solid[0]) ++surrounded;
    }
    double pn;
    if (surrounded > 1 && prev_f[u]) pn = c;   // MAC_Grid.cpp:924-925
    else pn = 2.0 * c - (double)pnext[idx] + lap * S.lam2;
    if (t.src)
        for (int s = 0; s < S.n_psrc; ++s) {
            const PSrcDev ps = cmem(S.psrc)[s];
            const AS_CONST float *g3 = cmem(S.gtab);
            pn += (double)(gmem(S.src_tab)[amp_row(step) * S.n_psrc + s] * g3[ps.goff[0] + i] * g3[ps.goff[1] + j] * g3[ps.goff[2] + k]);
        }
    if (t.src && S.n_spk) pn += (double)speaker_at(S, i, j, k);
    pnext[idx] = (float)pn;
}

// Tile of workgroup w.  Lists of 16 tiles or more are padded to a multiple of
// eight and arranged so that workgroup w (dispatched to XCD w % 8) runs chunk
// w % 8; the reversed order keeps that assignment and reverses each chunk.
// The first nhead tiles (a multiple of eight) are dispatched first on every
// step; the reversal applies to the rest.
__device__ __forceinline__ int tile_of(int w, int ntiles, int nhead, const unsigned long long *order_clock, int k)
{
    if (w < nhead || !order_clock || !((*order_clock + (unsigned long long)k) & 1ULL)) return w;
    const int r = w - nhead, nr = ntiles - nhead;
    if (nr >= 16 && (nr & 7) == 0) return nhead + (nr / 8 - 1 - r / 8) * 8 + (r & 7);
    return nhead + nr - 1 - r;
}

// Step advance (PML_WaveSolver::stepSystem's time update), one workgroup per
// scene after a block of steps: the counter moves n steps on (not past the
// target), and the amplitude ring (and a moving scene's pose ring) is
// refilled for the AMP_ROWS steps from there.  The kernel boundaries order it
// after every read of the block's amplitudes, poses and counter and before
// the next block reads them.
__global__ __launch_bounds__(256) void advance_kernel(const SceneDev *__restrict__ scenes, int n)
{
    const SceneDev &S = scenes[blockIdx.x];
    AS_GLOBAL unsigned long long *clk = gmem(S.clock);
    const long long step = (long long)clk[CLK_STEP], target = (long long)clk[CLK_TARGET];
    if (step >= target) return;
    const long long base = min(step + (long long)n, target);
    // the times of steps base .. base + AMP_ROWS - 1, accumulated from the
    // clock's time by one addition of dt per step (the reference's clock)
    __shared__ double tt[AMP_ROWS];
    if (threadIdx.x == 0) {
        double t = __longlong_as_double((long long)clk[CLK_TIME]);
        for (long long q = step; q < base; ++q) t += S.dt;
        for (int r = 0; r < AMP_ROWS; ++r) {
            tt[r] = t;
            t += S.dt;
        }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < AMP_ROWS; r += blockDim.x) gmem(S.t_tab)[amp_row(base + r)] = tt[r];
    // a time-parallel chunk's sources are off from step src_steps on
    const long long off = S.src_steps < 0 ? LLONG_MAX : S.src_steps;
    for (int q = threadIdx.x; q < AMP_ROWS * S.n_ch; q += blockDim.x) {
        const long long s = base + q / S.n_ch;
        const int ch = q % S.n_ch, o = ch / S.n_cpo, comp = ch % S.n_cpo;
        const double t = tt[q / S.n_ch];
        double a = 0.0;
        if (s < off) a = comp == 0 ? impulse_accel(S.isrc, S.n_isrc, o, t) : noise_accel(S.impacts, S.n_impact, o, comp - 1, t);
        gmem(S.acc_tab)[amp_row(s) * S.n_ch + ch] = a;
    }
    if (S.mv)
        for (int q = threadIdx.x; q < AMP_ROWS * S.n_obj; q += blockDim.x) {
            const long long s = base + q / S.n_obj;
            const int o = q % S.n_obj;
            const double t = tt[q / S.n_obj];
            const MoveDev &m = S.mobj[o];
            gmem(S.pose)[amp_row(s) * S.n_obj + o] =
                m.kin_n > 0 ? geom_keyframe(m.g0, S.kin + m.kin_off, m.kin_n, m.kin_dt, t) : geom_at(m.g0, m.v, m.w, t);
        }
    for (int q = threadIdx.x; q < AMP_ROWS * S.n_psrc; q += blockDim.x) {
        const long long s = base + q / S.n_psrc;
        const int p = q % S.n_psrc;
        gmem(S.src_tab)[amp_row(s) * S.n_psrc + p] = s < off ? source_amp(S.psrc[p], tt[q / S.n_psrc], S.dt, S.scheme) : 0.f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        clk[CLK_STEP] = (unsigned long long)base;
        clk[CLK_TIME] = (unsigned long long)__double_as_longlong(tt[0]);
    }
}

// One kernel per tile class, so each class gets its own register allocation
// (the FP64 object path would otherwise cap the occupancy of the streaming
// paths), or one launch for all classes (KIND_ALL).  The classes write
// disjoint cells and can run in any order.
template <int KIND>
__global__ __launch_bounds__(256) void step_kernel(const SceneDev *__restrict__ scenes, const TileDesc *__restrict__ tiles,
                                                   int ntiles, int nhead, const TermRef *__restrict__ terms,
                                                   const unsigned long long *__restrict__ order_clock, int k)
{
    const TileDesc t = gmem(tiles)[tile_of(blockIdx.x, ntiles, nhead, order_clock, k)];
    if (t.type == TILE_EMPTY) return;
    const SceneDev &S = scenes[t.scene];
    const long long step = (long long)S.clock[CLK_STEP] + k;
    if (step >= (long long)S.clock[CLK_TARGET]) return;
    if constexpr (KIND == KIND_INTERIOR) {
        // the interior launch also carries the object-adjacent cells, whose
        // FP64 gathers then hide under the interior's streaming
        if (t.type == TILE_INTERIOR_VEC) interior_vec_tile(S, t, terms, step);
        else if (t.type == TILE_SPECIAL) special_tile(S, t, terms, step);
        else if (t.type == TILE_MAC) mac_tile<false>(S, t, terms, step);
        else if (t.type == TILE_MAC_VEC) mac_vec_tile(S, t, terms, step);
        else interior_tile(S, t, terms, step, t.type == TILE_INTERIOR_SKIP);
    } else if constexpr (KIND == KIND_PML) {
        if (t.type == TILE_MAC_PML) mac_tile<true>(S, t, terms, step);
        else if (t.type == TILE_ABC) abc_tile(S, t, terms, step);
        else pml_tile(S, t, terms, step);
    } else if constexpr (KIND == KIND_SPECIAL) {
        // with one launch per class the moving-object tiles run in this
        // launch of their own, so they do not set the register budget of the
        // streaming tiles
        if (t.type == TILE_MOVING) moving_tile(S, t, step);
        else special_tile(S, t, terms, step);
    } else {
        if (t.type == TILE_PML) pml_tile(S, t, terms, step);
        else if (t.type == TILE_INTERIOR_VEC) interior_vec_tile(S, t, terms, step);
        else if (t.type == TILE_MOVING) moving_tile(S, t, step);
        else if (t.type == TILE_SPECIAL) special_tile(S, t, terms, step);
        else if (t.type == TILE_MAC) mac_tile<false>(S, t, terms, step);
        else if (t.type == TILE_MAC_PML) mac_tile<true>(S, t, terms, step);
        else if (t.type == TILE_MAC_VEC) mac_vec_tile(S, t, terms, step);
        else if (t.type == TILE_ABC) abc_tile(S, t, terms, step);
        else interior_tile(S, t, terms, step, t.type == TILE_INTERIOR_SKIP);
    }
    if (t.te > t.tb) emit_ghost_terms(S, t, terms, step);
}

hipError_t launch_step(int kind, const SceneDev *scenes, const TileDesc *tiles, int ntiles, int nhead, const TermRef *terms,
                       const unsigned long long *order_clock, int k, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    switch (kind) {
    case KIND_INTERIOR: hipLaunchKernelGGL(step_kernel<KIND_INTERIOR>, dim3(ntiles), dim3(256), 0, s, scenes, tiles, ntiles, nhead, terms, order_clock, k); break;
    case KIND_PML: hipLaunchKernelGGL(step_kernel<KIND_PML>, dim3(ntiles), dim3(256), 0, s, scenes, tiles, ntiles, nhead, terms, order_clock, k); break;
    case KIND_SPECIAL: hipLaunchKernelGGL(step_kernel<KIND_SPECIAL>, dim3(ntiles), dim3(256), 0, s, scenes, tiles, ntiles, nhead, terms, order_clock, k); break;
    default: hipLaunchKernelGGL(step_kernel<KIND_ALL>, dim3(ntiles), dim3(256), 0, s, scenes, tiles, ntiles, nhead, terms, order_clock, k); break;
    }
    return hipGetLastError();
}

hipError_t launch_advance(const SceneDev *scenes, int nscenes, int n, hipStream_t s)
{
    if (nscenes <= 0) return hipSuccess;
    hipLaunchKernelGGL(advance_kernel, dim3(nscenes), dim3(256), 0, s, scenes, n);
    return hipGetLastError();
}

// The speaker field of step clock + k over each listed scene's speaker box
// (workgroups of 256 cells along x, one scene per y), before the step kernels.
__global__ __launch_bounds__(256) void speaker_kernel(const SceneDev *__restrict__ scenes, const MotionDesc *__restrict__ list, int k)
{
    const SceneDev &S = scenes[list[blockIdx.y].scene];
    const long long step = (long long)S.clock[CLK_STEP] + k;
    if (step >= (long long)S.clock[CLK_TARGET]) return;
    const int ncell = S.sp_n[0] * S.sp_n[1] * S.sp_n[2];
    const int c = (int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x;
    if (c >= ncell) return;
    const int kk = c % S.sp_n[2], r = c / S.sp_n[2];
    const int j = S.sp_lo[1] + r % S.sp_n[1], i = S.sp_lo[0] + r / S.sp_n[1];
    S.spk_field[c] = (float)speaker_terms(S, step, i, j, S.sp_lo[2] + kk);
}

hipError_t launch_speaker(const SceneDev *scenes, const MotionDesc *list, int n, int max_cells, int k, hipStream_t s)
{
    if (n <= 0 || max_cells <= 0) return hipSuccess;
    hipLaunchKernelGGL(speaker_kernel, dim3((unsigned)((max_cells + 255) / 256), (unsigned)n), dim3(256), 0, s, scenes, list, k);
    return hipGetLastError();
}

hipError_t launch_motion(const SceneDev *scenes, const MotionDesc *list, int n, int max_cells, int k, hipStream_t s)
{
    if (n <= 0 || max_cells <= 0) return hipSuccess;
    const dim3 grid6((unsigned)((6 * max_cells + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(motion_kernel, grid6, dim3(256), 0, s, scenes, list, k);
    return hipGetLastError();
}


}  // namespace fdtd
