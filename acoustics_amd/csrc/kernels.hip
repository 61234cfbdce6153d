// kernels.hip -- CDNA4 (gfx950) kernels of the collocated FDTD step.
//
// One launch advances the whole grid by one step (PML_WaveSolver::stepCollocated,
// src/wavesolver/PML_WaveSolver.cpp:775-826):
//   * non-PML bulk cells: p^{n+1} = 2 p^n - p^{n-1} + (c dt/h)^2 L(p^n) + S dt
//     (MAC_Grid.cpp:924-944), FP32;
//   * PML cells: staggered face velocity v <- f1 v + f2 (p_R - p_L) on the
//     cell's faces, then pDir_d <- f3 pDir_d + f5/h (v_R - v_L) and
//     p^{n+1} = sum_d pDir_d (MAC_Grid.cpp:461-508), fused in the same thread;
//   * bulk cells next to rigid objects: the non-uniform ghost-cell Laplacian
//     (MAC_Grid.cpp:187-299) and the one-cell-cavity freeze, in FP64.
// Threads own a (j,k) column and march TI planes along i (slowest axis),
// keeping p^n(i-1), p^n(i), p^n(i+1) in registers; k is the fastest axis so
// a wavefront reads 64 consecutive floats of a row.  Face velocities are
// double-buffered (vin -> vout) so a thread may recompute the face it shares
// with a neighbour without racing the neighbour's write.  No MFMA: a 7-point
// stencil is not a contraction.
#include "kernels.h"

namespace fdtd {

__device__ __forceinline__ double impulse_accel(const ImpDev *isrc, int n, int obj, double t)
{
    double a = 0.0;
    for (int q = 0; q < n; ++q) {
        const ImpDev s = isrc[q];
        if (s.obj != obj) continue;
        const double u = t - s.ton;
        a += s.A * exp(-(u * u) / (2.0 * s.w * s.w));
    }
    return a;
}

__device__ __forceinline__ long long current_step(const long long *clock, long long off)
{
    return (clock ? *clock : 0LL) + off;
}

template <int TK, int TJ, int TI>
__global__ __launch_bounds__(TK * TJ) void step_kernel(const StepParams P)
{
    const int tx = threadIdx.x % TK, ty = threadIdx.x / TK;
    const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
    const int k = blockIdx.x * TK + tx;
    const int j = blockIdx.y * TJ + ty;
    const int i0 = blockIdx.z * TI;
    const int i1 = min(i0 + TI, nx);
    const bool active = (k < nz) && (j < ny);
    const int kk = min(k, nz - 1), jj = min(j, ny - 1);
    const int sx = P.sx, sy = P.sy;

    // does this tile touch a box of object-adjacent cells? (uniform branch)
    bool tile_special = false;
    for (int b = 0; b < P.n_boxes; ++b) {
        const SpecialBox bx = P.boxes[b];
        const int k0 = blockIdx.x * TK, j0 = blockIdx.y * TJ;
        tile_special |= (i0 < bx.hi[0]) && (i1 > bx.lo[0]) && (j0 < bx.hi[1]) && (j0 + TJ > bx.lo[1]) &&
                        (k0 < bx.hi[2]) && (k0 + TK > bx.lo[2]);
    }

    const long long step = current_step(P.clock, P.step_offset);
    const double tn = P.t_start + (double)step * P.dt;

    // pressure sources: S(x, t_n + dt/2) dt = T(t) gx[i] gy[j] gz[k]
    constexpr int kMaxRegSrc = 4;
    float src_jk[kMaxRegSrc];
    int src_lo[kMaxRegSrc], src_hi[kMaxRegSrc], src_goff[kMaxRegSrc];
    const int nsrc = min(P.n_psrc, kMaxRegSrc);
    for (int s = 0; s < nsrc; ++s) {
        const PSrcDev ps = P.psrc[s];
        const double u = (tn + 0.5 * P.dt - ps.t0) / ps.wt;
        const double T = exp(-(u * u) / 2.0) * ps.amp * P.dt;
        src_jk[s] = (float)T * P.gtab[ps.goff[1] + jj] * P.gtab[ps.goff[2] + kk];
        src_lo[s] = ps.lo[0]; src_hi[s] = ps.hi[0]; src_goff[s] = ps.goff[0];
    }

    const bool pml_j = P.cpml[1][jj], pml_k = P.cpml[2][kk];
    const int base_jk = jj * sy + kk;
    const int ym_off = (jj > 0) ? sy : 0, yp_off = (jj < ny - 1) ? sy : 0;
    const int zm_off = (kk > 0) ? 1 : 0, zp_off = (kk < nz - 1) ? 1 : 0;

    float pm = P.pcur[max(i0 - 1, 0) * sx + base_jk];
    float pc = P.pcur[i0 * sx + base_jk];
    for (int i = i0; i < i1; ++i) {
        const int idx = i * sx + base_jk;
        const float pp = P.pcur[min(i + 1, nx - 1) * sx + base_jk];
        if (active) {
            const float ym = P.pcur[idx - ym_off], yp = P.pcur[idx + yp_off];
            const float zm = P.pcur[idx - zm_off], zp = P.pcur[idx + zp_off];
            const bool pml_i = P.cpml[0][i];
            if (pml_i | pml_j | pml_k) {
                float pn = 0.f;
                // ---- x faces ----
                {
                    const bool oth = pml_j | pml_k;
                    const bool actL = (i >= 1) && (P.fpml[0][i] | oth);
                    const float vl = P.vin[0][idx];
                    const float vlN = actL ? P.f1[0][i] * vl + P.f2[0][i] * (pc - pm) : 0.f;
                    float vrN = 0.f;
                    if (i + 1 < nx) {
                        const bool actR = P.fpml[0][i + 1] | oth;
                        const float vr = P.vin[0][idx + sx];
                        vrN = actR ? P.f1[0][i + 1] * vr + P.f2[0][i + 1] * (pp - pc) : 0.f;
                        if (!(P.cpml[0][i + 1] | oth)) P.vout[0][idx + sx] = vrN;
                    }
                    P.vout[0][idx] = vlN;
                    float pd = P.pdir[0][idx];
                    pd = P.f3[0][i] * pd + P.f5h[0][i] * (vrN - vlN);
                    P.pdir[0][idx] = pd;
                    pn += pd;
                }
                // ---- y faces ----
                {
                    const bool oth = pml_i | pml_k;
                    const bool actL = (jj >= 1) && (P.fpml[1][jj] | oth);
                    const float vl = P.vin[1][idx];
                    const float vlN = actL ? P.f1[1][jj] * vl + P.f2[1][jj] * (pc - ym) : 0.f;
                    float vrN = 0.f;
                    if (jj + 1 < ny) {
                        const bool actR = P.fpml[1][jj + 1] | oth;
                        const float vr = P.vin[1][idx + sy];
                        vrN = actR ? P.f1[1][jj + 1] * vr + P.f2[1][jj + 1] * (yp - pc) : 0.f;
                        if (!(P.cpml[1][jj + 1] | oth)) P.vout[1][idx + sy] = vrN;
                    }
                    P.vout[1][idx] = vlN;
                    float pd = P.pdir[1][idx];
                    pd = P.f3[1][jj] * pd + P.f5h[1][jj] * (vrN - vlN);
                    P.pdir[1][idx] = pd;
                    pn += pd;
                }
                // ---- z faces ----
                {
                    const bool oth = pml_i | pml_j;
                    const bool actL = (kk >= 1) && (P.fpml[2][kk] | oth);
                    const float vl = P.vin[2][idx];
                    const float vlN = actL ? P.f1[2][kk] * vl + P.f2[2][kk] * (pc - zm) : 0.f;
                    float vrN = 0.f;
                    if (kk + 1 < nz) {
                        const bool actR = P.fpml[2][kk + 1] | oth;
                        const float vr = P.vin[2][idx + 1];
                        vrN = actR ? P.f1[2][kk + 1] * vr + P.f2[2][kk + 1] * (zp - pc) : 0.f;
                        if (!(P.cpml[2][kk + 1] | oth)) P.vout[2][idx + 1] = vrN;
                    }
                    P.vout[2][idx] = vlN;
                    float pd = P.pdir[2][idx];
                    pd = P.f3[2][kk] * pd + P.f5h[2][kk] * (vrN - vlN);
                    P.pdir[2][idx] = pd;
                    pn += pd;
                }
                P.pnext[idx] = pn;
            } else {
                int cls = -1;
                if (tile_special) {
                    for (int b = 0; b < P.n_boxes; ++b) {
                        const SpecialBox bx = P.boxes[b];
                        if (i >= bx.lo[0] && i < bx.hi[0] && jj >= bx.lo[1] && jj < bx.hi[1] && kk >= bx.lo[2] && kk < bx.hi[2]) {
                            cls = P.cls[bx.offset + ((i - bx.lo[0]) * (bx.hi[1] - bx.lo[1]) + (jj - bx.lo[1])) * (bx.hi[2] - bx.lo[2]) + (kk - bx.lo[2])];
                            break;
                        }
                    }
                }
                float src = 0.f;
                for (int s = 0; s < nsrc; ++s)
                    if (i >= src_lo[s] && i < src_hi[s]) src += src_jk[s] * P.gtab[src_goff[s] + i];
                for (int s = kMaxRegSrc; s < P.n_psrc; ++s) {
                    const PSrcDev ps = P.psrc[s];
                    const double u = (tn + 0.5 * P.dt - ps.t0) / ps.wt;
                    const double T = exp(-(u * u) / 2.0) * ps.amp * P.dt;
                    src += (float)T * P.gtab[ps.goff[0] + i] * P.gtab[ps.goff[1] + jj] * P.gtab[ps.goff[2] + kk];
                }
                const float prev = P.pprev[idx];
                if (cls == -1) {
                    const float lap = ((pm + pp) + (ym + yp) + (zm + zp)) - 6.f * pc;
                    P.pnext[idx] = 2.f * pc - prev + P.lam2f * lap + src;
                } else if (cls == -2) {
                    P.pnext[idx] = 0.f;  // solid cell: pressure held at zero
                } else {
                    const SpecialRec r = P.recs[cls];
                    const double c = pc;
                    const double nb[6] = {pm, pp, ym, yp, zm, zp};
                    double g[6];
                    for (int q = 0; q < 6; ++q) {
                        g[q] = 0.0;
                        const int gid = r.gid[q];
                        if (gid >= 0) {
                            if (P.raster) {
                                const GhostDev gd = P.ghosts[gid];
                                g[q] = c + gd.k_raster * impulse_accel(P.isrc, P.n_isrc, gd.obj, tn);
                            } else {
                                g[q] = P.gval[gid];
                            }
                        }
                    }
                    double pnext;
                    if (r.flags & 1) {
                        pnext = c;
                    } else {
                        double lap = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            const bool gn = r.gid[2 * d] >= 0, gp = r.gid[2 * d + 1] >= 0;
                            if (gp && gn)      lap += (g[2 * d + 1] + g[2 * d] - 2.0 * c) / (9.0 / 16.0);
                            else if (gp)       lap += (g[2 * d + 1] + 0.75 * nb[2 * d] - (7.0 / 4.0) * c) / (21.0 / 32.0);
                            else if (gn)       lap += (0.75 * nb[2 * d + 1] + g[2 * d] - (7.0 / 4.0) * c) / (21.0 / 32.0);
                            else               lap += (nb[2 * d + 1] + nb[2 * d] - 2.0 * c);
                        }
                        pnext = 2.0 * c - (double)prev + lap * P.lam2;
                    }
                    P.pnext[idx] = (float)(pnext + (double)src);
                }
            }
        }
        pm = pc;
        pc = pp;
    }
}

// Ghost values of the fully-coupled rows (MAC_Grid.cpp:1131-1441): Jacobi
// iteration x <- b + W x on the static coupling matrix, one workgroup.
__global__ __launch_bounds__(1024) void ghost_kernel(const GhostParams P)
{
    const long long step = current_step(P.clock, P.step_offset);
    const double tn = P.t_start + (double)step * P.dt;
    double *b = P.gtmp, *x1 = P.gtmp + P.G;
    for (int g = threadIdx.x; g < P.G; g += blockDim.x) {
        const GhostDev gd = P.ghosts[g];
        const double a = impulse_accel(P.isrc, P.n_isrc, gd.obj, tn);
        double v;
        if (gd.coupled) {
            v = gd.c_acc * a;
            for (int q = P.gk_ptr[g]; q < P.gk_ptr[g + 1]; ++q) v += P.gk_w[q] * (double)P.pcur[P.gk_idx[q]];
        } else {
            v = (double)P.pcur[gd.nbr] + gd.k_raster * a;
        }
        b[g] = v;
        P.gval[g] = v;
    }
    __syncthreads();
    for (int it = 0; it < P.iters; ++it) {
        const double *src = (it & 1) ? x1 : P.gval;
        double *dst = (it & 1) ? P.gval : x1;
        for (int g = threadIdx.x; g < P.G; g += blockDim.x) {
            double v = b[g];
            for (int q = P.gu_ptr[g]; q < P.gu_ptr[g + 1]; ++q) v += P.gu_w[q] * src[P.gu_idx[q]];
            dst[g] = v;
        }
        __syncthreads();
    }
    if (P.iters & 1) {
        for (int g = threadIdx.x; g < P.G; g += blockDim.x) P.gval[g] = x1[g];
    }
}

// Listener pressures (PML_WaveSolver::FetchScalarData): one wavefront per
// listener, FP64 dot product of precomputed MLS weights with the samples,
// reduced across the wavefront with DPP/permute shuffles.
__global__ __launch_bounds__(64) void listen_kernel(const ListenParams P)
{
    const int l = blockIdx.x;
    const int lane = threadIdx.x;
    const long long step = current_step(P.clock, P.step_offset);
    const double tn = P.t_start + (double)step * P.dt;
    double acc = 0.0;
    for (int q = P.ptr[l] + lane; q < P.ptr[l + 1]; q += 64) {
        const int kind = P.kind[q], idx = P.index[q];
        double v;
        if (kind == 0) {
            v = (double)P.pnew[idx];
        } else if (P.raster) {
            const GhostDev gd = P.ghosts[idx];
            v = (double)P.pold[gd.nbr] + gd.k_raster * impulse_accel(P.isrc, P.n_isrc, gd.obj, tn);
        } else {
            v = P.gval[idx];
        }
        acc += P.beta[q] * v;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) P.series[(size_t)(step % P.cap) * P.L + l] = acc;
}

__global__ void advance_clock_kernel(long long *clock, long long by)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) clock[0] += by;
}

__global__ void set_clock_kernel(long long *clock, long long v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) clock[0] = v;
}

TileConfig tile_config(int id)
{
    switch (id) {
    case 0: return {64, 4, 16};
    case 1: return {64, 4, 8};
    case 2: return {64, 8, 8};
    case 3: return {64, 2, 32};
    case 4: return {128, 2, 16};
    case 5: return {32, 8, 16};
    default: return {64, 4, 16};
    }
}

template <int TK, int TJ, int TI>
static hipError_t launch_t(const StepParams &p, hipStream_t s)
{
    dim3 grid((p.n[2] + TK - 1) / TK, (p.n[1] + TJ - 1) / TJ, (p.n[0] + TI - 1) / TI);
    hipLaunchKernelGGL((step_kernel<TK, TJ, TI>), grid, dim3(TK * TJ), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_step(const StepParams &p, int tile_id, hipStream_t s)
{
    switch (tile_id) {
    case 0: return launch_t<64, 4, 16>(p, s);
    case 1: return launch_t<64, 4, 8>(p, s);
    case 2: return launch_t<64, 8, 8>(p, s);
    case 3: return launch_t<64, 2, 32>(p, s);
    case 4: return launch_t<128, 2, 16>(p, s);
    case 5: return launch_t<32, 8, 16>(p, s);
    default: return launch_t<64, 4, 16>(p, s);
    }
}

hipError_t launch_ghost(const GhostParams &p, hipStream_t s)
{
    hipLaunchKernelGGL(ghost_kernel, dim3(1), dim3(1024), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_listen(const ListenParams &p, hipStream_t s)
{
    if (p.L == 0) return hipSuccess;
    hipLaunchKernelGGL(listen_kernel, dim3(p.L), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_advance_clock(long long *clock, long long by, hipStream_t s)
{
    hipLaunchKernelGGL(advance_clock_kernel, dim3(1), dim3(64), 0, s, clock, by);
    return hipGetLastError();
}

hipError_t launch_set_clock(long long *clock, long long v, hipStream_t s)
{
    hipLaunchKernelGGL(set_clock_kernel, dim3(1), dim3(64), 0, s, clock, v);
    return hipGetLastError();
}

}  // namespace fdtd
