// kernels.h -- device-side parameter blocks and launch entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.h"

namespace fdtd {

struct PSrcDev {
    double t0, wt, amp;        // amp = K / 3.6836
    int lo[3], hi[3];          // index box where the spatial factor is nonzero
    int goff[3];               // offsets into gtab
    int pad;
};

struct ImpDev {
    int obj, pad;
    double A, w, ton;
};

struct GhostDev {
    int nbr, obj, coupled, pad;
    double k_raster, c_acc;
};

// Everything the fused collocated step needs for one scene.
struct StepParams {
    int n[3];
    int sx, sy;                // strides of i and j (z fastest)
    int ilo[3], ihi[3];        // non-PML box
    const float *pprev;        // p^{n-1}
    const float *pcur;         // p^n
    float *pnext;              // p^{n+1}
    float *pdir[3];            // PML split pressure, in place
    const float *vin[3];       // PML face velocity, step n-1/2 (left face of each cell)
    float *vout[3];            // PML face velocity, step n+1/2
    const float *f1[3], *f2[3];       // face tables, n_d + 1
    const float *f3[3], *f5h[3];      // cell tables, n_d
    const uint8_t *cpml[3], *fpml[3];
    float lam2f;               // (c dt / h)^2
    double lam2;
    // time: step = (clock ? *clock : 0) + step_offset, t_n = t_start + step dt
    const long long *clock;
    long long step_offset;
    double t_start, dt;
    int n_psrc, n_isrc;
    const PSrcDev *psrc;
    const float *gtab;
    const ImpDev *isrc;
    // cells next to rigid objects
    int n_boxes, raster;
    const SpecialBox *boxes;
    const int32_t *cls;
    const SpecialRec *recs;
    const GhostDev *ghosts;
    const double *gval;        // ghost values of this step (fully-coupled mode)
};

struct GhostParams {
    int G, iters;
    const float *pcur;
    double *gval;              // output, G values
    double *gtmp;              // scratch, G values
    const GhostDev *ghosts;
    const int32_t *gk_ptr, *gk_idx;
    const double *gk_w;
    const int32_t *gu_ptr, *gu_idx;
    const double *gu_w;
    const ImpDev *isrc;
    int n_isrc;
    const long long *clock;
    long long step_offset;
    double t_start, dt;
};

struct ListenParams {
    int L, raster;
    const float *pnew;         // p^{n+1}
    const float *pold;         // p^n (rasterised ghost values derive from it)
    const double *gval;        // ghost values of step n (fully-coupled mode)
    const GhostDev *ghosts;
    const int32_t *ptr;
    const int32_t *kind;
    const int32_t *index;
    const double *beta;
    const ImpDev *isrc;
    int n_isrc;
    double *series;            // [cap][L]
    int cap;
    const long long *clock;
    long long step_offset;     // output row = step (t_n is the time before the step)
    double t_start, dt;
};

struct TileConfig {
    int tk, tj, ti;            // threads along k, j; planes marched per thread
};
constexpr int kNumTileConfigs = 6;
TileConfig tile_config(int id);

hipError_t launch_step(const StepParams &p, int tile_id, hipStream_t s);
hipError_t launch_ghost(const GhostParams &p, hipStream_t s);
hipError_t launch_listen(const ListenParams &p, hipStream_t s);
hipError_t launch_advance_clock(long long *clock, long long by, hipStream_t s);
hipError_t launch_set_clock(long long *clock, long long v, hipStream_t s);

}  // namespace fdtd
