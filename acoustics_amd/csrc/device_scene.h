// device_scene.h -- device-resident state of one scene (used by the single
// scene handle, batches and the server).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.h"

namespace fdtd {

#define FDTD_HIP_TRY(expr)                                                     \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return e_ == hipErrorOutOfMemory ? FDTD_ERR_OOM : FDTD_ERR_DEVICE; \
        }                                                                      \
    } while (0)

// Device buffers and uploaded tables of one scene.  All memory comes from a
// single hipMalloc arena owned by the object.
struct DeviceScene {
    SceneSetup S;
    int device = 0;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    float *p[3] = {nullptr, nullptr, nullptr};
    float *pdir[3] = {nullptr, nullptr, nullptr};
    float *v[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    uint8_t *cpml[3], *fpml[3];
    float *f1[3], *f2[3], *f3[3], *f5h[3];
    float *gtab = nullptr;
    PSrcDev *psrc = nullptr;
    ImpDev *isrc = nullptr;
    SpecialBox *boxes = nullptr;
    int32_t *cls = nullptr;
    SpecialRec *recs = nullptr;
    GhostDev *ghosts = nullptr;
    int32_t *gk_ptr = nullptr, *gk_idx = nullptr, *gu_ptr = nullptr, *gu_idx = nullptr;
    double *gk_w = nullptr, *gu_w = nullptr;
    double *gval[2] = {nullptr, nullptr};
    double *gtmp = nullptr;
    int32_t *lptr = nullptr, *lkind = nullptr, *lindex = nullptr;
    double *lbeta = nullptr;
    double *series = nullptr;
    int series_cap = 0;
    long long *clock = nullptr;

    // Sizes the arena for a setup; used by admission control.
    static size_t bytes_needed(const SceneSetup &S, int series_cap);
    fdtd_status allocate_and_upload(hipStream_t stream, int series_cap);
    void release();

    // Kernel parameter blocks of global step s (0-based).  clock != nullptr
    // makes the kernels read the base step from device memory (graph replay)
    // and use `offset` relative to it; otherwise offset is the absolute step.
    StepParams step_params(long long s, const long long *clk, long long offset) const;
    GhostParams ghost_params(long long s, const long long *clk, long long offset) const;
    ListenParams listen_params(long long s, const long long *clk, long long offset) const;
    bool needs_ghost_kernel() const { return S.boundary == FDTD_BOUNDARY_FULLY_COUPLED && !S.ghosts.empty(); }
    // Enqueues the kernels of step s.
    fdtd_status enqueue_step(long long s, const long long *clk, long long offset, int tile, hipStream_t st,
                             hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr) const;
};

}  // namespace fdtd
