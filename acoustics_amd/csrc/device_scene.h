// device_scene.h -- device-resident state of one scene, the tile plan of its
// step, and the launch sequence shared by single scenes, batches and the
// server.
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

// Tile shapes of one plan (chosen by the autotuner).
struct TilePlan {
    int interior_tk_log2 = 6;  // 64 threads along k, 4 rows along j
    int interior_ti = 16;      // planes marched per interior tile
    int pml_ti = 16;           // planes marched per PML tile
    bool interior_vec = false; // whole-row float4 interior tiles where the layout allows
    bool fused = false;        // one launch for all tile classes
    int pml_tk_log2 = 6;       // PML slab columns: up to 1 << pml_tk_log2 along k, 256 >> it along j
    int band_group = 1;        // float4 row tiles: this many bands of rows advance along x together
};
constexpr int kNumTilePlans = 11;
TilePlan tile_plan(int id);
struct SceneSetup;
// The single-scene default: plan 0; plan 1 for the staggered scheme; plan 10
// for a grid of under 400,000 cells.
int default_tile_plan(const SceneSetup &S);

// Device buffers and uploaded tables of one scene, all in one hipMalloc arena.
struct DeviceScene {
    SceneSetup S;
    int device = 0;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    SceneDev dev{};            // host copy of the scene block (pointers into the arena)
    SceneDev *self = nullptr;  // device copy of dev (in the arena), for the scene's own launches
    long long npad = 0;        // floats per pressure slot (padded rows, rounded up to 64)
    int sy_dev = 0;            // row stride of the device grid (n2 rounded up to 4)
    size_t state_floats = 0;   // floats of time-stepping state from dev.field on
    int series_cap = 0;

    // Device bytes a scene needs (admission control uses this).
    static size_t bytes_needed(const SceneSetup &S, int series_cap);
    // The same split in two regions: the time-stepping state (pressure slots,
    // PML state, viscous Laplacian) and everything else (tables, series ring,
    // amplitude ring, clock, scene block).
    static void region_bytes(const SceneSetup &S, int series_cap, size_t *state, size_t *tables);
    // Allocates, zeroes the fields and uploads every table; target is the
    // number of steps after which the kernels stop advancing this scene.
    fdtd_status allocate_and_upload(hipStream_t stream, int series_cap, long long target);
    // As allocate_and_upload, in memory the caller provides (region_bytes
    // sizes both regions; each base 256-byte aligned); allocates nothing, and
    // release() then frees nothing.
    fdtd_status place(hipStream_t stream, int series_cap, long long target, char *state_base, char *table_base);
    void release();
    float *slot(int s) const { return dev.field + (size_t)s * (size_t)npad; }
    // device index of a dense host cell index (k fastest, rows of n2)
    int64_t to_dev(int64_t dense) const
    {
        const int64_t row = dense / S.n[2];
        return row * sy_dev + (dense - row * S.n[2]);
    }
    // a full-grid device array (padded rows) to / from a dense host array
    hipError_t download_grid(const float *src, float *dense) const;
    hipError_t upload_grid(float *dst, const float *dense) const;
    long long target = 0;      // steps after which the kernels stop advancing this scene
    // zeroes every time-stepping array (pressures, split pressures, velocities)
    hipError_t zero_state(hipStream_t st) const { return hipMemsetAsync(dev.field, 0, sizeof(float) * state_floats, st); }
    // puts the scene back at step 0: fields at rest, empty series ring, clock
    // and per-step amplitudes of step 0 (synchronises st)
    fdtd_status reset_to_start(hipStream_t st);
    // refills the amplitude (and pose) ring for the AMP_ROWS steps from the device clock
    hipError_t fill_amplitudes(hipStream_t st) const { return launch_advance(self, 1, 0, st); }

    // Appends this scene's tiles, tagged with scene_id, and the listener
    // terms they carry (TileDesc::tb/te index `terms`).
    void append_tiles(int scene_id, const TilePlan &plan, std::vector<TileDesc> &out, std::vector<TermRef> &terms) const;
    void assign_terms(int scene_id, size_t base, std::vector<TileDesc> &out, std::vector<TermRef> &terms) const;
    // Listener values of ring rows [ring_row, ring_row + nrows): FP64 [nrows][L],
    // each the sum of the listener's terms in term order.
    fdtd_status read_series(long long ring_row, long long nrows, double *out) const;
    bool needs_motion_kernel() const { return S.moving; }
};

// Reorders a spatially ordered tile list so workgroup w (dispatched to XCD
// w % 8) runs tile chunk (w % 8), keeping neighbouring tiles on one L2.
void xcd_permute(std::vector<TileDesc> &tiles, int nxcd = 8);

// Splits the tiles of a step into the lists of its launches (per class, or
// one list when fused) in launch order, and checks them on the host: each
// tile's class matches its launch and its reads stay inside the grid.
// nhead (optional) gets, per launch, the tiles at its head that every step
// dispatches first (the object-adjacent cells, padded to a multiple of eight).
void build_launch_lists(const std::vector<TileDesc> &tiles, bool fused_launch, std::vector<TileDesc> by_kind[NUM_KINDS],
                        int nhead[NUM_KINDS] = nullptr);
std::string check_launch_lists(const std::vector<TileDesc> by_kind[NUM_KINDS], bool fused_launch,
                               const std::vector<SceneDev> &scenes, size_t nterms);

// The launch sequence of a set of scenes, uploaded once.  Per step: object
// motion (scenes with moving objects) -> speaker pass (scenes with speakers)
// -> the step kernels, which also write the
// listener terms; per block of up to AMP_ROWS steps, one step advance (one
// workgroup per scene).
// The step kernels write disjoint cells.  They run back to back on one
// stream in an order that alternates with the step parity (object cells,
// interior, PML on even steps; the reverse on odd steps), and the tiles
// inside each kernel alternate direction too, so each step begins on the
// data the previous step touched last.
struct StepProgram {
    SceneDev *d_scenes = nullptr;
    TileDesc *d_tiles[NUM_KINDS] = {nullptr, nullptr, nullptr};
    MotionDesc *d_motion = nullptr;
    TermRef *d_terms = nullptr;
    int ntiles[NUM_KINDS] = {0, 0, 0};
    int head[NUM_KINDS] = {0, 0, 0};   // tiles dispatched first on every step (no odd-step reversal)
    int nmotion = 0, motion_cells = 0, nspeaker = 0, speaker_cells = 0, nscenes = 0;
    bool fused = false;        // everything in d_tiles[KIND_INTERIOR], launched as KIND_ALL
    const unsigned long long *order_clock = nullptr;  // step word of the first scene
    // tiles of every kind in one list; split by type here (or kept in one
    // launch when `fused_launch` is set)
    // (ext_base: place the program there, ext_bytes long, instead of
    // allocating it; release() then frees nothing)
    // (host_stage, with ext_base: pinned host memory of ext_bytes; the pieces
    // are assembled there and go over in one asynchronous copy on st, with no
    // synchronisation; the caller keeps the stage until that copy has run)
    fdtd_status upload(const std::vector<SceneDev> &scenes, const std::vector<TileDesc> &tiles,
                       const std::vector<TermRef> &terms, hipStream_t st,
                       bool fused_launch = false, const std::vector<MotionDesc> &motion = {}, char *ext_base = nullptr,
                       size_t ext_bytes = 0, char *host_stage = nullptr);
    // Bytes upload() places for these lists (npre: motion plus speaker entries).
    static size_t bytes_needed(size_t nscenes, const std::vector<TileDesc> &tiles, bool fused_launch, size_t nterms,
                               size_t npre);
    bool owns = true;          // the program's block came from hipMalloc
    void release();
    // Enqueues n steps from `step` (the host's count of the step the first
    // runs; its parity picks the kernel order of each step): the step
    // launches at offsets k = 0, 1, ... from the scenes' device counters, and
    // one advance after every AMP_ROWS steps and after the last.  ev, when
    // given, holds 2n events; ev[2q], ev[2q+1] bracket step q's kernels.
    fdtd_status enqueue(hipStream_t st, long long step, int n = 1, const hipEvent_t *ev = nullptr) const;
    // one step's launches at block offset k, without the advance
    fdtd_status enqueue_step(hipStream_t st, long long step, int k, hipEvent_t ev_begin = nullptr,
                             hipEvent_t ev_end = nullptr) const;
};

}  // namespace fdtd
