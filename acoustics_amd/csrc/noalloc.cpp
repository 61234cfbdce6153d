// noalloc.cpp -- fdtd_solve_noalloc: one scene advanced num_steps steps with
// no device allocation, for callers that cannot call hipMalloc (code under
// HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The caller passes the inputs (the scene) and the outputs: `fields`, device
// memory for the time-stepping state (pressure slots, PML state), which holds
// the final state on return, and `series`, host memory for the listener
// series.  Everything else the kernels touch -- the scene's tables, the
// listener term ring, the amplitude ring, the step counter and the step
// program's tile lists -- goes into `scratch`, a static device array of
// kScratchBytes per GPU that the library holds until the process ends.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents: each call rewrites every table it reads and clears what must
// start at zero (term ring, fields).  A per-device lock serialises the
// library's own calls, and each call waits for its work before returning;
// calls that overlap on one GPU from outside this library's lock (another
// process) would race on the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>

#include "device_scene.h"

using namespace fdtd;

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

// device address of `scratch` on each device (nullptr until looked up)
void *g_scratch[kMaxDevices] = {};
std::mutex g_mu[kMaxDevices];
std::mutex g_lookup;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Looks the array up on `device` (current) the first time; every later call
// on that device uses the address kept here.  The array's host-side name is
// only a handle for the runtime and is never passed to a kernel or copy.
static fdtd_status scratch_ptr(int device, char **out)
{
    std::lock_guard<std::mutex> lk(g_lookup);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return FDTD_ERR_DEVICE;
        }
        g_scratch[device] = p;
    }
    *out = static_cast<char *>(g_scratch[device]);
    return FDTD_OK;
}

extern "C" fdtd_status fdtd_solve_noalloc(const fdtd_scene *scene, int num_steps, void *fields, int64_t fields_bytes,
                                          void *stream, double *series, int64_t *fields_bytes_needed)
{
    if (!scene || num_steps < 0) { set_last_error("null scene or negative step count"); return FDTD_ERR_INVALID; }
    // ---- host only: the scene, the sizes of both regions and the limit ----
    DeviceScene D;
    std::string err;
    fdtd_status st = build_setup(*scene, D.S, err);
    if (st != FDTD_OK) { set_last_error(err); return st; }
    set_pml_state(D.S, true);
    if (D.S.moving && num_steps > D.S.mv_steps) {
        set_last_error("moving objects are planned up to stop_time (" + std::to_string(D.S.mv_steps) + " steps)");
        return FDTD_ERR_INVALID;
    }
    const int cap = AMP_ROWS;   // one step block of listener rows, read out after each block
    size_t state_b = 0, table_b = 0;
    DeviceScene::region_bytes(D.S, cap, &state_b, &table_b);
    if (fields_bytes_needed) *fields_bytes_needed = (int64_t)state_b;
    if (!fields) return FDTD_OK;   // size query
    if (fields_bytes < (int64_t)state_b) {
        set_last_error("fields holds " + std::to_string(fields_bytes) + " bytes, the scene's state needs " +
                       std::to_string(state_b));
        return FDTD_ERR_INVALID;
    }
    if (((uintptr_t)fields & 255u) != 0) { set_last_error("fields must be 256-byte aligned"); return FDTD_ERR_INVALID; }
    if (D.S.L > 0 && num_steps > 0 && !series) { set_last_error("null series buffer"); return FDTD_ERR_INVALID; }
    const int plan = D.S.scheme == FDTD_SCHEME_STAGGERED ? 3 : 0;   // fdtd_create's default plans
    const TilePlan tp = tile_plan(plan);
    D.sy_dev = row_stride(D.S);   // (the tile builder indexes device rows)
    std::vector<TileDesc> tiles;
    std::vector<TermRef> terms;
    D.append_tiles(0, tp, tiles, terms);
    std::vector<MotionDesc> motion;
    if (D.needs_motion_kernel()) motion.push_back({0});
    const size_t npre = motion.size() + (D.S.spk.empty() ? 0 : 1);
    const size_t table_region = (table_b + 255) & ~(size_t)255;
    const size_t prog_b = StepProgram::bytes_needed(1, tiles, tp.fused, terms.size(), npre);
    if (table_region + prog_b > kScratchBytes) {
        set_last_error("the scene's tables and step program need " + std::to_string(table_region + prog_b) +
                       " bytes of scratch; fdtd_solve_noalloc holds at most " + std::to_string(kScratchBytes));
        return FDTD_ERR_UNSUPPORTED;
    }
    // ---- device ----------------------------------------------------------
    int device = 0;
    FDTD_HIP_TRY(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) { set_last_error("device ordinal beyond the scratch table"); return FDTD_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(g_mu[device]);
    char *base = nullptr;
    st = scratch_ptr(device, &base);
    if (st != FDTD_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    // place() zeroes the fields and the term ring and writes every table,
    // the clock and the amplitude and pose rings over the previous call's
    // contents
    st = D.place(s, cap, LLONG_MAX, (char *)fields, base);
    if (st != FDTD_OK) return st;
    StepProgram prog;
    st = prog.upload({D.dev}, tiles, terms, s, tp.fused, motion, base + table_region, kScratchBytes - table_region);
    if (st != FDTD_OK) return st;
    const int L = D.S.L;
    for (long long step = 0; step < num_steps;) {
        const int blk = (int)std::min<long long>(num_steps - step, cap);
        st = prog.enqueue(s, step, blk);
        if (st != FDTD_OK) break;
        FDTD_HIP_TRY(hipStreamSynchronize(s));
        if (L) {
            st = D.read_series(step % cap, blk, series + (size_t)step * L);
            if (st != FDTD_OK) break;
        }
        step += blk;
    }
    if (st == FDTD_OK) FDTD_HIP_TRY(hipStreamSynchronize(s));
    prog.release();   // frees nothing: the program lives in scratch
    D.release();      // likewise
    return st;
}
