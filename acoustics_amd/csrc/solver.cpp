// solver.cpp -- single-scene handle and its C-ABI (fdtd_create / fdtd_step /
// fdtd_listen / ...), the MI355X counterpart of PML_WaveSolver driven by
// FDTD_AcousticSimulator::RunForSteps
// (src/wavesolver/FDTD_AcousticSimulator.cpp:528-544).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "device_scene.h"

using namespace fdtd;

namespace fdtd {
const std::string &get_last_error();
}

struct fdtd_handle {
    DeviceScene D;
    StepProgram prog;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int tile = 0;
    int graph_steps = 0;
    hipGraphExec_t graph_exec = nullptr;
    long long step = 0;          // steps taken
    long long flushed = 0;       // listener rows copied to host_series
    std::vector<double> host_series;
    bool timing = false;
    std::vector<hipEvent_t> ev;  // [2*i] begin, [2*i+1] end of step kernel i
    hipEvent_t call_begin = nullptr, call_end = nullptr;
    int timed_steps = 0;
};

extern "C" {

const char *fdtd_status_string(fdtd_status s)
{
    switch (s) {
    case FDTD_OK: return "ok";
    case FDTD_ERR_INVALID: return "invalid argument";
    case FDTD_ERR_NO_DEVICE: return "no device";
    case FDTD_ERR_DEVICE: return "device error";
    case FDTD_ERR_OOM: return "out of device memory";
    case FDTD_ERR_PARSE: return "parse error";
    case FDTD_ERR_UNSUPPORTED: return "unsupported";
    case FDTD_ERR_NONFINITE: return "non-finite result";
    case FDTD_ERR_BUSY: return "busy";
    case FDTD_ERR_CANCELLED: return "cancelled";
    case FDTD_ERR_NOT_FOUND: return "not found";
    }
    return "unknown";
}

const char *fdtd_last_error(void) { return get_last_error().c_str(); }

void fdtd_settings_default(fdtd_settings *s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->sound_speed = 343.0;   // ImpulseResponseParser.cpp:506
    s->density = 1.0;         // ImpulseResponseParser.cpp:507
    s->cell_size = 0.005;
    s->time_step = 1.0 / 120000.0;
    s->pml_width = 10.0;
    s->pml_strength = 1.0e6;
    s->boundary_condition_preset = 0;
    s->boundary_handling = FDTD_BOUNDARY_RASTERIZE;
    s->use_mesh = 1;
    for (int d = 0; d < 3; ++d) s->divisions[d] = 64;
    for (int d = 0; d < 3; ++d) s->min_bound[d] = -0.005 * 64 / 2.0;
}

void fdtd_options_default(fdtd_options *o)
{
    if (!o) return;
    o->stream = nullptr;
    o->device = -1;
    o->steps_per_graph = 0;
    o->series_capacity = 0;
    o->tile_config = -1;
}

fdtd_status fdtd_create(const fdtd_scene *scene, const fdtd_options *opt_in, fdtd_handle **out)
{
    if (!scene || !out) { set_last_error("null argument"); return FDTD_ERR_INVALID; }
    *out = nullptr;
    fdtd_options opt;
    fdtd_options_default(&opt);
    if (opt_in) opt = *opt_in;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_last_error("no HIP device"); return FDTD_ERR_NO_DEVICE; }
    std::unique_ptr<fdtd_handle> h(new (std::nothrow) fdtd_handle());
    if (!h) return FDTD_ERR_OOM;
    std::string err;
    fdtd_status st = build_setup(*scene, h->D.S, err);
    if (st != FDTD_OK) { set_last_error(err); return st; }
    if (opt.device >= 0) {
        if (opt.device >= ndev) { set_last_error("device ordinal out of range"); return FDTD_ERR_INVALID; }
        FDTD_HIP_TRY(hipSetDevice(opt.device));
    }
    FDTD_HIP_TRY(hipGetDevice(&h->D.device));
    if (opt.stream) h->stream = (hipStream_t)opt.stream;
    else { FDTD_HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)); h->own_stream = true; }
    h->tile = (opt.tile_config >= 0 && opt.tile_config < kNumTilePlans) ? opt.tile_config : 0;
    h->graph_steps = std::max(0, opt.steps_per_graph);
    int cap = opt.series_capacity > 0 ? opt.series_capacity : 1024;
    cap = std::max(cap, h->graph_steps + 1);
    st = h->D.allocate_and_upload(h->stream, cap, LLONG_MAX);
    if (st == FDTD_OK) {
        std::vector<TileDesc> tiles;
        std::vector<TermRef> terms;
        std::vector<GhostDesc> ghost;
        h->D.append_tiles(0, tile_plan(h->tile), tiles, terms);
        if (h->D.needs_ghost_kernel()) ghost.push_back({0});
        st = h->prog.upload({h->D.dev}, tiles, terms, ghost, h->stream, tile_plan(h->tile).fused);
    }
    if (st != FDTD_OK) {
        h->prog.release();
        h->D.release();
        if (h->own_stream) (void)hipStreamDestroy(h->stream);
        return st;
    }
    *out = h.release();
    return FDTD_OK;
}

static void drop_graph(fdtd_handle *h)
{
    if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
}

fdtd_status fdtd_destroy(fdtd_handle *h)
{
    if (!h) return FDTD_ERR_INVALID;
    (void)hipSetDevice(h->D.device);
    (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    if (h->call_begin) (void)hipEventDestroy(h->call_begin);
    if (h->call_end) (void)hipEventDestroy(h->call_end);
    h->prog.release();
    h->D.release();
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return FDTD_OK;
}

// copies listener rows [flushed, step) from the device ring to host_series
static fdtd_status flush_series(fdtd_handle *h)
{
    const int L = h->D.S.L;
    if (h->flushed == h->step) return FDTD_OK;
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    if (L == 0) { h->flushed = h->step; return FDTD_OK; }
    const long long rows = h->step - h->flushed;
    const int cap = h->D.series_cap;
    const size_t old = h->host_series.size();
    h->host_series.resize(old + (size_t)rows * L);
    long long r = h->flushed;
    while (r < h->step) {
        const long long ring = r % cap;
        const long long n = std::min<long long>(h->step - r, cap - ring);
        fdtd_status st = h->D.read_series(ring, n, h->host_series.data() + old + (size_t)(r - h->flushed) * L);
        if (st != FDTD_OK) return st;
        r += n;
    }
    h->flushed = h->step;
    return FDTD_OK;
}

static fdtd_status build_graph(fdtd_handle *h)
{
    hipGraph_t g = nullptr;
    FDTD_HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    fdtd_status st = FDTD_OK;
    for (int o = 0; o < h->graph_steps && st == FDTD_OK; ++o) st = h->prog.enqueue(h->stream, o);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (st != FDTD_OK) { if (g) (void)hipGraphDestroy(g); return st; }
    if (e != hipSuccess) { set_last_error(std::string("graph capture: ") + hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { set_last_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    return FDTD_OK;
}

fdtd_status fdtd_step(fdtd_handle *h, int nsteps)
{
    if (!h || nsteps < 0) return FDTD_ERR_INVALID;
    FDTD_HIP_TRY(hipSetDevice(h->D.device));
    const int cap = h->D.series_cap;
    if (h->timing) {
        if ((int)h->ev.size() < 2 * nsteps) {
            while ((int)h->ev.size() < 2 * nsteps) {
                hipEvent_t e;
                FDTD_HIP_TRY(hipEventCreate(&e));
                h->ev.push_back(e);
            }
        }
        if (!h->call_begin) { FDTD_HIP_TRY(hipEventCreate(&h->call_begin)); FDTD_HIP_TRY(hipEventCreate(&h->call_end)); }
        FDTD_HIP_TRY(hipEventRecord(h->call_begin, h->stream));
        h->timed_steps = nsteps;
    }
    int left = nsteps, done = 0;
    while (left > 0) {
        // graphs are captured for a block that starts on an even step (the
        // kernel order alternates with the step parity)
        const bool use_graph = !h->timing && h->graph_steps > 0 && left >= h->graph_steps && (h->step & 1) == 0;
        const int block = use_graph ? h->graph_steps : 1;
        if (h->step + block - h->flushed > cap) {
            fdtd_status st = flush_series(h);
            if (st != FDTD_OK) return st;
        }
        if (use_graph) {
            if (!h->graph_exec) {
                fdtd_status st = build_graph(h);
                if (st != FDTD_OK) return st;
            }
            FDTD_HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
        } else {
            hipEvent_t b = nullptr, e = nullptr;
            if (h->timing) { b = h->ev[2 * done]; e = h->ev[2 * done + 1]; }
            fdtd_status st = h->prog.enqueue(h->stream, h->step, b, e);
            if (st != FDTD_OK) return st;
        }
        h->step += block;
        left -= block;
        done += block;
    }
    if (h->timing) FDTD_HIP_TRY(hipEventRecord(h->call_end, h->stream));
    return FDTD_OK;
}

fdtd_status fdtd_synchronize(fdtd_handle *h)
{
    if (!h) return FDTD_ERR_INVALID;
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

fdtd_status fdtd_set_timing(fdtd_handle *h, int enable)
{
    if (!h) return FDTD_ERR_INVALID;
    h->timing = enable != 0;
    return FDTD_OK;
}

fdtd_status fdtd_last_kernel_ms(fdtd_handle *h, double *stencil_ms, double *step_ms)
{
    if (!h) return FDTD_ERR_INVALID;
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    double ks = 0, cs = 0;
    if (h->timed_steps > 0) {
        for (int i = 0; i < h->timed_steps; ++i) {
            float ms = 0;
            FDTD_HIP_TRY(hipEventElapsedTime(&ms, h->ev[2 * i], h->ev[2 * i + 1]));
            ks += ms;
        }
        float ms = 0;
        FDTD_HIP_TRY(hipEventElapsedTime(&ms, h->call_begin, h->call_end));
        cs = ms;
        ks /= h->timed_steps;
        cs /= h->timed_steps;
    }
    if (stencil_ms) *stencil_ms = ks;
    if (step_ms) *step_ms = cs;
    return FDTD_OK;
}

fdtd_status fdtd_listen(fdtd_handle *h, double *out)
{
    if (!h || (!out && h->D.S.L)) return FDTD_ERR_INVALID;
    if (h->step == 0) { std::fill(out, out + h->D.S.L, 0.0); return FDTD_OK; }
    fdtd_status st = flush_series(h);
    if (st != FDTD_OK) return st;
    const int L = h->D.S.L;
    std::memcpy(out, h->host_series.data() + (size_t)(h->step - 1) * L, sizeof(double) * L);
    for (int l = 0; l < L; ++l)
        if (!std::isfinite(out[l])) { set_last_error("listener signal is not finite"); return FDTD_ERR_NONFINITE; }
    return FDTD_OK;
}

fdtd_status fdtd_read_listener_series(fdtd_handle *h, int64_t first, int64_t n, double *out)
{
    if (!h || first < 0 || n < 0 || first + n > h->step) { set_last_error("rows out of range"); return FDTD_ERR_INVALID; }
    fdtd_status st = flush_series(h);
    if (st != FDTD_OK) return st;
    const int L = h->D.S.L;
    if (n && L) std::memcpy(out, h->host_series.data() + (size_t)first * L, sizeof(double) * (size_t)n * L);
    return FDTD_OK;
}

fdtd_status fdtd_read_pressure(fdtd_handle *h, double *out)
{
    if (!h || !out) return FDTD_ERR_INVALID;
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    const int slot = slot_cur(h->step);  // p^{step}: the newest field
    std::vector<float> tmp((size_t)h->D.S.N);
    FDTD_HIP_TRY(hipMemcpy(tmp.data(), h->D.slot(slot), tmp.size() * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < tmp.size(); ++i) out[i] = tmp[i];
    return FDTD_OK;
}

fdtd_status fdtd_get_info(fdtd_handle *h, fdtd_info *o)
{
    if (!h || !o) return FDTD_ERR_INVALID;
    const SceneSetup &S = h->D.S;
    o->n_cells = S.N;
    o->n_pml_cells = S.n_pml;
    o->n_ghost_cells = (int64_t)S.ghosts.size();
    o->n_solid_cells = S.n_solid;
    o->n_listeners = S.L;
    o->step_index = h->step;
    o->time = S.t_start + (double)h->step * S.dt;
    o->bytes_per_step = algorithmic_bytes(S);
    o->device_bytes = (int64_t)h->D.arena_bytes;
    return FDTD_OK;
}

}  // extern "C"
