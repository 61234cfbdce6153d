// solver.cpp -- single-scene handle and its C-ABI (fdtd_create / fdtd_step /
// fdtd_listen / ...), the MI355X counterpart of PML_WaveSolver driven by
// FDTD_AcousticSimulator::RunForSteps
// (src/wavesolver/FDTD_AcousticSimulator.cpp:528-544).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
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
    int timing = 0;             // FDTD_TIMING_*
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
    s->time_parallel = 0;
    s->num_time_chunks = 1;
    s->sampling = FDTD_SAMPLE_CENTER;   // the reference's compiled default (ONLY_SAMPLE_CENTER)
    s->overlap_time = 6e-3;             // ImpulseResponseParser.cpp:531
    s->early_termination_ratio = 1e-3;  // ImpulseResponseParser.cpp:532
    s->stop_boundary_acc_time = 0.0;
}

void fdtd_options_default(fdtd_options *o)
{
    if (!o) return;
    o->stream = nullptr;
    o->device = -1;
    o->steps_per_graph = 0;
    o->series_capacity = 0;
    o->tile_config = -1;
    o->pml_state = FDTD_PML_STATE_MERGED;
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
    if (opt.pml_state != FDTD_PML_STATE_MERGED && opt.pml_state != FDTD_PML_STATE_SPLIT) {
        set_last_error("pml_state must be FDTD_PML_STATE_MERGED or FDTD_PML_STATE_SPLIT");
        return FDTD_ERR_INVALID;
    }
    set_pml_state(h->D.S, opt.pml_state == FDTD_PML_STATE_MERGED);
    if (opt.device >= 0) {
        if (opt.device >= ndev) { set_last_error("device ordinal out of range"); return FDTD_ERR_INVALID; }
        FDTD_HIP_TRY(hipSetDevice(opt.device));
    }
    FDTD_HIP_TRY(hipGetDevice(&h->D.device));
    if (opt.stream) h->stream = (hipStream_t)opt.stream;
    else { FDTD_HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)); h->own_stream = true; }
    h->tile = (opt.tile_config >= 0 && opt.tile_config < kNumTilePlans) ? opt.tile_config : default_tile_plan(h->D.S);
    h->graph_steps = std::max(0, opt.steps_per_graph);
    int cap = opt.series_capacity > 0 ? opt.series_capacity : 1024;
    cap = std::max(cap, h->graph_steps + 1);
    st = h->D.allocate_and_upload(h->stream, cap, LLONG_MAX);
    if (st == FDTD_OK) {
        std::vector<TileDesc> tiles;
        std::vector<TermRef> terms;
        h->D.append_tiles(0, tile_plan(h->tile), tiles, terms);
        std::vector<MotionDesc> motion;
        if (h->D.needs_motion_kernel()) motion.push_back({0});
        st = h->prog.upload({h->D.dev}, tiles, terms, h->stream, tile_plan(h->tile).fused, motion);
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
    st = h->prog.enqueue(h->stream, 0, h->graph_steps);
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
    if (h->D.S.moving && h->step + nsteps > h->D.S.mv_steps) {
        set_last_error("moving objects are planned up to stop_time (" + std::to_string(h->D.S.mv_steps) + " steps)");
        return FDTD_ERR_INVALID;
    }
    FDTD_HIP_TRY(hipSetDevice(h->D.device));
    const int cap = h->D.series_cap;
    const bool per_step = h->timing == FDTD_TIMING_PER_STEP;
    if (h->timing) {
        if (per_step && (int)h->ev.size() < 2 * nsteps) {
            // timing-only events: no system-scope fence, so recording one does
            // not write back and invalidate the caches under the next kernel
            // (read only after a stream synchronisation)
            while ((int)h->ev.size() < 2 * nsteps) {
                hipEvent_t e;
                FDTD_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
                h->ev.push_back(e);
            }
        }
        if (!h->call_begin) {
            FDTD_HIP_TRY(hipEventCreateWithFlags(&h->call_begin, hipEventDisableSystemFence));
            FDTD_HIP_TRY(hipEventCreateWithFlags(&h->call_end, hipEventDisableSystemFence));
        }
        FDTD_HIP_TRY(hipEventRecord(h->call_begin, h->stream));
        h->timed_steps = nsteps;
    }
    int left = nsteps, done = 0;
    while (left > 0) {
        // graphs are captured for a block that starts on an even step (the
        // kernel order alternates with the step parity)
        const bool use_graph = !per_step && h->graph_steps > 0 && left >= h->graph_steps && (h->step & 1) == 0;
        // eager blocks share one advance launch (at most AMP_ROWS steps, and
        // no more than the series ring holds)
        const int block = use_graph ? h->graph_steps : std::min({left, (int)AMP_ROWS, cap});
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
            fdtd_status st = h->prog.enqueue(h->stream, h->step, block, per_step ? &h->ev[2 * done] : nullptr);
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
    if (enable < FDTD_TIMING_OFF || enable > FDTD_TIMING_SPAN) { set_last_error("unknown timing mode"); return FDTD_ERR_INVALID; }
    h->timing = enable;
    return FDTD_OK;
}

fdtd_status fdtd_last_kernel_ms(fdtd_handle *h, double *stencil_ms, double *step_ms)
{
    if (!h) return FDTD_ERR_INVALID;
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    double ks = 0, cs = 0;
    if (h->timed_steps > 0) {
        for (int i = 0; h->timing == FDTD_TIMING_PER_STEP && i < h->timed_steps; ++i) {
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
    FDTD_HIP_TRY(h->D.download_grid(h->D.slot(slot), tmp.data()));
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
    // the device clock's time: one addition of dt per step from the start
    double t = S.t_start;
    for (long long q = 0; q < h->step; ++q) t += S.dt;
    o->time = t;
    o->bytes_per_step = algorithmic_bytes(S);
    o->device_bytes = (int64_t)h->D.arena_bytes;
    o->n_pml_merged_cells = S.n_pml_merged;
    o->state_bytes_per_step = state_bytes(S);
    return FDTD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// snapshots (FDTD_AcousticSimulator::_SaveSimulationSnapshot layout)
// ---------------------------------------------------------------------------
namespace {

std::string snap_name(const char *pattern, const char *id, int numbering, const std::string &what)
{
    const std::string name = "snapshot_" + std::to_string(numbering) + "_" + what;
    std::vector<char> buf(name.size() + (pattern ? std::strlen(pattern) : 2) + (id ? std::strlen(id) : 0) + 16);
    fdtd_composite_filename(pattern, id, name.c_str(), buf.data(), (int)buf.size());
    return std::string(buf.data());
}

fdtd_status read_matrix(const std::string &path, long long rows, int cols, std::vector<double> &out)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) { set_last_error("cannot open " + path); return FDTD_ERR_INVALID; }
    int r = 0, c = 0;
    in.read(reinterpret_cast<char *>(&r), sizeof(int));
    in.read(reinterpret_cast<char *>(&c), sizeof(int));
    if (!in || r != rows || c != cols) { set_last_error(path + ": matrix shape does not match the scene"); return FDTD_ERR_INVALID; }
    out.resize((size_t)rows * cols);
    in.read(reinterpret_cast<char *>(out.data()), sizeof(double) * out.size());
    if (!in) { set_last_error(path + ": short file"); return FDTD_ERR_INVALID; }
    return FDTD_OK;
}

// index of face (i, j, k) of axis d in the full face grid, n_d + 1 faces along d
inline long long face_index(const int *n, int d, long long i, long long j, long long k)
{
    const long long ny = n[1] + (d == 1), nz = n[2] + (d == 2);
    return (i * ny + j) * nz + k;
}

// Walks every compact PML cell (fn(cell_offset, flat_cell)) or face of axis d
// (fn(face_offset, full_face_index)).
template <class F>
void for_pml_cells(const SceneSetup &S, F fn)
{
    for (const PmlSlab &sl : S.slabs)
        for (int li = 0; li < sl.e[0]; ++li)
            for (int lj = 0; lj < sl.e[1]; ++lj)
                for (int lk = 0; lk < sl.e[2]; ++lk) {
                    const long long c = sl.cell_base + ((long long)li * sl.e[1] + lj) * sl.e[2] + lk;
                    const long long g = ((long long)(sl.lo[0] + li) * S.n[1] + sl.lo[1] + lj) * S.n[2] + sl.lo[2] + lk;
                    fn(c, g);
                }
}
template <class F>
void for_pml_faces(const SceneSetup &S, int d, F fn)
{
    for (const PmlSlab &sl : S.slabs) {
        const int E[3] = {sl.e[0] + (d == 0), sl.e[1] + (d == 1), sl.e[2] + (d == 2)};
        for (int li = 0; li < E[0]; ++li)
            for (int lj = 0; lj < E[1]; ++lj)
                for (int lk = 0; lk < E[2]; ++lk) {
                    const long long c = sl.v_base[d] + ((long long)li * E[1] + lj) * E[2] + lk;
                    fn(c, face_index(S.n, d, sl.lo[0] + li, sl.lo[1] + lj, sl.lo[2] + lk));
                }
    }
}

// A cell of a merging slab is merged when both axes other than the slab axis
// are plain there (the kernel's rule; x is plain over y and z slabs that merge).
inline bool merged_cell(const SceneSetup &S, const PmlSlab &sl, int j, int k)
{
    return sl.merge && (sl.axis == 1 || axis_plain(S, 1, j)) && (sl.axis == 2 || axis_plain(S, 2, k));
}

// The engine's own state file: a header, then the whole time-stepping state
// region of the device as it lies there (pressure slots, PML state, viscous
// Laplacian), so a handle of the same scene and state representation resumes
// bit for bit.
struct EngineStateHeader {
    char magic[8];
    int32_t version, pml_state, scheme, pad;
    int64_t step, n_cells, state_floats, npad;
};
constexpr char kStateMagic[8] = {'F', 'D', 'T', 'D', 'S', 'T', 'A', 'T'};

}  // namespace

extern "C" {

fdtd_status fdtd_save_snapshot(fdtd_handle *h, const char *pattern, const char *simulator_id, int numbering)
{
    if (!h) return FDTD_ERR_INVALID;
    if (h->D.S.moving) {
        set_last_error("snapshots do not hold the per-step classification of moving objects");
        return FDTD_ERR_UNSUPPORTED;
    }
    FDTD_HIP_TRY(hipSetDevice(h->D.device));
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    const SceneSetup &S = h->D.S;
    const long long N = S.N, s = h->step;
    const bool mac = S.scheme == FDTD_SCHEME_STAGGERED;
    const bool merged = S.n_pml_merged > 0;
    std::vector<float> f(N);
    std::vector<double> cur(N), prev(N), zero(N, 0.0);
    FDTD_HIP_TRY(h->D.download_grid(h->D.slot(slot_cur(s)), f.data()));
    for (long long c = 0; c < N; ++c) cur[c] = f[c];
    FDTD_HIP_TRY(h->D.download_grid(h->D.slot(slot_next(s)), f.data()));
    for (long long c = 0; c < N; ++c) prev[c] = f[c];
    auto put = [&](const std::string &what, const std::vector<double> &v, long long rows) {
        return fdtd_write_matrix(snap_name(pattern, simulator_id, numbering, what).c_str(), v.data(), (int)rows, 1);
    };
    fdtd_status st;
    // the reference's ring: p_collocated[ind] is p^n, p_collocated[(ind + 2) % 3] is p^{n-1}
    if ((st = put("p_collocated_0.dat", mac ? zero : prev, N)) != FDTD_OK) return st;
    if ((st = put("p_collocated_1.dat", mac ? zero : cur, N)) != FDTD_OK) return st;
    if ((st = put("p_collocated_2.dat", zero, N)) != FDTD_OK) return st;
    if ((st = put("p_pml_full.dat", mac ? cur : zero, N)) != FDTD_OK) return st;
    {
        std::ofstream of(snap_name(pattern, simulator_id, numbering, "p_collocated_ind.dat"));
        of << 1;
        std::ofstream os(snap_name(pattern, simulator_id, numbering, "step.txt"));
        os << s << "\n";
        if (!of || !os) { set_last_error("cannot write snapshot text files"); return FDTD_ERR_INVALID; }
    }
    // split pressures; a merged cell keeps pDir_A and the increment of the sum
    // q of its transverse splits, so its files hold pDir_A on the slab axis,
    // q = p^n - pDir_A on the lower transverse axis and zero on the other
    std::vector<float> comp[3];
    for (int d = 0; d < 3; ++d) {
        comp[d].resize(std::max<int64_t>(1, S.n_pml));
        if (S.n_pml) FDTD_HIP_TRY(hipMemcpy(comp[d].data(), h->D.dev.pd[d], S.n_pml * 4, hipMemcpyDeviceToHost));
    }
    std::vector<double> full[3];
    for (int d = 0; d < 3; ++d) full[d].assign(N, 0.0);
    for (const PmlSlab &sl : S.slabs)
        for (int li = 0; li < sl.e[0]; ++li)
            for (int lj = 0; lj < sl.e[1]; ++lj)
                for (int lk = 0; lk < sl.e[2]; ++lk) {
                    const long long c = sl.cell_base + ((long long)li * sl.e[1] + lj) * sl.e[2] + lk;
                    const int j = sl.lo[1] + lj, k = sl.lo[2] + lk;
                    const long long g = ((long long)(sl.lo[0] + li) * S.n[1] + j) * S.n[2] + k;
                    if (!merged_cell(S, sl, j, k)) {
                        for (int d = 0; d < 3; ++d) full[d][g] = comp[d][c];
                    } else {
                        const int A = sl.axis, lower = A == 0 ? 1 : 0;
                        full[A][g] = comp[A][c];
                        full[lower][g] = cur[g] - (double)comp[A][c];
                    }
                }
    for (int d = 0; d < 3; ++d)
        if ((st = put("p_pml_" + std::to_string(d) + ".dat", full[d], N)) != FDTD_OK) return st;
    const int vb = (int)(s & 1);  // the buffer the next step reads
    std::vector<float> vc;
    for (int d = 0; d < 3; ++d) {
        const long long NV = (long long)(S.n[0] + (d == 0)) * (S.n[1] + (d == 1)) * (S.n[2] + (d == 2));
        std::vector<double> fv(NV, 0.0);
        if (mac) {
            vc.resize(N);
            FDTD_HIP_TRY(h->D.download_grid(h->D.dev.vel[vb][d], vc.data()));
            for (long long i = 0; i < S.n[0]; ++i)
                for (long long j = 0; j < S.n[1]; ++j)
                    for (long long k = 0; k < S.n[2]; ++k)
                        fv[face_index(S.n, d, i, j, k)] = vc[(i * S.n[1] + j) * S.n[2] + k];
        } else {
            vc.resize(std::max<int64_t>(1, S.n_pml_faces[d]));
            if (S.n_pml_faces[d])
                FDTD_HIP_TRY(hipMemcpy(vc.data(), h->D.dev.vel[vb][d], S.n_pml_faces[d] * 4, hipMemcpyDeviceToHost));
            // a transverse face between two merged cells is not kept by the
            // merged state (the files hold zero there); every other face is
            for (const PmlSlab &sl : S.slabs) {
                const int E[3] = {sl.e[0] + (d == 0), sl.e[1] + (d == 1), sl.e[2] + (d == 2)};
                for (int li = 0; li < E[0]; ++li)
                    for (int lj = 0; lj < E[1]; ++lj)
                        for (int lk = 0; lk < E[2]; ++lk) {
                            const int i = sl.lo[0] + li, j = sl.lo[1] + lj, k = sl.lo[2] + lk;
                            if (d != sl.axis && sl.merge) {
                                // the cells on either side of the face inside this box
                                const int lo_l = (d == 1 ? lj : lk) - 1, hi_l = (d == 1 ? lj : lk);
                                const int ext = d == 1 ? sl.e[1] : sl.e[2];
                                bool kept = false;
                                for (int side = 0; side < 2; ++side) {
                                    const int l = side ? hi_l : lo_l;
                                    if (l < 0 || l >= ext) continue;
                                    const int jj = d == 1 ? sl.lo[1] + l : j, kk = d == 2 ? sl.lo[2] + l : k;
                                    kept = kept || !merged_cell(S, sl, jj, kk);
                                }
                                if (!kept) continue;
                            }
                            const long long c = sl.v_base[d] + ((long long)li * E[1] + lj) * E[2] + lk;
                            fv[face_index(S.n, d, i, j, k)] = vc[c];
                        }
            }
        }
        if ((st = put("v_pml_" + std::to_string(d) + ".dat", fv, NV)) != FDTD_OK) return st;
    }
    // the engine's own state, for a bit-exact resume of any state representation
    {
        EngineStateHeader hd{};
        std::memcpy(hd.magic, kStateMagic, 8);
        hd.version = 1;
        hd.pml_state = merged ? FDTD_PML_STATE_MERGED : FDTD_PML_STATE_SPLIT;
        hd.scheme = S.scheme;
        hd.step = s;
        hd.n_cells = N;
        hd.state_floats = (int64_t)h->D.state_floats;
        hd.npad = h->D.npad;
        std::vector<float> raw(h->D.state_floats);
        FDTD_HIP_TRY(hipMemcpy(raw.data(), h->D.dev.field, raw.size() * 4, hipMemcpyDeviceToHost));
        std::ofstream os(snap_name(pattern, simulator_id, numbering, "engine_state.bin"), std::ios::binary);
        os.write(reinterpret_cast<const char *>(&hd), sizeof hd);
        os.write(reinterpret_cast<const char *>(raw.data()), (std::streamsize)(raw.size() * 4));
        if (!os) { set_last_error("cannot write the engine state file"); return FDTD_ERR_INVALID; }
    }
    return FDTD_OK;
}

// Sets the clock of step s and the host's series bookkeeping after a load.
static fdtd_status resume_at(fdtd_handle *h, long long s)
{
    unsigned long long clk[CLK_WORDS];
    clock_words(clk, s, (unsigned long long)LLONG_MAX, h->D.S.t_start, h->D.S.dt);
    FDTD_HIP_TRY(hipMemcpy(h->D.dev.clock, clk, sizeof clk, hipMemcpyHostToDevice));
    FDTD_HIP_TRY(h->D.fill_amplitudes(h->stream));
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    h->step = s;
    h->flushed = s;
    h->host_series.assign((size_t)s * h->D.S.L, 0.0);
    return FDTD_OK;
}

fdtd_status fdtd_load_snapshot(fdtd_handle *h, const char *pattern, const char *simulator_id, int numbering)
{
    if (!h) return FDTD_ERR_INVALID;
    if (h->D.S.moving) {
        set_last_error("snapshots do not hold the per-step classification of moving objects");
        return FDTD_ERR_UNSUPPORTED;
    }
    FDTD_HIP_TRY(hipSetDevice(h->D.device));
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    const SceneSetup &S = h->D.S;
    const long long N = S.N;
    const bool mac = S.scheme == FDTD_SCHEME_STAGGERED;
    const bool merged = S.n_pml_merged > 0;
    // the engine's own state file, when it matches this handle, resumes bit for bit
    {
        std::ifstream is(snap_name(pattern, simulator_id, numbering, "engine_state.bin"), std::ios::binary);
        EngineStateHeader hd{};
        if (is && is.read(reinterpret_cast<char *>(&hd), sizeof hd) && std::memcmp(hd.magic, kStateMagic, 8) == 0 &&
            hd.version == 1 && hd.n_cells == N && hd.scheme == S.scheme && hd.state_floats == (int64_t)h->D.state_floats &&
            hd.npad == h->D.npad && hd.pml_state == (merged ? FDTD_PML_STATE_MERGED : FDTD_PML_STATE_SPLIT) && hd.step >= 0) {
            std::vector<float> raw(h->D.state_floats);
            if (!is.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)(raw.size() * 4))) {
                set_last_error("engine state file: short file");
                return FDTD_ERR_INVALID;
            }
            FDTD_HIP_TRY(hipMemcpy(h->D.dev.field, raw.data(), raw.size() * 4, hipMemcpyHostToDevice));
            return resume_at(h, hd.step);
        }
        if (std::memcmp(hd.magic, kStateMagic, 8) == 0 && hd.pml_state == FDTD_PML_STATE_MERGED && !merged) {
            set_last_error("snapshot written by a handle with the merged PML state, whose reference files lack the "
                           "transverse face velocities of merged cells; load it into a merged handle of the same scene");
            return FDTD_ERR_UNSUPPORTED;
        }
    }
    if (S.visc) {
        set_last_error("the reference's snapshot files do not hold the previous step's Laplacian that air viscosity "
                       "needs; resume a viscous scene from the engine state file of a handle like this one");
        return FDTD_ERR_UNSUPPORTED;
    }
    long long s = -1;
    int ind = -1;
    {
        std::ifstream is(snap_name(pattern, simulator_id, numbering, "step.txt"));
        std::ifstream ii(snap_name(pattern, simulator_id, numbering, "p_collocated_ind.dat"));
        if (!(is >> s) || s < 0 || !(ii >> ind) || ind < 0 || ind > 2) {
            set_last_error("snapshot step or ring index missing");
            return FDTD_ERR_INVALID;
        }
    }
    std::vector<double> cur, prev;
    fdtd_status st;
    if (mac) {
        if ((st = read_matrix(snap_name(pattern, simulator_id, numbering, "p_pml_full.dat"), N, 1, cur)) != FDTD_OK) return st;
        prev.assign(N, 0.0);
    } else {
        if ((st = read_matrix(snap_name(pattern, simulator_id, numbering, "p_collocated_" + std::to_string(ind) + ".dat"), N, 1, cur)) != FDTD_OK)
            return st;
        if ((st = read_matrix(snap_name(pattern, simulator_id, numbering, "p_collocated_" + std::to_string((ind + 2) % 3) + ".dat"), N, 1, prev)) != FDTD_OK)
            return st;
    }
    // the Engquist boundary keeps p^{n-1} of each face cell's inner neighbour
    // at the face cell's exterior cell of the current slot
    abc_pack_state(S, prev.data(), cur.data());
    // the whole state starts from zero: faces and cells the snapshot does not hold stay at rest
    FDTD_HIP_TRY(h->D.zero_state(h->stream));
    FDTD_HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<float> f(N);
    for (long long c = 0; c < N; ++c) f[c] = (float)cur[c];
    FDTD_HIP_TRY(h->D.upload_grid(h->D.slot(slot_cur(s)), f.data()));
    for (long long c = 0; c < N; ++c) f[c] = (float)prev[c];
    FDTD_HIP_TRY(h->D.upload_grid(h->D.slot(slot_next(s)), f.data()));
    std::vector<double> pml[3], vel[3];
    for (int d = 0; d < 3; ++d)
        if ((st = read_matrix(snap_name(pattern, simulator_id, numbering, "p_pml_" + std::to_string(d) + ".dat"), N, 1, pml[d])) != FDTD_OK)
            return st;
    for (int d = 0; d < 3; ++d) {
        const long long NV = (long long)(S.n[0] + (d == 0)) * (S.n[1] + (d == 1)) * (S.n[2] + (d == 2));
        if ((st = read_matrix(snap_name(pattern, simulator_id, numbering, "v_pml_" + std::to_string(d) + ".dat"), NV, 1, vel[d])) != FDTD_OK)
            return st;
    }
    // split pressures; a merged cell takes pDir_A and the increment of the
    // transverse sum q, d = sum_b f5/h_b (v_b(high face) - v_b(low face)) at
    // the newest velocities (exact: with f1 = f3 = 1 on both transverse axes
    // q changes by that sum every step)
    std::vector<float> comp[3];
    for (int d = 0; d < 3; ++d) comp[d].assign(std::max<int64_t>(1, S.n_pml), 0.f);
    for (const PmlSlab &sl : S.slabs)
        for (int li = 0; li < sl.e[0]; ++li)
            for (int lj = 0; lj < sl.e[1]; ++lj)
                for (int lk = 0; lk < sl.e[2]; ++lk) {
                    const long long c = sl.cell_base + ((long long)li * sl.e[1] + lj) * sl.e[2] + lk;
                    const int i = sl.lo[0] + li, j = sl.lo[1] + lj, k = sl.lo[2] + lk;
                    const long long g = ((long long)i * S.n[1] + j) * S.n[2] + k;
                    if (!merged_cell(S, sl, j, k)) {
                        for (int d = 0; d < 3; ++d) comp[d][c] = (float)pml[d][g];
                        continue;
                    }
                    const int A = sl.axis, lower = A == 0 ? 1 : 0;
                    comp[A][c] = (float)pml[A][g];
                    double dd = 0.0;
                    const int ijk[3] = {i, j, k};
                    for (int b = 0; b < 3; ++b) {
                        if (b == A) continue;
                        int hi[3] = {ijk[0], ijk[1], ijk[2]};
                        hi[b] += 1;
                        dd += (double)S.f5h[b][ijk[b]] *
                              (vel[b][face_index(S.n, b, hi[0], hi[1], hi[2])] - vel[b][face_index(S.n, b, i, j, k)]);
                    }
                    comp[lower][c] = (float)dd;
                }
    for (int d = 0; d < 3; ++d)
        if (S.n_pml) FDTD_HIP_TRY(hipMemcpy(h->D.dev.pd[d], comp[d].data(), S.n_pml * 4, hipMemcpyHostToDevice));
    const int vb = (int)(s & 1);
    std::vector<float> vc;
    for (int d = 0; d < 3; ++d) {
        if (mac) {
            vc.assign(N, 0.f);
            for (long long i = 0; i < S.n[0]; ++i)
                for (long long j = 0; j < S.n[1]; ++j)
                    for (long long k = 0; k < S.n[2]; ++k)
                        vc[(i * S.n[1] + j) * S.n[2] + k] = (float)vel[d][face_index(S.n, d, i, j, k)];
            FDTD_HIP_TRY(h->D.upload_grid(h->D.dev.vel[vb][d], vc.data()));
        } else {
            vc.assign(std::max<int64_t>(1, S.n_pml_faces[d]), 0.f);
            for_pml_faces(S, d, [&](long long c, long long g) { vc[c] = (float)vel[d][g]; });
            if (S.n_pml_faces[d])
                FDTD_HIP_TRY(hipMemcpy(h->D.dev.vel[vb][d], vc.data(), S.n_pml_faces[d] * 4, hipMemcpyHostToDevice));
        }
    }
    return resume_at(h, s);
}

}  // extern "C"
