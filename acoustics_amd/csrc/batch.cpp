// batch.cpp -- many independent scenes advanced together: one work list of
// tiles over every scene, one set of step launches per time step for the
// whole batch, captured as hipGraph step blocks.  Scenes carry their own
// device step counters and targets, so uneven grids and uneven step counts
// share the launches; finished scenes are dropped from the work list at the
// next block boundary.  The tile plan is chosen by an autotuner that times
// every plan on the batch itself before the first step.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>

#include "batch_internal.h"

namespace fdtd {

fdtd_status BatchRunner::rebuild(hipStream_t st)
{
    drop_graph();
    std::vector<SceneDev> devs;
    std::vector<TileDesc> tiles;
    std::vector<TermRef> terms;
    std::vector<MotionDesc> motion;
    members.clear();
    slot_tiles.assign(scenes.size(), 0);
    for (size_t i = 0; i < scenes.size(); ++i) {
        if (!scenes[i] || done[i]) continue;
        const int id = (int)devs.size();
        devs.push_back(scenes[i]->dev);
        const size_t before = tiles.size();
        scenes[i]->append_tiles(id, tile_plan(plan), tiles, terms);
        slot_tiles[i] = tiles.size() - before;
        if (scenes[i]->needs_motion_kernel()) motion.push_back({id});
        members.push_back((int)i);
    }
    prog_tiles = tiles.size();
    dead_tiles = 0;
    if (devs.empty()) { prog.release(); return FDTD_OK; }
    const bool fused = tile_plan(plan).fused;
    size_t npre = motion.size();
    for (const SceneDev &d : devs) npre += d.n_spk ? 1 : 0;
    const size_t need = StepProgram::bytes_needed(devs.size(), tiles, fused, terms.size(), npre);
    if (need > pbytes) {
        // the lists only shrink while a batch runs, so this happens on the
        // first builds (the autotuner's plans), with nothing in flight
        FDTD_HIP_TRY(hipStreamSynchronize(st));
        prog.release();
        free_buffers();
        for (int b = 0; b < 2; ++b) {
            FDTD_HIP_TRY(hipMalloc(&pbuf[b], need));
            FDTD_HIP_TRY(hipHostMalloc(&pstage[b], need));
            FDTD_HIP_TRY(hipEventCreateWithFlags(&pdone[b], hipEventDisableTiming));
        }
        pbytes = need;
        pnext = 0;
    }
    const int b = pnext;
    pnext ^= 1;
    FDTD_HIP_TRY(hipEventSynchronize(pdone[b]));   // the stage's previous copy has run
    prog.release();
    const fdtd_status s = prog.upload(devs, tiles, terms, st, fused, motion, pbuf[b], pbytes, pstage[b]);
    if (s != FDTD_OK) return s;
    FDTD_HIP_TRY(hipEventRecord(pdone[b], st));
    return FDTD_OK;
}

void BatchRunner::drop_graph()
{
    if (exec) { retired.push_back(exec); exec = nullptr; }
}

void BatchRunner::reap()
{
    for (hipGraphExec_t e : retired) (void)hipGraphExecDestroy(e);
    retired.clear();
}

void BatchRunner::free_buffers()
{
    for (int b = 0; b < 2; ++b) {
        if (pbuf[b]) (void)hipFree(pbuf[b]);
        if (pstage[b]) (void)hipHostFree(pstage[b]);
        if (pdone[b]) (void)hipEventDestroy(pdone[b]);
        pbuf[b] = pstage[b] = nullptr;
        pdone[b] = nullptr;
    }
    pbytes = 0;
}

fdtd_status BatchRunner::build_graph(hipStream_t st)
{
    drop_graph();
    hipGraph_t g = nullptr;
    FDTD_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    fdtd_status s = FDTD_OK;
    s = prog.enqueue(st, 0, graph_steps);
    hipError_t e = hipStreamEndCapture(st, &g);
    if (s != FDTD_OK) { if (g) (void)hipGraphDestroy(g); return s; }
    if (e != hipSuccess) { set_last_error(std::string("graph capture: ") + hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { set_last_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    return FDTD_OK;
}

// Times each tile plan's step program (three steps after one untimed step)
// on the batch and keeps the fastest.  The timed steps advance the scenes,
// so every scene is put back at step 0 afterwards.
fdtd_status BatchRunner::autotune(hipStream_t st)
{
    float best = 1e30f;
    int best_plan = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    struct Events {
        hipEvent_t *e;
        ~Events() { for (int i = 0; i < 2; ++i) if (e[i]) (void)hipEventDestroy(e[i]); }
    } guard{ev};
    FDTD_HIP_TRY(hipEventCreate(&ev[0]));
    FDTD_HIP_TRY(hipEventCreate(&ev[1]));
    for (int p = 0; p < kNumTilePlans; ++p) {
        plan = p;
        fdtd_status s = rebuild(st);
        if (s != FDTD_OK) return s;
        if ((s = prog.enqueue(st, 0)) != FDTD_OK) return s;
        FDTD_HIP_TRY(hipEventRecord(ev[0], st));
        if ((s = prog.enqueue(st, 1, 3)) != FDTD_OK) return s;
        FDTD_HIP_TRY(hipEventRecord(ev[1], st));
        FDTD_HIP_TRY(hipEventSynchronize(ev[1]));
        float ms = 0;
        FDTD_HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        timings[p] = ms / 3.0f;
        if (ms < best) { best = ms; best_plan = p; }
        for (auto &sc : scenes)
            if (sc) { fdtd_status z = sc->reset_to_start(st); if (z != FDTD_OK) return z; }
    }
    plan = best_plan;
    return rebuild(st);
}

// A finished scene's tiles stay in the work list and return at once on its
// device clock; the list is rebuilt without them only once they are at least
// this share of it.  Each rebuild costs host time (new lists, an upload, a new
// graph) and the early-returning workgroups of a few finished scenes cost the
// device little: on the 64-scene batch a quarter measured faster than a half,
// than never rebuilding, and than rebuilding at every finish.
constexpr double kRebuildDeadShare = 0.25;

fdtd_status BatchRunner::advance(hipStream_t st, int nsteps)
{
    int left = nsteps;
    while (left > 0) {
        long long remaining = 0;
        bool cut = false;
        for (size_t i = 0; i < scenes.size(); ++i) {
            if (!scenes[i] || done[i]) continue;
            if (taken >= targets[i]) {
                done[i] = 1;
                if (i < slot_tiles.size()) dead_tiles += slot_tiles[i];
                // a target cut on the host (a time chunk's early termination)
                // is not on the device clock: only a rebuild stops that scene
                cut = cut || targets[i] < scenes[i]->target;
                continue;
            }
            remaining = std::max(remaining, targets[i] - taken);
        }
        if (remaining == 0) break;
        if (cut || (dead_tiles > 0 && (double)dead_tiles >= kRebuildDeadShare * (double)prog_tiles)) {
            fdtd_status s = rebuild(st);
            if (s != FDTD_OK) return s;
        }
        const int want = (int)std::min<long long>(left, remaining);
        if (graph_steps > 0 && want >= graph_steps) {
            if (!exec) {
                fdtd_status s = build_graph(st);
                if (s != FDTD_OK) return s;
            }
            FDTD_HIP_TRY(hipGraphLaunch(exec, st));
            taken += graph_steps;
            left -= graph_steps;
            ++blocks;
        } else {
            fdtd_status s = prog.enqueue(st, taken);
            if (s != FDTD_OK) return s;
            taken += 1;
            left -= 1;
        }
    }
    return FDTD_OK;
}

BatchRunner::~BatchRunner()
{
    drop_graph();
    reap();
    prog.release();
    free_buffers();
}

}  // namespace fdtd

struct fdtd_batch {
    fdtd::BatchRunner R;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
};

namespace fdtd {

fdtd_status run_time_chunks(const fdtd_scene &scene, int64_t num_steps, const fdtd_batch_options *opt,
                            std::vector<ChunkRun> &out)
{
    const fdtd_settings &s0 = scene.settings;
    if (!s0.time_parallel || s0.num_time_chunks < 1) {
        set_last_error("time-parallel runs need time_parallel = 1 and num_time_chunks >= 1");
        return FDTD_ERR_INVALID;
    }
    const int K = s0.num_time_chunks, L = scene.n_listeners;
    std::vector<fdtd_scene> chunks(K, scene);
    std::vector<int> steps(K);
    out.assign(K, ChunkRun{});
    for (int c = 0; c < K; ++c) {
        int64_t n = 0;
        fdtd_status st = fdtd_time_chunk(&s0, num_steps, c, &chunks[c].settings, &out[c].first, &n);
        if (st != FDTD_OK) return st;
        // rows past the end of the run are not part of its series
        n = std::min<int64_t>(n, num_steps - out[c].first);
        if (n > INT_MAX) { set_last_error("a chunk is longer than 2^31 steps"); return FDTD_ERR_INVALID; }
        steps[c] = (int)n;
        out[c].steps_run = n;
        out[c].start = chunks[c].settings.start_time;
        out[c].stop = chunks[c].settings.stop_boundary_acc_time;
    }
    fdtd_batch *b = nullptr;
    fdtd_status st = fdtd_batch_create(chunks.data(), steps.data(), K, opt, &b);
    if (st != FDTD_OK) return st;
    BatchRunner &R = b->R;
    // SimWorld::TerminationMonitor (SimWorld.h:35-42; SimWorld.cpp:556-590):
    // after the chunk's sources are off, the step's peak listener pressure
    // (the running max from 0 of the signed values) counts as loud above
    // ratio times the chunk's peak and 2e-5 Pa; 0.05 s after the last loud
    // step the chunk stops (that step included).  The monitor arms only once
    // a loud step has been seen after the gate.
    const double window = 0.05, floor_pa = 2e-5, ratio = s0.early_termination_ratio;
    struct Monitor { double p_chunk = 0, last = DBL_MAX, time = 0; int64_t seen = 0; bool live = true; };
    std::vector<Monitor> mon(K);
    for (int c = 0; c < K; ++c) mon[c].time = out[c].start;
    const int block = R.graph_steps > 0 ? R.graph_steps : 64;
    std::vector<double> rows;
    for (;;) {
        bool any = false;
        for (int c = 0; c < K; ++c) any = any || (mon[c].live && R.taken < R.targets[c]);
        if (!any) break;
        st = R.advance(b->stream, block);
        if (st != FDTD_OK) break;
        if (hipStreamSynchronize(b->stream) != hipSuccess) { set_last_error("stream synchronize failed"); st = FDTD_ERR_DEVICE; break; }
        for (int c = 0; c < K && st == FDTD_OK; ++c) {
            Monitor &m = mon[c];
            if (!m.live) continue;
            const int64_t upto = std::min<int64_t>(R.taken, R.targets[c]);
            if (upto > m.seen && L) {
                rows.resize((size_t)(upto - m.seen) * L);
                st = R.scenes[c]->read_series(m.seen, upto - m.seen, rows.data());
            }
            for (int64_t q = m.seen; q < upto && st == FDTD_OK && m.live; ++q) {
                double peak = 0.0;
                for (int l = 0; l < L; ++l) peak = std::max(peak, rows[(size_t)(q - m.seen) * L + l]);
                m.time += chunks[c].settings.time_step;
                m.p_chunk = std::max(m.p_chunk, peak);
                if (m.time > out[c].stop) {
                    if (peak > ratio * m.p_chunk && peak > floor_pa) m.last = m.time;
                    if (m.time - m.last > window) {
                        out[c].steps_run = q + 1;
                        R.targets[c] = std::min<long long>(R.targets[c], q + 1);
                        m.live = false;
                    }
                }
            }
            m.seen = upto;
            if (upto >= R.targets[c]) m.live = false;
        }
        if (st != FDTD_OK) break;
    }
    for (int c = 0; c < K && st == FDTD_OK; ++c) {
        out[c].series.assign((size_t)out[c].steps_run * L, 0.0);
        if (out[c].steps_run > 0 && L) st = R.scenes[c]->read_series(0, out[c].steps_run, out[c].series.data());
    }
    fdtd_batch_destroy(b);
    return st;
}

}  // namespace fdtd

using namespace fdtd;


extern "C" {

void fdtd_batch_options_default(fdtd_batch_options *o)
{
    if (!o) return;
    o->stream = nullptr;
    o->steps_per_graph = 32;
    o->tile_config = -1;
}

}  // extern "C"

// Builds the batch's stream, scenes and work list into b; on failure the
// caller destroys b, which frees whatever was built.
static fdtd_status batch_fill(fdtd_batch *b, const fdtd_scene *scenes, const int *num_steps, int n,
                              const fdtd_batch_options &opt)
{
    FDTD_HIP_TRY(hipGetDevice(&b->device));
    if (opt.stream) b->stream = (hipStream_t)opt.stream;
    else { FDTD_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)); b->own_stream = true; }
    BatchRunner &R = b->R;
    R.graph_steps = std::max(0, opt.steps_per_graph);
    for (int i = 0; i < n; ++i) {
        if (num_steps[i] < 0) { set_last_error("negative step count"); return FDTD_ERR_INVALID; }
        auto ds = std::make_unique<DeviceScene>();
        std::string err;
        fdtd_status st = build_setup(scenes[i], ds->S, err);
        if (st != FDTD_OK) { set_last_error("scene " + std::to_string(i) + ": " + err); return st; }
        if (ds->S.moving && num_steps[i] > ds->S.mv_steps) {
            set_last_error("scene " + std::to_string(i) + ": moving objects are planned up to stop_time");
            return FDTD_ERR_INVALID;
        }
        st = ds->allocate_and_upload(b->stream, std::max(1, num_steps[i]), num_steps[i]);
        if (st != FDTD_OK) { ds->release(); return st; }
        R.scenes.push_back(std::move(ds));
        R.targets.push_back(num_steps[i]);
        R.done.push_back(num_steps[i] == 0);
    }
    if (opt.tile_config >= 0 && opt.tile_config < kNumTilePlans) { R.plan = opt.tile_config; return R.rebuild(b->stream); }
    return R.autotune(b->stream);
}

extern "C" {

fdtd_status fdtd_batch_create(const fdtd_scene *scenes, const int *num_steps, int n, const fdtd_batch_options *opt_in,
                              fdtd_batch **out)
{
    if (!scenes || !num_steps || n <= 0 || !out) { set_last_error("bad batch arguments"); return FDTD_ERR_INVALID; }
    *out = nullptr;
    fdtd_batch_options opt;
    fdtd_batch_options_default(&opt);
    if (opt_in) opt = *opt_in;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_last_error("no HIP device"); return FDTD_ERR_NO_DEVICE; }
    fdtd_batch *b = new (std::nothrow) fdtd_batch();
    if (!b) return FDTD_ERR_OOM;
    const fdtd_status st = batch_fill(b, scenes, num_steps, n, opt);
    if (st != FDTD_OK) {
        // scenes already uploaded, the program buffers and an owned stream
        (void)fdtd_batch_destroy(b);
        return st;
    }
    *out = b;
    return FDTD_OK;
}

fdtd_status fdtd_batch_step(fdtd_batch *b, int nsteps)
{
    if (!b || nsteps < 0) return FDTD_ERR_INVALID;
    FDTD_HIP_TRY(hipSetDevice(b->device));
    return b->R.advance(b->stream, nsteps);
}

fdtd_status fdtd_batch_run(fdtd_batch *b)
{
    if (!b) return FDTD_ERR_INVALID;
    long long mx = 0;
    for (long long t : b->R.targets) mx = std::max(mx, t);
    fdtd_status st = fdtd_batch_step(b, (int)std::max<long long>(0, mx - b->R.taken));
    if (st != FDTD_OK) return st;
    FDTD_HIP_TRY(hipStreamSynchronize(b->stream));
    b->R.reap();
    return FDTD_OK;
}

fdtd_status fdtd_batch_steps_done(fdtd_batch *b, int scene, int64_t *steps)
{
    if (!b || !steps || scene < 0 || scene >= (int)b->R.scenes.size()) return FDTD_ERR_INVALID;
    *steps = std::min<long long>(b->R.taken, b->R.targets[scene]);
    return FDTD_OK;
}

fdtd_status fdtd_batch_read_listener_series(fdtd_batch *b, int scene, double *out)
{
    if (!b || scene < 0 || scene >= (int)b->R.scenes.size()) return FDTD_ERR_INVALID;
    FDTD_HIP_TRY(hipStreamSynchronize(b->stream));
    b->R.reap();
    const DeviceScene &D = *b->R.scenes[scene];
    const long long rows = std::min<long long>(b->R.taken, b->R.targets[scene]);
    return D.read_series(0, rows, out);
}

fdtd_status fdtd_time_chunk(const fdtd_settings *in, int64_t num_steps, int chunk, fdtd_settings *out,
                            int64_t *first_step, int64_t *chunk_steps)
{
    if (!in || !out || num_steps < 0) { set_last_error("bad time-chunk arguments"); return FDTD_ERR_INVALID; }
    if (!in->time_parallel || in->num_time_chunks < 1) {
        set_last_error("time chunks need time_parallel = 1 and num_time_chunks >= 1");
        return FDTD_ERR_INVALID;
    }
    if (chunk < 0 || chunk >= in->num_time_chunks || !(in->time_step > 0) || !(in->overlap_time >= 0)) {
        set_last_error("chunk index, time_step or overlap_time out of range");
        return FDTD_ERR_INVALID;
    }
    // SimWorld.cpp:151-166: the chunk's start and source stop times come from
    // adding time_step one step at a time, as the solver's clock does
    const int64_t per = num_steps / in->num_time_chunks;
    double start = in->start_time;
    for (int64_t i = 0; i < per * chunk; ++i) start += in->time_step;
    double stop = start;
    for (int64_t i = 0; i < per; ++i) stop += in->time_step;
    *out = *in;
    out->start_time = start;
    out->stop_boundary_acc_time = stop;
    const double steps = (double)per + in->overlap_time / in->time_step;   // int += double, truncated
    if (first_step) *first_step = per * chunk;
    if (chunk_steps) *chunk_steps = (int64_t)steps;
    return FDTD_OK;
}

fdtd_status fdtd_run_time_parallel(const fdtd_scene *scene, int64_t num_steps, const fdtd_batch_options *opt,
                                   double *out, int64_t *steps_run)
{
    if (!scene || num_steps <= 0 || (!out && scene->n_listeners > 0)) { set_last_error("bad time-parallel arguments"); return FDTD_ERR_INVALID; }
    std::vector<ChunkRun> chunks;
    fdtd_status st = run_time_chunks(*scene, num_steps, opt, chunks);
    if (st != FDTD_OK) return st;
    const int L = scene->n_listeners;
    if (L) {
        // linear superposition, chunk by chunk in order
        std::fill(out, out + (size_t)num_steps * L, 0.0);
        for (const ChunkRun &c : chunks)
            for (int64_t q = 0; q < c.steps_run; ++q)
                for (int l = 0; l < L; ++l) out[(size_t)(c.first + q) * L + l] += c.series[(size_t)q * L + l];
    }
    if (steps_run)
        for (size_t c = 0; c < chunks.size(); ++c) steps_run[c] = chunks[c].steps_run;
    return FDTD_OK;
}

fdtd_status fdtd_batch_destroy(fdtd_batch *b)
{
    if (!b) return FDTD_ERR_INVALID;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->R.drop_graph();
    b->R.reap();
    b->R.prog.release();
    b->R.free_buffers();
    b->R.release_scenes();
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FDTD_OK;
}

}  // extern "C"
