// batch_internal.h -- the multi-scene step driver shared by batches and the
// serving loop.
#pragma once

#include <memory>
#include <vector>

#include "device_scene.h"

namespace fdtd {

struct BatchRunner {
    std::vector<std::unique_ptr<DeviceScene>> scenes;  // slot i (nullptr: freed)
    std::vector<long long> targets;                   // steps each scene takes
    std::vector<char> done;
    std::vector<int> members;                         // program scene id -> slot
    std::vector<size_t> slot_tiles;                   // tiles slot i holds in the program (0: not in it)
    size_t prog_tiles = 0, dead_tiles = 0;            // tiles in the program; of finished scenes
    StepProgram prog;
    // Rebuilds without a drain: the program alternates between two device
    // buffers, uploaded from two pinned host stages in one asynchronous copy
    // each (a buffer's next upload is ordered after every kernel that read it
    // by the stream; a stage is reused once its copy's event has passed), and
    // replaced graphs are destroyed only after the stream has been
    // synchronised (reap).
    char *pbuf[2] = {nullptr, nullptr};
    char *pstage[2] = {nullptr, nullptr};
    hipEvent_t pdone[2] = {nullptr, nullptr};
    size_t pbytes = 0;
    int pnext = 0;
    std::vector<hipGraphExec_t> retired;
    int plan = 0;
    int graph_steps = 32;
    hipGraphExec_t exec = nullptr;
    long long taken = 0;                              // batch steps launched
    long long blocks = 0;
    float timings[kNumTilePlans] = {};

    fdtd_status rebuild(hipStream_t st);              // work list of unfinished scenes
    fdtd_status build_graph(hipStream_t st);
    void drop_graph();                                // retires the graph (see reap)
    void reap();                                      // destroys retired graphs: call with the stream idle
    void free_buffers();                              // with the stream idle
    fdtd_status autotune(hipStream_t st);
    fdtd_status advance(hipStream_t st, int nsteps);
    void release_scenes()
    {
        for (auto &s : scenes)
            if (s) s->release();
        scenes.clear();
    }
    ~BatchRunner();
};

// One chunk of a time-parallel run (SimWorld.cpp:141-221): its start and
// source stop times, its first step in the whole run, the steps it ran and
// its listener series [steps_run][L].
struct ChunkRun {
    double start = 0, stop = 0;
    int64_t first = 0, steps_run = 0;
    std::vector<double> series;
};
// Runs every chunk of a time-parallel scene as one batch, with the
// reference's early-termination monitor per chunk.
fdtd_status run_time_chunks(const fdtd_scene &scene, int64_t num_steps, const fdtd_batch_options *opt,
                            std::vector<ChunkRun> &chunks);

}  // namespace fdtd
