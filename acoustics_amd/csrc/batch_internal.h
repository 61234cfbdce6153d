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
    StepProgram prog;
    int plan = 0;
    int graph_steps = 32;
    hipGraphExec_t exec = nullptr;
    long long taken = 0;                              // batch steps launched
    long long blocks = 0;
    float timings[kNumTilePlans] = {};

    fdtd_status rebuild(hipStream_t st);              // work list of unfinished scenes
    fdtd_status build_graph(hipStream_t st);
    void drop_graph();
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
