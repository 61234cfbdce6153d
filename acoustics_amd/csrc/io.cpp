// io.cpp -- the reference tool's files around a solve: output naming
// (FDTD_AcousticSimulator::CompositeFilename), binary matrices
// (IO::writeMatrixX, BINARY), the solver-settings dump with the copied
// control file, the listening positions, and the all_audio.dat stream of
// SimWorld (raw FP64 rows).  Everything here goes through the public C-ABI.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "batch_internal.h"
#include "engine.h"

namespace {

std::string composite(const char *pattern, const char *simulator_id, const std::string &name)
{
    // FDTD_AcousticSimulator.cpp:110-124: "<id>_" + name, formatted into the
    // pattern's %s (time chunks are not used here)
    const std::string file = (simulator_id && *simulator_id) ? std::string(simulator_id) + "_" + name : name;
    const std::string pat = pattern ? pattern : "%s";
    const size_t at = pat.find("%s");
    if (at == std::string::npos) return pat + file;
    return pat.substr(0, at) + file + pat.substr(at + 2);
}

fdtd_status write_matrix(const std::string &path, const double *data, int rows, int cols)
{
    // IO.h:107-139: int32 rows, int32 cols, then the entries row by row
    std::ofstream of(path, std::ios::binary);
    if (!of) { fdtd::set_last_error("cannot open " + path + " for writing"); return FDTD_ERR_INVALID; }
    of.write(reinterpret_cast<const char *>(&rows), sizeof(int));
    of.write(reinterpret_cast<const char *>(&cols), sizeof(int));
    if (rows > 0 && cols > 0) of.write(reinterpret_cast<const char *>(data), sizeof(double) * (size_t)rows * cols);
    if (!of) { fdtd::set_last_error("write failed: " + path); return FDTD_ERR_INVALID; }
    return FDTD_OK;
}

// Steps PML_WaveSolver::stepSystem takes from start to end: it adds dt to the
// clock after each step and stops once the clock reaches the end time
// (PML_WaveSolver.cpp:605-606); end <= 0 means no end.
long long steps_to_end(double start, double end, double dt)
{
    if (!(end > 0.0) || !(dt > 0.0)) return 0;
    long long k = 0;
    double t = start;
    do {
        t += dt;
        ++k;
    } while (t < end);
    return k;
}

}  // namespace

extern "C" {

fdtd_status fdtd_composite_filename(const char *pattern, const char *simulator_id, const char *name, char *out, int cap)
{
    if (!name || !out || cap <= 0) return FDTD_ERR_INVALID;
    const std::string s = composite(pattern, simulator_id, name);
    if ((int)s.size() + 1 > cap) { fdtd::set_last_error("output buffer too small"); return FDTD_ERR_INVALID; }
    std::memcpy(out, s.c_str(), s.size() + 1);
    return FDTD_OK;
}

fdtd_status fdtd_write_matrix(const char *filename, const double *data, int rows, int cols)
{
    if (!filename || rows < 0 || cols < 0 || (!data && rows * cols > 0)) return FDTD_ERR_INVALID;
    return write_matrix(filename, data, rows, cols);
}

fdtd_status fdtd_run_xml(const char *xml_path, const char *simulator_id, const fdtd_options *opt, int64_t *steps_run)
{
    if (!xml_path) return FDTD_ERR_INVALID;
    if (steps_run) *steps_run = 0;
    fdtd_parsed *P = nullptr;
    fdtd_status st = fdtd_parse_xml(xml_path, &P);
    if (st != FDTD_OK) return st;
    const fdtd_scene *sc = fdtd_parsed_scene(P);
    const std::string pattern = fdtd_parsed_output_pattern(P);
    const int L = sc->n_listeners;
    const fdtd_settings &s = sc->settings;
    long long steps = fdtd_parsed_num_steps(P) > 0 ? fdtd_parsed_num_steps(P) : steps_to_end(s.start_time, s.stop_time, s.time_step);
    if (steps <= 0) {
        fdtd_parsed_free(P);
        fdtd::set_last_error("the scene has neither number_of_timesteps nor a stop_time after its start");
        return FDTD_ERR_INVALID;
    }

    // FDTD_AcousticSimulator::SaveSolverConfig (FDTD_AcousticSimulator.cpp:646-662)
    {
        std::ofstream of(composite(pattern.c_str(), simulator_id, "solver_settings.txt"));
        if (of) {
            of << "scheme " << (s.scheme == FDTD_SCHEME_STAGGERED ? "staggered" : "collocated") << "\n"
               << "sound_speed " << s.sound_speed << "\ndensity " << s.density << "\ncellsize " << s.cell_size
               << "\ntimestep " << s.time_step << "\nstop_time " << s.stop_time << "\nstart_time " << s.start_time
               << "\nsteps " << steps << "\nPML_width " << s.pml_width << "\nPML_strength " << s.pml_strength
               << "\nboundary_condition_preset " << s.boundary_condition_preset << "\nboundary_handling "
               << (s.boundary_handling == FDTD_BOUNDARY_FULLY_COUPLED ? "fully_coupled" : "rasterize")
               << "\ndivisions " << s.divisions[0] << " " << s.divisions[1] << " " << s.divisions[2]
               << "\nmin_bound " << s.min_bound[0] << " " << s.min_bound[1] << " " << s.min_bound[2]
               << "\nobjects " << sc->n_objects << "\nimpulse_sources " << sc->n_impulse_sources
               << "\npressure_sources " << sc->n_pressure_sources << "\nlisteners " << L << "\n";
        }
        std::ifstream in(xml_path, std::ios::binary);
        std::ofstream cp(composite(pattern.c_str(), simulator_id, "input_control_file.xml"), std::ios::binary);
        if (in && cp) cp << in.rdbuf();
    }
    if (L > 0) {
        st = write_matrix(composite(pattern.c_str(), simulator_id, "listening_position.dat"), sc->listeners, L, 3);
        if (st != FDTD_OK) { fdtd_parsed_free(P); return st; }
    }

    if (s.time_parallel) {
        // SimWorld.cpp:354-405, 870-911: each chunk writes its own audio
        // stream "%05d_all_audio.dat" and "%05d_start_time", "%05d_time_range"
        // and "%05d_sim-done" under the output pattern; all chunks run here as
        // one batch (the reference runs a process per chunk)
        std::vector<fdtd::ChunkRun> chunks;
        fdtd_batch_options bo;
        fdtd_batch_options_default(&bo);
        if (opt) bo.tile_config = opt->tile_config;
        st = fdtd::run_time_chunks(*sc, steps, &bo, chunks);
        char tag[32];
        for (size_t c = 0; c < chunks.size() && st == FDTD_OK; ++c) {
            const fdtd::ChunkRun &ch = chunks[c];
            std::snprintf(tag, sizeof tag, "%05d_", (int)c);
            const std::string audio = composite(pattern.c_str(), nullptr, std::string(tag) + "all_audio.dat");
            std::ofstream out(audio, std::ios::binary);
            if (!out) { fdtd::set_last_error("cannot open " + audio + " for writing"); st = FDTD_ERR_INVALID; break; }
            if (L) out.write(reinterpret_cast<const char *>(ch.series.data()), sizeof(double) * ch.series.size());
            double end = ch.start;
            for (int64_t q = 0; q < ch.steps_run; ++q) end += s.time_step;
            std::ofstream t0(composite(pattern.c_str(), nullptr, std::string(tag) + "start_time"));
            t0 << std::setprecision(18) << std::fixed << ch.start << std::endl;
            std::ofstream tr(composite(pattern.c_str(), nullptr, std::string(tag) + "time_range"));
            tr << std::setprecision(18) << std::fixed << ch.start << std::endl << ch.stop << std::endl << end << std::endl;
            std::ofstream done(composite(pattern.c_str(), nullptr, std::string(tag) + "sim-done"));
            if (steps_run) *steps_run += ch.steps_run;
        }
        fdtd_parsed_free(P);
        return st;
    }

    fdtd_handle *h = nullptr;
    st = fdtd_create(sc, opt, &h);
    if (st != FDTD_OK) { fdtd_parsed_free(P); return st; }
    // SimWorld.cpp:354-360: the audio stream is the pattern applied to "all_audio.dat"
    const std::string audio = composite(pattern.c_str(), nullptr, "all_audio.dat");
    std::ofstream out(audio, std::ios::binary);
    if (!out) {
        fdtd_destroy(h);
        fdtd_parsed_free(P);
        fdtd::set_last_error("cannot open " + audio + " for writing");
        return FDTD_ERR_INVALID;
    }
    const long long block = 1024;
    std::vector<double> rows;
    long long done = 0;
    while (done < steps && st == FDTD_OK) {
        const int n = (int)std::min(block, steps - done);
        st = fdtd_step(h, n);
        if (st != FDTD_OK) break;
        rows.resize((size_t)n * L);
        if (L) st = fdtd_read_listener_series(h, done, n, rows.data());
        if (st != FDTD_OK) break;
        // AudioOutput::WriteAndResetBuffer (AudioOutput.cpp:55-64): one row of L doubles per step
        if (L) out.write(reinterpret_cast<const char *>(rows.data()), sizeof(double) * rows.size());
        done += n;
    }
    out.close();
    fdtd_destroy(h);
    fdtd_parsed_free(P);
    if (steps_run) *steps_run = done;
    return st;
}

}  // extern "C"
