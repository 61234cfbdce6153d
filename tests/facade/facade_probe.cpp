// Drives the C++ facade: builds a small free-field scene, steps it and prints
// the last listener value (GPU), or reports the C-ABI's error (no GPU).
#include <cstdio>
#include <cmath>
#include "../../acoustics_amd/facade/wavesolver.hpp"
using namespace fdtd_amd;
int main(int argc, char **argv)
{
    try {
        if (argc > 1) {  // XML path
            FDTD_AcousticSimulator sim(argv[1]);
            sim.InitializeSolver();
            sim.Run();
            std::vector<double> series = sim.ListenerSeries();
            printf("xml steps=%d last=%.17g t=%.9g\n", sim.NumberOfTimesteps(), series.empty() ? 0.0 : series.back(), sim.GetSimulationTime());
            return 0;
        }
        auto settings = std::make_shared<PML_WaveSolver_Settings>();
        settings->cellSize = 0.005;
        settings->airDensity = 1.2;
        settings->timeStepSize = 0.5 * 0.005 / (343.0 * std::sqrt(3.0));
        settings->PML_width = 6;
        settings->listeningPoints = {Vector3d(0.02, 0.01, -0.01)};
        auto objects = std::make_shared<FDTD_Objects>();
        const double ws = 343.0 * 2e-4;
        objects->AddPressureSource(Vector3d(0.001, 0.002, -0.003), 2e-4, ws, 8e-4, 1.0 / std::pow(std::sqrt(2 * M_PI) * ws, 3));
        PML_WaveSolver solver(BoundingBox(0.005, 32, Vector3d(0, 0, 0)), settings, objects);
        for (int i = 0; i < 100; ++i) solver.stepSystem();
        std::vector<double> out;
        solver.FetchPressureData(settings->listeningPoints, out);
        Tuple3i d = solver.GetGrid().pressureFieldDivisions();
        printf("facade cells=%d divs=%d,%d,%d pml=%lld p=%.17g t=%.9g\n", solver.GetGrid().numPressureCells(), d.x, d.y, d.z,
               solver.GetGrid().numPMLCells(), out[0], solver.currentSimTime());
        return 0;
    } catch (const std::runtime_error &e) {
        printf("runtime_error: %s\n", e.what());
        return 3;
    }
}
