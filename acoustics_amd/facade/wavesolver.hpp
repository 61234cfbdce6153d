// wavesolver.hpp -- C++ facade with the class surface of the reference
// wavesolver (names and method meanings kept), forwarding to the C-ABI in
// include/fdtd_acoustics.h.  Header only; link against libfdtd_acoustics.so.
//
//   PML_WaveSolver_Settings   src/wavesolver/PML_WaveSolver_Settings.h:17-99
//   FDTD_Objects              src/wavesolver/FDTD_Objects.h:103-146 (analytic
//                             sphere/ellipsoid objects, Gaussian-impulse,
//                             harmonic and acceleration-noise vibrational
//                             sources, Gaussian pressure sources)
//   ScalarField               src/field/ScalarField.cpp:29-140 (geometry only)
//   MAC_Grid                  src/wavesolver/MAC_Grid.h:437-470
//   PML_WaveSolver            src/wavesolver/PML_WaveSolver.h:165-248
//   FDTD_AcousticSimulator    src/wavesolver/FDTD_AcousticSimulator.h:30-119
//
// Errors surface as std::runtime_error("**ERROR** ...") like the reference
// (e.g. PML_WaveSolver.cpp:157); the C-ABI underneath never throws.
#pragma once

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fdtd_acoustics.h"

namespace fdtd_amd {

struct Vector3d {
    double x = 0, y = 0, z = 0;
    Vector3d() = default;
    Vector3d(double a, double b, double c) : x(a), y(b), z(c) {}
};
using Vector3Array = std::vector<Vector3d>;
struct Tuple3i {
    int x = 0, y = 0, z = 0;
    Tuple3i() = default;
    Tuple3i(int a, int b, int c) : x(a), y(b), z(c) {}
    int operator[](int d) const { return d == 0 ? x : d == 1 ? y : z; }
};

// Dense row-major matrix with the (row, col) access and resize of the
// Eigen::MatrixXd the reference fills (PML_WaveSolver.cpp:210-216).
struct Matrix {
    int rows_ = 0, cols_ = 0;
    std::vector<double> v;
    void resize(int r, int c) { rows_ = r; cols_ = c; v.assign((size_t)r * c, 0.0); }
    int rows() const { return rows_; }
    int cols() const { return cols_; }
    double &operator()(int r, int c) { return v[(size_t)r * cols_ + c]; }
    double operator()(int r, int c) const { return v[(size_t)r * cols_ + c]; }
};

inline void check(fdtd_status s, const char *what)
{
    if (s != FDTD_OK)
        throw std::runtime_error(std::string("**ERROR** ") + what + ": " + fdtd_status_string(s) + " (" + fdtd_last_error() + ")");
}

// BoundingBox(cellSize, cellDivisions, center) of src/geometry/BoundingBox.h:90-100
struct BoundingBox {
    Vector3d minBound, maxBound;
    BoundingBox(double cellSize, int cellDivisions, const Vector3d &center)
        : minBound(center.x - cellSize * (double)cellDivisions / 2.0, center.y - cellSize * (double)cellDivisions / 2.0,
                   center.z - cellSize * (double)cellDivisions / 2.0),
          maxBound(center.x + cellSize * (double)cellDivisions / 2.0, center.y + cellSize * (double)cellDivisions / 2.0,
                   center.z + cellSize * (double)cellDivisions / 2.0) {}
    BoundingBox(const Vector3d &mn, const Vector3d &mx) : minBound(mn), maxBound(mx) {}
};

// airDensity defaults to the XML reader's 1.0 (ImpulseResponseParser.cpp:509)
struct PML_WaveSolver_Settings {
    double soundSpeed = 343.0, airDensity = 1.0, cellSize = 0.005, timeEnd = 0.0, timeStepSize = 1.0 / 120000.0;
    double PML_width = 10.0, PML_strength = 1e6, alpha = 0.0, fastForwardToEventTime = 0.0;
    int boundaryConditionPreset = 0;
    bool useMesh = true;
    enum BoundaryHandling { RASTERIZE = 0, FULLY_COUPLED = 1 } boundaryHandlingType = RASTERIZE;
    Vector3Array listeningPoints;
};
using PML_WaveSolver_Settings_Ptr = std::shared_ptr<PML_WaveSolver_Settings>;

class FDTD_Objects {
  public:
    int AddSphere(const Vector3d &c, double radius)
    {
        fdtd_object o{};
        o.type = FDTD_OBJ_SPHERE;
        o.center[0] = c.x; o.center[1] = c.y; o.center[2] = c.z;
        o.semi_axes[0] = o.semi_axes[1] = o.semi_axes[2] = radius;
        o.rotation[0] = o.rotation[4] = o.rotation[8] = 1.0;
        _objects.push_back(o);
        return (int)_objects.size() - 1;
    }
    int AddEllipsoid(const Vector3d &c, const Vector3d &semiAxes, const double *rotationRowMajor)
    {
        fdtd_object o{};
        o.type = FDTD_OBJ_ELLIPSOID;
        o.center[0] = c.x; o.center[1] = c.y; o.center[2] = c.z;
        o.semi_axes[0] = semiAxes.x; o.semi_axes[1] = semiAxes.y; o.semi_axes[2] = semiAxes.z;
        for (int i = 0; i < 9; ++i) o.rotation[i] = rotationRowMajor[i];
        _objects.push_back(o);
        return (int)_objects.size() - 1;
    }
    void AddImpulseSource(int object, double amplitude, double width, double onset)
    {
        fdtd_impulse_source s{};
        s.object = object; s.kind = FDTD_VIB_GAUSSIAN; s.amplitude = amplitude; s.width = width; s.onset = onset;
        _impulses.push_back(s);
    }
    // HarmonicVibrationalSource(owner, omega, phase[, decayRate, decayCenter])
    // with its unit pulsation velocity (HarmonicVibrationalSource.h:25-31)
    void AddHarmonicSource(int object, double omega, double phase, double decayRate = -1.0, double decayCenter = 0.0)
    {
        fdtd_impulse_source s{};
        s.object = object; s.kind = FDTD_VIB_HARMONIC; s.amplitude = 1.0; s.omega = omega; s.phase = phase;
        s.decay_rate = decayRate; s.decay_center = decayCenter;
        _impulses.push_back(s);
    }
    // AccelerationNoiseVibrationalSource of a static object with its impacts
    void AddAccelerationNoiseSource(int object, double mass, const Vector3d &centerOfMass, const double *invInertiaRowMajor,
                                    const std::vector<fdtd_impact> &impacts)
    {
        auto a = std::make_shared<std::vector<fdtd_impact>>(impacts);
        fdtd_acc_noise_source s{};
        s.object = object; s.mass = mass;
        s.center_of_mass[0] = centerOfMass.x; s.center_of_mass[1] = centerOfMass.y; s.center_of_mass[2] = centerOfMass.z;
        for (int i = 0; i < 9; ++i) s.inv_inertia[i] = invInertiaRowMajor[i];
        s.n_impacts = (int)a->size();
        s.impacts = a->data();
        _impactLists.push_back(a);
        _accNoise.push_back(s);
    }
    void AddPressureSource(const Vector3d &pos, double widthTime, double widthSpace, double offsetTime, double normalizeConstant)
    {
        fdtd_pressure_source s{};
        s.position[0] = pos.x; s.position[1] = pos.y; s.position[2] = pos.z;
        s.width_time = widthTime; s.width_space = widthSpace; s.offset_time = offsetTime; s.normalize_constant = normalizeConstant;
        _pressure.push_back(s);
    }
    int N() const { return (int)_objects.size(); }
    bool HasExternalPressureSources() const { return !_pressure.empty(); }
    std::vector<fdtd_object> _objects;
    std::vector<fdtd_impulse_source> _impulses;
    std::vector<fdtd_pressure_source> _pressure;
    std::vector<fdtd_acc_noise_source> _accNoise;
    std::vector<std::shared_ptr<std::vector<fdtd_impact>>> _impactLists;
};
using FDTD_Objects_Ptr = std::shared_ptr<FDTD_Objects>;

// Geometry of a cell-centred field (ScalarField.cpp:29-140): cells of size
// h from minBound, flat index (i ny + j) nz + k, centres at
// minBound + (index + 1/2) h.
class ScalarField {
  public:
    ScalarField() = default;
    ScalarField(const Vector3d &minBound, const Tuple3i &divisions, double cellSize)
        : _min(minBound), _div(divisions), _h(cellSize) {}
    double cellSize() const { return _h; }
    const Tuple3i &cellDivisions() const { return _div; }
    int numCells() const { return _div.x * _div.y * _div.z; }
    int cellIndex(const Tuple3i &i) const { return (i.x * _div.y + i.y) * _div.z + i.z; }
    Tuple3i cellIndex(int flat) const { return Tuple3i(flat / (_div.y * _div.z), (flat / _div.z) % _div.y, flat % _div.z); }
    Vector3d cellPosition(const Tuple3i &i) const
    {
        return Vector3d(_min.x + (0.5 + i.x) * _h, _min.y + (0.5 + i.y) * _h, _min.z + (0.5 + i.z) * _h);
    }
    Vector3d cellPosition(int flat) const { return cellPosition(cellIndex(flat)); }
    BoundingBox bbox() const
    {
        return BoundingBox(_min, Vector3d(_min.x + _div.x * _h, _min.y + _div.y * _h, _min.z + _div.z * _h));
    }

  private:
    Vector3d _min;
    Tuple3i _div;
    double _h = 0.0;
};

class PML_WaveSolver;

// The grid of a solver: the pressure field's geometry and the staggered
// velocity fields' (velocity field d starts h/2 lower along d and has one
// more layer there, MAC_Grid.cpp:101-113), cell counts, and the pressure
// values read back from the device.
class MAC_Grid {
  public:
    explicit MAC_Grid(PML_WaveSolver *s) : _s(s) {}
    inline int numPressureCells() const;
    inline Tuple3i pressureFieldDivisions() const;
    const ScalarField &pressureField() const { return _p; }
    const ScalarField &velocityField(int ind) const { return _v[ind]; }
    inline std::vector<double> pressureValues() const;  // FP64 copy of the newest pressure, z fastest
    inline long long numPMLCells() const;
    inline long long numGhostCells() const;

  private:
    friend class PML_WaveSolver;
    void setGeometry(const Vector3d &mn, const Tuple3i &div, double h)
    {
        _p = ScalarField(mn, div, h);
        _v[0] = ScalarField(Vector3d(mn.x - h / 2, mn.y, mn.z), Tuple3i(div.x + 1, div.y, div.z), h);
        _v[1] = ScalarField(Vector3d(mn.x, mn.y - h / 2, mn.z), Tuple3i(div.x, div.y + 1, div.z), h);
        _v[2] = ScalarField(Vector3d(mn.x, mn.y, mn.z - h / 2), Tuple3i(div.x, div.y, div.z + 1), h);
    }
    PML_WaveSolver *_s;
    ScalarField _p, _v[3];
};

class PML_WaveSolver {
  public:
    PML_WaveSolver(const BoundingBox &bbox, PML_WaveSolver_Settings_Ptr settings, std::shared_ptr<FDTD_Objects> objects,
                   int stepsPerGraph = 0)
        : _settings(settings), _objects(objects), _grid(this)
    {
        fdtd_scene sc{};
        fdtd_settings &s = sc.settings;
        fdtd_settings_default(&s);
        s.sound_speed = settings->soundSpeed;
        s.density = settings->airDensity;
        s.cell_size = settings->cellSize;
        s.time_step = settings->timeStepSize;
        s.stop_time = settings->timeEnd;
        s.pml_width = settings->PML_width;
        s.pml_strength = settings->PML_strength;
        s.air_viscosity_alpha = settings->alpha;
        s.start_time = settings->fastForwardToEventTime;
        s.boundary_condition_preset = settings->boundaryConditionPreset;
        s.boundary_handling = (int)settings->boundaryHandlingType;
        s.use_mesh = settings->useMesh ? 1 : 0;
        const double mn[3] = {bbox.minBound.x, bbox.minBound.y, bbox.minBound.z};
        const double mx[3] = {bbox.maxBound.x, bbox.maxBound.y, bbox.maxBound.z};
        for (int d = 0; d < 3; ++d) {
            s.min_bound[d] = mn[d];
            s.divisions[d] = (int)std::ceil(((mx[d] - mn[d]) - 1e-10) / settings->cellSize);  // ScalarField.cpp:37-46
        }
        _divs.x = s.divisions[0]; _divs.y = s.divisions[1]; _divs.z = s.divisions[2];
        _listeners.clear();
        for (const Vector3d &p : settings->listeningPoints) { _listeners.push_back(p.x); _listeners.push_back(p.y); _listeners.push_back(p.z); }
        sc.n_objects = objects ? objects->N() : 0;
        sc.objects = objects ? objects->_objects.data() : nullptr;
        sc.n_impulse_sources = objects ? (int)objects->_impulses.size() : 0;
        sc.impulse_sources = objects ? objects->_impulses.data() : nullptr;
        sc.n_pressure_sources = objects ? (int)objects->_pressure.size() : 0;
        sc.pressure_sources = objects ? objects->_pressure.data() : nullptr;
        sc.n_acc_noise_sources = objects ? (int)objects->_accNoise.size() : 0;
        sc.acc_noise_sources = objects ? objects->_accNoise.data() : nullptr;
        sc.n_listeners = (int)settings->listeningPoints.size();
        sc.listeners = _listeners.data();
        create(sc, stepsPerGraph);
    }
    // A solver for a scene read through the C-ABI (the XML path), with its
    // settings and listening points mirrored into settings.
    PML_WaveSolver(const fdtd_scene &scene, PML_WaveSolver_Settings_Ptr settings, int stepsPerGraph = 0)
        : _settings(settings), _grid(this)
    {
        _divs = Tuple3i(scene.settings.divisions[0], scene.settings.divisions[1], scene.settings.divisions[2]);
        create(scene, stepsPerGraph);
    }
    // One time step; false once the settings' end time is reached
    // (PML_WaveSolver.cpp:579-609).
    bool stepSystem()
    {
        check(fdtd_step(_h.get(), 1), "stepSystem");
        return !(_settings->timeEnd > 0.0 && currentSimTime() >= _settings->timeEnd);
    }
    bool stepSystem(int n)
    {
        check(fdtd_step(_h.get(), n), "stepSystem");
        return !(_settings->timeEnd > 0.0 && currentSimTime() >= _settings->timeEnd);
    }
    // Listener pressures after the newest step, one row per point
    // (PML_WaveSolver.cpp:210-216, 319-373); the points must be the settings'
    // listening points (their MLS weights are precomputed).  dim = -1 reads
    // the pressure; the split pressures (dim 0..2) are not kept per listener
    // and are refused.
    void FetchPressureData(const Vector3Array &listeningPoints, Matrix &data, const int dim = -1)
    {
        std::vector<double> v;
        FetchPressureData(listeningPoints, v, dim);
        data.resize((int)v.size(), 1);
        for (size_t i = 0; i < v.size(); ++i) data((int)i, 0) = v[i];
    }
    void FetchPressureData(const Vector3Array &listeningPoints, std::vector<double> &data, const int dim = -1)
    {
        if (dim != -1) throw std::runtime_error("**ERROR** FetchPressureData: split pressures (dim >= 0) are not fetched per listener");
        if (listeningPoints.size() != _settings->listeningPoints.size())
            throw std::runtime_error("**ERROR** FetchPressureData: points differ from the configured listening points");
        data.resize(listeningPoints.size());
        if (data.empty()) return;
        check(fdtd_listen(_h.get(), data.data()), "FetchPressureData");
    }
    double currentSimTime() const
    {
        fdtd_info i;
        check(fdtd_get_info(_h.get(), &i), "currentSimTime");
        return i.time;
    }
    int numCells() const
    {
        fdtd_info i;
        check(fdtd_get_info(_h.get(), &i), "numCells");
        return (int)i.n_cells;
    }
    MAC_Grid &GetGrid() { return _grid; }
    const Tuple3i &fieldDivisions() const { return _divs; }
    fdtd_handle *handle() const { return _h.get(); }

  private:
    void create(const fdtd_scene &sc, int stepsPerGraph)
    {
        fdtd_options o;
        fdtd_options_default(&o);
        o.steps_per_graph = stepsPerGraph;
        fdtd_handle *h = nullptr;
        check(fdtd_create(&sc, &o, &h), "PML_WaveSolver");
        _h.reset(h);
        const fdtd_settings &s = sc.settings;
        _grid.setGeometry(Vector3d(s.min_bound[0], s.min_bound[1], s.min_bound[2]), _divs, s.cell_size);
    }
    struct Deleter { void operator()(fdtd_handle *h) const { if (h) fdtd_destroy(h); } };
    PML_WaveSolver_Settings_Ptr _settings;
    std::shared_ptr<FDTD_Objects> _objects;
    std::vector<double> _listeners;
    Tuple3i _divs;
    std::unique_ptr<fdtd_handle, Deleter> _h;
    MAC_Grid _grid;
    friend class MAC_Grid;
};

int MAC_Grid::numPressureCells() const { return _s->numCells(); }
Tuple3i MAC_Grid::pressureFieldDivisions() const { return _s->fieldDivisions(); }
std::vector<double> MAC_Grid::pressureValues() const
{
    std::vector<double> p((size_t)_s->numCells());
    check(fdtd_read_pressure(_s->handle(), p.data()), "pressureValues");
    return p;
}
long long MAC_Grid::numPMLCells() const
{
    fdtd_info i;
    check(fdtd_get_info(_s->handle(), &i), "numPMLCells");
    return i.n_pml_cells;
}
long long MAC_Grid::numGhostCells() const
{
    fdtd_info i;
    check(fdtd_get_info(_s->handle(), &i), "numGhostCells");
    return i.n_ghost_cells;
}

// XML-driven simulator (FDTD_AcousticSimulator.cpp:419-544): parse, build,
// RunForSteps, and the FP64 listener series that AudioOutput writes.  Like
// the reference it owns (and deletes) the simulator id it is given.
class FDTD_AcousticSimulator {
  public:
    explicit FDTD_AcousticSimulator(std::string *simulatorID = nullptr) : FDTD_AcousticSimulator("", simulatorID) {}
    FDTD_AcousticSimulator(const std::string &configFile, std::string *simulatorID = nullptr)
        : _configFile(configFile), _simulatorID(simulatorID) {}
    ~FDTD_AcousticSimulator() { delete _simulatorID; }
    FDTD_AcousticSimulator(const FDTD_AcousticSimulator &) = delete;
    FDTD_AcousticSimulator &operator=(const FDTD_AcousticSimulator &) = delete;

    // parse, then build the solver on the scene's own box and settings
    void InitializeSolver(int stepsPerGraph = 0)
    {
        Parsed p(_configFile);
        _numSteps = fdtd_parsed_num_steps(p.p);
        build(*fdtd_parsed_scene(p.p), stepsPerGraph);
    }
    // parse, then build the solver on solverBox with settings (InitializeSolver(
    // const BoundingBox &, const PML_WaveSolver_Settings_Ptr &),
    // FDTD_AcousticSimulator.cpp:435-453): the scene's objects, sources and
    // listeners, the given grid and solver settings
    void InitializeSolver(const BoundingBox &solverBox, const PML_WaveSolver_Settings_Ptr &settings, int stepsPerGraph = 0)
    {
        Parsed p(_configFile);
        _numSteps = fdtd_parsed_num_steps(p.p);
        fdtd_scene sc = *fdtd_parsed_scene(p.p);
        fdtd_settings &s = sc.settings;
        s.sound_speed = settings->soundSpeed;
        s.density = settings->airDensity;
        s.cell_size = settings->cellSize;
        s.time_step = settings->timeStepSize;
        s.stop_time = settings->timeEnd;
        s.pml_width = settings->PML_width;
        s.pml_strength = settings->PML_strength;
        s.air_viscosity_alpha = settings->alpha;
        s.start_time = settings->fastForwardToEventTime;
        s.boundary_condition_preset = settings->boundaryConditionPreset;
        s.boundary_handling = (int)settings->boundaryHandlingType;
        s.use_mesh = settings->useMesh ? 1 : 0;
        const double mn[3] = {solverBox.minBound.x, solverBox.minBound.y, solverBox.minBound.z};
        const double mx[3] = {solverBox.maxBound.x, solverBox.maxBound.y, solverBox.maxBound.z};
        for (int d = 0; d < 3; ++d) {
            s.min_bound[d] = mn[d];
            s.divisions[d] = (int)std::ceil(((mx[d] - mn[d]) - 1e-10) / settings->cellSize);  // ScalarField.cpp:37-46
        }
        std::vector<double> lst;
        if (!settings->listeningPoints.empty()) {
            for (const Vector3d &q : settings->listeningPoints) { lst.push_back(q.x); lst.push_back(q.y); lst.push_back(q.z); }
            sc.n_listeners = (int)settings->listeningPoints.size();
            sc.listeners = lst.data();
        }
        _settings = settings;
        build(sc, stepsPerGraph);
    }
    bool RunForSteps(int n)
    {
        check(fdtd_step(handle(), n), "RunForSteps");
        return true;
    }
    void Run() { RunForSteps(_numSteps); }
    double GetSimulationTime() const { return _solver ? _solver->currentSimTime() : 0.0; }
    const std::shared_ptr<PML_WaveSolver> &GetSolver() const { return _solver; }
    const PML_WaveSolver_Settings_Ptr &GetSolverSettings() const { return _settings; }
    MAC_Grid &GetGrid() { return _solver->GetGrid(); }
    std::string *GetSimulatorID() { return _simulatorID; }
    // CompositeFilename (FDTD_AcousticSimulator.cpp:110-124)
    std::string CompositeFilename(const std::string &filename, const std::string &pattern = "%s") const
    {
        char buf[4096];
        check(fdtd_composite_filename(pattern.c_str(), _simulatorID ? _simulatorID->c_str() : nullptr, filename.c_str(), buf,
                                      (int)sizeof buf),
              "CompositeFilename");
        return buf;
    }
    // [steps][listeners] FP64, the content of all_audio.dat (AudioOutput.cpp:55-64)
    std::vector<double> ListenerSeries() const
    {
        fdtd_info i;
        check(fdtd_get_info(handle(), &i), "ListenerSeries");
        std::vector<double> out((size_t)i.step_index * (size_t)_listeners);
        if (!out.empty()) check(fdtd_read_listener_series(handle(), 0, i.step_index, out.data()), "ListenerSeries");
        return out;
    }
    int NumberOfTimesteps() const { return _numSteps; }

  private:
    struct Parsed {
        fdtd_parsed *p = nullptr;
        explicit Parsed(const std::string &path) { check(fdtd_parse_xml(path.c_str(), &p), "InitializeSolver (parse)"); }
        ~Parsed() { fdtd_parsed_free(p); }
    };
    void build(const fdtd_scene &sc, int stepsPerGraph)
    {
        if (!_settings) {
            _settings = std::make_shared<PML_WaveSolver_Settings>();
            const fdtd_settings &s = sc.settings;
            _settings->soundSpeed = s.sound_speed; _settings->airDensity = s.density; _settings->cellSize = s.cell_size;
            _settings->timeEnd = s.stop_time; _settings->timeStepSize = s.time_step; _settings->PML_width = s.pml_width;
            _settings->PML_strength = s.pml_strength; _settings->alpha = s.air_viscosity_alpha;
            _settings->fastForwardToEventTime = s.start_time; _settings->boundaryConditionPreset = s.boundary_condition_preset;
            _settings->useMesh = s.use_mesh != 0;
            _settings->boundaryHandlingType = (PML_WaveSolver_Settings::BoundaryHandling)s.boundary_handling;
            _settings->listeningPoints.clear();
            for (int l = 0; l < sc.n_listeners; ++l)
                _settings->listeningPoints.push_back(Vector3d(sc.listeners[3 * l], sc.listeners[3 * l + 1], sc.listeners[3 * l + 2]));
        }
        _listeners = sc.n_listeners;
        _solver = std::make_shared<PML_WaveSolver>(sc, _settings, stepsPerGraph);
    }
    fdtd_handle *handle() const
    {
        if (!_solver) throw std::runtime_error("**ERROR** the solver is not initialized");
        return _solver->handle();
    }
    std::string _configFile;
    std::string *_simulatorID = nullptr;
    PML_WaveSolver_Settings_Ptr _settings;
    std::shared_ptr<PML_WaveSolver> _solver;
    int _numSteps = 0, _listeners = 0;
};

}  // namespace fdtd_amd
