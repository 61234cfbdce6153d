/*
 * fdtd_acoustics.h -- C-ABI of the MI355X FDTD acoustic wave engine.
 *
 * This is the drop-in boundary for the time-stepping path of the wavesolver
 * (MAC_Grid / PML_WaveSolver / FDTD_AcousticSimulator).  Plain C types only:
 * pointers, sizes and PODs.  All device memory is owned by the handles; all
 * host output buffers are owned by the caller.
 *
 * Reference interfaces each group replaces (paths relative to the reference
 * repository root):
 *   fdtd_settings        <- PML_WaveSolver_Settings
 *                           (src/wavesolver/PML_WaveSolver_Settings.h:17-99)
 *   fdtd_object          <- FDTD_RigidObject geometry queries
 *                           (src/wavesolver/FDTD_RigidObject.h:135-147); analytic
 *                           sphere / ellipsoid instead of a mesh + distance field
 *   fdtd_impulse_source  <- VibrationalSource::Evaluate
 *                           (src/wavesolver/VibrationalSource.h:17-37): a new
 *                           Gaussian-impulse normal acceleration, and
 *                           HarmonicVibrationalSource
 *                           (src/wavesolver/HarmonicVibrationalSource.cpp:5-28)
 *   fdtd_acc_noise_source <- AccelerationNoiseVibrationalSource
 *                           (src/wavesolver/AccelerationNoiseVibrationalSource.cpp:
 *                           19-90) with its ImpactRecords
 *                           (src/modal_model/ImpulseSeriesObject.h:23-32)
 *   fdtd_kinematics      <- FDTD_RigidObject_Blender_Animator (kinematics files,
 *                           src/wavesolver/FDTD_RigidObject_Animator.cpp:78-160)
 *   fdtd_pressure_source <- GaussianPressureSource
 *                           (src/wavesolver/GaussianPressureSource.h:31-39)
 *   fdtd_create          <- PML_WaveSolver(bbox, settings, objects)
 *                           (src/wavesolver/PML_WaveSolver.cpp:78-100) and
 *                           FDTD_AcousticSimulator::InitializeSolver
 *                           (src/wavesolver/FDTD_AcousticSimulator.cpp:435-453);
 *                           opaque-handle style of wave_sim_init
 *                           (src/wavesolver/gpusolver/cuda/cuda_PAT_wave_3d.h:11-53)
 *   fdtd_create_from_xml <- FDTD_AcousticSimulator(configFile) + InitializeSolver()
 *                           (src/wavesolver/FDTD_AcousticSimulator.cpp:419-431)
 *   fdtd_step            <- FDTD_AcousticSimulator::RunForSteps
 *                           (src/wavesolver/FDTD_AcousticSimulator.cpp:528-544)
 *                           -> PML_WaveSolver::stepSystem / stepCollocated
 *                           (src/wavesolver/PML_WaveSolver.cpp:579-609, 775-826)
 *   fdtd_listen          <- PML_WaveSolver::FetchPressureData
 *                           (src/wavesolver/PML_WaveSolver.cpp:319-373)
 *   fdtd_read_listener_series
 *                        <- AudioOutput::AccumulateBuffer/WriteAndResetBuffer
 *                           (src/wavesolver/AudioOutput.cpp:26-64): FP64 [step][L]
 *   fdtd_read_pressure   <- PML_WaveSolver::vertexPressure / pressure field
 *                           (src/wavesolver/PML_WaveSolver.cpp:641-651)
 *   fdtd_destroy         <- wave_sim_free
 *                           (src/wavesolver/gpusolver/cuda/cuda_PAT_wave_3d.h:11-53)
 *
 * Errors are status codes; nothing here calls exit() (the reference's CUDA
 * helpers do, src/wavesolver/gpusolver/cuda/cuda_helper.h:4-12) and nothing
 * throws across this boundary.
 */
#ifndef FDTD_ACOUSTICS_H
#define FDTD_ACOUSTICS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* status codes                                                              */
/* ------------------------------------------------------------------------- */
typedef enum fdtd_status {
    FDTD_OK              = 0,
    FDTD_ERR_INVALID     = 1,  /* bad argument or scene                     */
    FDTD_ERR_NO_DEVICE   = 2,  /* no usable gfx950 device                   */
    FDTD_ERR_DEVICE      = 3,  /* HIP runtime error                         */
    FDTD_ERR_OOM         = 4,  /* device allocation failed / over budget    */
    FDTD_ERR_PARSE       = 5,  /* XML scene could not be parsed             */
    FDTD_ERR_UNSUPPORTED = 6,  /* scene uses a feature this engine lacks    */
    FDTD_ERR_NONFINITE   = 7,  /* NaN/Inf detected in a listener signal     */
    FDTD_ERR_BUSY        = 8,  /* request not finished yet                  */
    FDTD_ERR_CANCELLED   = 9,  /* request cancelled or evicted              */
    FDTD_ERR_NOT_FOUND   = 10  /* unknown request id                        */
} fdtd_status;

const char *fdtd_status_string(fdtd_status s);

/* ------------------------------------------------------------------------- */
/* scene description                                                         */
/* ------------------------------------------------------------------------- */
enum { FDTD_BOUNDARY_RASTERIZE = 0, FDTD_BOUNDARY_FULLY_COUPLED = 1 };
enum { FDTD_OBJ_SPHERE = 0, FDTD_OBJ_ELLIPSOID = 1 };

/* Solver settings, field by field from PML_WaveSolver_Settings and the XML
 * <solver> / <solver_control_policy type="static"> attributes
 * (src/parser/ImpulseResponseParser.cpp:492-641). */
/* Time-stepping scheme.  The reference compiles one of two: the collocated
 * pressure scheme (PML_WaveSolver::stepCollocated, its default because
 * USE_COLLOCATED is defined in Wavesolver_ConstantsAndTypes.h:37) or the fully
 * staggered MAC leapfrog (PML_WaveSolver::stepLeapfrog, PML_WaveSolver.cpp:710-773).
 * The staggered scheme runs free-field scenes (pressure sources, PML, no
 * objects); its Gaussian pressure source is the time integral of the pulse
 * (GaussianPressureSource.cpp:24-27). */
enum {
    FDTD_SCHEME_COLLOCATED = 0,
    FDTD_SCHEME_STAGGERED  = 1
};

/* How object cells are sampled: the reference's compile-time choice
 * (Wavesolver_ConstantsAndTypes.h:46-47).  CENTER is its compiled default
 * (ONLY_SAMPLE_CENTER and STAGGER_ACC_SAMPLE defined): a cell is solid when
 * an object holds its centre (MAC_Grid.cpp:3558-3564); a rasterised ghost
 * sits on the face between the solid cell and its fluid neighbour and takes
 * p_g = p_nbr - h grad_p . n_cell (MAC_Grid.cpp:1277-1278, 4322-4331), a
 * fully-coupled ghost sits at the solid cell's centre; the Laplacian takes
 * ghost values with the plain 7-point weights (MAC_Grid.cpp:247-266).
 * SEVEN is the other configuration (ONLY_SAMPLE_CENTER undefined): solid
 * when the centre or any of the six points h/4 from it is inside an object
 * (3566-3577), ghosts h/4 from the solid cell's centre, p_g = p_nbr -
 * 0.75 h grad_p . n_cell (1280), and the non-uniform weights for a ghost at
 * 3h/4 (268-286). */
enum {
    FDTD_SAMPLE_CENTER = 0,
    FDTD_SAMPLE_SEVEN  = 1
};

typedef struct fdtd_settings {
    double sound_speed;               /* sound_speed (m/s), default 343        */
    double density;                   /* density (kg/m^3)                      */
    double cell_size;                 /* cellsize h (m)                        */
    double time_step;                 /* 1 / timestepfrequency (s)             */
    double stop_time;                 /* stop_time (s); <= 0: no limit         */
    double pml_width;                 /* PML_width in cells                    */
    double pml_strength;              /* PML_strength                          */
    double air_viscosity_alpha;       /* air_viscosity_alpha (0: lossless)     */
    double start_time;                /* fast_forward_to_event_time            */
    int    boundary_condition_preset; /* 0 no wall, 1, 2 (MAC_Grid::InsidePML) */
    int    boundary_handling;         /* FDTD_BOUNDARY_*                       */
    int    use_mesh;                  /* 1: classify PML and object cells; 0:
                                         no PML and no objects, the Engquist
                                         boundary on layer 1 (MAC_Grid.cpp:
                                         521-919, 2457-2473).  A PML width of
                                         0 also selects that boundary.       */
    int    divisions[3];              /* pressure cells per axis               */
    double min_bound[3];              /* minimum corner of the pressure grid   */
    int    scheme;                    /* FDTD_SCHEME_*                         */
    /* time-parallel chunks (ImpulseResponseParser.cpp:525-533, SimWorld.cpp:
     * 151-166): with time_parallel set, every source (pressure sources and
     * boundary accelerations) acts only in the steps that start before
     * stop_boundary_acc_time (counted from start_time by repeated addition of
     * time_step, as the reference's clock runs; SimWorld.cpp:474). */
    int    time_parallel;             /* time_parallel (0: off)                */
    int    num_time_chunks;           /* num_time_chunks                       */
    int    sampling;                  /* FDTD_SAMPLE_* (object cells, ghosts)  */
    double overlap_time;              /* overlap_time (s), default 6e-3        */
    double early_termination_ratio;   /* early_termination_ratio, default 1e-3 */
    double stop_boundary_acc_time;    /* set per chunk by fdtd_time_chunk      */
} fdtd_settings;

/* Rigid object with an analytic signed distance.  The rotation maps object
 * coordinates to world coordinates (row-major 3x3, orthonormal).  center and
 * rotation give the pose at simulation time 0; an object with a nonzero
 * velocity or angular velocity moves rigidly: c(t) = center + velocity t,
 * R(t) = Rot(angular_velocity, |angular_velocity| t) rotation, about its
 * centre (stands in for the rigid-body animator, FDTD_Objects.cpp:551-582).
 * Moving objects are reclassified every step, with fresh cells filled by
 * MAC_Grid::InterpolateFreshPressureCell's rule. */
typedef struct fdtd_object {
    int    type;          /* FDTD_OBJ_SPHERE | FDTD_OBJ_ELLIPSOID        */
    int    reserved;
    double center[3];     /* world position of the centre at t = 0 (m)   */
    double semi_axes[3];  /* sphere: radius in semi_axes[0]              */
    double rotation[9];   /* world-from-object rotation, row-major       */
    double velocity[3];   /* m/s, world frame                            */
    double angular_velocity[3]; /* rad/s, world frame, about the centre   */
} fdtd_object;

/* Kinds of fdtd_impulse_source. */
enum { FDTD_VIB_GAUSSIAN = 0, FDTD_VIB_HARMONIC = 1, FDTD_VIB_CHIRP = 2 };

/* Vibrational source on the whole surface of one object (VibrationalSource,
 * VibrationalSource.h:17-37) whose acceleration vector at a boundary point is
 * a_n(t) times the outward normal there:
 *   FDTD_VIB_GAUSSIAN (new): a_n = A exp(-(t - t_on)^2 / (2 w^2));
 *   FDTD_VIB_HARMONIC (HarmonicVibrationalSource, HarmonicVibrationalSource.
 *   cpp:5-28, a pulsating surface): a_n = A omega sin(omega t + phase), times
 *   exp(-decay_rate (t - decay_center)) when decay_rate >= 0, and zero
 *   outside [t0, t1] when t1 > t0.  A is the amplitude of the pulsation
 *   velocity (the reference fixes it at 1 m/s; XML has_harmonic_source with
 *   harmonic_source_f = omega / 2 pi, phase 0, no decay, no range);
 *   FDTD_VIB_CHIRP (ChirpVibrationalSource, ChirpVibrationalSource.cpp:5-19):
 *   zero before t0, else a_n = sin(w t + phi) with alpha = 0.5 (t - t0) /
 *   (t1 - t0) before t1 and 1 from t1 on, w = (1 - alpha) omega + alpha
 *   omega1, phi = 0 while alpha < 0.52 and t1 (omega1 - omega) / 2 after
 *   (the reference's unit amplitude; XML has_chirp_source with
 *   chirp_frequency_0/1 = omega / 2 pi, omega1 / 2 pi, chirp_time_0/1). */
typedef struct fdtd_impulse_source {
    int    object;        /* index into fdtd_scene.objects               */
    int    kind;          /* FDTD_VIB_GAUSSIAN | _HARMONIC | _CHIRP      */
    double amplitude;     /* Gaussian: A (m/s^2); harmonic: A (m/s)      */
    double width;         /* Gaussian: w (s)                             */
    double onset;         /* Gaussian: t_on (s)                          */
    double omega;         /* harmonic, chirp: angular frequency (rad/s)  */
    double phase;         /* harmonic: phase (rad)                       */
    double decay_rate;    /* harmonic: < 0 none (1/s)                    */
    double decay_center;  /* harmonic (s)                                */
    double t0, t1;        /* harmonic: active range when t1 > t0 (s);
                             chirp: the sweep's start and end (s)        */
    double omega1;        /* chirp: final angular frequency (rad/s)      */
} fdtd_impulse_source;

/* One impact of an acceleration-noise source (ImpulseSeriesObject::
 * ImpactRecord, ImpulseSeriesObject.h:23-32): impulse J applied at a point
 * of the object, spread over a half-sine contact force of duration tau. */
typedef struct fdtd_impact {
    double timestamp;     /* t_0 (s)                                     */
    double support;       /* tau (s); impacts with tau < 1e-12 add nothing */
    double impulse[3];    /* J (N s), object frame                       */
    double position[3];   /* impact point (m), object frame              */
} fdtd_impact;

/* Acceleration noise of a rigid object (AccelerationNoiseVibrationalSource,
 * AccelerationNoiseVibrationalSource.cpp:63-90): one acceleration vector for
 * the whole surface, a(t) = R sum over impacts with t_0 <= t < t_0 + tau of
 *   pi S / (2 tau) (J / m + (I^-1 (r x J)) x r),   r = position - centre of mass,
 * S = sin(pi (t - t_0) / tau), R the object's rotation (object to world);
 * the boundary value at a point is a(t) . normal.  Static objects only. */
typedef struct fdtd_acc_noise_source {
    int                object;             /* index into fdtd_scene.objects */
    int                n_impacts;
    double             mass;               /* m (kg)                        */
    double             center_of_mass[3];  /* object frame (m)              */
    double             inv_inertia[9];     /* I^-1, object frame, row-major */
    const fdtd_impact *impacts;            /* n_impacts (the caller owns them) */
} fdtd_acc_noise_source;

/* GaussianPressureSource: exp(-|x-x0|^2/(2 ws^2)) exp(-((t-t0)/wt)^2/2) K/3.6836,
 * zero outside the open box |x - x0|_inf < 2 ws. */
typedef struct fdtd_pressure_source {
    double position[3];
    double width_time;          /* source_width_time wt                  */
    double width_space;         /* source_width_space ws (default c*wt)  */
    double offset_time;         /* source_offset_time t0                 */
    double normalize_constant;  /* K (default 1/(sqrt(2 pi) ws)^3)       */
} fdtd_pressure_source;

/* Speaker pressure source (SpeakerPressureSource.cpp:91-113, static
 * positioning): S(x, t) = exp(-|x - p|^2 / (2 w^2)) * samples[idx] with
 * idx = (int)((t - |x - p| / c - start_time) * sample period^-1), zero outside
 * [0, n_samples); added as S dt to every non-PML fluid cell like the Gaussian
 * source.  samples: channel 0 of the speaker file, scaled to [-1, 1) as
 * libsndfile reads integer PCM. */
typedef struct fdtd_speaker_source {
    double        position[3];   /* position_x/y/z                          */
    double        width_space;   /* width_space w (m)                       */
    double        start_time;    /* start_time (s)                          */
    double        sample_rate;   /* samples per second of the file          */
    int           n_samples;
    int           reserved;
    const double *samples;       /* n_samples values (the caller owns them) */
} fdtd_speaker_source;

/* Plane constraint (FDTD_PlaneConstraint, FDTD_PlaneConstraint.h:9-45; XML
 * <constraint direction sign height>, ImpulseResponseParser.cpp:434-449): the
 * half-space (x[direction] - height) * sign < 0 is solid.  A cell whose
 * centre lies in it is solid (centre test only, MAC_Grid.cpp:3566-3577), its
 * pressure is held at zero, and a fluid neighbour's ghost value is the
 * neighbour's own pressure (a perfectly reflecting wall, MAC_Grid.cpp:
 * 1198-1204).  Collocated scheme with a PML only. */
typedef struct fdtd_plane_constraint {
    int    direction;   /* normal axis 0, 1 or 2 (XML default 1)              */
    int    sign;        /* +1 or -1 (XML default 1)                           */
    double height;      /* position of the plane along that axis (m), XML 0  */
} fdtd_plane_constraint;

/* Keyframe animation of one object from a kinematics file
 * (FDTD_RigidObject_Blender_Animator, FDTD_RigidObject_Animator.cpp:78-160;
 * XML file_kinematics + kinematics_frame_rate, ImpulseResponseParser.cpp:
 * 248-263).  Frame f holds the object's centre and rotation quaternion as the
 * file writes them (tx ty tz qw qx qy qz; a Blender z-up rotation, which is
 * turned by -90 degrees about x, q <- q * Rot(x, -90), with qw clamped to
 * [-1, 1] first).  At time t the frames f0 = min(max(0, (int)(t / step)),
 * n - 1) and f1 = min(f0 + 1, n - 1) are blended with alpha = (t - f0 step) /
 * step (1 when f0 = f1): the centre linearly, the rotation by the reference's
 * slerp; the pose replaces the object's own centre and rotation (its
 * velocities must be zero).  Collocated scheme, rasterized ghost cells. */
typedef struct fdtd_kinematics {
    int           object;     /* index into fdtd_scene.objects              */
    int           n_frames;
    double        step_size;  /* 1 / kinematics_frame_rate (s)              */
    const double *frames;     /* n_frames x 7 (the caller owns them)        */
} fdtd_kinematics;

typedef struct fdtd_scene {
    fdtd_settings                settings;
    int                          n_objects;
    const fdtd_object           *objects;
    int                          n_impulse_sources;
    const fdtd_impulse_source   *impulse_sources;
    int                          n_pressure_sources;
    const fdtd_pressure_source  *pressure_sources;
    int                          n_listeners;
    const double                *listeners;          /* [n_listeners][3] */
    int                          n_speaker_sources;
    const fdtd_speaker_source   *speaker_sources;
    int                          n_constraints;
    const fdtd_plane_constraint *constraints;
    int                          n_acc_noise_sources;
    const fdtd_acc_noise_source *acc_noise_sources;
    int                          n_kinematics;
    const fdtd_kinematics       *kinematics;
} fdtd_scene;

/* Fills every field of a settings block with the parser defaults
 * (src/parser/ImpulseResponseParser.cpp:506-556). */
void fdtd_settings_default(fdtd_settings *s);

/* Parsed XML scene (owns the arrays that fdtd_parsed_scene points at). */
typedef struct fdtd_parsed fdtd_parsed;
fdtd_status        fdtd_parse_xml(const char *path, fdtd_parsed **out);
const fdtd_scene  *fdtd_parsed_scene(const fdtd_parsed *p);
int                fdtd_parsed_num_steps(const fdtd_parsed *p); /* number_of_timesteps */
/* solver output_pattern (ImpulseResponseParser.cpp:521), "%s" when absent. */
const char        *fdtd_parsed_output_pattern(const fdtd_parsed *p);
void               fdtd_parsed_free(fdtd_parsed *p);
/* Error text of the last failed call on this thread (parse errors name the
 * element and attribute). */
const char        *fdtd_last_error(void);

/* ------------------------------------------------------------------------- */
/* single scene                                                              */
/* ------------------------------------------------------------------------- */
typedef struct fdtd_handle fdtd_handle;

/* How a PML cell of the collocated scheme keeps its state.  SPLIT is the
 * reference's state (MAC_Grid.cpp:383-508): three split pressures pDir_d and
 * the face velocities of every PML cell.  MERGED (default) gives the same
 * pressures up to FP32 rounding with fewer bytes: a cell damped along one
 * axis A only keeps that axis's split and faces plus the increment d of the
 * sum q of the two undamped splits, since with f1 = f3 = 1 the split updates
 * of those axes reduce exactly to
 *   d^{n+1} = d^n + sum_b f5/h_b (f2_b(R) dp_R - f2_b(L) dp_L),
 *   p^{n+1} = p^n + (pDir_A^{n+1} - pDir_A^n) + d^{n+1}
 * (32 B per such cell and step instead of 56).  Snapshots need SPLIT (they
 * write every pDir_d and velocity). */
enum {
    FDTD_PML_STATE_MERGED = 0,
    FDTD_PML_STATE_SPLIT  = 1
};

typedef struct fdtd_options {
    void *stream;            /* hipStream_t to run on; NULL: handle-owned stream */
    int   device;            /* HIP device ordinal (-1: current)               */
    int   steps_per_graph;   /* >0: capture blocks of this many steps in a hipGraph */
    int   series_capacity;   /* device rows of the listener series (0: 1024)   */
    int   tile_config;       /* kernel tile id from the autotuner (-1: default) */
    int   pml_state;         /* FDTD_PML_STATE_*: how PML cells keep their state   */
} fdtd_options;

void fdtd_options_default(fdtd_options *o);

typedef struct fdtd_info {
    int64_t n_cells;           /* N = nx*ny*nz                                 */
    int64_t n_pml_cells;       /* cells inside the PML                         */
    int64_t n_ghost_cells;     /* ghost-cell entries (owner, fluid neighbour)  */
    int64_t n_solid_cells;     /* cells occupied by objects                    */
    int64_t n_listeners;
    int64_t step_index;        /* steps taken since fdtd_create                */
    double  time;              /* simulation time (s): start + dt added per step */
    double  bytes_per_step;    /* algorithmic HBM bytes: 12 N + 44 N_pml      */
    int64_t device_bytes;      /* device memory held by the handle             */
    int64_t n_pml_merged_cells;   /* PML cells in the merged state (pml_state)  */
    double  state_bytes_per_step; /* bytes the state representation moves per
                                     step: bytes_per_step - 24 per merged cell */
} fdtd_info;

fdtd_status fdtd_create(const fdtd_scene *scene, const fdtd_options *opt, fdtd_handle **out);
fdtd_status fdtd_create_from_xml(const char *path, const fdtd_options *opt, fdtd_handle **out);
fdtd_status fdtd_destroy(fdtd_handle *h);
/* Advances nsteps time steps.  Asynchronous on the handle's stream. */
fdtd_status fdtd_step(fdtd_handle *h, int nsteps);
fdtd_status fdtd_synchronize(fdtd_handle *h);
/* Listener pressures after the newest step (FP64, n_listeners values). */
fdtd_status fdtd_listen(fdtd_handle *h, double *out);
/* Listener time series rows [first_step, first_step + nsteps): FP64 [nsteps][L].
 * Row s holds the listener pressures after step s + 1. */
fdtd_status fdtd_read_listener_series(fdtd_handle *h, int64_t first_step, int64_t nsteps, double *out);
/* Current pressure field (FP64 copy of the FP32 field, z fastest).  With the
 * Engquist boundary (a scene without PML cells) the outermost cell layer is
 * never updated (the reference skips it too) and holds boundary state instead
 * of pressures. */
fdtd_status fdtd_read_pressure(fdtd_handle *h, double *out);
fdtd_status fdtd_get_info(fdtd_handle *h, fdtd_info *out);
/* Device timing of the next fdtd_step calls with HIP events on the handle's
 * stream (timing-only events without a system-scope fence).  PER_STEP
 * brackets each step's kernels (stencil_ms: their mean duration; the event
 * records add a few microseconds between kernels); SPAN brackets the call
 * only (step_ms: device time per step, nothing recorded between kernels;
 * stencil_ms 0).  Both set step_ms.  1 = PER_STEP for older callers. */
enum {
    FDTD_TIMING_OFF      = 0,
    FDTD_TIMING_PER_STEP = 1,
    FDTD_TIMING_SPAN     = 2
};
fdtd_status fdtd_set_timing(fdtd_handle *h, int mode);
fdtd_status fdtd_last_kernel_ms(fdtd_handle *h, double *stencil_ms, double *step_ms);

/* ------------------------------------------------------------------------- */
/* batches of independent scenes (one launch per step for the whole batch)    */
/* ------------------------------------------------------------------------- */
typedef struct fdtd_batch fdtd_batch;

typedef struct fdtd_batch_options {
    void *stream;            /* NULL: batch-owned stream                          */
    int   steps_per_graph;   /* hipGraph step block (0: 32)                       */
    int   tile_config;       /* -1: autotune once, cached per process             */
} fdtd_batch_options;

void fdtd_batch_options_default(fdtd_batch_options *o);
fdtd_status fdtd_batch_create(const fdtd_scene *scenes, const int *num_steps, int n_scenes,
                              const fdtd_batch_options *opt, fdtd_batch **out);
/* Runs step blocks until every scene has taken its num_steps steps. */
fdtd_status fdtd_batch_run(fdtd_batch *b);
/* Advances every unfinished scene by up to nsteps steps. */
fdtd_status fdtd_batch_step(fdtd_batch *b, int nsteps);
fdtd_status fdtd_batch_steps_done(fdtd_batch *b, int scene, int64_t *steps);
fdtd_status fdtd_batch_read_listener_series(fdtd_batch *b, int scene, double *out); /* [num_steps][L] */
fdtd_status fdtd_batch_destroy(fdtd_batch *b);

/* ------------------------------------------------------------------------- */
/* time-parallel chunks (SimWorld.cpp:141-221, 474-478, 556-590)              */
/* ------------------------------------------------------------------------- */
/* Settings of chunk `chunk` of a num_steps run split in settings->
 * num_time_chunks chunks, as SimWorld sets them up for its chunk index
 * (SimWorld.cpp:151-166): S = num_steps / K steps per chunk; the chunk starts
 * at start_time + S * chunk steps and keeps its sources on for S steps (both
 * times by repeated addition of time_step), and runs S + overlap_time /
 * time_step steps.  first_step gets the chunk's first step in the whole run,
 * chunk_steps its step count.  Host only. */
fdtd_status fdtd_time_chunk(const fdtd_settings *settings, int64_t num_steps, int chunk, fdtd_settings *out,
                            int64_t *first_step, int64_t *chunk_steps);
/* Runs every chunk of a time-parallel scene (settings.time_parallel = 1) as
 * one batch on the device (the reference runs one process per chunk) and sums
 * the chunks' listener series into the run's series by linear superposition:
 * out [num_steps][L], row r = listener pressures after step r + 1 of the
 * whole run.  A chunk stops early, after its sources are off, once its
 * listeners' peak pressure stays below early_termination_ratio times the
 * chunk's running peak (and below 2e-5 Pa) for 0.05 s of simulated time
 * (SimWorld.cpp:556-590).  steps_run (NULL or num_time_chunks entries) gets
 * the steps each chunk ran. */
fdtd_status fdtd_run_time_parallel(const fdtd_scene *scene, int64_t num_steps, const fdtd_batch_options *opt,
                                   double *out, int64_t *steps_run);

/* ------------------------------------------------------------------------- */
/* serving: continuous batching of impact-sound solve requests                */
/* ------------------------------------------------------------------------- */
typedef struct fdtd_server fdtd_server;

typedef struct fdtd_server_options {
    int64_t hbm_budget_bytes;  /* admission budget for resident request state     */
    int     max_active;        /* slot cap (0: no cap beyond the budget)          */
    int     steps_per_block;   /* steps advanced per batched launch block (0: 64) */
    int     device;            /* -1: current                                     */
} fdtd_server_options;

typedef struct fdtd_server_stats {
    int64_t submitted, completed, failed, cancelled, active, queued;
    int64_t blocks_run;
    int64_t peak_resident_bytes;
} fdtd_server_stats;

void fdtd_server_options_default(fdtd_server_options *o);
fdtd_status fdtd_server_create(const fdtd_server_options *opt, fdtd_server **out);
/* Queues a request (thread-safe).  The scene is copied. */
fdtd_status fdtd_server_submit(fdtd_server *s, const fdtd_scene *scene, int num_steps, int64_t *request_id);
/* FDTD_OK when finished (result ready), FDTD_ERR_BUSY while queued/running,
 * or the request's error status. */
fdtd_status fdtd_server_poll(fdtd_server *s, int64_t request_id);
/* Blocks until the request finishes, or timeout_ms passes (<0: no timeout). */
fdtd_status fdtd_server_wait(fdtd_server *s, int64_t request_id, int timeout_ms);
/* Copies the FP64 [num_steps][L] listener series of a finished request and
 * releases it; elapsed_ms gets submit-to-finish wall time. */
fdtd_status fdtd_server_take_result(fdtd_server *s, int64_t request_id, double *out, double *elapsed_ms);
fdtd_status fdtd_server_cancel(fdtd_server *s, int64_t request_id);
fdtd_status fdtd_server_get_stats(fdtd_server *s, fdtd_server_stats *out);
/* Bytes of device state a scene needs when admitted (used by admission). */
int64_t     fdtd_scene_device_bytes(const fdtd_scene *scene);
/* ------------------------------------------------------------------------- */
/* the reference tool's files                                                */
/* ------------------------------------------------------------------------- */
/* Output name as FDTD_AcousticSimulator::CompositeFilename
 * (FDTD_AcousticSimulator.cpp:110-124): pattern's %s replaced by
 * "<simulator_id>_<name>" (just <name> when simulator_id is NULL or empty). */
fdtd_status fdtd_composite_filename(const char *pattern, const char *simulator_id, const char *name,
                                    char *out, int capacity);
/* Binary matrix as IO::writeMatrixX(..., BINARY) (src/utils/IO/IO.h:107-139):
 * int32 rows, int32 cols, then rows*cols doubles row by row; the reader is
 * scripts/binary_io.py:readMatrixXdBinary. */
fdtd_status fdtd_write_matrix(const char *filename, const double *data, int rows, int cols);
/* Runs an XML scene the way the reference's tool does
 * (FDTD_AcousticSimulator::InitializeSolver / SaveSolverConfig / Run and
 * SimWorld's audio output): writes solver_settings.txt, a copy of the control
 * file (input_control_file.xml) and listening_position.dat under the scene's
 * output_pattern, steps until stop_time (or number_of_timesteps), and writes
 * all_audio.dat as one row of L doubles per step (AudioOutput.cpp:55-64,
 * read by scripts/wavesolver_results.py).  A time_parallel scene runs all
 * its chunks as one batch and writes, per chunk c, what the reference's chunk
 * process writes (SimWorld.cpp:354-405, 870-911): %05d_all_audio.dat,
 * %05d_start_time, %05d_time_range and %05d_sim-done; steps_run then counts
 * the steps of all chunks. */
fdtd_status fdtd_run_xml(const char *xml_path, const char *simulator_id, const fdtd_options *opt, int64_t *steps_run);
/* Snapshot of the time-stepping state in the file set of
 * FDTD_AcousticSimulator::_SaveSimulationSnapshot (FDTD_AcousticSimulator.cpp:
 * 381-415), each file a MATRIX::write matrix (int32 rows, int32 cols, doubles):
 * snapshot_<n>_p_collocated_{0,1,2}.dat, _p_collocated_ind.dat (text),
 * _p_pml_{0,1,2}.dat (split pressures, full grid), _v_pml_{0,1,2}.dat (face
 * velocities, (n_d+1) along axis d), _p_pml_full.dat, plus _step.txt with the
 * step index and _engine_state.bin, this engine's whole state region (bit-
 * exact resume of the same scene in a handle of the same PML state).  Names
 * follow fdtd_composite_filename.  A handle with the merged PML state keeps,
 * per merged cell, pDir_A and the increment of the sum q of its transverse
 * splits: its p_pml files hold pDir_A, q on the lower transverse axis and
 * zero on the other, and its v_pml files zero on faces between two merged
 * cells (velocities that state does not keep). */
fdtd_status fdtd_save_snapshot(fdtd_handle *h, const char *pattern, const char *simulator_id, int numbering);
/* Restores a snapshot into a handle created for the same scene (the
 * reference declares LoadSolverResult, FDTD_AcousticSimulator.h:108, without
 * a definition): from the engine state file when it matches the handle (bit
 * for bit), else from the reference's files -- exact for a split handle; a
 * merged handle converts the split state (pDir_A, and the transverse
 * increment sum_b f5/h_b (v_b(high) - v_b(low)) of the stored velocities).  A
 * split handle refuses files written by a merged one; air viscosity needs the
 * engine state file (the reference's files lack the previous Laplacian).
 * Listener rows before the loaded step read as zero. */
fdtd_status fdtd_load_snapshot(fdtd_handle *h, const char *pattern, const char *simulator_id, int numbering);

/* Host-only dry run of the device planner (no counterpart in the reference,
 * which has no tiling): builds the scene's tiles for tile_config (-1: default)
 * and the launch lists of its step, and checks on the host that every tile's
 * class matches its launch, every read stays inside the grid and the launch
 * lists carry every tile exactly once.  No device is touched.  n_tiles (may
 * be NULL) gets the number of workgroups per step. */
fdtd_status fdtd_plan_check(const fdtd_scene *scene, int tile_config, int64_t *n_tiles);
fdtd_status fdtd_server_destroy(fdtd_server *s);

/* ------------------------------------------------------------------------- */
/* no device allocation                                                      */
/* ------------------------------------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the work of fdtd_create (default options) +
 * fdtd_step(num_steps) + fdtd_read_listener_series + fdtd_destroy, with the
 * same kernels and the same results bit for bit, calling no hipMalloc,
 * hipFree or other allocating routine.
 *   fields   device memory (256-byte aligned) for the time-stepping state,
 *            zeroed here and holding the final state (engine layout) on
 *            return; NULL only queries its size into *fields_bytes_needed
 *            (host work only, no device call)
 *   stream   hipStream_t to run on (NULL: the null stream); the call waits
 *            for its work before it returns
 *   series   host buffer [num_steps][n_listeners] (FP64)
 * Everything else -- tables, listener term ring, amplitude ring, counters,
 * tile lists -- lives in a static device array of 64 MiB per GPU, held until
 * the process ends; a scene that needs more scratch returns
 * FDTD_ERR_UNSUPPORTED before any device work.  Calls on one GPU share that
 * array: the library serialises its own calls per device, and calls that
 * overlap on one GPU from elsewhere would race on it. */
fdtd_status fdtd_solve_noalloc(const fdtd_scene *scene, int num_steps, void *fields, int64_t fields_bytes,
                               void *stream, double *series, int64_t *fields_bytes_needed);

#ifdef __cplusplus
}
#endif

#endif /* FDTD_ACOUSTICS_H */
