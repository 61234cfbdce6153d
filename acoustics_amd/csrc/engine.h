// engine.h -- internal structures of the MI355X FDTD engine.
//
// Host side: SceneSetup is everything that is static for a scene (PML
// coefficient tables, rasterised objects, ghost-cell stencils, listener
// interpolation weights), precomputed once in FP64.  Device side: DeviceScene
// holds the FP32 fields and the uploaded tables; StepParams is the POD handed
// to the fused step kernel.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fdtd_acoustics.h"
#include "geom.h"

namespace fdtd {

// --------------------------------------------------------------------------
// geometry (analytic rigid objects; stand-in for FDTD_RigidObject queries)
// --------------------------------------------------------------------------
enum : int { MAX_MOVING_OBJ = 16 };   // objects of a scene with moving objects
Geom make_geom(const fdtd_object &o);
void geom_aabb(const Geom &g, double *mn, double *mx);

// --------------------------------------------------------------------------
// host precompute
// --------------------------------------------------------------------------
enum : int { DIR_XM = 0, DIR_XP = 1, DIR_YM = 2, DIR_YP = 3, DIR_ZM = 4, DIR_ZP = 5 };

// The boundary acceleration of object o at a boundary point with outward
// normal n is a_s(o, t) n + a_v(o, t): a_s sums the sources whose vector is
// a normal one (Gaussian impulse, harmonic), a_v the acceleration-noise
// vector.  The device keeps them as amplitude channels, per object
// [a_s] or, in a scene with acceleration noise, [a_s, a_v.x, a_v.y, a_v.z].
struct GhostEntry {
    int owner;        // solid cell (flat index)
    int nbr;          // bulk neighbour (flat index)
    int topology;     // +-1, +-2, +-3 (direction owner -> nbr)
    int obj;          // closest object
    double pos[3];    // owner centre + h/4 towards nbr
    // rasterised value p_g = p[nbr] - 0.75 h grad_p . cell_n with
    // grad_p = -rho (a_s n_b + a_v):  p[nbr] + k_raster a_s + k_vec a_v[|topology| - 1]
    double k_raster;  // 0.75 h rho (n_b . cell_n)
    double k_vec;     // 0.75 h rho sign(topology)
    // fully-coupled row: x_g = sum_k w_k p[k] + c_acc a_s + c_vec . a_v + sum_u w_u x_u
    int coupled;      // 1: coupled row, 0: rasterised value
    double c_acc;
    double c_vec[3];
    int type;         // 0: object ghost, 1: plane-constraint ghost (p_g = p[nbr])
};

struct SpecialRec {      // bulk cell with at least one ghost neighbour
    int32_t cell;
    int32_t flags;       // bit0: one-cell cavity (p^{n+1} = p^n)
    int32_t gid[6];      // ghost entry per direction (-1: regular neighbour)
};

struct SpecialBox {      // index box holding every special cell of one object
    int32_t lo[3], hi[3];  // inclusive lower, exclusive upper
    int32_t offset;        // into the concatenated class array
    int32_t pad;
};

struct PSrcHost {
    double t0, wt, amp;    // amp = K / 3.6836 (collocated), K (staggered)
    int lo[3], hi[3];      // index range with a nonzero spatial factor
    int goff[3];           // offsets of the 1D factor tables in gtab
};

struct ImpulseHost {
    int obj, kind;                 // FDTD_VIB_*
    double A, w, ton;              // Gaussian
    double omega, phase, decay, dcenter, t0, t1;   // harmonic (t1 <= t0: no range); chirp: omega, t0, t1
    double omega1;                                 // chirp
};

// One acceleration-noise impact: a_v(t) += S(t) v over t0 <= t < t0 + tau,
// S = sin(pi (t - t0) / tau), v = R pi / (2 tau) (J / m + (I^-1 (r x J)) x r)
// in the world frame.
struct ImpactHost {
    int obj;
    double t0, tau, v[3];
};

struct ListenerTerm {
    int kind;      // 0: FP32 field cell, 1: ghost entry
    int index;
    double beta;
};

// One box of PML cells with its own compact storage: split pressures at
// cell_base + (li * e1 + lj) * e2 + lk (local indices), and for each axis d
// face velocities in a box one layer longer along d, at v_base[d] + the same
// formula with e[d] + 1 in place of e[d].  Face li along d is the face on the
// low side of local cell li; the extra layer holds the high faces of the last
// layer of cells.  A face shared by two boxes is stored in both and given the
// same update by each.
//
// Merged transverse state (merge = 1, collocated scheme only): a cell of the
// box that is plain along both axes other than the box axis A (no damping on
// those axes: f1 = f3 = 1 on its faces and at the cell, no wall faces) keeps
// only the last increment d^n = q^n - q^{n-1} of the sum q of its two
// transverse split pressures, instead of the two splits and their four face
// velocities.  With f1 = f3 = 1 the split updates give exactly
//   d^{n+1} = d^n + sum_{b transverse} f5h_b (f2_b(R) (p_R - p) - f2_b(L) (p - p_L)),
//   q^{n+1} = q^n + d^{n+1},   p^{n+1} = pDir_A^{n+1} + q^{n+1},
// (d is f5h times the transverse velocity divergence), so q need not be kept:
//   p^{n+1} = p^n + (pDir_A^{n+1} - pDir_A^n) + d^{n+1}.
// Only increments are rounded before the final addition, so rounding stays of
// order eps |p| per step, integrated once.  (Forming q^n = p^n - pDir_A^n
// rounds at the scale of pDir_A, which deep in the layer is much larger than
// p; keeping q^{n-1} and forming 2 q^n - q^{n-1} integrates rounding twice and
// lets it drift.)  d lives in the split array of the lower transverse axis.  The box-axis split and its faces are kept as
// before.  Transverse faces of a merged cell are updated only where a
// non-merged cell of the same box reads them.
struct PmlSlab {
    int32_t lo[3], e[3];
    int32_t axis;                // the box's damped axis (x slabs 0, y slabs 1, z slabs 2)
    int32_t merge;               // 1: merged transverse state for eligible cells
    int64_t cell_base;
    int64_t v_base[3];
};

struct SceneSetup {
    int n[3] = {0, 0, 0};
    int64_t N = 0;
    double h = 0, dt = 0, c = 0, rho = 0, t_start = 0, stop_time = 0;
    double lam2 = 0;             // (c dt / h)^2
    bool visc = false;           // air viscosity: the bulk update keeps the previous step's Laplacian
    // time-parallel chunk: every source acts only in steps s < src_steps
    // (-1: always), the first step whose start time, counted by repeated
    // addition of dt, reaches stop_boundary_acc_time (SimWorld.cpp:474)
    int64_t src_steps = -1;
    double kcalp = 0, c2k2 = 0;  // c alpha dt and (c dt)^2
    int boundary = 0;
    int scheme = 0;                 // FDTD_SCHEME_*
    // FDTD_SAMPLE_*: CENTER classifies by the cell centre, puts a
    // rasterised ghost on the face (a fully-coupled one at the solid
    // centre) with p_g = p_nbr - h grad_p . n_cell, and takes ghost values
    // with the 7-point weights; SEVEN adds the six points h/4 from the
    // centre, puts ghosts h/4 from it with 0.75 h and the 3h/4 weights
    int sampling = FDTD_SAMPLE_CENTER;
    // distance of a ghost from its solid cell's centre, in cells, and the
    // factor of h in the rasterised ghost value
    double ghost_offset() const { return sampling == FDTD_SAMPLE_SEVEN ? 0.25 : (boundary == FDTD_BOUNDARY_RASTERIZE ? 0.5 : 0.0); }
    double raster_factor() const { return sampling == FDTD_SAMPLE_SEVEN ? 0.75 : 1.0; }
    // PML, separable per axis
    std::vector<uint8_t> cpml[3];   // cell predicate, n_d entries
    std::vector<uint8_t> fpml[3];   // face predicate, n_d + 1 entries (walls included)
    std::vector<float> f1[3], f2[3];   // face tables (n_d + 1)
    std::vector<float> f3[3], f5h[3];  // cell tables (n_d), f5h = f5 / h
    std::vector<uint8_t> sig0[3];      // cell: sigma_d <= 1e-12 (staggered source goes into pDir_d)
    int inner_lo[3], inner_hi[3];      // non-PML index box (with the Engquist boundary: [2, n - 2))
    // plane constraints; per axis and index, 1 where a constraint holds the
    // cells' centres (solid whatever the other coordinates); the interior
    // tiles cover [tile_lo, tile_hi), the non-PML box without the solid
    // layers and the fluid layer next to them
    std::vector<fdtd_plane_constraint> cons;
    std::vector<uint8_t> csolid[3];
    int tile_lo[3], tile_hi[3];
    bool has_solid = false;            // objects or constraints: solid cells, ghosts, a skip mask
    // Engquist boundary (no PML cell anywhere: use_mesh = 0, or a PML width
    // of zero): the layer-1 cells take the absorbing update and layer 0 is
    // never updated (MAC_Grid.cpp:521-919)
    bool abc = false;
    int64_t n_pml = 0;
    std::vector<PmlSlab> slabs;        // the PML as up to six disjoint boxes
    int64_t n_pml_faces[3] = {0, 0, 0};  // compact face-velocity entries per axis
    int64_t n_pml_merged = 0;           // PML cells holding merged transverse state
    // sources
    std::vector<PSrcHost> psrc;
    std::vector<float> gtab;
    std::vector<ImpulseHost> isrc;
    std::vector<ImpactHost> impacts;
    int n_cpo = 1;                     // amplitude channels per object: 1, or 4 with acceleration noise
    // speaker sources (collocated scheme): parameters and samples, and the
    // index box within 8 widths of each, where tiles evaluate them
    struct SpeakerHost { double p[3], w, t0, period; int off, n; int lo[3], hi[3]; };
    std::vector<SpeakerHost> spk;
    std::vector<double> spk_samples;
    int spk_lo[3] = {0, 0, 0}, spk_hi[3] = {0, 0, 0};   // union of the speakers' boxes
    // objects
    std::vector<Geom> objs;             // poses at t = 0
    // moving objects (some object has a velocity or angular velocity): the
    // cells any object can reach up to stop_time, with a two-cell margin,
    // form the box [mv_lo, mv_hi); the device reclassifies them every step
    // (MAC_Grid::classifyCells_FAST) and fills fresh cells
    // (InterpolateFreshPressureCell).  mv_cls0: bulk flags of the box at the
    // start pose, z fastest.
    double bmin[3] = {0, 0, 0};         // minimum corner of the pressure grid
    bool moving = false;
    int mv_lo[3] = {0, 0, 0}, mv_hi[3] = {0, 0, 0};
    std::vector<double> mv_v, mv_w;     // 3 per object
    // keyframe animation (kinematics files): per object the offset and count
    // of its frames in kin (0 frames: constant motion) and the frame step;
    // kin holds 8 doubles per frame, centre and corrected rotation (w x y z)
    std::vector<double> kin;
    std::vector<int32_t> kin_off, kin_n;
    std::vector<double> kin_dt;
    int64_t mv_steps = 0;               // steps the planned path covers (start to stop_time)
    std::vector<uint8_t> mv_cls0;
    std::vector<uint8_t> bulk;          // full grid (host only)
    std::vector<GhostEntry> ghosts;     // sorted by (owner, key)
    std::vector<SpecialRec> special;
    std::vector<SpecialBox> boxes;
    std::vector<int32_t> cls;           // per box cell: -1 regular, -2 solid, >=0 special record
    int64_t n_solid = 0;
    int64_t n_cavity = 0;
    // coupled ghost system (CSR over ghosts): known-cell terms and unknown-ghost terms
    std::vector<int32_t> gk_ptr, gk_idx;  std::vector<double> gk_w;
    std::vector<int32_t> gu_ptr, gu_idx;  std::vector<double> gu_w;
    // the system's solution operator x = M p + m a (fully-coupled scenes):
    // per ghost, cells and weights of M (CSR, dense cell indices) and
    // amplitude channels (object * n_cpo + component) and weights of m;
    // op_max_cycle is the largest set of mutually coupled
    // ghosts (1: the coupling graph is acyclic)
    std::vector<int32_t> op_ptr, op_idx;  std::vector<double> op_w;
    std::vector<int32_t> oc_ptr, oc_ch;  std::vector<double> oc_w;
    int op_max_cycle = 0;
    int64_t n_coupled = 0;
    // listeners
    int L = 0;
    std::vector<double> listeners;
    std::vector<int32_t> lst_ptr;
    std::vector<ListenerTerm> lst_terms;
};

// Builds the full setup; returns FDTD_OK or an error with a message in err.
fdtd_status build_setup(const fdtd_scene &scene, SceneSetup &out, std::string &err);
// The solution operator of the fully-coupled ghost system (op_* and oc_*).
fdtd_status build_coupled_operator(SceneSetup &S, std::string &err);
// Axis d is plain at cell i: no damping on the cell or its two faces
// (f3 = f1 = 1 in the FP32 tables) and neither face is a wall.
bool axis_plain(const SceneSetup &s, int d, int i);
// Chooses the PML state representation: merged (the default of build_setup
// for the collocated scheme) or the reference's per-axis split state, which
// snapshots need.  Sets PmlSlab::merge and n_pml_merged.
void set_pml_state(SceneSetup &s, bool merged);
// Engquist boundary state on the device: a face cell of layer 1 reads
// p^{n-1} of its inner neighbour (two cells in from the exterior cell),
// which the interior update overwrites in place during the same step, so the
// face cell keeps a copy in the slot p^{n+1} goes to, at its exterior cell
// (that slot's exterior value, p^{n-1} there, is read first).  Between steps
// the exterior cell of every face cell therefore holds p^{n-1} of the inner
// neighbour in the current slot and the extrapolated exterior pressure in the
// previous slot.  Writes that copy into cur from prev (full grid, dense
// indices) for a state loaded from elsewhere; a no-op without the boundary.
void abc_pack_state(const SceneSetup &s, const double *prev, double *cur);
// Row stride of the device grid: n2 rounded up to a multiple of four.
inline int row_stride(const SceneSetup &s) { return (s.n[2] + 3) & ~3; }
// Algorithmic bytes per step (FP32 fields): collocated 12 N + 44 N_pml;
// staggered 32 N + 24 N_pml.
double algorithmic_bytes(const SceneSetup &s);
// Bytes per step of this engine's state representation: a merged PML cell
// moves 32 B instead of 56 (it drops two splits and four face velocities,
// read and written, and reads and writes d).
double state_bytes(const SceneSetup &s);

// thread-local error text for fdtd_last_error
void set_last_error(const std::string &msg);

}  // namespace fdtd
