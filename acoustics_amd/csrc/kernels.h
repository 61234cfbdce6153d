// kernels.h -- device-side scene blocks, work lists and launch entry points.
//
// Every launch works from a list of tiles; a tile belongs to one scene and is
// homogeneous (interior leapfrog, PML slab, or a run of object-adjacent
// cells), so a wavefront never mixes code paths.  Scene parameters live in
// device memory (SceneDev) and are read with uniform loads; the step index of
// each scene is a device counter, so a captured graph can be replayed at any
// step and a batch can hold scenes at different steps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.h"

namespace fdtd {

struct PSrcDev {
    double t0, wt, amp;        // amp = K / 3.6836
    int lo[3], hi[3];          // index box where the spatial factor is nonzero
    int goff[3];               // offsets into gtab
    int pad;
};

// vibrational source with a normal acceleration vector (FDTD_VIB_*)
struct ImpDev {
    int obj, kind;
    double A, w, ton;                              // Gaussian
    double omega, phase, decay, dcenter, t0, t1;   // harmonic; chirp: omega, t0, t1
    double omega1;                                 // chirp
};

// acceleration-noise impact: S(t) v over [t0, t0 + tau), world frame
struct ImpactDev {
    int obj, pad;
    double t0, tau, v[3];
};

// rasterised ghost value p[nbr] + k_raster a_s(obj) + k_vec a_v(obj)[vaxis]
struct GhostDev {
    int nbr, obj, coupled, vaxis;
    double k_raster, k_vec;
};

// pressure slots inside SceneDev::field, each npad floats
enum : int { NUM_P_SLOTS = 2 };

// Pressure slots of step s: p^s is read from slot_cur, and p^{s+1} is written
// over p^{s-1} in slot_next (the leapfrog reads p^{s-1} only at the cell it
// writes, so two slots suffice).
__host__ __device__ inline int slot_cur(long long s) { return (int)((s + 1) & 1); }
__host__ __device__ inline int slot_next(long long s) { return (int)(s & 1); }

// clock words of a scene; CLK_TIME holds the bits of the simulation time of
// step clock[CLK_STEP], accumulated from the start time by adding dt once per
// step, as the reference's clock runs (_simulationTime += dt)
enum : int { CLK_STEP = 0, CLK_TARGET = 1, CLK_DONE = 2, CLK_TIME = 3, CLK_WORDS = 4 };

// the clock words of step s of a scene starting at t_start (host side)
inline void clock_words(unsigned long long *clk, long long s, unsigned long long target, double t_start, double dt)
{
    double t = t_start;
    for (long long q = 0; q < s; ++q) t += dt;
    unsigned long long tb;
    static_assert(sizeof tb == sizeof t, "time word");
    __builtin_memcpy(&tb, &t, sizeof t);
    clk[CLK_STEP] = (unsigned long long)s;
    clk[CLK_TARGET] = target;
    clk[CLK_DONE] = 0ULL;
    clk[CLK_TIME] = tb;
}

// Steps are launched in blocks: the kernels of the k-th step of a block run
// step clock[CLK_STEP] + k, and one advance kernel after the block moves the
// counter on.  The per-step amplitudes live in a ring of AMP_ROWS rows (step
// s in row s & (AMP_ROWS - 1)) that the advance kernel refills for the next
// AMP_ROWS steps, so a block is at most AMP_ROWS steps long.
enum : int { AMP_ROWS = 64 };
__host__ __device__ inline int amp_row(long long s) { return (int)(s & (AMP_ROWS - 1)); }

// a moving object: its pose at t = 0 and its constant velocity and angular velocity
// Motion of one object: constant velocity v and angular velocity w from
// the pose g0 at t = 0, or (kin_n > 0) keyframes kin_off .. kin_off + kin_n
// of SceneDev::kin, kin_dt apart, with g0 giving the shape
struct MoveDev {
    Geom g0;
    double v[3], w[3];
    int kin_off, kin_n;
    double kin_dt;
};

struct SceneDev {
    int n[3];
    int sx, sy;                 // strides of i and j (k fastest)
    int raster;                 // 1: rasterized ghost values, 0: fully coupled
    int scheme;                 // FDTD_SCHEME_*
    long long npad;             // floats per pressure slot
    float *field;               // NUM_P_SLOTS pressure slots, full grid, k fastest
    // PML state in compact per-slab storage (PmlSlab): split pressures per
    // PML cell, face velocities per axis, double-buffered by step parity
    const PmlSlab *slabs;
    float *pd[3];
    float *vel[2][3];           // collocated: compact PML faces; staggered: every face, full grid
                                // (face d of cell c, on its low side, at c's index)
    const float4 *coef[3];      // per axis, index 0..n: {f1(face), f2(face), f3(cell), f5/h(cell)}
    const int *flags[3];        // per axis, index 0..n: 1 cell in PML, 2 face in PML, 4 sigma zero, 8 plain, 16 behind a plane constraint
    const uint8_t *skip;        // per cell: 1 = object cell (solid or ghost-adjacent)
    double lam2, t_start, dt;
    float lam2f;
    double lam;                 // c dt / h (the Engquist boundary's coefficients)
    // air viscosity (MAC_Grid.cpp:926-930): the bulk cells' Laplacian of the
    // previous step, per cell (null without viscosity), and the coefficients
    // c alpha dt, (c dt)^2, 1/h^2 in FP64 and FP32
    float *lapv;
    double kcalp, c2k2, inv_h2;
    float kcalpf, c2k2_kcalpf, inv_h2f;
    int n_psrc, n_isrc, n_special, n_impact;
    const PSrcDev *psrc;
    const float *gtab;
    const ImpDev *isrc;
    const ImpactDev *impacts;
    const SpecialRec *recs;
    const GhostDev *ghosts;
    // fully-coupled mode: the coupled system's solution operator, per ghost g
    // x_g = sum_q op_w[q] p[op_idx[q]] (q in [op_ptr[g], op_ptr[g+1]))
    //     + sum_q oc_w[q] acc[oc_ch[q]]  (q in [oc_ptr[g], oc_ptr[g+1]))
    int G;
    const int32_t *op_ptr, *op_idx, *oc_ptr, *oc_ch;
    const double *op_w, *oc_w;
    // listener series: row s of the ring holds beta_q * sample_q of every MLS
    // term q of every listener after step s; a listener's value is the sum
    // of its terms, formed when the rows are read out
    int T, cap;                 // terms per row, rows of the ring
    double *terms;              // [cap][T]
    unsigned long long *clock;  // CLK_WORDS words
    int n_obj;
    // 1: the seven-sample configuration (ghosts at 3h/4 from the fluid cell,
    // non-uniform Laplacian weights, 0.75 h); 0: centre sampling (ghosts on
    // the face, plain weights, h)
    int seven;
    int n_cpo, n_ch;            // amplitude channels per object (1, or 4 with acceleration noise), n_obj n_cpo
    // time-dependent amplitudes, one ring row per step (amp_row):
    double *t_tab;              // [AMP_ROWS] the accumulated time of the step (CLK_TIME)
    double *acc_tab;            // [AMP_ROWS][n_ch] channel o n_cpo: a_s(o, t_s); o n_cpo + 1 + d: a_v(o, t_s)[d]
    float *src_tab;             // [AMP_ROWS][n_psrc] S(t_s + dt/2) dt without the spatial factor
    long long src_steps;        // time-parallel chunk: sources act in steps < src_steps (-1: always)
    // speaker sources: per source {px, py, pz, w, t0, period, off, n} and the samples
    int n_spk;
    const double *spk;          // [n_spk][8]
    const double *spk_samples;
    double c;                   // sound speed
    int sp_lo[3], sp_n[3];      // box of cells within 8 widths of any speaker
    float *spk_field;           // [sp_n0 sp_n1 sp_n2]: the speakers' S dt of the step being run
    // moving objects: the box of cells reclassified every step, their bulk
    // flags by step parity (row s & 1 holds step s's classification), the
    // objects' motion and a ring of their poses (row amp_row(s) holds the
    // poses at t_s; the advance kernel refills it with the amplitudes)
    int mv;
    int u_lo[3], u_n[3];
    uint8_t *ucls;              // [2][u_n0 u_n1 u_n2], z fastest: 1 bulk, 0 solid
    const MoveDev *mobj;        // [n_obj]
    const double *kin;          // keyframes, 8 doubles each
    Geom *pose;                 // [AMP_ROWS][n_obj]
    double *ughost;             // [6][u cells]: ghost term of each solid neighbour of a fluid cell (-x, +x, ..., +z)
    double h, rho, bmin[3];
};

// A listener term handled by the tile that produces its sample: kind 0 is
// the new pressure of flat cell `cell`; kind 1 the value of ghost entry
// `cell` (raster: p^n of its fluid neighbour plus the acceleration term;
// fully coupled: the solved ghost value).  `slot` is the term's column in
// the scene's series ring.
struct TermRef {
    int scene, slot, kind, cell;
    double beta;
};

enum : int {
    TILE_INTERIOR = 0, TILE_INTERIOR_SKIP = 1, TILE_PML = 2, TILE_SPECIAL = 3, TILE_EMPTY = 4, TILE_INTERIOR_VEC = 5,
    TILE_MAC = 6,      // staggered scheme, cells outside the PML
    TILE_MAC_PML = 7,  // staggered scheme, a PML slab (split pressures in the slab's compact storage)
    TILE_ABC = 8,      // Engquist boundary: a box of layer-1 cells, one thread per cell (k fastest)
    TILE_MOVING = 9,   // a box (at most 256 cells) of a moving scene's reclassified cells, one thread per cell
    TILE_MAC_VEC = 10, // staggered scheme, cells outside the PML, whole rows with four cells per lane (as TILE_INTERIOR_VEC)
};

struct TileDesc {
    int scene;
    int type;
    int i0, i1, j0, j1, k0, k1;  // TILE_SPECIAL: [i0, i1) is a range of special records; TILE_ABC: at most 256 cells
    int tk_log2;                 // TK = 1 << tk_log2, TJ = 256 / TK (TILE_INTERIOR_VEC: whole rows, see rpw)
    int src;                     // tile intersects a pressure-source box
    int slab;                    // TILE_PML: index into SceneDev::slabs
    int tb, te;                  // listener terms [tb, te) of the program's TermRef table
    int rpw;                     // TILE_INTERIOR_VEC: rows per wavefront (the tile holds 4 * rpw rows)
    int lead;                    // TILE_INTERIOR_VEC: cells at the start of each row's first group (from k0,
                                 // a multiple of four) that belong to another tile and are not written
};

struct MotionDesc {
    int scene;
};

// KIND_ALL: one launch carrying every tile class (dispatch per workgroup on
// the tile type); the other kinds carry one class each.
enum : int { KIND_INTERIOR = 0, KIND_PML = 1, KIND_SPECIAL = 2, NUM_KINDS = 3, KIND_ALL = 3 };
// order_clock: when non-null, the step word of a scene clock whose parity
// (plus k) selects the tile order -- odd steps walk each XCD's chunk of the
// tile list backwards, so a step starts on the data the previous step touched
// last and the Infinity Cache still holds.  k: the step's offset in its block.
// nhead: the first nhead tiles (a multiple of eight) keep their place on odd
// steps too.
hipError_t launch_step(int kind, const SceneDev *scenes, const TileDesc *tiles, int ntiles, int nhead, const TermRef *terms,
                       const unsigned long long *order_clock, int k, hipStream_t s);
// Moves every scene of the launch n steps on (clamped to its target) and
// refills its amplitude and pose rings for the AMP_ROWS steps from there;
// n = 0 only fills the rings.
hipError_t launch_advance(const SceneDev *scenes, int nscenes, int n, hipStream_t s);
// Moving objects, before the step kernels of step clock + k, in one launch:
// per-step reclassification of each scene's box, zeroed new solid cells,
// filled fresh cells and the ghost terms of the box's fluid cells; six
// threads per cell along x (max_cells: the largest box of the list), one
// scene per y.
hipError_t launch_motion(const SceneDev *scenes, const MotionDesc *list, int n, int max_cells, int k, hipStream_t s);
// Speaker sources, before the step kernels of step clock + k: the speakers'
// S dt over each listed scene's speaker box.
hipError_t launch_speaker(const SceneDev *scenes, const MotionDesc *list, int n, int max_cells, int k, hipStream_t s);

}  // namespace fdtd
