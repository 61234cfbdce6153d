/*
 * fdtd_oracle.h -- FP64 CPU restatement of the reference wavesolver's default
 * (USE_COLLOCATED) time-stepping path, and of its fully staggered MAC
 * leapfrog for free-field scenes (settings.scheme = FDTD_SCHEME_STAGGERED).
 *
 * TEST INFRASTRUCTURE.  This library is the checker: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it.  The product (acoustics_amd/) never links or calls it.
 *
 * It follows, in FP64 and in the reference's order of operations:
 *   src/wavesolver/MAC_Grid.cpp          90-150, 187-299, 461-946, 1131-1441,
 *                                        2397-2693, 3455-3891, 4027-4102,
 *                                        4310-4340, 4444-4510
 *   src/wavesolver/MAC_Grid.h            97-145, 528-571
 *   src/wavesolver/PML_WaveSolver.cpp    102-146, 210-373, 579-609, 775-826
 *   src/field/ScalarField.cpp            29-140, 164-286
 *   src/wavesolver/GaussianPressureSource.{h,cpp}
 *   src/wavesolver/FDTD_Objects.cpp      35-49, 224-257, 356-367
 *   src/wavesolver/FDTD_RigidObject.cpp  256-285, 330-368, 429-536 (semantics;
 *                                        analytic sphere/ellipsoid geometry)
 *   src/wavesolver/HarmonicVibrationalSource.cpp          5-28
 *   src/wavesolver/AccelerationNoiseVibrationalSource.cpp 19-90
 *   src/modal_model/ImpulseSeriesObject.cpp               88-101
 *   src/wavesolver/FDTD_RigidObject_Animator.cpp          78-160
 *   src/wavesolver/FDTD_RigidObject.cpp                   550-566
 *   src/linearalgebra/Quaternion.hpp                      82-89, 186-208, 365-420
 *   src/math/MLSModeInterpolator.hpp     282-429, 588-597, 624-660, 712-738
 * and, for the staggered scheme:
 *   src/wavesolver/PML_WaveSolver.cpp    710-773 (stepLeapfrog)
 *   src/wavesolver/MAC_Grid.cpp          301-370, 948-1020, 1102-1129
 *   src/wavesolver/GaussianPressureSource.cpp 24-27 (erf time factor)
 *   src/linearalgebra/MATRIX_FAST.cpp    555-569 (pFull = pDir0 + pDir1 + pDir2)
 * Third-party arithmetic the reference takes from Eigen 3.3.7 (JacobiSVD,
 * BiCGSTAB) is restated here from the published algorithms (one-sided
 * Jacobi SVD with Eigen's rank threshold; BiCGSTAB with a diagonal
 * preconditioner).
 */
#ifndef FDTD_ORACLE_H
#define FDTD_ORACLE_H

#include "../include/fdtd_acoustics.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

/* reclassify_each_step: 1 re-runs the per-step rasterisation (classifyCells_FAST)
 * every step as the reference does; 0 classifies once (static scenes give the
 * same result either way). */
oracle *oracle_create(const fdtd_scene *scene, int reclassify_each_step);
void    oracle_free(oracle *o);
/* Advances nsteps steps; rows (optional) receives [nsteps][L] listener values. */
int     oracle_step(oracle *o, int nsteps, double *rows);
void    oracle_get_pressure(const oracle *o, double *out);          /* newest p, N */
/* Overwrites p^{n-1} and p^n (N values each) before the next step. */
void    oracle_set_state(oracle *o, const double *p_last, const double *p_curr);
/* p^{n-1} and p^n (N values each; either may be NULL). */
void    oracle_get_state(const oracle *o, double *p_last, double *p_curr);
/* Split pressures pDir_d (3 N) and face velocities (axis d: (n_d + 1) faces
 * along d, index ((i ny' + j) nz' + k), one axis after another). */
void    oracle_get_pml_state(const oracle *o, double *pdir /*3N*/, double *v /*3 faces*/);
/* counts: [0] N, [1] N_pml cells, [2] N_pml faces, [3] ghost entries,
 *         [4] solid cells, [5] bulk cells, [6] cavity cells,
 *         [7] ghosts on the fully-coupled path, [8] step index */
void    oracle_counts(const oracle *o, long long *counts);
double  oracle_time(const oracle *o);
const char *oracle_error(void);

/* --- unit-level entry points used by the oracle's own pinning tests --- */
/* One-sided Jacobi SVD of an m x n row-major matrix (m >= n): singular values
 * in decreasing order. */
int     oracle_svd_values(int m, int n, const double *A, double *sv);
/* Min-norm least squares solve (pseudo-inverse with Eigen's JacobiSVD rank
 * threshold); returns the rank. */
int     oracle_lsq_solve(int m, int n, const double *A, const double *b, double *x);
/* Order-1 MLS value at p from npts points (x,y,z) and values. */
double  oracle_mls_lookup(const double *p, int npts, const double *pts, const double *vals);
/* Pose (centre, row-major world-from-object rotation) at time t of an object
 * animated by the keyframes k. */
void    oracle_keyframe_pose(const fdtd_kinematics *k, double t, double *center, double *rotation);
/* Boundary acceleration vector of object obj at time t at a surface point
 * with outward normal nrm (the sum over its vibrational sources). */
void    oracle_object_accel(const oracle *o, int obj, double t, const double *nrm, double *out);
/* Signed distance, closest boundary point and outward unit normal there. */
double  oracle_object_closest(const fdtd_object *obj, const double *x, double *xb, double *n);
/* ReflectAgainstBoundary semantics: image point, boundary point, erected
 * normal; returns the signed distance travelled. */
double  oracle_reflect(const fdtd_object *obj, const double *x, double *image, double *xb, double *erected);
/* Fully-coupled Vandermonde weights of MAC_Grid::PML_pressureUpdateGhostCells:
 * rows[8][3] sample positions, row s_i (or -1) is the Neumann row at
 * boundary point bp with normal nrm; returns cond and writes w[8]. */
double  oracle_vandermonde_weights(const double *rows, int s_i, const double *bp, const double *nrm,
                                   const double *image, const double *origin, double h, double *w);
/* PML coefficient helpers (MAC_Grid.h:528-571) at a position. */
double  oracle_pml_sigma(const fdtd_settings *s, const double *x, int dim);
int     oracle_inside_pml(const fdtd_settings *s, const double *x);

/* Ghost-cell entries of the current classification. */
int     oracle_ghost_count(const oracle *o);
/* Test hook of PML_pressureUpdateGhostCells: the ghost values the update
 * gives for pressure field p (N values) at time t (val), with each entry's
 * position (pos, 3 per entry) and fluid neighbour (nbr); any output may be
 * NULL.  Returns the number of entries. */
int     oracle_ghost_update_at(oracle *o, const double *p, double t, double *val, double *pos, long long *nbr);
/* The last assembled ghost system A x = rhs (CSR, duplicates summed): sizes
 * into n and nnz, then the arrays when non-NULL; -1 before any update. */
int     oracle_ghost_system(const oracle *o, int *n, int *nnz, int *rp, int *ci, double *vv, double *rhs);

/* Pose at time t of an object moving with its velocity and angular velocity
 * (test hook of the per-step object motion). */
void oracle_object_pose(const fdtd_object *obj, double t, fdtd_object *out);

#ifdef __cplusplus
}
#endif
#endif
