// setup.cpp -- FP64 host precompute of everything static in a scene.
//
// What the reference recomputes every step (classifyCells_FAST, the ghost
// reflection and Vandermonde solves, the MLS listener fit) depends only on
// the geometry when the objects are static, so it is done once here and the
// per-step device work reduces to gathers and dot products:
//   * PML: MAC_Grid::classifyCells + InsidePML + PML_absorptionCoefficient
//     (src/wavesolver/MAC_Grid.cpp:2397-2693, 4027-4102; coefficient forms
//     MAC_Grid.h:528-571).  sigma_d depends only on coordinate d, so the
//     f1/f2 (faces) and f3/f5 (cells) coefficients are 1D tables per axis.
//   * objects: classifyCells_FAST's 7-sample occupancy rule and MakeGhostCell
//     (MAC_Grid.cpp:3455-3891, 4310-4340), LowestObjectDistance
//     (FDTD_Objects.cpp:240-257).
//   * ghost values: PML_pressureUpdateGhostCells (MAC_Grid.cpp:1131-1441):
//     rasterize -> p_g = p[nbr] + k * a_n(t); fully_coupled -> a static
//     sparse row x_g = sum w_k p_k + c * a_n(t) + sum w_u x_u.
//   * listeners: FetchScalarData + MLS lookup (PML_WaveSolver.cpp:210-317,
//     MLSModeInterpolator.hpp:282-429): the order-1 MLS value is linear in
//     the sample values, c0 = sum beta_i v_i, with beta from the pseudo-inverse
//     of the weighted design matrix.
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "engine.h"
#include "linalg.h"

namespace fdtd {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
const std::string &get_last_error() { return g_last_error; }

namespace {

constexpr double kPmlBuffer = 1e-12;      // MAC_Grid.cpp:4031
constexpr double kCheckBound = 2.0;       // GAUSSIAN_CHECK_BOUND
constexpr double kCondLimit = 25.0;       // CONDN_LIM
constexpr double kMlsRadius = 0.018;      // MLS_FLEISHMAN_RAD
constexpr double kRefEps = 1e-10;         // EPS, config.h

// ScalarField::enclosingNeighbours (src/field/ScalarField.cpp:261-286)
int enclosing_cells(const SceneSetup &S, const double *bmin, const double *p, int64_t *out)
{
    int a[3], b[3];
    for (int d = 0; d < 3; ++d) {
        double diff = p[d] - bmin[d];
        diff -= S.h / 2.0;
        int v0 = (int)(diff / S.h);
        v0 = std::min(std::max(v0, 0), S.n[d] - 1);
        a[d] = v0;
        b[d] = diff < 0.0 ? 0 : std::min(v0 + 1, S.n[d] - 1);
    }
    int64_t tmp[8];
    int m = 0;
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
            for (int z = 0; z < 2; ++z)
                tmp[m++] = ((int64_t)(x ? b[0] : a[0]) * S.n[1] + (y ? b[1] : a[1])) * S.n[2] + (z ? b[2] : a[2]);
    std::sort(tmp, tmp + 8);
    int u = 0;
    for (int i = 0; i < 8; ++i)
        if (u == 0 || tmp[i] != out[u - 1]) out[u++] = tmp[i];
    return u;
}

inline uint64_t pair_key(int64_t a_, int64_t b_)
{
    const uint64_t a = (uint64_t)std::min(a_, b_), b = (uint64_t)std::max(a_, b_);
    return a >= b ? a * a + a + b : a + b * b;
}

}  // namespace

double algorithmic_bytes(const SceneSetup &s)
{
    // collocated: interior cell reads p^n, p^{n-1}, writes p^{n+1}; PML cell adds
    // v[3] and pDir[3] read and written.  staggered: every cell reads and
    // writes v[3] and reads p^n, writes p^{n+1}; PML cells add pDir[3].
    if (s.scheme == FDTD_SCHEME_STAGGERED) return 32.0 * (double)s.N + 24.0 * (double)s.n_pml;
    return 12.0 * (double)s.N + 44.0 * (double)s.n_pml;
}

double state_bytes(const SceneSetup &s)
{
    return algorithmic_bytes(s) - 24.0 * (double)s.n_pml_merged;
}

void abc_pack_state(const SceneSetup &s, const double *prev, double *cur)
{
    if (!s.abc) return;
    const int *n = s.n;
    auto flat = [&](int i, int j, int k) { return ((int64_t)i * n[1] + j) * n[2] + k; };
    // exterior cell e of the face cell one step in along axis d, and the inner
    // neighbour two steps in; face cells have both other coordinates in [2, n - 3]
    for (int d = 0; d < 3; ++d) {
        const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
        for (int side = 0; side < 2; ++side)
            for (int u = 2; u < n[e1] - 2; ++u)
                for (int v = 2; v < n[e2] - 2; ++v) {
                    int ext[3], in[3];
                    ext[d] = side ? n[d] - 1 : 0;
                    in[d] = side ? n[d] - 3 : 2;
                    ext[e1] = in[e1] = u;
                    ext[e2] = in[e2] = v;
                    cur[flat(ext[0], ext[1], ext[2])] = prev[flat(in[0], in[1], in[2])];
                }
    }
}

bool axis_plain(const SceneSetup &s, int d, int i)
{
    const int n = s.n[d];
    if (i < 1 || i > n - 2) return false;
    return s.f3[d][i] == 1.0f && s.f1[d][i] == 1.0f && s.f1[d][i + 1] == 1.0f;
}

void set_pml_state(SceneSetup &s, bool merged)
{
    s.n_pml_merged = 0;
    for (PmlSlab &sl : s.slabs) {
        sl.merge = 0;
        if (!merged || s.scheme != FDTD_SCHEME_COLLOCATED) continue;
        // the march runs along x: y and z slabs merge only when x is plain
        // over their whole x range, so a column is merged or split throughout
        bool ok = true;
        if (sl.axis != 0)
            for (int i = sl.lo[0]; i < sl.lo[0] + sl.e[0] && ok; ++i) ok = axis_plain(s, 0, i);
        if (!ok) continue;
        sl.merge = 1;
        int64_t ny = 0, nz = 0;
        for (int j = sl.lo[1]; j < sl.lo[1] + sl.e[1]; ++j) ny += (sl.axis == 1 || axis_plain(s, 1, j));
        for (int k = sl.lo[2]; k < sl.lo[2] + sl.e[2]; ++k) nz += (sl.axis == 2 || axis_plain(s, 2, k));
        s.n_pml_merged += (int64_t)sl.e[0] * ny * nz;
    }
}

namespace {

// a sparse row: (column, weight), sorted by column, no repeats
using SparseRow = std::vector<std::pair<int32_t, double>>;

// sum_q f_q R_q of sparse rows
SparseRow combine(const std::vector<std::pair<double, const SparseRow *>> &terms)
{
    std::vector<std::pair<int32_t, double>> all;
    for (const auto &t : terms)
        for (const auto &e : *t.second) all.push_back({e.first, t.first * e.second});
    std::sort(all.begin(), all.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    SparseRow out;
    for (const auto &e : all) {
        if (!out.empty() && out.back().first == e.first) out.back().second += e.second;
        else out.push_back(e);
    }
    return out;
}

// Strongly connected components of the digraph g -> u (W_gu != 0), in the
// order Tarjan's algorithm closes them: every component after all the
// components it reaches (iterative, no recursion depth limit).
std::vector<std::vector<int>> coupling_components(const SceneSetup &S)
{
    const int G = (int)S.ghosts.size();
    std::vector<int> index(G, -1), low(G, 0), stack, comp_of(G, -1);
    std::vector<char> on(G, 0);
    std::vector<std::vector<int>> comps;
    int counter = 0;
    std::vector<std::pair<int, int>> call;   // (vertex, next edge)
    for (int root = 0; root < G; ++root) {
        if (index[root] >= 0) continue;
        call.push_back({root, S.gu_ptr[root]});
        index[root] = low[root] = counter++;
        stack.push_back(root); on[root] = 1;
        while (!call.empty()) {
            int v = call.back().first;
            int &e = call.back().second;
            if (e < S.gu_ptr[v + 1]) {
                const int w = S.gu_idx[e++];
                if (index[w] < 0) {
                    index[w] = low[w] = counter++;
                    stack.push_back(w); on[w] = 1;
                    call.push_back({w, S.gu_ptr[w]});
                } else if (on[w]) {
                    low[v] = std::min(low[v], index[w]);
                }
                continue;
            }
            if (low[v] == index[v]) {
                std::vector<int> c;
                for (;;) {
                    const int w = stack.back(); stack.pop_back(); on[w] = 0;
                    c.push_back(w);
                    if (w == v) break;
                }
                std::sort(c.begin(), c.end());
                comps.push_back(std::move(c));
            }
            call.pop_back();
            if (!call.empty()) { const int p = call.back().first; low[p] = std::min(low[p], low[v]); }
        }
    }
    return comps;
}

}  // namespace

// Rows of M (cells) and m (amplitude channels) of the coupled ghost operator
// x = M p + m a, from the coupled rows (gk: known cells, gu: unknown ghosts,
// c_acc a_s + c_vec . a_v) and the rasterized rows (p[nbr] + k_raster a_s +
// k_vec a_v along the ghost's axis).
fdtd_status build_coupled_operator(SceneSetup &S, std::string &err)
{
    const int G = (int)S.ghosts.size();
    std::vector<SparseRow> K(G), C(G), M(G), Mo(G);
    for (int g = 0; g < G; ++g) {
        const GhostEntry &e = S.ghosts[g];
        if (e.coupled) {
            std::vector<std::pair<double, const SparseRow *>> none;
            SparseRow r;
            for (int q = S.gk_ptr[g]; q < S.gk_ptr[g + 1]; ++q) r.push_back({S.gk_idx[q], S.gk_w[q]});
            K[g] = combine({{1.0, &r}});
            C[g] = {{e.obj * S.n_cpo, e.c_acc}};
            if (S.n_cpo > 1)
                for (int d = 0; d < 3; ++d)
                    if (e.c_vec[d] != 0.0) C[g].push_back({e.obj * S.n_cpo + 1 + d, e.c_vec[d]});
        } else {
            K[g] = {{e.nbr, 1.0}};
            C[g] = {{e.obj * S.n_cpo, e.k_raster}};
            if (S.n_cpo > 1 && e.k_vec != 0.0) C[g].push_back({e.obj * S.n_cpo + 1 + std::abs(e.topology) - 1, e.k_vec});
        }
    }
    const std::vector<std::vector<int>> comps = coupling_components(S);
    std::vector<int> comp_of(G, -1);
    for (size_t c = 0; c < comps.size(); ++c)
        for (int g : comps[c]) comp_of[g] = (int)c;
    for (size_t c = 0; c < comps.size(); ++c) {
        const std::vector<int> &cg = comps[c];
        const int m = (int)cg.size();
        // B_g = K_g + sum over ghosts u outside the component of W_gu M_u
        std::vector<SparseRow> BM(m), BO(m);
        std::vector<double> A((size_t)m * m, 0.0);   // I - W restricted to the component
        for (int a = 0; a < m; ++a) {
            const int g = cg[a];
            A[(size_t)a * m + a] = 1.0;
            std::vector<std::pair<double, const SparseRow *>> tm = {{1.0, &K[g]}}, to = {{1.0, &C[g]}};
            for (int q = S.gu_ptr[g]; q < S.gu_ptr[g + 1]; ++q) {
                const int u = S.gu_idx[q];
                if (comp_of[u] == (int)c) {
                    const int b = (int)(std::lower_bound(cg.begin(), cg.end(), u) - cg.begin());
                    A[(size_t)a * m + b] -= S.gu_w[q];
                } else {
                    tm.push_back({S.gu_w[q], &M[u]});
                    to.push_back({S.gu_w[q], &Mo[u]});
                }
            }
            BM[a] = combine(tm);
            BO[a] = combine(to);
        }
        if (m == 1) {
            const double d = A[0];
            if (!(std::fabs(d) > 1e-12)) { err = "fully_coupled ghost system is singular (a ghost couples to itself with weight 1)"; return FDTD_ERR_UNSUPPORTED; }
            M[cg[0]] = combine({{1.0 / d, &BM[0]}});
            Mo[cg[0]] = combine({{1.0 / d, &BO[0]}});
            continue;
        }
        if (m > 512) {
            err = "fully_coupled ghost system: a cycle of " + std::to_string(m) + " mutually coupled ghosts (the direct solve "
                  "takes cycles of at most 512)";
            return FDTD_ERR_UNSUPPORTED;
        }
        S.op_max_cycle = std::max(S.op_max_cycle, m);
        // inverse of A by Gauss-Jordan elimination with partial pivoting
        std::vector<double> Inv((size_t)m * m, 0.0);
        for (int a = 0; a < m; ++a) Inv[(size_t)a * m + a] = 1.0;
        for (int col = 0; col < m; ++col) {
            int piv = col;
            for (int r = col + 1; r < m; ++r)
                if (std::fabs(A[(size_t)r * m + col]) > std::fabs(A[(size_t)piv * m + col])) piv = r;
            if (!(std::fabs(A[(size_t)piv * m + col]) > 1e-12)) { err = "fully_coupled ghost system is singular"; return FDTD_ERR_UNSUPPORTED; }
            if (piv != col)
                for (int k = 0; k < m; ++k) {
                    std::swap(A[(size_t)piv * m + k], A[(size_t)col * m + k]);
                    std::swap(Inv[(size_t)piv * m + k], Inv[(size_t)col * m + k]);
                }
            const double inv_p = 1.0 / A[(size_t)col * m + col];
            for (int k = 0; k < m; ++k) { A[(size_t)col * m + k] *= inv_p; Inv[(size_t)col * m + k] *= inv_p; }
            for (int r = 0; r < m; ++r) {
                if (r == col) continue;
                const double f = A[(size_t)r * m + col];
                if (f == 0.0) continue;
                for (int k = 0; k < m; ++k) { A[(size_t)r * m + k] -= f * A[(size_t)col * m + k]; Inv[(size_t)r * m + k] -= f * Inv[(size_t)col * m + k]; }
            }
        }
        for (int a = 0; a < m; ++a) {
            std::vector<std::pair<double, const SparseRow *>> tm, to;
            for (int b = 0; b < m; ++b) {
                const double f = Inv[(size_t)a * m + b];
                if (f != 0.0) { tm.push_back({f, &BM[b]}); to.push_back({f, &BO[b]}); }
            }
            M[cg[a]] = combine(tm);
            Mo[cg[a]] = combine(to);
        }
    }
    for (int g = 0; g < G; ++g) {
        for (const auto &e : M[g]) if (e.second != 0.0) { S.op_idx.push_back(e.first); S.op_w.push_back(e.second); }
        S.op_ptr.push_back((int32_t)S.op_idx.size());
        for (const auto &e : Mo[g]) if (e.second != 0.0) { S.oc_ch.push_back(e.first); S.oc_w.push_back(e.second); }
        S.oc_ptr.push_back((int32_t)S.oc_ch.size());
    }
    return FDTD_OK;
}

fdtd_status build_setup(const fdtd_scene &sc, SceneSetup &S, std::string &err)
{
    const fdtd_settings &st = sc.settings;
    for (int d = 0; d < 3; ++d) {
        if (st.divisions[d] < 8) { err = "grid divisions must be >= 8 on every axis"; return FDTD_ERR_INVALID; }
        S.n[d] = st.divisions[d];
    }
    if (!(st.cell_size > 0) || !(st.time_step > 0) || !(st.sound_speed > 0) || !(st.density > 0)) {
        err = "cell_size, time_step, sound_speed and density must be positive";
        return FDTD_ERR_INVALID;
    }
    if (!st.use_mesh && sc.n_objects > 0) {
        // MAC_Grid::classifyCells returns before any object is rasterised
        err = "use_mesh=0 classifies no object cells; objects need use_mesh=1";
        return FDTD_ERR_UNSUPPORTED;
    }
    if (!(st.air_viscosity_alpha >= 0.0)) { err = "air_viscosity_alpha must be >= 0"; return FDTD_ERR_INVALID; }
    if (st.boundary_handling != FDTD_BOUNDARY_RASTERIZE && st.boundary_handling != FDTD_BOUNDARY_FULLY_COUPLED) {
        err = "boundary_handling must be rasterize or fully_coupled";
        return FDTD_ERR_INVALID;
    }
    if (sc.n_objects < 0 || sc.n_objects > 256 || sc.n_listeners < 0 || sc.n_pressure_sources < 0 || sc.n_impulse_sources < 0) {
        err = "bad object/source/listener counts";
        return FDTD_ERR_INVALID;
    }
    const int64_t N64 = (int64_t)S.n[0] * S.n[1] * S.n[2];
    if (N64 > (int64_t)INT32_MAX - 1) { err = "grid too large for 32-bit cell indices"; return FDTD_ERR_INVALID; }
    S.N = N64;
    S.h = st.cell_size; S.dt = st.time_step; S.c = st.sound_speed; S.rho = st.density;
    S.t_start = st.start_time; S.stop_time = st.stop_time;
    S.lam2 = (S.c * S.dt / S.h) * (S.c * S.dt / S.h);
    // ImpulseResponseParser.cpp:540-549: alpha > 0 turns the viscous bulk
    // update on; the staggered scheme (stepLeapfrog) has no viscous term
    S.visc = st.air_viscosity_alpha > 0.0 && st.scheme == FDTD_SCHEME_COLLOCATED;
    S.kcalp = S.c * st.air_viscosity_alpha * S.dt;   // MAC_Grid.cpp:486
    S.c2k2 = (S.c * S.dt) * (S.c * S.dt);
    S.boundary = st.boundary_handling;
    S.scheme = st.scheme;
    if (st.sampling != FDTD_SAMPLE_CENTER && st.sampling != FDTD_SAMPLE_SEVEN) {
        err = "sampling must be FDTD_SAMPLE_CENTER or FDTD_SAMPLE_SEVEN";
        return FDTD_ERR_INVALID;
    }
    S.sampling = st.sampling;
    S.src_steps = -1;
    if (st.time_parallel) {
        // SimWorld::UpdateObjectState switches every source off once the
        // clock (start time plus dt, added step by step) reaches the stop time
        if (!std::isfinite(st.stop_boundary_acc_time)) { err = "time_parallel needs a finite stop_boundary_acc_time"; return FDTD_ERR_INVALID; }
        double t = st.start_time;
        int64_t k = 0;
        while (!(t >= st.stop_boundary_acc_time)) {
            t += st.time_step;
            if (++k > (int64_t)INT32_MAX) { err = "stop_boundary_acc_time lies more than 2^31 steps after the start"; return FDTD_ERR_INVALID; }
        }
        S.src_steps = k;
    }
    if (st.scheme != FDTD_SCHEME_COLLOCATED && st.scheme != FDTD_SCHEME_STAGGERED) { err = "unknown scheme"; return FDTD_ERR_INVALID; }
    if (sc.n_constraints < 0 || (sc.n_constraints > 0 && !sc.constraints)) { err = "bad constraint list"; return FDTD_ERR_INVALID; }
    S.cons.assign(sc.constraints, sc.constraints + sc.n_constraints);
    for (const fdtd_plane_constraint &c : S.cons)
        if (c.direction < 0 || c.direction > 2 || (c.sign != 1 && c.sign != -1) || !std::isfinite(c.height)) {
            err = "a plane constraint needs direction 0, 1 or 2, sign +1 or -1 and a finite height";
            return FDTD_ERR_INVALID;
        }
    if (!S.cons.empty() && (st.scheme != FDTD_SCHEME_COLLOCATED || !st.use_mesh)) {
        err = "plane constraints are classified with use_mesh = 1 in the collocated scheme";
        return FDTD_ERR_UNSUPPORTED;
    }
    if (st.scheme == FDTD_SCHEME_STAGGERED && (sc.n_objects > 0 || sc.n_impulse_sources > 0 || sc.n_acc_noise_sources > 0)) {
        err = "the staggered scheme runs free-field scenes only (the reference's staggered object path mixes two "
              "cell masks and a stubbed ghost lookup); use the collocated scheme, the reference's default";
        return FDTD_ERR_UNSUPPORTED;
    }
    const double h = S.h, dt = S.dt, rho = S.rho, c = S.c;
    const double W = st.pml_width * h;         // MAC_Grid::setPMLBoundaryWidth
    const double Sg = st.pml_strength;
    const int preset = st.boundary_condition_preset;

    double bmin[3], bmax[3];
    for (int d = 0; d < 3; ++d) { bmin[d] = st.min_bound[d]; bmax[d] = bmin[d]; bmax[d] += h * (double)S.n[d]; }
    for (int d = 0; d < 3; ++d) S.bmin[d] = bmin[d];

    // ---- PML tables -------------------------------------------------------
    auto inside_axis = [&](int d, double x) {
        const double pw = W + kPmlBuffer, hMin = x - bmin[d], hMax = bmax[d] - x;
        switch (preset) {
        case 0: return hMin <= pw || hMax <= pw;
        case 1: return hMin <= pw || (d == 1 && hMax <= pw);
        case 2: return d == 2 && hMax <= pw;
        default: return false;
        }
    };
    auto sigma_axis = [&](int d, double x) {
        const double hMin = x - bmin[d], hMax = bmax[d] - x, a2 = W * W;
        const double dMin = W - hMin, dMax = W - hMax;
        switch (preset) {
        case 0:
            if (hMin <= W) return Sg * std::pow(dMin, 2) / a2;
            if (hMax <= W) return Sg * std::pow(dMax, 2) / a2;
            return 0.0;
        case 1:
            if (hMin <= W) return Sg * std::pow(dMin, 2) / a2;
            if (d == 1 && hMax <= W) return Sg * std::pow(dMax, 2) / a2;
            return 0.0;
        case 2:
            if (d != 2 || hMin <= W) return 0.0;
            if (hMax <= W) return Sg * std::pow(dMax, 2) / a2;
            return 0.0;
        default: return 0.0;
        }
    };
    for (int d = 0; d < 3; ++d) {
        const int n = S.n[d];
        S.cpml[d].assign(n, 0); S.f3[d].assign(n, 0.f); S.f5h[d].assign(n, 0.f); S.sig0[d].assign(n, 1);
        S.fpml[d].assign(n + 1, 0); S.f1[d].assign(n + 1, 0.f); S.f2[d].assign(n + 1, 0.f);
        for (int i = 0; i < n; ++i) {
            double x = bmin[d];
            x += (0.5 + (double)i) * h;
            S.cpml[d][i] = st.use_mesh && inside_axis(d, x);
            const double sg = sigma_axis(d, x);
            S.sig0[d][i] = sg > 1e-12 ? 0 : 1;   // MAC_Grid.cpp:966-967 inPML test
            const double D = 1.0 / dt + sg / 2.0;
            S.f3[d][i] = (float)((1.0 / dt - sg / 2.0) / D);
            S.f5h[d][i] = (float)((-1.0 * rho * c * c / D) * (1.0 / h));
        }
        const double vmin = bmin[d] - h / 2.0;
        for (int i = 0; i <= n; ++i) {
            double x = vmin;
            x += (0.5 + (double)i) * h;
            S.fpml[d][i] = st.use_mesh && inside_axis(d, x);
            const double sg = sigma_axis(d, x);
            double f1 = 1.0 / dt - sg / 2.0;
            f1 /= 1.0 / dt + sg / 2.0;
            double f2 = 1.0 / dt + sg / 2.0;
            f2 *= rho * h;
            S.f1[d][i] = (float)f1;
            S.f2[d][i] = (float)(-1.0 / f2);
        }
    }
    // no PML cell at all: the Engquist boundary on layer 1 (MAC_Grid.cpp:521-919)
    S.abc = true;
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < S.n[d]; ++i) S.abc = S.abc && !S.cpml[d][i];
    if (S.abc && st.scheme != FDTD_SCHEME_COLLOCATED) {
        err = "the Engquist boundary (no PML) runs with the collocated scheme; give the staggered scheme a PML";
        return FDTD_ERR_UNSUPPORTED;
    }
    for (int d = 0; d < 3 && !S.abc; ++d) {
        const int n = S.n[d];
        // non-PML cells along d must form one interval
        int lo = -1, hi = -1;
        for (int i = 0; i < n; ++i)
            if (!S.cpml[d][i]) { if (lo < 0) lo = i; hi = i + 1; }
        if (lo < 0) { err = "PML covers the whole grid"; return FDTD_ERR_INVALID; }
        for (int i = lo; i < hi; ++i)
            if (S.cpml[d][i]) { err = "non-contiguous PML interior"; return FDTD_ERR_UNSUPPORTED; }
        if (lo < 2 || hi > n - 2) {
            err = "the PML must cover the two outermost cell layers on every face it damps (PML_width >= 2 with "
                  "boundary preset 0), or no face at all (the Engquist boundary)";
            return FDTD_ERR_UNSUPPORTED;
        }
        S.inner_lo[d] = lo; S.inner_hi[d] = hi;
        // a PML face between two non-PML cells makes classifyCells throw
        for (int i = 1; i < n; ++i)
            if (S.fpml[d][i] && !S.cpml[d][i - 1] && !S.cpml[d][i]) {
                err = "PML face between two interior cells (reference classifyCells throws)";
                return FDTD_ERR_UNSUPPORTED;
            }
    }
    if (S.abc)
        for (int d = 0; d < 3; ++d) { S.inner_lo[d] = 2; S.inner_hi[d] = S.n[d] - 2; }
    // plane constraints: a cell is solid when a constraint holds its centre
    // (FDTD_PlaneConstraint::Inside), which depends on one coordinate only
    for (int d = 0; d < 3; ++d) {
        S.csolid[d].assign(S.n[d], 0);
        S.tile_lo[d] = S.inner_lo[d];
        S.tile_hi[d] = S.inner_hi[d];
    }
    if (!S.cons.empty() && S.abc) {
        err = "plane constraints need a PML (the Engquist boundary's layer-1 cells would lie behind the plane)";
        return FDTD_ERR_UNSUPPORTED;
    }
    for (const fdtd_plane_constraint &c : S.cons) {
        const int d = c.direction;
        for (int i = 0; i < S.n[d]; ++i) {
            double x = bmin[d];
            x += (0.5 + (double)i) * h;
            if ((x - c.height) * (double)c.sign < 0.0) S.csolid[d][i] = 1;
        }
    }
    for (int d = 0; d < 3; ++d) {
        // Along an axis the constraints leave the fluid layers [L, U) (each
        // solid set is a prefix or a suffix).  The interior tiles start past
        // the solid layers and the fluid layer next to them, whose cells take
        // the ghost-cell update when they lie outside the PML.
        int L = 0, U = S.n[d];
        while (L < S.n[d] && S.csolid[d][L]) ++L;
        while (U > L && S.csolid[d][U - 1]) --U;
        for (int i = L; i < U; ++i)
            if (S.csolid[d][i]) { err = "internal: constraint layers are not a prefix and a suffix"; return FDTD_ERR_INVALID; }
        int lo = S.inner_lo[d], hi = S.inner_hi[d];
        if (L > 0 && L >= lo) lo = L + 1;
        if (U < S.n[d] && U <= hi) hi = U - 1;
        if (hi - lo < 2) {
            err = "plane constraints leave fewer than two interior cell layers along an axis";
            return FDTD_ERR_UNSUPPORTED;
        }
        S.tile_lo[d] = lo;
        S.tile_hi[d] = hi;
    }
    S.n_pml = S.abc ? 0 : S.N - (int64_t)(S.inner_hi[0] - S.inner_lo[0]) * (S.inner_hi[1] - S.inner_lo[1]) * (S.inner_hi[2] - S.inner_lo[2]);
    S.slabs.clear();
    for (int d = 0; d < 3; ++d) S.n_pml_faces[d] = 0;
    S.n_pml_merged = 0;
    if (!S.abc) {
        // x slabs span the grid, y slabs the inner x range, z slabs the inner x and y ranges
        const int *n = S.n, *a = S.inner_lo, *b = S.inner_hi;
        const int boxes[6][6] = {
            {0, a[0], 0, n[1], 0, n[2]},    {b[0], n[0], 0, n[1], 0, n[2]},
            {a[0], b[0], 0, a[1], 0, n[2]}, {a[0], b[0], b[1], n[1], 0, n[2]},
            {a[0], b[0], a[1], b[1], 0, a[2]}, {a[0], b[0], a[1], b[1], b[2], n[2]},
        };
        S.slabs.clear();
        int64_t cells = 0, faces[3] = {0, 0, 0};
        for (const auto &bx : boxes) {
            PmlSlab sl{};
            for (int d = 0; d < 3; ++d) { sl.lo[d] = bx[2 * d]; sl.e[d] = bx[2 * d + 1] - bx[2 * d]; }
            sl.axis = (int)(&bx - &boxes[0]) / 2;
            if (sl.e[0] <= 0 || sl.e[1] <= 0 || sl.e[2] <= 0) continue;
            sl.cell_base = cells;
            cells += (int64_t)sl.e[0] * sl.e[1] * sl.e[2];
            for (int d = 0; d < 3; ++d) {
                sl.v_base[d] = faces[d];
                faces[d] += (int64_t)(sl.e[0] + (d == 0)) * (sl.e[1] + (d == 1)) * (sl.e[2] + (d == 2));
            }
            S.slabs.push_back(sl);
        }
        if (cells != S.n_pml) { err = "internal: PML slab decomposition does not cover the PML"; return FDTD_ERR_INVALID; }
        for (int d = 0; d < 3; ++d) S.n_pml_faces[d] = faces[d];
        set_pml_state(S, true);
    }

    // ---- pressure sources (GaussianPressureSource) ------------------------
    S.psrc.clear(); S.gtab.clear();
    for (int q = 0; q < sc.n_pressure_sources; ++q) {
        const fdtd_pressure_source &ps = sc.pressure_sources[q];
        if (!(ps.width_time > 0) || !(ps.width_space > 0)) { err = "pressure source widths must be positive"; return FDTD_ERR_INVALID; }
        PSrcHost P;
        P.t0 = ps.offset_time; P.wt = ps.width_time;
        // GaussianPressureSource.cpp:21-27: the collocated build scales the pulse
        // by K / 3.6836, the staggered build integrates it in time and scales by K
        P.amp = st.scheme == FDTD_SCHEME_STAGGERED ? ps.normalize_constant : 1. / 3.6836 * ps.normalize_constant;
        for (int d = 0; d < 3; ++d) {
            const double mn = ps.position[d] - ps.width_space * kCheckBound;
            const double mx = ps.position[d] + ps.width_space * kCheckBound;
            const double ctr = (mx + mn) / 2.0;
            P.goff[d] = (int)S.gtab.size();
            P.lo[d] = S.n[d]; P.hi[d] = 0;
            for (int i = 0; i < S.n[d]; ++i) {
                double x = bmin[d];
                x += (0.5 + (double)i) * h;
                double g = 0.0;
                if (x < mx && x > mn) {
                    g = std::exp(-((x - ctr) * (x - ctr)) / (2.0 * ps.width_space * ps.width_space));
                    P.lo[d] = std::min(P.lo[d], i); P.hi[d] = std::max(P.hi[d], i + 1);
                }
                S.gtab.push_back((float)g);
            }
        }
        S.psrc.push_back(P);
    }
    S.spk.clear(); S.spk_samples.clear();
    if (sc.n_speaker_sources < 0 || (sc.n_speaker_sources > 0 && !sc.speaker_sources)) { err = "bad speaker source list"; return FDTD_ERR_INVALID; }
    if (sc.n_speaker_sources > 0 && st.scheme != FDTD_SCHEME_COLLOCATED) {
        err = "speaker sources run with the collocated scheme";
        return FDTD_ERR_UNSUPPORTED;
    }
    for (int q = 0; q < sc.n_speaker_sources; ++q) {
        const fdtd_speaker_source &sp = sc.speaker_sources[q];
        if (!(sp.width_space > 0) || !(sp.sample_rate > 0) || sp.n_samples < 0 || (sp.n_samples > 0 && !sp.samples)) {
            err = "speaker source: width_space and sample_rate must be positive, samples given";
            return FDTD_ERR_INVALID;
        }
        SceneSetup::SpeakerHost H{};
        for (int d = 0; d < 3; ++d) H.p[d] = sp.position[d];
        H.w = sp.width_space; H.t0 = sp.start_time;
        H.period = 1.0 / sp.sample_rate;   // WavReader::SampleRate returns the sample period
        H.off = (int)S.spk_samples.size(); H.n = sp.n_samples;
        S.spk_samples.insert(S.spk_samples.end(), sp.samples, sp.samples + sp.n_samples);
        for (int d = 0; d < 3; ++d) {
            // cells within 8 w of the speaker (exp(-32) of its peak beyond)
            H.lo[d] = std::max(0, (int)std::floor((H.p[d] - 8.0 * H.w - bmin[d]) / h - 0.5));
            H.hi[d] = std::min(S.n[d], (int)std::ceil((H.p[d] + 8.0 * H.w - bmin[d]) / h - 0.5) + 1);
            if (H.hi[d] <= H.lo[d]) { H.lo[d] = 0; H.hi[d] = 0; }
        }
        S.spk.push_back(H);
        for (int d = 0; d < 3; ++d) {
            S.spk_lo[d] = S.spk.size() == 1 ? H.lo[d] : std::min(S.spk_lo[d], H.lo[d]);
            S.spk_hi[d] = S.spk.size() == 1 ? H.hi[d] : std::max(S.spk_hi[d], H.hi[d]);
        }
    }
    S.isrc.clear();
    for (int q = 0; q < sc.n_impulse_sources; ++q) {
        const fdtd_impulse_source &is = sc.impulse_sources[q];
        if (is.object < 0 || is.object >= sc.n_objects) { err = "impulse source names a missing object"; return FDTD_ERR_INVALID; }
        ImpulseHost H{};
        H.obj = is.object;
        H.kind = is.kind;
        if (is.kind == FDTD_VIB_GAUSSIAN) {
            if (!(is.width > 0)) { err = "impulse width must be positive"; return FDTD_ERR_INVALID; }
            H.A = is.amplitude; H.w = is.width; H.ton = is.onset;
        } else if (is.kind == FDTD_VIB_HARMONIC) {
            if (!std::isfinite(is.omega) || !std::isfinite(is.phase) || !std::isfinite(is.amplitude)) {
                err = "a harmonic source needs a finite amplitude, omega and phase";
                return FDTD_ERR_INVALID;
            }
            H.A = is.amplitude; H.omega = is.omega; H.phase = is.phase;
            H.decay = is.decay_rate; H.dcenter = is.decay_center; H.t0 = is.t0; H.t1 = is.t1;
        } else if (is.kind == FDTD_VIB_CHIRP) {
            if (!std::isfinite(is.omega) || !std::isfinite(is.omega1) || !(is.t1 > is.t0)) {
                err = "a chirp source needs finite frequencies and t1 > t0";
                return FDTD_ERR_INVALID;
            }
            H.omega = is.omega; H.omega1 = is.omega1; H.t0 = is.t0; H.t1 = is.t1;
        } else {
            err = "vibrational source kind must be FDTD_VIB_GAUSSIAN, FDTD_VIB_HARMONIC or FDTD_VIB_CHIRP";
            return FDTD_ERR_INVALID;
        }
        S.isrc.push_back(H);
    }
    // acceleration noise: the constant vector of each impact, in the world
    // frame (AccelerationNoiseVibrationalSource.cpp:70-89; impacts with
    // tau < SMALL_NUM = 1e-12 are skipped there)
    S.impacts.clear();
    S.n_cpo = sc.n_acc_noise_sources > 0 ? 4 : 1;
    if (sc.n_acc_noise_sources < 0) { err = "negative acceleration-noise source count"; return FDTD_ERR_INVALID; }
    for (int q = 0; q < sc.n_acc_noise_sources; ++q) {
        const fdtd_acc_noise_source &an = sc.acc_noise_sources[q];
        if (an.object < 0 || an.object >= sc.n_objects) { err = "acceleration-noise source names a missing object"; return FDTD_ERR_INVALID; }
        if (an.n_impacts < 0 || (an.n_impacts > 0 && !an.impacts) || !(an.mass > 0)) {
            err = "an acceleration-noise source needs a positive mass and its impacts";
            return FDTD_ERR_INVALID;
        }
        const fdtd_object &ob = sc.objects[an.object];
        for (int d = 0; d < 3; ++d)
            if (ob.velocity[d] != 0.0 || ob.angular_velocity[d] != 0.0) {
                err = "acceleration noise is supported on static objects only";
                return FDTD_ERR_UNSUPPORTED;
            }
        const double *Iv = an.inv_inertia, *R = ob.rotation;
        for (int m = 0; m < an.n_impacts; ++m) {
            const fdtd_impact &im = an.impacts[m];
            if (!(im.support >= 1e-12)) continue;
            const double *J = im.impulse;
            double r[3], rxJ[3], al[3], a[3];
            for (int d = 0; d < 3; ++d) r[d] = im.position[d] - an.center_of_mass[d];
            rxJ[0] = r[1] * J[2] - r[2] * J[1]; rxJ[1] = r[2] * J[0] - r[0] * J[2]; rxJ[2] = r[0] * J[1] - r[1] * J[0];
            for (int d = 0; d < 3; ++d) al[d] = Iv[3 * d] * rxJ[0] + Iv[3 * d + 1] * rxJ[1] + Iv[3 * d + 2] * rxJ[2];
            // J / m + (I^-1 (r x J)) x r, then pi / (2 tau), in the object frame
            a[0] = J[0] / an.mass + (al[1] * r[2] - al[2] * r[1]);
            a[1] = J[1] / an.mass + (al[2] * r[0] - al[0] * r[2]);
            a[2] = J[2] / an.mass + (al[0] * r[1] - al[1] * r[0]);
            const double f = 3.141592653589793 / (2.0 * im.support);
            ImpactHost H{};
            H.obj = an.object;
            H.t0 = im.timestamp;
            H.tau = im.support;
            for (int d = 0; d < 3; ++d) H.v[d] = f * (R[3 * d] * a[0] + R[3 * d + 1] * a[1] + R[3 * d + 2] * a[2]);
            S.impacts.push_back(H);
        }
    }

    // ---- objects: rasterisation -------------------------------------------
    S.objs.clear();
    for (int q = 0; q < sc.n_objects; ++q) S.objs.push_back(make_geom(sc.objects[q]));
    const int64_t N = S.N;
    const int nx = S.n[0], ny = S.n[1], nz = S.n[2];
    auto flat = [&](int i, int j, int k) { return ((int64_t)i * ny + j) * nz + k; };
    auto cpos = [&](int i, int j, int k, double *x) {
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) { x[d] = bmin[d]; x[d] += (0.5 + (double)ijk[d]) * h; }
    };
    S.moving = false;
    S.mv_v.clear(); S.mv_w.clear(); S.mv_cls0.clear();
    for (int q = 0; q < sc.n_objects; ++q)
        for (int d = 0; d < 3; ++d)
            S.moving = S.moving || sc.objects[q].velocity[d] != 0.0 || sc.objects[q].angular_velocity[d] != 0.0;
    // keyframe animation (FDTD_RigidObject_Blender_Animator::ReadObjectKinData,
    // FDTD_RigidObject_Animator.cpp:78-131): qw clamped to [-1, 1], the
    // rotation turned from Blender's z-up by q <- q * Rot(x, -90 degrees)
    S.kin.clear();
    S.kin_off.assign(sc.n_objects, 0);
    S.kin_n.assign(sc.n_objects, 0);
    S.kin_dt.assign(sc.n_objects, 0.0);
    if (sc.n_kinematics < 0 || (sc.n_kinematics > 0 && !sc.kinematics)) { err = "bad kinematics list"; return FDTD_ERR_INVALID; }
    for (int q = 0; q < sc.n_kinematics; ++q) {
        const fdtd_kinematics &km = sc.kinematics[q];
        if (km.object < 0 || km.object >= sc.n_objects) { err = "kinematics names a missing object"; return FDTD_ERR_INVALID; }
        if (km.n_frames < 1 || !km.frames || !(km.step_size > 0.0)) {
            err = "kinematics need at least one frame and a positive frame step";
            return FDTD_ERR_INVALID;
        }
        if (S.kin_n[km.object] > 0) { err = "an object has two kinematics"; return FDTD_ERR_INVALID; }
        for (int d = 0; d < 3; ++d)
            if (sc.objects[km.object].velocity[d] != 0.0 || sc.objects[km.object].angular_velocity[d] != 0.0) {
                err = "an object animated by kinematics has no velocity or angular velocity of its own";
                return FDTD_ERR_INVALID;
            }
        const double half = -90.0 * M_PI / 180.0 / 2.0;
        const double cw = std::cos(half), cx = std::sin(half);   // corr = (cw, cx, 0, 0)
        S.kin_off[km.object] = (int32_t)S.kin.size();
        S.kin_n[km.object] = km.n_frames;
        S.kin_dt[km.object] = km.step_size;
        for (int f = 0; f < km.n_frames; ++f) {
            const double *r = km.frames + 7 * (size_t)f;
            for (int e = 0; e < 7; ++e)
                if (!std::isfinite(r[e])) { err = "kinematics frames must be finite"; return FDTD_ERR_INVALID; }
            const double w = std::max(-1.0, std::min(1.0, r[3])), x = r[4], y = r[5], z = r[6];
            // Hamilton product (w, x, y, z) * (cw, cx, 0, 0) (Quaternion.hpp:82-89)
            const double qw = w * cw - x * cx, qx = w * cx + x * cw, qy = y * cw + z * cx, qz = z * cw - y * cx;
            const double row[8] = {r[0], r[1], r[2], qw, qx, qy, qz, 0.0};
            S.kin.insert(S.kin.end(), row, row + 8);
        }
        S.moving = true;
    }
    // pose of object q at time t: its keyframes, or its constant motion
    auto pose_at = [&](int q, double t) {
        if (S.kin_n[q] > 0) return geom_keyframe(S.objs[q], &S.kin[S.kin_off[q]], S.kin_n[q], S.kin_dt[q], t);
        return geom_at(S.objs[q], &S.mv_v[3 * q], &S.mv_w[3 * q], t);
    };
    if (S.moving) {
        if (st.scheme != FDTD_SCHEME_COLLOCATED || S.boundary != FDTD_BOUNDARY_RASTERIZE || S.visc) {
            err = "moving objects run with the collocated scheme, rasterized ghost cells and no air viscosity";
            return FDTD_ERR_UNSUPPORTED;
        }
        if (S.n_cpo > 1) { err = "acceleration noise is supported in scenes without moving objects"; return FDTD_ERR_UNSUPPORTED; }
        if (sc.n_objects > MAX_MOVING_OBJ) {
            err = "a scene with moving objects holds at most " + std::to_string((int)MAX_MOVING_OBJ) + " objects";
            return FDTD_ERR_UNSUPPORTED;
        }
        if (!(st.stop_time > st.start_time)) {
            err = "moving objects need a stop_time after the start: the cells they sweep are planned up to it";
            return FDTD_ERR_INVALID;
        }
        for (int q = 0; q < sc.n_objects; ++q)
            for (int d = 0; d < 3; ++d) { S.mv_v.push_back(sc.objects[q].velocity[d]); S.mv_w.push_back(sc.objects[q].angular_velocity[d]); }
        // steps whose start time t_start + s dt lies within [start, stop]
        S.mv_steps = (int64_t)std::floor((st.stop_time - st.start_time) / S.dt) + 1;
        // swept box: the centre's path from start to stop, widened by the
        // largest semi-axis (the shape turns about its centre); a keyframed
        // centre moves piecewise linearly, so its positions at every step
        // and at every frame inside the run bound it
        double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        auto widen = [&](const double *c, double r) {
            for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], c[d] - r); hi[d] = std::max(hi[d], c[d] + r); }
        };
        for (int q = 0; q < sc.n_objects; ++q) {
            const Geom &g = S.objs[q];
            const double r = std::max(g.a[0], std::max(g.a[1], g.a[2]));
            if (S.kin_n[q] > 0) {
                for (int64_t s = 0; s < S.mv_steps; ++s) widen(pose_at(q, st.start_time + (double)s * S.dt).c, r);
                for (int f = 0; f < S.kin_n[q]; ++f)
                    if (f * S.kin_dt[q] >= st.start_time && f * S.kin_dt[q] <= st.stop_time) widen(&S.kin[S.kin_off[q] + 8 * f], r);
                continue;
            }
            double c0[3], c1[3];
            for (int d = 0; d < 3; ++d) { c0[d] = g.c[d] + S.mv_v[3 * q + d] * st.start_time; c1[d] = g.c[d] + S.mv_v[3 * q + d] * st.stop_time; }
            widen(c0, r);
            widen(c1, r);
        }
        const int edge = S.abc ? 3 : 0;
        for (int d = 0; d < 3; ++d) {
            // cells whose centre lies within two cells of the swept box
            S.mv_lo[d] = (int)std::floor((lo[d] - bmin[d]) / h - 0.5) - 2;
            S.mv_hi[d] = (int)std::ceil((hi[d] - bmin[d]) / h - 0.5) + 3;
            if (S.mv_lo[d] < std::max(S.inner_lo[d], edge) || S.mv_hi[d] > std::min(S.inner_hi[d], S.n[d] - edge)) {
                err = "a moving object's path (with a two-cell margin) leaves the non-PML interior before stop_time";
                return FDTD_ERR_UNSUPPORTED;
            }
        }
        // bulk flags at the start pose (7-sample rule, as the device reclassifies)
        std::vector<Geom> g0(sc.n_objects);
        for (int q = 0; q < sc.n_objects; ++q) g0[q] = pose_at(q, st.start_time);
        const double q4 = 0.25 * h;
        for (int i = S.mv_lo[0]; i < S.mv_hi[0]; ++i)
            for (int j = S.mv_lo[1]; j < S.mv_hi[1]; ++j)
                for (int k = S.mv_lo[2]; k < S.mv_hi[2]; ++k) {
                    double x[3];
                    cpos(i, j, k, x);
                    bool fluid = true;
                    const int nsmp = S.sampling == FDTD_SAMPLE_SEVEN ? 7 : 1;
                    for (int smp = 0; smp < nsmp && fluid; ++smp) {
                        double y[3] = {x[0], x[1], x[2]};
                        if (smp > 0) y[(smp - 1) / 2] += ((smp - 1) % 2 ? q4 : -q4);
                        for (const Geom &g : g0) fluid = fluid && !geom_inside(g, y);
                    }
                    S.mv_cls0.push_back(fluid ? 1 : 0);
                }
    }
    if (!S.cons.empty() && S.moving) {
        err = "plane constraints are classified statically; moving objects are not combined with them";
        return FDTD_ERR_UNSUPPORTED;
    }
    S.has_solid = !S.objs.empty() || !S.cons.empty();
    const bool static_objs = S.has_solid && !S.moving;
    S.bulk.assign(S.has_solid ? N : 0, 1);
    S.ghosts.clear(); S.special.clear(); S.boxes.clear(); S.cls.clear();
    S.n_solid = 0; S.n_cavity = 0;
    std::vector<int64_t> solid;
    std::vector<uint8_t> solid_type;   // per solid cell: 1 when a constraint holds its centre (gctype)
    std::vector<std::array<int, 6>> itbox(S.objs.size());
    if (static_objs) {
        std::vector<uint8_t> classified(N, 0);
        auto occupied = [&](const double *x) {
            for (const Geom &g : S.objs)
                if (geom_inside(g, x)) return true;
            return false;
        };
        auto constrained = [&](const double *x) {
            for (const fdtd_plane_constraint &c : S.cons)
                if ((x[c.direction] - c.height) * (double)c.sign < 0.0) return true;
            return false;
        };
        // iteration boxes (MAC_Grid.cpp:3511-3550): each object's box with
        // the volume behind every constraint cut away, then each constraint's
        // solid half-space
        const size_t nbox = S.objs.size() + S.cons.size();
        std::vector<std::array<int, 6>> boxes(nbox);
        for (size_t ob = 0; ob < nbox; ++ob) {
            double mn[3], mx[3];
            if (ob < S.objs.size()) {
                geom_aabb(S.objs[ob], mn, mx);
                for (const fdtd_plane_constraint &c : S.cons) {
                    if (c.sign > 0) mn[c.direction] = std::max(mn[c.direction], c.height);
                    else mx[c.direction] = std::min(mx[c.direction], c.height);
                }
            } else {
                const fdtd_plane_constraint &c = S.cons[ob - S.objs.size()];
                for (int d = 0; d < 3; ++d) { mn[d] = -DBL_MAX; mx[d] = DBL_MAX; }
                if (c.sign > 0) mx[c.direction] = c.height;
                else mn[c.direction] = c.height;
            }
            // ScalarField::GetIterationBox (src/field/ScalarField.cpp:288-314)
            bool in_min = true, in_max = true;
            for (int d = 0; d < 3; ++d) {
                if (!(mn[d] >= bmin[d] && mn[d] <= bmax[d])) in_min = false;
                if (!(mx[d] >= bmin[d] && mx[d] <= bmax[d])) in_max = false;
            }
            int st3[3], rg3[3];
            for (int d = 0; d < 3; ++d) {
                const double lo_ = in_min ? mn[d] : std::max(bmin[d], mn[d]);
                const double hi_ = in_max ? mx[d] : std::min(bmax[d], mx[d]);
                st3[d] = (int)std::floor((lo_ - bmin[d]) / h);
                const int end = (int)std::ceil((hi_ - bmin[d]) / h);
                rg3[d] = std::max(end - st3[d], 1);
            }
            for (int d = 0; d < 3; ++d) {
                boxes[ob][d] = std::max(st3[d], 0);
                boxes[ob][3 + d] = std::min(st3[d] + rg3[d], S.n[d]);
            }
            if (ob < S.objs.size()) itbox[ob] = boxes[ob];
            const double q4 = 0.25 * h;
            for (int i = boxes[ob][0]; i < boxes[ob][3]; ++i)
                for (int j = boxes[ob][1]; j < boxes[ob][4]; ++j)
                    for (int k = boxes[ob][2]; k < boxes[ob][5]; ++k) {
                        const int64_t ci = flat(i, j, k);
                        if (classified[ci]) continue;
                        classified[ci] = 1;
                        double x[3];
                        cpos(i, j, k, x);
                        // isFluid (MAC_Grid.cpp:3558-3577): the centre outside
                        // every constraint and every object, and with the
                        // seven-sample rule the six points h/4 from it too
                        const bool in_con = constrained(x);
                        bool fluid = !in_con && !occupied(x);
                        for (int d = 0; d < 3 && fluid && S.sampling == FDTD_SAMPLE_SEVEN; ++d)
                            for (int sg = -1; sg <= 1 && fluid; sg += 2) {
                                double y[3] = {x[0], x[1], x[2]};
                                y[d] += (double)sg * q4;
                                if (occupied(y)) fluid = false;
                            }
                        if (!fluid) { S.bulk[ci] = 0; solid.push_back(ci); }
                    }
        }
        std::sort(solid.begin(), solid.end());
        S.n_solid = (int64_t)solid.size();
        solid_type.resize(solid.size());
        for (size_t q = 0; q < solid.size(); ++q) {
            const int64_t ci = solid[q];
            const int ijk[3] = {(int)(ci / ((int64_t)ny * nz)), (int)((ci / nz) % ny), (int)(ci % nz)};
            solid_type[q] = (S.csolid[0][ijk[0]] || S.csolid[1][ijk[1]] || S.csolid[2][ijk[2]]) ? 1 : 0;
        }
        // a solid cell of an object in the PML or on the outer layers is
        // outside what the device path restates; with the Engquist boundary
        // the cells next to an object stay off layer 1 (layers 0-2 hold no
        // solid cell).  Cells behind a plane constraint may lie in the PML:
        // their pressure is held at zero there.
        for (size_t q = 0; q < solid.size(); ++q) {
            if (solid_type[q]) continue;
            const int64_t ci = solid[q];
            const int i = (int)(ci / ((int64_t)ny * nz)), j = (int)((ci / nz) % ny), k = (int)(ci % nz);
            if (S.cpml[0][i] || S.cpml[1][j] || S.cpml[2][k]) {
                err = "an object reaches into the PML; objects must lie inside the non-PML interior";
                return FDTD_ERR_UNSUPPORTED;
            }
            const int ijk[3] = {i, j, k};
            for (int d = 0; d < 3 && S.abc; ++d)
                if (ijk[d] < 3 || ijk[d] > S.n[d] - 4) {
                    err = "an object comes within three cells of the grid boundary; the Engquist boundary needs "
                          "layers 0-2 free of solid cells";
                    return FDTD_ERR_UNSUPPORTED;
                }
        }
        // ghost entries (MakeGhostCell), topology order -x,+x,-y,+y,-z,+z
        std::vector<std::pair<uint64_t, GhostEntry>> tmp;
        for (size_t q = 0; q < solid.size(); ++q) {
            const int64_t ci = solid[q];
            const int ijk[3] = {(int)(ci / ((int64_t)ny * nz)), (int)((ci / nz) % ny), (int)(ci % nz)};
            for (int d = 0; d < 3; ++d)
                for (int sg = -1; sg <= 1; sg += 2) {
                    int t[3] = {ijk[0], ijk[1], ijk[2]};
                    t[d] += sg;
                    if (t[d] < 0 || t[d] >= S.n[d]) continue;
                    const int64_t nb = flat(t[0], t[1], t[2]);
                    if (!S.bulk[nb]) continue;
                    GhostEntry g{};
                    g.owner = (int)ci; g.nbr = (int)nb; g.topology = sg * (d + 1);
                    g.type = solid_type[q];
                    double x[3];
                    cpos(ijk[0], ijk[1], ijk[2], x);
                    // LowestObjectDistance from the owner centre
                    double best = DBL_MAX; int bo = 0;
                    for (size_t ob = 0; ob < S.objs.size(); ++ob) {
                        double xb[3], nn[3];
                        const double dd = geom_closest(S.objs[ob], x, xb, nn);
                        if (std::fabs(dd) < std::fabs(best)) { best = dd; bo = (int)ob; }
                    }
                    g.obj = bo;
                    for (int e = 0; e < 3; ++e) g.pos[e] = x[e];
                    // MakeGhostCell (MAC_Grid.cpp:4310-4340)
                    g.pos[d] += (g.topology > 0 ? 1.0 : -1.0) * S.ghost_offset() * h;
                    tmp.push_back({pair_key(ci, nb), g});
                }
        }
        std::sort(tmp.begin(), tmp.end(), [](const auto &a, const auto &b) {
            if (a.second.owner != b.second.owner) return a.second.owner < b.second.owner;
            return a.first < b.first;
        });
        for (auto &t : tmp) S.ghosts.push_back(t.second);
    }
    const int G = (int)S.ghosts.size();
    std::map<uint64_t, int> key2g;
    std::vector<int> first_child;  // owner -> first ghost index
    for (int g = 0; g < G; ++g) key2g[pair_key(S.ghosts[g].owner, S.ghosts[g].nbr)] = g;
    auto children = [&](int64_t owner, int &first) {
        auto it = std::lower_bound(S.ghosts.begin(), S.ghosts.end(), owner,
                                   [](const GhostEntry &e, int64_t o) { return e.owner < o; });
        first = (int)(it - S.ghosts.begin());
        int cnt = 0;
        while (first + cnt < G && S.ghosts[first + cnt].owner == owner) ++cnt;
        return cnt;
    };
    std::vector<uint8_t> isghost(S.has_solid ? N : 0, 0);
    for (const GhostEntry &g : S.ghosts) isghost[g.owner] = 1;

    // ---- ghost values (PML_pressureUpdateGhostCells) -----------------------
    S.gk_ptr.assign(1, 0); S.gk_idx.clear(); S.gk_w.clear();
    S.gu_ptr.assign(1, 0); S.gu_idx.clear(); S.gu_w.clear();
    S.n_coupled = 0;
    const bool coupled = S.boundary == FDTD_BOUNDARY_FULLY_COUPLED;
    for (int g = 0; g < G; ++g) {
        GhostEntry &gc = S.ghosts[g];
        if (gc.type == 1) {
            // a plane constraint reflects perfectly: p_g = p[nbr] (MAC_Grid.cpp:1198-1204)
            gc.k_raster = 0.0;
            gc.k_vec = 0.0;
            gc.coupled = 0;
            gc.c_acc = 0.0;
            gc.c_vec[0] = gc.c_vec[1] = gc.c_vec[2] = 0.0;
            S.gk_ptr.push_back((int32_t)S.gk_idx.size());
            S.gu_ptr.push_back((int32_t)S.gu_idx.size());
            continue;
        }
        const Geom &ob = S.objs[gc.obj];
        double bp[3], nb[3];
        const double dist = geom_closest(ob, gc.pos, bp, nb);
        // FDTD_RigidObject::ReflectAgainstBoundary
        double en[3], image[3];
        const double diff[3] = {bp[0] - gc.pos[0], bp[1] - gc.pos[1], bp[2] - gc.pos[2]};
        if (std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]) > kRefEps) {
            if (dist < 0) for (int e = 0; e < 3; ++e) { en[e] = bp[e] - gc.pos[e]; image[e] = bp[e] + en[e]; }
            else          for (int e = 0; e < 3; ++e) { en[e] = gc.pos[e] - bp[e]; image[e] = gc.pos[e] + en[e]; }
        } else {
            for (int e = 0; e < 3; ++e) { en[e] = nb[e]; image[e] = gc.pos[e]; }
        }
        const double enn = std::sqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
        for (int e = 0; e < 3; ++e) en[e] /= enn;
        double cell_n[3] = {0, 0, 0};
        cell_n[std::abs(gc.topology) - 1] = gc.topology > 0 ? 1.0 : -1.0;
        const double ndc = nb[0] * cell_n[0] + nb[1] * cell_n[1] + nb[2] * cell_n[2];
        const double nde = nb[0] * en[0] + nb[1] * en[1] + nb[2] * en[2];
        gc.k_raster = S.raster_factor() * h * rho * ndc;
        gc.k_vec = S.raster_factor() * h * rho * (gc.topology > 0 ? 1.0 : -1.0);
        gc.coupled = 0;
        gc.c_acc = 0;
        gc.c_vec[0] = gc.c_vec[1] = gc.c_vec[2] = 0.0;
        bool done = false;
        if (coupled) {
            int64_t cells[8];
            const int ncell = enclosing_cells(S, bmin, image, cells);
            if (ncell == 8) {
                double origin[3] = {0, 0, 0};
                for (int q = 0; q < 8; ++q) {
                    const int64_t ci = cells[q];
                    double x[3];
                    cpos((int)(ci / ((int64_t)ny * nz)), (int)((ci / nz) % ny), (int)(ci % nz), x);
                    for (int e = 0; e < 3; ++e) origin[e] += x[e];
                }
                for (int e = 0; e < 3; ++e) origin[e] /= 8.0;
                double V[64], v[8];
                int s_i = -1, kcell[8], ughost[8];
                auto scaled = [&](const double *p, double *np) {
                    for (int e = 0; e < 3; ++e) { np[e] = p[e]; np[e] -= origin[e]; np[e] /= h; np[e] += 0.5; }
                };
                auto fill_reg = [&](const double *p, double *row) {
                    double np[3]; scaled(p, np);
                    row[0] = np[0] * np[1] * np[2]; row[1] = np[0] * np[1]; row[2] = np[0] * np[2]; row[3] = np[1] * np[2];
                    row[4] = np[0]; row[5] = np[1]; row[6] = np[2]; row[7] = 1;
                };
                bool bad = false;
                for (int q = 0; q < 8; ++q) {
                    kcell[q] = -1; ughost[q] = -1;
                    const int64_t ci = cells[q];
                    if (ci == gc.owner) {
                        double np[3]; scaled(bp, np);
                        const double x = np[0], y = np[1], z = np[2];
                        double *row = &V[q * 8];
                        row[0] = en[0] * y * z + en[1] * x * z + en[2] * x * y; row[1] = en[0] * y + en[1] * x;
                        row[2] = en[0] * z + en[2] * x; row[3] = en[1] * z + en[2] * y;
                        row[4] = en[0]; row[5] = en[1]; row[6] = en[2]; row[7] = 0;
                        s_i = q;
                    } else if (!isghost[ci]) {
                        double x[3];
                        cpos((int)(ci / ((int64_t)ny * nz)), (int)((ci / nz) % ny), (int)(ci % nz), x);
                        fill_reg(x, &V[q * 8]);
                        kcell[q] = (int)ci;
                    } else {
                        int f0; const int cnt = children(ci, f0);
                        if (cnt == 0) { bad = true; break; }
                        double best = DBL_MAX; int bi = f0;
                        for (int u = f0; u < f0 + cnt; ++u) {
                            const double *gp = S.ghosts[u].pos;
                            const double d2 = (gp[0] - image[0]) * (gp[0] - image[0]) + (gp[1] - image[1]) * (gp[1] - image[1]) + (gp[2] - image[2]) * (gp[2] - image[2]);
                            if (d2 < best) { best = d2; bi = u; }
                        }
                        fill_reg(S.ghosts[bi].pos, &V[q * 8]);
                        ughost[q] = bi;
                    }
                }
                if (bad) { err = "ghost cell without children in a fully-coupled stencil"; return FDTD_ERR_INVALID; }
                fill_reg(image, v);
                // w solves V^T w = v: w = (V V^T)^+ V v; cond of V from V V^T
                double M[64], Mp[64], Vv[8];
                for (int a = 0; a < 8; ++a)
                    for (int b = 0; b < 8; ++b) {
                        double s = 0;
                        for (int q = 0; q < 8; ++q) s += V[a * 8 + q] * V[b * 8 + q];
                        M[a * 8 + b] = s;
                    }
                const double cond = gram_pinv<8>(M, 8.0, Mp);
                for (int a = 0; a < 8; ++a) {
                    double s = 0;
                    for (int q = 0; q < 8; ++q) s += V[a * 8 + q] * v[q];
                    Vv[a] = s;
                }
                double w[8];
                for (int a = 0; a < 8; ++a) {
                    double s = 0;
                    for (int b = 0; b < 8; ++b) s += Mp[a * 8 + b] * Vv[b];
                    w[a] = s;
                }
                if (!(cond > kCondLimit && s_i != -1)) {
                    const double weights = (dist < 0.0 ? -2.0 * std::fabs(dist) : -std::fabs(dist));
                    gc.coupled = 1;
                    gc.c_acc = -rho * nde * (weights + (s_i >= 0 ? w[s_i] * h : 0.0));
                    for (int e = 0; e < 3; ++e) gc.c_vec[e] = -rho * en[e] * (weights + (s_i >= 0 ? w[s_i] * h : 0.0));
                    for (int q = 0; q < 8; ++q) {
                        if (kcell[q] >= 0 && S.bulk[kcell[q]]) { S.gk_idx.push_back(kcell[q]); S.gk_w.push_back(w[q]); }
                        if (ughost[q] >= 0) { S.gu_idx.push_back(ughost[q]); S.gu_w.push_back(w[q]); }
                    }
                    ++S.n_coupled;
                    done = true;
                }
            }
        }
        (void)done;
        S.gk_ptr.push_back((int32_t)S.gk_idx.size());
        S.gu_ptr.push_back((int32_t)S.gu_idx.size());
    }
    // The coupled system (MAC_Grid.cpp:1400-1425: A(r,r) = -1, A(r,u) = w_u,
    // solved by BiCGSTAB every step) is x = b + W x with a static W and a
    // right-hand side b = K p + c a_n that is linear in the pressures and the
    // accelerations.  Its solution is therefore a static linear operator,
    //   x = (I - W)^{-1} (K p + c a) = M p + m a,
    // evaluated on the device in one pass.  W couples a ghost to the ghosts in
    // its image point's stencil, further along the surface; its graph is
    // almost always acyclic (the rare cycles are pairs), so the rows of M
    // follow by substitution in dependency order: M_g = K_g + sum_u W_gu M_u,
    // with a small dense inverse of I - W over each cycle.  The result is the
    // exact solution (to rounding) whether or not an iteration would converge.
    S.op_ptr.assign(1, 0); S.op_idx.clear(); S.op_w.clear();
    S.oc_ptr.assign(1, 0); S.oc_ch.clear(); S.oc_w.clear();
    S.op_max_cycle = 0;
    if (S.n_coupled > 0) {
        fdtd_status ost = build_coupled_operator(S, err);
        if (ost != FDTD_OK) return ost;
    }

    // ---- special cells and their boxes ------------------------------------
    if (G > 0) {
        std::vector<int> rec_of(N, -1);  // sized to N only when objects exist
        for (int g = 0; g < G; ++g) {
            const GhostEntry &gc = S.ghosts[g];
            const int c_ = gc.nbr;
            {
                // a fluid neighbour in the PML takes the PML update, which
                // reads the solid cell's zero pressure; its ghost serves
                // listener interpolation only
                const int i = (int)(c_ / ((int64_t)ny * nz)), j = (int)((c_ / nz) % ny), k = (int)(c_ % nz);
                if (S.cpml[0][i] || S.cpml[1][j] || S.cpml[2][k]) continue;
            }
            if (rec_of[c_] < 0) {
                SpecialRec r{};
                r.cell = c_; r.flags = 0;
                for (int q = 0; q < 6; ++q) r.gid[q] = -1;
                rec_of[c_] = (int)S.special.size();
                S.special.push_back(r);
            }
            // direction from the bulk cell towards the ghost owner
            const int d = std::abs(gc.topology) - 1;
            const int towards_owner_positive = gc.topology < 0;  // owner -> nbr negative => owner on + side
            S.special[rec_of[c_]].gid[2 * d + (towards_owner_positive ? 1 : 0)] = g;
        }
        for (SpecialRec &r : S.special) {
            int both = 0;
            for (int d = 0; d < 3; ++d) both += (r.gid[2 * d] >= 0 && r.gid[2 * d + 1] >= 0);
            if (both > 1) { r.flags |= 1; ++S.n_cavity; }
        }
        std::sort(S.special.begin(), S.special.end(), [](const SpecialRec &a, const SpecialRec &b) { return a.cell < b.cell; });
        for (size_t r = 0; r < S.special.size(); ++r) rec_of[S.special[r].cell] = (int)r;
        for (size_t ob = 0; ob < S.objs.size(); ++ob) {
            SpecialBox b{};
            for (int d = 0; d < 3; ++d) { b.lo[d] = std::max(itbox[ob][d] - 1, 0); b.hi[d] = std::min(itbox[ob][3 + d] + 1, S.n[d]); }
            b.offset = (int32_t)S.cls.size();
            for (int i = b.lo[0]; i < b.hi[0]; ++i)
                for (int j = b.lo[1]; j < b.hi[1]; ++j)
                    for (int k = b.lo[2]; k < b.hi[2]; ++k) {
                        const int64_t ci = flat(i, j, k);
                        S.cls.push_back(!S.bulk[ci] ? -2 : rec_of[ci]);
                    }
            S.boxes.push_back(b);
        }
        // every object's solid cell must sit inside a box (cells behind a
        // constraint lie outside the interior tiles or in the PML)
        for (size_t q = 0; q < solid.size(); ++q) {
            if (solid_type[q]) continue;
            const int64_t ci = solid[q];
            const int ijk[3] = {(int)(ci / ((int64_t)ny * nz)), (int)((ci / nz) % ny), (int)(ci % nz)};
            bool in = false;
            for (const SpecialBox &b : S.boxes)
                in |= ijk[0] >= b.lo[0] && ijk[0] < b.hi[0] && ijk[1] >= b.lo[1] && ijk[1] < b.hi[1] && ijk[2] >= b.lo[2] && ijk[2] < b.hi[2];
            if (!in) { err = "internal: solid cell outside every object box"; return FDTD_ERR_INVALID; }
        }
    } else if (S.n_solid > 0 && static_objs) {
        // solid cells with no bulk neighbour at all cannot happen: the
        // object would fill the grid
        err = "object covers the whole interior";
        return FDTD_ERR_INVALID;
    }

    // ---- listeners: MLS weights --------------------------------------------
    S.L = sc.n_listeners;
    S.listeners.assign(sc.listeners, sc.listeners + 3 * (size_t)S.L);
    S.lst_ptr.assign(1, 0); S.lst_terms.clear();
    for (int l = 0; l < S.L; ++l) {
        const double *p = &S.listeners[3 * l];
        for (int d = 0; d < 3; ++d)
            if (!(p[d] >= bmin[d] && p[d] <= bmax[d])) {
                err = "listener " + std::to_string(l) + " lies outside the grid";
                return FDTD_ERR_INVALID;
            }
        int64_t cells[8];
        const int ne = enclosing_cells(S, bmin, p, cells);
        for (int q = 0; q < ne && S.moving; ++q) {
            const int ijk[3] = {(int)(cells[q] / ((int64_t)ny * nz)), (int)((cells[q] / nz) % ny), (int)(cells[q] % nz)};
            bool in = true;
            for (int d = 0; d < 3; ++d) in = in && ijk[d] >= S.mv_lo[d] && ijk[d] < S.mv_hi[d];
            if (in) {
                err = "listener " + std::to_string(l) + " reads a cell that a moving object's path reaches";
                return FDTD_ERR_UNSUPPORTED;
            }
        }
        for (int q = 0; q < ne && S.abc; ++q) {
            const int ijk[3] = {(int)(cells[q] / ((int64_t)ny * nz)), (int)((cells[q] / nz) % ny), (int)(cells[q] % nz)};
            for (int d = 0; d < 3; ++d)
                if (ijk[d] == 0 || ijk[d] == S.n[d] - 1) {
                    err = "listener " + std::to_string(l) + " reads a cell of the outermost layer, which the Engquist "
                          "boundary never updates";
                    return FDTD_ERR_UNSUPPORTED;
                }
        }
        struct Pt { double x[3]; int kind; int index; };
        std::vector<Pt> pts;
        for (int q = 0; q < ne; ++q) {
            const int64_t ci = cells[q];
            const bool gh = S.has_solid && isghost[ci];
            if (!gh) {
                Pt P; cpos((int)(ci / ((int64_t)ny * nz)), (int)((ci / nz) % ny), (int)(ci % nz), P.x);
                const bool isbulk = !S.has_solid || S.bulk[ci];
                P.kind = isbulk ? 0 : -1;  // solid non-ghost cells hold p = 0
                P.index = (int)ci;
                pts.push_back(P);
            } else {
                int f0; const int cnt = children(ci, f0);
                for (int u = f0; u < f0 + cnt; ++u) {
                    Pt P; for (int e = 0; e < 3; ++e) P.x[e] = S.ghosts[u].pos[e];
                    P.kind = 1; P.index = u;
                    pts.push_back(P);
                }
            }
        }
        if (ne < 4) {
            int64_t best = -1; double bd = DBL_MAX;
            for (int q = 0; q < ne; ++q) {
                double x[3]; const int64_t ci = cells[q];
                cpos((int)(ci / ((int64_t)ny * nz)), (int)((ci / nz) % ny), (int)(ci % nz), x);
                const double d2 = (x[0] - p[0]) * (x[0] - p[0]) + (x[1] - p[1]) * (x[1] - p[1]) + (x[2] - p[2]) * (x[2] - p[2]);
                if (d2 < bd) { bd = d2; best = ci; }
            }
            if (best >= 0 && (!S.has_solid || S.bulk[best])) S.lst_terms.push_back({0, (int)best, 1.0});
        } else {
            const int m = (int)pts.size();
            // weighted design matrix, columns 1..3 scaled by 1/h for conditioning
            std::vector<double> A(4 * m), wgt(m);
            const double r2 = kMlsRadius * kMlsRadius;
            for (int i = 0; i < m; ++i) {
                const double dx = pts[i].x[0] - p[0], dy = pts[i].x[1] - p[1], dz = pts[i].x[2] - p[2];
                const double w = std::exp(-(dx * dx + dy * dy + dz * dz) / (2. * r2));
                wgt[i] = w;
                A[4 * i + 0] = w; A[4 * i + 1] = w * dx / h; A[4 * i + 2] = w * dy / h; A[4 * i + 3] = w * dz / h;
            }
            double Gm[16] = {0}, Gp[16];
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    double s = 0;
                    for (int i = 0; i < m; ++i) s += A[4 * i + a] * A[4 * i + b];
                    Gm[4 * a + b] = s;
                }
            gram_pinv<4>(Gm, (double)std::min(m, 4), Gp);
            for (int i = 0; i < m; ++i) {
                double s = 0;
                for (int b = 0; b < 4; ++b) s += Gp[b] * A[4 * i + b];  // row 0 of G^+ A^T
                const double beta = s * wgt[i];
                if (pts[i].kind >= 0) S.lst_terms.push_back({pts[i].kind, pts[i].index, beta});
            }
        }
        S.lst_ptr.push_back((int32_t)S.lst_terms.size());
    }
    return FDTD_OK;
}

}  // namespace fdtd
