// device_scene.cpp -- arena allocation, table upload, tile plans and the
// per-step launch sequence.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "device_scene.h"

namespace fdtd {

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Two passes over the same layout: first to size the arena, then to carve it.
struct Carver {
    char *base = nullptr;
    size_t off = 0;
    template <class T>
    T *take(size_t count)
    {
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += align_up(sizeof(T) * (count ? count : 1));
        return p;
    }
};

struct Layout {
    float *field;               // pressure slots, then the compact PML state
    size_t state_floats;
    PmlSlab *slabs;
    float *pd[3];
    float *vel[2][3];
    float *lapv;
    float4 *coef[3];
    int *flags[3];
    uint8_t *skip;
    float *gtab;
    PSrcDev *psrc;
    ImpDev *isrc;
    ImpactDev *impacts;
    SpecialRec *recs;
    GhostDev *ghosts;
    int32_t *op_ptr, *op_idx, *oc_ptr, *oc_ch;
    double *op_w, *oc_w;
    double *terms;
    unsigned long long *clock;
    double *t_tab;
    double *acc_tab;
    float *src_tab;
    double *spk, *spk_samples;
    float *spk_field;
    uint8_t *ucls;
    MoveDev *mobj;
    double *kin;
    Geom *pose;
    double *ughost;
    SceneDev *self;
};

// cells of a moving scene's reclassified box
static size_t moving_cells(const SceneSetup &S)
{
    if (!S.moving) return 0;
    return (size_t)(S.mv_hi[0] - S.mv_lo[0]) * (S.mv_hi[1] - S.mv_lo[1]) * (S.mv_hi[2] - S.mv_lo[2]);
}

long long pad_cells(long long N) { return (N + 63) / 64 * 64; }

// Cells of one pressure slot in device storage: rows along k are padded to a
// multiple of four floats, so every row starts 16-byte aligned and the
// four-cells-per-lane kernels can serve any grid width.
long long slot_cells(const SceneSetup &S) { return pad_cells((long long)S.n[0] * S.n[1] * row_stride(S)); }

// Every piece of time-stepping state in one zero-initialised block: the two
// pressure slots, the split pressures of the PML cells, both buffers of the
// PML face velocities, and (air viscosity) the bulk cells' previous
// Laplacian.  Each array starts on a 256-byte boundary.
size_t state_layout(const SceneSetup &S, long long npad, size_t off[1 + 3 + 6 + 1])
{
    size_t o = (size_t)NUM_P_SLOTS * (size_t)npad;
    auto al = [](size_t x) { return (x + 63) / 64 * 64; };
    off[10] = o;
    if (S.visc) o += al((size_t)npad);
    for (int d = 0; d < 3; ++d) { off[1 + d] = o; o += al((size_t)S.n_pml); }
    // face velocities: the PML's compact faces (collocated) or every face of the grid (staggered)
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d) {
            off[4 + 3 * b + d] = o;
            o += al(S.scheme == FDTD_SCHEME_STAGGERED ? (size_t)npad : (size_t)S.n_pml_faces[d]);
        }
    off[0] = 0;
    return o;
}

// The time-stepping state (L.field) comes from `cs`, every table, ring and
// counter from `c`; both are the same carver for a single arena.
void carve(const SceneSetup &S, int cap, Carver &cs, Carver &c, Layout &L)
{
    size_t off[11];
    L.state_floats = state_layout(S, slot_cells(S), off);
    L.field = cs.take<float>(L.state_floats);
    L.lapv = (L.field && S.visc) ? L.field + off[10] : nullptr;
    for (int d = 0; d < 3; ++d) L.pd[d] = L.field ? L.field + off[1 + d] : nullptr;
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d) L.vel[b][d] = L.field ? L.field + off[4 + 3 * b + d] : nullptr;
    L.slabs = c.take<PmlSlab>(S.slabs.size());
    for (int d = 0; d < 3; ++d) {
        L.coef[d] = c.take<float4>(S.n[d] + 1);
        L.flags[d] = c.take<int>(S.n[d] + 1);
    }
    L.skip = c.take<uint8_t>(!S.has_solid ? 0 : (size_t)slot_cells(S));
    L.gtab = c.take<float>(S.gtab.size());
    L.psrc = c.take<PSrcDev>(S.psrc.size());
    L.isrc = c.take<ImpDev>(S.isrc.size());
    L.impacts = c.take<ImpactDev>(S.impacts.size());
    L.recs = c.take<SpecialRec>(S.special.size());
    L.ghosts = c.take<GhostDev>(S.ghosts.size());
    L.op_ptr = c.take<int32_t>(S.op_ptr.size());
    L.op_idx = c.take<int32_t>(S.op_idx.size());
    L.op_w = c.take<double>(S.op_w.size());
    L.oc_ptr = c.take<int32_t>(S.oc_ptr.size());
    L.oc_ch = c.take<int32_t>(S.oc_ch.size());
    L.oc_w = c.take<double>(S.oc_w.size());
    L.terms = c.take<double>((size_t)cap * std::max<size_t>(1, S.lst_terms.size()));
    L.clock = c.take<unsigned long long>(CLK_WORDS);
    L.t_tab = c.take<double>(AMP_ROWS);
    L.acc_tab = c.take<double>((size_t)AMP_ROWS * S.objs.size() * S.n_cpo);
    L.src_tab = c.take<float>((size_t)AMP_ROWS * S.psrc.size());
    L.spk = c.take<double>(8 * S.spk.size());
    L.spk_samples = c.take<double>(S.spk_samples.size());
    L.spk_field = c.take<float>(S.spk.empty() ? 0 : (size_t)(S.spk_hi[0] - S.spk_lo[0]) * (S.spk_hi[1] - S.spk_lo[1]) * (S.spk_hi[2] - S.spk_lo[2]));
    L.ucls = c.take<uint8_t>(2 * moving_cells(S));
    L.mobj = c.take<MoveDev>(S.moving ? S.objs.size() : 0);
    L.kin = c.take<double>(S.kin.size());
    L.pose = c.take<Geom>(S.moving ? (size_t)AMP_ROWS * S.objs.size() : 0);
    L.ughost = c.take<double>(6 * moving_cells(S));
    L.self = c.take<SceneDev>(1);
}

}  // namespace

TilePlan tile_plan(int id)
{
    // {interior k threads (log2), interior planes, PML planes, whole-row float4 interior, one launch,
    //  PML k threads (log2)[, bands of float4 rows taken together]}
    static const TilePlan plans[kNumTilePlans] = {
        {6, 2, 8, true, true, 6, 4}, {6, 4, 4, true, true, 6, 8}, {6, 8, 8, true, true, 6},   {6, 4, 8, true, false, 6},
        {6, 8, 8, false, false, 6}, {6, 4, 8, true, true, 5},   {6, 4, 16, true, true, 6},  {6, 4, 8, true, true, 6},
        {6, 4, 4, false, false, 6}, {6, 4, 8, true, true, 4},   {6, 2, 2, true, true, 6},
    };
    return plans[(id >= 0 && id < kNumTilePlans) ? id : 0];
}

// One fused launch; the collocated scheme marches 2 interior and 8 PML
// planes (plan 0), the staggered scheme 4 and 4 (plan 1: 199 us per step at
// 256^3, against 207 us with plan 0 and 227 us as two launches).  A grid this
// small is latency-bound and takes 2-plane PML marches (plan 10: 6.2 us per
// step at 64^3 against 9.8; the same at 80^3, slower from 128^3 on).
int default_tile_plan(const SceneSetup &S)
{
    if (S.scheme == FDTD_SCHEME_STAGGERED) return 1;
    return S.N < 400000 ? 10 : 0;
}

size_t DeviceScene::bytes_needed(const SceneSetup &S, int series_cap)
{
    Carver c;
    Layout L;
    carve(S, series_cap, c, c, L);
    return c.off;
}

void DeviceScene::region_bytes(const SceneSetup &S, int series_cap, size_t *state, size_t *tables)
{
    Carver cs, ct;
    Layout L;
    carve(S, series_cap, cs, ct, L);
    *state = cs.off;
    *tables = ct.off;
}

fdtd_status DeviceScene::reset_to_start(hipStream_t st)
{
    FDTD_HIP_TRY(zero_state(st));
    FDTD_HIP_TRY(hipMemsetAsync(dev.terms, 0, sizeof(double) * (size_t)series_cap * std::max<size_t>(1, S.lst_terms.size()), st));
    unsigned long long clk[CLK_WORDS];
    clock_words(clk, 0, (unsigned long long)target, S.t_start, S.dt);
    FDTD_HIP_TRY(hipMemcpyAsync(dev.clock, clk, sizeof clk, hipMemcpyHostToDevice, st));
    FDTD_HIP_TRY(fill_amplitudes(st));
    if (S.moving)
        FDTD_HIP_TRY(hipMemcpyAsync(dev.ucls + moving_cells(S), S.mv_cls0.data(), S.mv_cls0.size(), hipMemcpyHostToDevice, st));
    FDTD_HIP_TRY(hipStreamSynchronize(st));
    return FDTD_OK;
}

fdtd_status DeviceScene::allocate_and_upload(hipStream_t st, int cap, long long target_steps)
{
    size_t state_b = 0, table_b = 0;
    region_bytes(S, cap, &state_b, &table_b);
    arena_bytes = state_b + table_b;
    FDTD_HIP_TRY(hipMalloc(&arena, arena_bytes));
    return place(st, cap, target_steps, arena, arena + state_b);
}

fdtd_status DeviceScene::place(hipStream_t st, int cap, long long target_steps, char *state_base, char *table_base)
{
    const long long target = target_steps;
    this->target = target_steps;
    series_cap = cap;
    npad = slot_cells(S);
    sy_dev = row_stride(S);
    Carver cs, c;
    cs.base = state_base;
    c.base = table_base;
    Layout L;
    carve(S, cap, cs, c, L);
    state_floats = L.state_floats;
    // fields start at rest: pressures, split pressures and both face-velocity buffers
    FDTD_HIP_TRY(hipMemsetAsync(L.field, 0, sizeof(float) * state_floats, st));
    // terms no tile writes (samples of cells that never change) read as zero
    FDTD_HIP_TRY(hipMemsetAsync(L.terms, 0, sizeof(double) * (size_t)cap * std::max<size_t>(1, S.lst_terms.size()), st));
    // the acceleration ring reads as zero where no object refills it (a
    // constraint's ghost entries weight it by zero)
    FDTD_HIP_TRY(hipMemsetAsync(L.acc_tab, 0, sizeof(double) * std::max<size_t>(1, (size_t)AMP_ROWS * S.objs.size() * S.n_cpo), st));
    std::vector<char> staging;  // pageable copies are synchronised before return
    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        if (!bytes) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    std::vector<float4> coef[3];
    std::vector<int> flags[3];
    for (int d = 0; d < 3; ++d) {
        const int n = S.n[d];
        coef[d].resize(n + 1);
        flags[d].resize(n + 1);
        for (int i = 0; i <= n; ++i) {
            coef[d][i] = make_float4(S.f1[d][i], S.f2[d][i], i < n ? S.f3[d][i] : 0.f, i < n ? S.f5h[d][i] : 0.f);
            flags[d][i] = (i < n && S.cpml[d][i] ? 1 : 0) | (S.fpml[d][i] ? 2 : 0) | (i < n && S.sig0[d][i] ? 4 : 0) |
                          (axis_plain(S, d, i) ? 8 : 0) | (i < n && S.csolid[d][i] ? 16 : 0);
        }
        FDTD_HIP_TRY(up(L.coef[d], coef[d].data(), coef[d].size() * sizeof(float4)));
        FDTD_HIP_TRY(up(L.flags[d], flags[d].data(), flags[d].size() * sizeof(int)));
    }
    // the host's cell indices are dense (k fastest, row length n2); device
    // arrays index padded rows
    std::vector<uint8_t> skip;
    if (S.has_solid) {
        skip.assign((size_t)npad, 0);
        for (int64_t i = 0; i < S.N; ++i) skip[to_dev(i)] = S.bulk[i] ? 0 : 1;
        for (const SpecialRec &r : S.special) skip[to_dev(r.cell)] = 1;
        // a moving scene's reclassified box belongs to its own tiles
        if (S.moving)
            for (int i = S.mv_lo[0]; i < S.mv_hi[0]; ++i)
                for (int j = S.mv_lo[1]; j < S.mv_hi[1]; ++j)
                    for (int k = S.mv_lo[2]; k < S.mv_hi[2]; ++k) skip[(size_t)i * S.n[1] * sy_dev + (size_t)j * sy_dev + k] = 1;
        FDTD_HIP_TRY(up(L.skip, skip.data(), skip.size()));
    }
    if (!S.spk.empty()) {
        std::vector<double> sp;
        for (const SceneSetup::SpeakerHost &q : S.spk) {
            const double v[8] = {q.p[0], q.p[1], q.p[2], q.w, q.t0, q.period, (double)q.off, (double)q.n};
            sp.insert(sp.end(), v, v + 8);
        }
        FDTD_HIP_TRY(up(L.spk, sp.data(), sp.size() * 8));
        FDTD_HIP_TRY(up(L.spk_samples, S.spk_samples.data(), S.spk_samples.size() * 8));
    }
    if (S.moving) {
        std::vector<MoveDev> mv(S.objs.size());
        for (size_t q = 0; q < S.objs.size(); ++q) {
            mv[q].g0 = S.objs[q];
            for (int d = 0; d < 3; ++d) { mv[q].v[d] = S.mv_v[3 * q + d]; mv[q].w[d] = S.mv_w[3 * q + d]; }
            mv[q].kin_off = S.kin_off[q];
            mv[q].kin_n = S.kin_n[q];
            mv[q].kin_dt = S.kin_dt[q];
        }
        FDTD_HIP_TRY(up(L.mobj, mv.data(), mv.size() * sizeof(MoveDev)));
        FDTD_HIP_TRY(up(L.kin, S.kin.data(), S.kin.size() * sizeof(double)));
        // the flags of step -1 (row 1): the classification at the start pose
        FDTD_HIP_TRY(up(L.ucls + moving_cells(S), S.mv_cls0.data(), S.mv_cls0.size()));
    }
    FDTD_HIP_TRY(up(L.gtab, S.gtab.data(), S.gtab.size() * 4));
    FDTD_HIP_TRY(up(L.slabs, S.slabs.data(), S.slabs.size() * sizeof(PmlSlab)));
    std::vector<PSrcDev> ps;
    for (const PSrcHost &h : S.psrc) {
        PSrcDev d{};
        d.t0 = h.t0; d.wt = h.wt; d.amp = h.amp;
        for (int e = 0; e < 3; ++e) { d.lo[e] = h.lo[e]; d.hi[e] = h.hi[e]; d.goff[e] = h.goff[e]; }
        ps.push_back(d);
    }
    FDTD_HIP_TRY(up(L.psrc, ps.data(), ps.size() * sizeof(PSrcDev)));
    std::vector<ImpDev> is;
    for (const ImpulseHost &h : S.isrc)
        is.push_back({h.obj, h.kind, h.A, h.w, h.ton, h.omega, h.phase, h.decay, h.dcenter, h.t0, h.t1, h.omega1});
    FDTD_HIP_TRY(up(L.isrc, is.data(), is.size() * sizeof(ImpDev)));
    std::vector<ImpactDev> im;
    for (const ImpactHost &h : S.impacts) im.push_back({h.obj, 0, h.t0, h.tau, {h.v[0], h.v[1], h.v[2]}});
    FDTD_HIP_TRY(up(L.impacts, im.data(), im.size() * sizeof(ImpactDev)));
    std::vector<SpecialRec> recs(S.special);
    for (SpecialRec &r : recs) r.cell = (int32_t)to_dev(r.cell);
    FDTD_HIP_TRY(up(L.recs, recs.data(), recs.size() * sizeof(SpecialRec)));
    std::vector<GhostDev> gd;
    for (const GhostEntry &g : S.ghosts)
        gd.push_back({(int)to_dev(g.nbr), g.obj, g.coupled, std::abs(g.topology) - 1, g.k_raster, g.k_vec});
    FDTD_HIP_TRY(up(L.ghosts, gd.data(), gd.size() * sizeof(GhostDev)));
    FDTD_HIP_TRY(up(L.op_ptr, S.op_ptr.data(), S.op_ptr.size() * 4));
    std::vector<int32_t> op_idx(S.op_idx);
    for (int32_t &c : op_idx) c = (int32_t)to_dev(c);
    FDTD_HIP_TRY(up(L.op_idx, op_idx.data(), op_idx.size() * 4));
    FDTD_HIP_TRY(up(L.op_w, S.op_w.data(), S.op_w.size() * 8));
    FDTD_HIP_TRY(up(L.oc_ptr, S.oc_ptr.data(), S.oc_ptr.size() * 4));
    FDTD_HIP_TRY(up(L.oc_ch, S.oc_ch.data(), S.oc_ch.size() * 4));
    FDTD_HIP_TRY(up(L.oc_w, S.oc_w.data(), S.oc_w.size() * 8));
    unsigned long long clk[CLK_WORDS];
    clock_words(clk, 0, (unsigned long long)target, S.t_start, S.dt);
    FDTD_HIP_TRY(up(L.clock, clk, sizeof clk));
    FDTD_HIP_TRY(hipStreamSynchronize(st));

    SceneDev &D = dev;
    std::memset(&D, 0, sizeof D);
    for (int d = 0; d < 3; ++d) D.n[d] = S.n[d];
    D.sy = sy_dev;
    D.sx = S.n[1] * sy_dev;
    D.raster = S.boundary == FDTD_BOUNDARY_RASTERIZE ? 1 : 0;
    D.scheme = S.scheme;
    D.npad = npad;
    D.field = L.field;
    D.slabs = L.slabs;
    for (int d = 0; d < 3; ++d) {
        D.pd[d] = L.pd[d];
        D.vel[0][d] = L.vel[0][d];
        D.vel[1][d] = L.vel[1][d];
    }
    for (int d = 0; d < 3; ++d) { D.coef[d] = L.coef[d]; D.flags[d] = L.flags[d]; }
    D.skip = !S.has_solid ? nullptr : L.skip;
    D.lam2 = S.lam2;
    D.lam2f = (float)S.lam2;
    D.lam = S.c * S.dt / S.h;
    D.src_steps = S.src_steps;
    D.lapv = L.lapv;
    D.kcalp = S.kcalp;
    D.c2k2 = S.c2k2;
    D.inv_h2 = 1.0 / (S.h * S.h);
    D.kcalpf = (float)S.kcalp;
    D.c2k2_kcalpf = (float)(S.c2k2 + S.kcalp);
    D.inv_h2f = (float)(1.0 / (S.h * S.h));
    D.t_start = S.t_start;
    D.dt = S.dt;
    D.n_psrc = (int)S.psrc.size();
    D.n_isrc = (int)S.isrc.size();
    D.n_special = (int)S.special.size();
    D.psrc = L.psrc;
    D.gtab = L.gtab;
    D.isrc = L.isrc;
    D.n_impact = (int)S.impacts.size();
    D.impacts = L.impacts;
    D.recs = L.recs;
    D.ghosts = L.ghosts;
    D.G = (int)S.ghosts.size();
    D.op_ptr = L.op_ptr; D.op_idx = L.op_idx; D.op_w = L.op_w;
    D.oc_ptr = L.oc_ptr; D.oc_ch = L.oc_ch; D.oc_w = L.oc_w;
    D.T = (int)S.lst_terms.size();
    D.cap = cap;
    D.terms = L.terms;
    D.clock = L.clock;
    D.n_obj = (int)S.objs.size();
    D.seven = S.sampling == FDTD_SAMPLE_SEVEN ? 1 : 0;
    D.n_cpo = S.n_cpo;
    D.n_ch = D.n_obj * S.n_cpo;
    D.t_tab = L.t_tab;
    D.acc_tab = L.acc_tab;
    D.src_tab = L.src_tab;
    D.n_spk = (int)S.spk.size();
    D.spk = L.spk;
    D.spk_samples = L.spk_samples;
    D.c = S.c;
    for (int d = 0; d < 3; ++d) { D.sp_lo[d] = S.spk_lo[d]; D.sp_n[d] = S.spk.empty() ? 0 : S.spk_hi[d] - S.spk_lo[d]; }
    D.spk_field = L.spk_field;
    D.mv = S.moving ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
        D.u_lo[d] = S.mv_lo[d];
        D.u_n[d] = S.moving ? S.mv_hi[d] - S.mv_lo[d] : 0;
        D.bmin[d] = S.bmin[d];
    }
    D.ucls = L.ucls;
    D.mobj = L.mobj;
    D.kin = L.kin;
    D.pose = L.pose;
    D.ughost = L.ughost;
    D.h = S.h;
    D.rho = S.rho;
    self = L.self;
    FDTD_HIP_TRY(hipMemcpyAsync(self, &D, sizeof D, hipMemcpyHostToDevice, st));
    FDTD_HIP_TRY(fill_amplitudes(st));
    FDTD_HIP_TRY(hipStreamSynchronize(st));
    return FDTD_OK;
}

hipError_t DeviceScene::download_grid(const float *src, float *dense) const
{
    return hipMemcpy2D(dense, sizeof(float) * S.n[2], src, sizeof(float) * sy_dev, sizeof(float) * S.n[2],
                       (size_t)S.n[0] * S.n[1], hipMemcpyDeviceToHost);
}

hipError_t DeviceScene::upload_grid(float *dst, const float *dense) const
{
    return hipMemcpy2D(dst, sizeof(float) * sy_dev, dense, sizeof(float) * S.n[2], sizeof(float) * S.n[2],
                       (size_t)S.n[0] * S.n[1], hipMemcpyHostToDevice);
}

fdtd_status DeviceScene::read_series(long long ring_row, long long nrows, double *out) const
{
    const size_t T = S.lst_terms.size();
    if (nrows <= 0 || S.L == 0) return FDTD_OK;
    std::vector<double> rows((size_t)nrows * T);
    if (T) FDTD_HIP_TRY(hipMemcpy(rows.data(), dev.terms + (size_t)ring_row * T, rows.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (long long r = 0; r < nrows; ++r)
        for (int l = 0; l < S.L; ++l) {
            double v = 0.0;
            for (int q = S.lst_ptr[l]; q < S.lst_ptr[l + 1]; ++q) v += rows[(size_t)r * T + q];
            out[(size_t)r * S.L + l] = v;
        }
    return FDTD_OK;
}

void DeviceScene::release()
{
    if (arena) {
        (void)hipFree(arena);
        arena = nullptr;
    }
}

static int pow2_ceil_log2(int x)
{
    int l = 0;
    while ((1 << l) < x) ++l;
    return l;
}

static bool box_overlap(const int *lo, const int *hi, const int *blo, const int *bhi)
{
    for (int d = 0; d < 3; ++d)
        if (!(lo[d] < bhi[d] && hi[d] > blo[d])) return false;
    return true;
}

void DeviceScene::append_tiles(int scene_id, const TilePlan &plan, std::vector<TileDesc> &out, std::vector<TermRef> &terms) const
{
    const size_t base = out.size();
    const int *n = S.n;
    const int *a = S.tile_lo, *b = S.tile_hi;
    auto src_hit = [&](const int *lo, const int *hi) {
        for (const PSrcHost &p : S.psrc)
            if (box_overlap(lo, hi, p.lo, p.hi)) return 1;
        for (const SceneSetup::SpeakerHost &p : S.spk)
            if (box_overlap(lo, hi, p.lo, p.hi)) return 1;
        return 0;
    };
    auto special_hit = [&](const int *lo, const int *hi) {
        for (const SpecialBox &bx : S.boxes)
            if (box_overlap(lo, hi, bx.lo, bx.hi)) return true;
        return S.moving && box_overlap(lo, hi, S.mv_lo, S.mv_hi);
    };
    // PML slabs first (heaviest tiles start first)
    std::vector<size_t> row_end;   // positions in out of tiles that ride with the row tiles
    for (int q = 0; q < (int)S.slabs.size(); ++q) {
        const PmlSlab &sl = S.slabs[q];
        const int tk = std::min(plan.pml_tk_log2, pow2_ceil_log2(sl.e[2]));
        const int TK = 1 << tk, TJ = 256 >> tk;
        for (int i0 = sl.lo[0]; i0 < sl.lo[0] + sl.e[0]; i0 += plan.pml_ti)
            for (int j0 = sl.lo[1]; j0 < sl.lo[1] + sl.e[1]; j0 += TJ)
                for (int k0 = sl.lo[2]; k0 < sl.lo[2] + sl.e[2]; k0 += TK) {
                    TileDesc t{};
                    t.scene = scene_id;
                    t.type = S.scheme == FDTD_SCHEME_STAGGERED ? TILE_MAC_PML : TILE_PML;
                    t.i0 = i0; t.i1 = std::min(i0 + plan.pml_ti, sl.lo[0] + sl.e[0]);
                    t.j0 = j0; t.j1 = std::min(j0 + TJ, sl.lo[1] + sl.e[1]);
                    t.k0 = k0; t.k1 = std::min(k0 + TK, sl.lo[2] + sl.e[2]);
                    t.tk_log2 = tk;
                    t.slab = q;
                    if (S.scheme == FDTD_SCHEME_STAGGERED) {
                        const int lo[3] = {t.i0, t.j0, t.k0}, hi[3] = {t.i1, t.j1, t.k1};
                        t.src = src_hit(lo, hi);
                    }
                    if (sl.axis == 2 && sl.e[2] >= 12) row_end.push_back(out.size());
                    out.push_back(t);
                }
    }
    // Engquist boundary: the layer-1 shell as six disjoint boxes (the x faces
    // whole, the y faces inside x, the z faces inside x and y), cut into
    // tiles of at most 256 cells.  A z-face cell is one float of a row's
    // first or last line: when the float4 rows run in the fused launch, the
    // z faces are cut into tiles of 64 planes by 4 rows that follow the row
    // tile holding their last column, while those lines are in the L2
    // (256^3 free field 45.8 -> 41.0-41.4 us per step; the same tiles in
    // the boundary class measured 46.4)
    if (S.abc) {
        const bool ride_rows = plan.fused && plan.interior_vec && !S.visc && S.scheme != FDTD_SCHEME_STAGGERED &&
                               b[2] - (a[2] & ~3) <= 256;
        const int boxes[6][6] = {
            {1, 2, 1, n[1] - 1, 1, n[2] - 1},        {n[0] - 2, n[0] - 1, 1, n[1] - 1, 1, n[2] - 1},
            {2, n[0] - 2, 1, 2, 1, n[2] - 1},        {2, n[0] - 2, n[1] - 2, n[1] - 1, 1, n[2] - 1},
            {2, n[0] - 2, 2, n[1] - 2, 1, 2},        {2, n[0] - 2, 2, n[1] - 2, n[2] - 2, n[2] - 1},
        };
        for (int f = 0; f < 6; ++f) {
            const int *bx = boxes[f];
            const bool zface = ride_rows && f >= 4;
            const int ek = bx[5] - bx[4], ej = bx[3] - bx[2], ei = bx[1] - bx[0];
            const int tk = std::min(ek, 256), tj = zface ? std::min(ej, 4) : std::min(ej, std::max(1, 256 / tk));
            const int ti = zface ? std::min(ei, 64) : std::min(ei, std::max(1, 256 / (tk * tj)));
            for (int i0 = bx[0]; i0 < bx[1]; i0 += ti)
                for (int j0 = bx[2]; j0 < bx[3]; j0 += tj)
                    for (int k0 = bx[4]; k0 < bx[5]; k0 += tk) {
                        TileDesc t{};
                        t.scene = scene_id;
                        t.type = TILE_ABC;
                        t.i0 = i0; t.i1 = std::min(i0 + ti, bx[1]);
                        t.j0 = j0; t.j1 = std::min(j0 + tj, bx[3]);
                        t.k0 = k0; t.k1 = std::min(k0 + tk, bx[5]);
                        const int lo[3] = {t.i0, t.j0, t.k0}, hi[3] = {t.i1, t.j1, t.k1};
                        t.src = src_hit(lo, hi);
                        if (zface) row_end.push_back(out.size());
                        out.push_back(t);
                    }
        }
    }
    // moving objects: the reclassified box in tiles of at most 256 cells
    if (S.moving) {
        const int *bl = S.mv_lo, *bh = S.mv_hi;
        const int ek = bh[2] - bl[2], ej = bh[1] - bl[1], ei = bh[0] - bl[0];
        const int tk = std::min(ek, 256), tj = std::min(ej, std::max(1, 256 / tk));
        const int ti = std::min(ei, std::max(1, 256 / (tk * tj)));
        for (int i0 = bl[0]; i0 < bh[0]; i0 += ti)
            for (int j0 = bl[1]; j0 < bh[1]; j0 += tj)
                for (int k0 = bl[2]; k0 < bh[2]; k0 += tk) {
                    TileDesc t{};
                    t.scene = scene_id;
                    t.type = TILE_MOVING;
                    t.i0 = i0; t.i1 = std::min(i0 + ti, bh[0]);
                    t.j0 = j0; t.j1 = std::min(j0 + tj, bh[1]);
                    t.k0 = k0; t.k1 = std::min(k0 + tk, bh[2]);
                    const int lo[3] = {t.i0, t.j0, t.k0}, hi[3] = {t.i1, t.j1, t.k1};
                    t.src = src_hit(lo, hi);
                    out.push_back(t);
                }
    }
    // object-adjacent cells
    for (int r0 = 0; r0 < (int)S.special.size(); r0 += 256) {
        TileDesc t{};
        t.scene = scene_id;
        t.type = TILE_SPECIAL;
        t.i0 = r0;
        t.i1 = std::min(r0 + 256, (int)S.special.size());
        out.push_back(t);
    }
    // interior: whole-row float4 tiles of four rows where the layout allows
    // and the rows hold no source and no object cell; 2D chunks otherwise
    const int tk = std::min(plan.interior_tk_log2, std::max(0, pow2_ceil_log2(b[2] - a[2])));
    const int TK = 1 << tk, TJ = 256 >> tk;
    const int kw = b[2] - a[2];
    const bool mac = S.scheme == FDTD_SCHEME_STAGGERED;
    // float4 rows start at the multiple of four at or below a[2] (the cells
    // before a[2] are read and not written); the row stride is a multiple of
    // four too, and a row's last group may be partial
    const int kv0 = a[2] & ~3, lead = a[2] - kv0, kwv = b[2] - kv0;
    // staggered rows take one coefficient entry per axis: every face of the
    // box (a .. b inclusive) and every cell must hold the same table values
    bool mac_uniform = true;
    if (mac)
        for (int d = 0; d < 3; ++d)
            for (int i = a[d]; i <= b[d]; ++i) {
                mac_uniform = mac_uniform && S.f1[d][i] == S.f1[d][a[d]] && S.f2[d][i] == S.f2[d][a[d]];
                if (i < b[d]) mac_uniform = mac_uniform && S.f5h[d][i] == S.f5h[d][a[d]];
            }
    const bool vec = !S.visc && plan.interior_vec && kwv <= 256 && (!mac || mac_uniform);
    auto scalar_tiles = [&](int i0, int i1, int jlo, int jhi) {
        for (int j0 = jlo; j0 < jhi; j0 += TJ)
            for (int k0 = a[2]; k0 < b[2]; k0 += TK) {
                TileDesc t{};
                t.scene = scene_id;
                t.i0 = i0; t.i1 = i1;
                t.j0 = j0; t.j1 = std::min(j0 + TJ, jhi);
                t.k0 = k0; t.k1 = std::min(k0 + TK, b[2]);
                t.tk_log2 = tk;
                const int lo[3] = {t.i0, t.j0, t.k0}, hi[3] = {t.i1, t.j1, t.k1};
                t.type = mac ? TILE_MAC : special_hit(lo, hi) ? TILE_INTERIOR_SKIP : TILE_INTERIOR;
                t.src = src_hit(lo, hi);
                out.push_back(t);
            }
    };
    if (!vec) {
        for (int i0 = a[0]; i0 < b[0]; i0 += plan.interior_ti) scalar_tiles(i0, std::min(i0 + plan.interior_ti, b[0]), a[1], b[1]);
    } else {
        // short rows share a wavefront: rpw rows of kwv/4 lanes each.  The
        // tiles of a band of rows follow each other along x, so a tile's x
        // halo (two planes of its rows, twice its y halo) was just read by the
        // tile before it and is still in the L2 (at 256^3: 73.4 us per step
        // against 75.6 with the y-neighbours adjacent; the Engquist free field
        // 46.0 against 51.9).  With PML slabs, G = band_group bands advance
        // along x together (G tiles per x step), so the rows between them are
        // read once for both neighbours' y halos: 72.4 us against 73.2 at
        // 256^3 with G = 4 (2 and 8: 73.1, 72.9); the staggered scheme (plan
        // 1, marches of four planes) takes G = 8: 189.7 us against 194.2 with
        // 4 and 207.4 with 2.  The Engquist free field, all rows, measured
        // 46.7 against 45.8 with groups, so it keeps single bands.
        const int rpw = std::max(1, std::min(8, 64 / ((kwv + 3) / 4)));
        const int G = S.slabs.empty() ? 1 : std::max(1, plan.band_group);
        // A z slab's row segment and the row's interior part share a 128-byte
        // line at each end of the row (16 PML cells, 64 bytes, at 256^3).  In
        // the fused launch each z-slab tile (and each Engquist z-face tile)
        // follows the row tile that holds its last (i, j) column, so both
        // halves of those lines are fetched
        // once: 256^3 step 72.5 -> 71.0 us, the staggered scheme 0.424 ->
        // 0.436 of peak, 128^3 with a 12-cell PML 19.5 -> 19.3 us.  Slabs
        // under 12 cells deep keep their place: the 64-scene batch (8-cell
        // slabs) measured 124.7 -> 121.8 Gcell/s with them moved.  (Listed
        // after the scene's interior tiles, a slab tile keeps its place in the
        // launch; build_launch_lists.)
        auto at = [](int i, int j) { return ((uint64_t)(uint32_t)i << 32) | (uint32_t)j; };
        std::unordered_map<uint64_t, std::vector<TileDesc>> ride;
        if (plan.fused && !row_end.empty()) {
            std::vector<char> moved(out.size(), 0);
            for (size_t q : row_end) {
                const TileDesc &z = out[q];
                const int li = z.i1 - 1, lj = z.j1 - 1;
                if (li < a[0] || li >= b[0] || lj < a[1] || lj >= b[1]) continue;
                const int i0t = a[0] + (li - a[0]) / plan.interior_ti * plan.interior_ti;
                const int j0t = a[1] + (lj - a[1]) / (4 * rpw) * (4 * rpw);
                ride[at(i0t, j0t)].push_back(z);
                moved[q] = 1;
            }
            size_t w = base;
            for (size_t q = base; q < out.size(); ++q)
                if (!moved[q]) out[w++] = out[q];
            out.resize(w);
        }
        for (int jg = a[1]; jg < b[1]; jg += G * 4 * rpw) {
            for (int i0 = a[0]; i0 < b[0]; i0 += plan.interior_ti) {
                const int i1 = std::min(i0 + plan.interior_ti, b[0]);
                for (int j0 = jg; j0 < std::min(b[1], jg + G * 4 * rpw); j0 += 4 * rpw) {
                    const int j1 = std::min(j0 + 4 * rpw, b[1]);
                    const int lo[3] = {i0, j0, a[2]}, hi[3] = {i1, j1, b[2]};
                    if (special_hit(lo, hi) || src_hit(lo, hi)) {
                        scalar_tiles(i0, i1, j0, j1);
                    } else {
                        TileDesc t{};
                        t.scene = scene_id;
                        t.type = mac ? TILE_MAC_VEC : TILE_INTERIOR_VEC;
                        t.i0 = i0; t.i1 = i1; t.j0 = j0; t.j1 = j1; t.k0 = kv0; t.k1 = b[2];
                        t.tk_log2 = 6;
                        t.rpw = rpw;
                        t.lead = lead;
                        out.push_back(t);
                    }
                    if (ride.empty()) continue;
                    auto it = ride.find(at(i0, j0));
                    if (it == ride.end()) continue;
                    out.insert(out.end(), it->second.begin(), it->second.end());
                    ride.erase(it);
                }
            }
        }
        for (auto &kv : ride) out.insert(out.end(), kv.second.begin(), kv.second.end());
    }
    assign_terms(scene_id, base, out, terms);
}

// Gives every listener term to the tile that produces its sample: a cell's
// new pressure comes from the object-adjacent tile holding its record, else
// from the PML or interior tile whose box holds it; ghost-entry terms go to
// the tile of the entry's fluid neighbour.
void DeviceScene::assign_terms(int scene_id, size_t base, std::vector<TileDesc> &out, std::vector<TermRef> &terms) const
{
    const int nterm = (int)S.lst_terms.size();
    std::vector<int> owner(nterm, -1);
    std::unordered_map<int, int> record_of;
    for (int r = 0; r < (int)S.special.size(); ++r) record_of[S.special[r].cell] = r;
    const int sy = S.n[2], sx = S.n[1] * S.n[2];
    // the cell tiles by (i, j) column, and the record tiles by their first
    // record: each term then looks at the few tiles over its column
    std::vector<std::vector<int>> column((size_t)S.n[0] * S.n[1]);
    std::vector<std::pair<int, int>> rec_tiles;   // (first record, tile)
    for (size_t q = base; q < out.size(); ++q) {
        const TileDesc &t = out[q];
        if (t.type == TILE_SPECIAL) { rec_tiles.push_back({t.i0, (int)q}); continue; }
        if (t.type == TILE_EMPTY) continue;
        for (int i = t.i0; i < t.i1; ++i)
            for (int j = t.j0; j < t.j1; ++j) column[(size_t)i * S.n[1] + j].push_back((int)q);
    }
    std::sort(rec_tiles.begin(), rec_tiles.end());
    auto tile_of_cell = [&](int cell) -> int {
        auto it = record_of.find(cell);
        if (it != record_of.end()) {
            // the record tile whose range [i0, i1) holds the record
            auto r = std::upper_bound(rec_tiles.begin(), rec_tiles.end(), std::make_pair(it->second, INT_MAX));
            if (r == rec_tiles.begin()) return -1;
            --r;
            const TileDesc &t = out[r->second];
            return (it->second >= t.i0 && it->second < t.i1) ? r->second : -1;
        }
        const int i = cell / sx, j = (cell - i * sx) / sy, k = cell - i * sx - j * sy;
        int best = -1;
        for (int q : column[(size_t)i * S.n[1] + j]) {
            const TileDesc &t = out[q];
            const int kfirst = t.k0 + ((t.type == TILE_INTERIOR_VEC || t.type == TILE_MAC_VEC) ? t.lead : 0);   // cells the tile writes
            if (k >= kfirst && k < t.k1 && (best < 0 || q < best)) best = q;
        }
        return best;
    };
    for (int q = 0; q < nterm; ++q) {
        const ListenerTerm &lt = S.lst_terms[q];
        owner[q] = tile_of_cell(lt.kind == 0 ? lt.index : S.ghosts[lt.index].nbr);
    }
    std::vector<std::vector<int>> by_tile(out.size() - base);
    for (int q = 0; q < nterm; ++q)
        if (owner[q] >= 0) by_tile[owner[q] - base].push_back(q);
    for (size_t q = base; q < out.size(); ++q) {
        TileDesc &t = out[q];
        t.tb = (int)terms.size();
        for (int qq : by_tile[q - base]) {
            const ListenerTerm &lt = S.lst_terms[qq];
            terms.push_back({scene_id, qq, lt.kind, lt.kind == 0 ? (int)to_dev(lt.index) : lt.index, lt.beta});
        }
        t.te = (int)terms.size();
    }
}

void xcd_permute(std::vector<TileDesc> &tiles, int nxcd)
{
    if (tiles.size() < (size_t)(2 * nxcd)) return;
    TileDesc empty{};
    empty.type = TILE_EMPTY;
    while (tiles.size() % nxcd) tiles.push_back(empty);
    const size_t per = tiles.size() / nxcd;
    std::vector<TileDesc> out(tiles.size());
    for (size_t w = 0; w < tiles.size(); ++w) out[w] = tiles[(w % nxcd) * per + w / nxcd];
    tiles.swap(out);
}

// Host check of a launch list before anything runs on the device: every tile
// belongs to a class the list's kernel handles, and its box keeps every
// neighbour read of that class inside the grid (interior tiles read one cell
// beyond the box on each side; whole-row float4 tiles also need aligned,
// full-width rows).  Returns an error text, or empty.
static std::string check_tiles(const std::vector<TileDesc> &list, int kind, const std::vector<SceneDev> &scenes,
                               size_t nterms)
{
    for (const TileDesc &t : list) {
        if (t.type == TILE_EMPTY) continue;
        if (t.scene < 0 || t.scene >= (int)scenes.size()) return "tile scene out of range";
        const SceneDev &d = scenes[t.scene];
        if (t.tb < 0 || t.te < t.tb || (size_t)t.te > nterms) return "tile term range out of range";
        const bool pml_class = t.type == TILE_PML || t.type == TILE_MAC_PML || t.type == TILE_ABC;
        const bool moving = t.type == TILE_MOVING;
        const bool ok_kind = moving ? (kind == KIND_SPECIAL || kind == KIND_ALL)
                                    : (kind == KIND_ALL || (kind == KIND_PML && pml_class) ||
                                       (kind == KIND_INTERIOR && !pml_class) || (kind == KIND_SPECIAL && t.type == TILE_SPECIAL));
        if ((t.type == TILE_MAC || t.type == TILE_MAC_PML || t.type == TILE_MAC_VEC) != (d.scheme == FDTD_SCHEME_STAGGERED))
            return "tile class does not match the scene's scheme";
        if (t.type == TILE_MAC_VEC && kind != KIND_INTERIOR && kind != KIND_ALL)
            return "staggered float4 rows outside the interior and fused launches";
        if (!ok_kind) return "tile class does not match its launch";
        if (t.type == TILE_SPECIAL) {
            if (t.i0 < 0 || t.i1 > d.n_special || t.i1 - t.i0 > 256) return "special tile range out of range";
            continue;
        }
        if (moving) {
            // inside the scene's reclassified box, which keeps a cell from the grid's edge
            const int lo[3] = {t.i0, t.j0, t.k0}, hi[3] = {t.i1, t.j1, t.k1};
            for (int a = 0; a < 3; ++a)
                if (!d.mv || lo[a] < d.u_lo[a] || hi[a] > d.u_lo[a] + d.u_n[a] || lo[a] >= hi[a] || lo[a] < 1 || hi[a] > d.n[a] - 1)
                    return "moving-object tile outside the scene's reclassified box";
            if ((int64_t)(t.i1 - t.i0) * (t.j1 - t.j0) * (t.k1 - t.k0) > 256) return "moving-object tile too large";
            continue;
        }
        if (t.type == TILE_ABC) {
            // a box inside [1, n - 2] lying in one layer-1 plane: every cell is
            // a boundary cell and every neighbour it reads or writes is in the grid
            bool plane = false;
            const int lo[3] = {t.i0, t.j0, t.k0}, hi[3] = {t.i1, t.j1, t.k1};
            for (int a = 0; a < 3; ++a) {
                if (lo[a] < 1 || hi[a] > d.n[a] - 1 || lo[a] >= hi[a]) return "boundary tile outside layers 1 .. n-2";
                plane = plane || (hi[a] - lo[a] == 1 && (lo[a] == 1 || lo[a] == d.n[a] - 2));
            }
            if (!plane) return "boundary tile off the layer-1 shell";
            if ((int64_t)(t.i1 - t.i0) * (t.j1 - t.j0) * (t.k1 - t.k0) > 256) return "boundary tile too large";
            continue;
        }
        // a float4 row tile reads from k0 and writes from k0 + lead
        const bool vec_row = t.type == TILE_INTERIOR_VEC || t.type == TILE_MAC_VEC;
        const int lo[3] = {t.i0, t.j0, t.k0 + (vec_row ? t.lead : 0)}, hi[3] = {t.i1, t.j1, t.k1};
        for (int a = 0; a < 3; ++a) {
            if (lo[a] >= hi[a]) return "empty tile box";
            if (pml_class) {
                if (lo[a] < 0 || hi[a] > d.n[a]) return "PML tile outside the grid";
            } else if (lo[a] < 1 || hi[a] > d.n[a] - 1) {
                return "interior tile reads outside the grid";
            }
        }
        if (pml_class) {
            if ((t.k1 - t.k0) > (1 << t.tk_log2) || (t.j1 - t.j0) * (1 << t.tk_log2) > 256) return "PML tile too large";
        } else if (vec_row) {
            // a partial last group reads up to k0 + 4 ceil((k1 - k0) / 4) <= the row stride
            if (t.rpw < 1 || t.j1 - t.j0 > 4 * t.rpw || (t.k0 & 3) || t.lead < 0 || t.lead > 3 || t.k1 - t.k0 > 256 ||
                t.rpw * ((t.k1 - t.k0 + 3) >> 2) > 64 || (d.sy & 3) || t.k0 + ((t.k1 - t.k0 + 3) & ~3) > d.sy)
                return "float4 row tile misaligned, too wide or with too many rows";
        } else if ((t.k1 - t.k0) > (1 << t.tk_log2) || (t.j1 - t.j0) * (1 << t.tk_log2) > 256) {
            return "interior tile too large";
        }
    }
    return std::string();
}

void build_launch_lists(const std::vector<TileDesc> &tiles, bool fused_launch, std::vector<TileDesc> by_kind[NUM_KINDS],
                        int nhead[NUM_KINDS])
{
    std::vector<TileDesc> special, moving;
    for (int k = 0; k < NUM_KINDS; ++k) by_kind[k].clear();
    if (nhead)
        for (int k = 0; k < NUM_KINDS; ++k) nhead[k] = 0;
    // a scene lists its PML slab and boundary tiles before its interior
    // tiles; one listed after them is a z-slab or z-face tile that
    // DeviceScene::append_tiles placed beside the row tile it shares lines
    // with, and keeps that place
    int rows_scene = -1;
    for (const TileDesc &t : tiles) {
        if (t.type == TILE_EMPTY) continue;
        if (t.type == TILE_SPECIAL) special.push_back(t);
        else if (t.type == TILE_MOVING) moving.push_back(t);
        else if (t.type == TILE_PML || t.type == TILE_MAC_PML || t.type == TILE_ABC)
            by_kind[fused_launch && t.scene == rows_scene ? KIND_INTERIOR : KIND_PML].push_back(t);
        else { by_kind[KIND_INTERIOR].push_back(t); rows_scene = t.scene; }
    }
    // neighbouring tiles of a scene share an XCD (and its L2)
    xcd_permute(by_kind[KIND_INTERIOR]);
    xcd_permute(by_kind[KIND_PML]);
    // the object-adjacent cells (FP64 gathers, latency-bound) ride at the
    // head of the launch that carries them, one tile per XCD in turn, and
    // are dispatched first on every step (the kernel keeps this head out of
    // the odd-step reversal), so their latency hides under the streaming
    // tiles; the head is padded to a multiple of eight so every later chunk
    // still lands on a single XCD.  The moving-object boxes join that head
    // in the fused launch (their register budget is below the fused
    // kernel's); with one launch per class they keep a launch of their own,
    // which leaves the interior launch's budget to the streaming tiles.
    if (fused_launch) special.insert(special.end(), moving.begin(), moving.end());
    else by_kind[KIND_SPECIAL] = moving;
    if (!special.empty()) {
        TileDesc empty{};
        empty.type = TILE_EMPTY;
        while (special.size() % 8) special.push_back(empty);
    }
    if (fused_launch) {
        // then the PML and interior tiles interleaved in proportion, a group of
        // eight (one tile per XCD) at a time, so each XCD walks its chunk of
        // each class in order and both classes run side by side for the whole
        // step: each class alone reaches about 3.6 TB/s, together more (at
        // 256^3, 73.8-75.5 us per step against 77.0 with the PML tiles first;
        // the 64-scene batch 116.6 Gcell/s against 110.6)
        std::vector<TileDesc> all = special;
        const std::vector<TileDesc> &P = by_kind[KIND_PML], &I = by_kind[KIND_INTERIOR];
        if (P.size() % 8 || I.size() % 8) {
            all.insert(all.end(), P.begin(), P.end());
            all.insert(all.end(), I.begin(), I.end());
        } else {
            const size_t gP = P.size() / 8, gI = I.size() / 8;
            size_t p = 0, q = 0;
            while (p < gP || q < gI) {
                // a PML group while the PML share so far is at most its overall share
                const bool takeP = q >= gI || (p < gP && (double)p * (double)(gP + gI) <= (double)(p + q) * (double)gP);
                const std::vector<TileDesc> &L = takeP ? P : I;
                const size_t g = takeP ? p++ : q++;
                all.insert(all.end(), L.begin() + 8 * g, L.begin() + 8 * g + 8);
            }
        }
        by_kind[KIND_INTERIOR].swap(all);
        by_kind[KIND_PML].clear();
    } else {
        by_kind[KIND_INTERIOR].insert(by_kind[KIND_INTERIOR].begin(), special.begin(), special.end());
    }
    if (nhead) nhead[KIND_INTERIOR] = (int)special.size();
}

std::string check_launch_lists(const std::vector<TileDesc> by_kind[NUM_KINDS], bool fused_launch,
                               const std::vector<SceneDev> &scenes, size_t nterms)
{
    for (int k = 0; k < NUM_KINDS; ++k) {
        const std::string e = check_tiles(by_kind[k], (fused_launch && k != KIND_SPECIAL) ? KIND_ALL : k, scenes, nterms);
        if (!e.empty()) return e;
    }
    return std::string();
}

// Bytes of the device block of a step program: scene blocks, listener term
// references, ghost and pre-step (motion, speaker) lists, and the tile list
// of each launch.
static void program_bytes(size_t nscenes, size_t nterms, size_t npre,
                          const std::vector<TileDesc> by_kind[NUM_KINDS], size_t bytes[3 + NUM_KINDS], size_t *total)
{
    bytes[0] = align_up(sizeof(SceneDev) * std::max<size_t>(1, nscenes));
    bytes[1] = align_up(sizeof(TermRef) * std::max<size_t>(1, nterms));
    bytes[2] = align_up(sizeof(MotionDesc) * std::max<size_t>(1, npre));
    *total = bytes[0] + bytes[1] + bytes[2];
    for (int k = 0; k < NUM_KINDS; ++k) {
        bytes[3 + k] = align_up(sizeof(TileDesc) * std::max<size_t>(1, by_kind[k].size()));
        *total += bytes[3 + k];
    }
}

size_t StepProgram::bytes_needed(size_t nscenes, const std::vector<TileDesc> &tiles, bool fused_launch, size_t nterms,
                                 size_t npre)
{
    std::vector<TileDesc> by_kind[NUM_KINDS];
    build_launch_lists(tiles, fused_launch, by_kind);
    size_t bytes[3 + NUM_KINDS], total = 0;
    program_bytes(nscenes, nterms, npre, by_kind, bytes, &total);
    return total;
}

fdtd_status StepProgram::upload(const std::vector<SceneDev> &scenes, const std::vector<TileDesc> &tiles,
                                const std::vector<TermRef> &terms, hipStream_t st,
                                bool fused_launch, const std::vector<MotionDesc> &motion, char *ext_base, size_t ext_bytes,
                                char *host_stage)
{
    // scenes with speaker sources get the speaker pass before their step kernels
    std::vector<MotionDesc> speaker;
    for (size_t q = 0; q < scenes.size(); ++q)
        if (scenes[q].n_spk) speaker.push_back({(int)q});
    release();
    std::vector<TileDesc> by_kind[NUM_KINDS];
    int nhead_k[NUM_KINDS];
    build_launch_lists(tiles, fused_launch, by_kind, nhead_k);
    fused = fused_launch;
    {
        const std::string e = check_launch_lists(by_kind, fused, scenes, terms.size());
        if (!e.empty()) { set_last_error("step program: " + e); return FDTD_ERR_INVALID; }
    }
    const std::vector<SceneDev> &devs = scenes;
    nscenes = (int)devs.size();
    for (const MotionDesc &m : motion)
        if (m.scene < 0 || m.scene >= (int)devs.size() || !devs[m.scene].mv) {
            set_last_error("step program: motion list names a scene without moving objects");
            return FDTD_ERR_INVALID;
        }
    size_t bytes[3 + NUM_KINDS], total = 0;
    program_bytes(devs.size(), terms.size(), motion.size() + speaker.size(), by_kind, bytes, &total);
    char *base = nullptr;
    if (ext_base) {
        if (total > ext_bytes) {
            set_last_error("step program needs " + std::to_string(total) + " bytes, the memory given holds " +
                           std::to_string(ext_bytes));
            return FDTD_ERR_OOM;
        }
        base = ext_base;
        owns = false;
    } else {
        FDTD_HIP_TRY(hipMalloc(&base, total));
        owns = true;
    }
    d_scenes = reinterpret_cast<SceneDev *>(base);
    d_terms = reinterpret_cast<TermRef *>(base + bytes[0]);
    d_motion = reinterpret_cast<MotionDesc *>(base + bytes[0] + bytes[1]);
    // each piece goes straight to the device (pageable copies, then a stream
    // synchronisation), or into the pinned host stage at the same offset, which
    // then goes over in one asynchronous copy (the caller keeps the stage
    // intact until that copy has run)
    const bool staged = ext_base && host_stage;
    auto put = [&](void *dst, const void *src, size_t n) -> hipError_t {
        if (!n) return hipSuccess;
        if (staged) { std::memcpy(host_stage + ((char *)dst - base), src, n); return hipSuccess; }
        return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st);
    };
    size_t off = bytes[0] + bytes[1] + bytes[2];
    for (int k = 0; k < NUM_KINDS; ++k) {
        d_tiles[k] = reinterpret_cast<TileDesc *>(base + off);
        off += bytes[3 + k];
        ntiles[k] = (int)by_kind[k].size();
        head[k] = std::min(nhead_k[k], ntiles[k]);
        FDTD_HIP_TRY(put(d_tiles[k], by_kind[k].data(), sizeof(TileDesc) * ntiles[k]));
    }
    nmotion = (int)motion.size();
    motion_cells = 0;
    for (const MotionDesc &m : motion)
        motion_cells = std::max(motion_cells, devs[m.scene].u_n[0] * devs[m.scene].u_n[1] * devs[m.scene].u_n[2]);
    nspeaker = (int)speaker.size();
    speaker_cells = 0;
    for (const MotionDesc &m : speaker)
        speaker_cells = std::max(speaker_cells, devs[m.scene].sp_n[0] * devs[m.scene].sp_n[1] * devs[m.scene].sp_n[2]);
    std::vector<MotionDesc> pre(motion);
    pre.insert(pre.end(), speaker.begin(), speaker.end());
    FDTD_HIP_TRY(put(d_motion, pre.data(), sizeof(MotionDesc) * pre.size()));
    FDTD_HIP_TRY(put(d_scenes, devs.data(), sizeof(SceneDev) * devs.size()));
    FDTD_HIP_TRY(put(d_terms, terms.data(), sizeof(TermRef) * terms.size()));
    if (staged) FDTD_HIP_TRY(hipMemcpyAsync(base, host_stage, total, hipMemcpyHostToDevice, st));
    else FDTD_HIP_TRY(hipStreamSynchronize(st));
    order_clock = scenes.empty() ? nullptr : scenes[0].clock + CLK_STEP;
    return FDTD_OK;
}

void StepProgram::release()
{
    if (d_scenes && owns) (void)hipFree(d_scenes);
    d_scenes = nullptr;
    owns = true;
    for (int k = 0; k < NUM_KINDS; ++k) { d_tiles[k] = nullptr; ntiles[k] = 0; head[k] = 0; }
    d_terms = nullptr;
    d_motion = nullptr;
    nmotion = 0;
    motion_cells = 0;
    nspeaker = 0;
    speaker_cells = 0;
    nscenes = 0;
    order_clock = nullptr;
    fused = false;
}

fdtd_status StepProgram::enqueue_step(hipStream_t st, long long step, int k, hipEvent_t ev_begin, hipEvent_t ev_end) const
{
    if (nmotion) FDTD_HIP_TRY(launch_motion(d_scenes, d_motion, nmotion, motion_cells, k, st));
    if (nspeaker) FDTD_HIP_TRY(launch_speaker(d_scenes, d_motion + nmotion, nspeaker, speaker_cells, k, st));
    if (ev_begin) FDTD_HIP_TRY(hipEventRecord(ev_begin, st));
    auto launch = [&](int kind, hipStream_t s) {
        // fused: every class in the interior list, except the moving-object tiles
        if (fused && kind == KIND_INTERIOR)
            return launch_step(KIND_ALL, d_scenes, d_tiles[kind], ntiles[kind], head[kind], d_terms, order_clock, k, s);
        if (fused && kind == KIND_PML) return hipSuccess;
        return launch_step(kind, d_scenes, d_tiles[kind], ntiles[kind], head[kind], d_terms, order_clock, k, s);
    };
    // (running the PML slabs on a second stream beside the interior was
    // measured slower than one stream and is not offered)
    if ((step & 1) == 0) {
        FDTD_HIP_TRY(launch(KIND_SPECIAL, st));
        FDTD_HIP_TRY(launch(KIND_INTERIOR, st));
        FDTD_HIP_TRY(launch(KIND_PML, st));
    } else {
        FDTD_HIP_TRY(launch(KIND_PML, st));
        FDTD_HIP_TRY(launch(KIND_INTERIOR, st));
        FDTD_HIP_TRY(launch(KIND_SPECIAL, st));
    }
    if (ev_end) FDTD_HIP_TRY(hipEventRecord(ev_end, st));
    return FDTD_OK;
}

fdtd_status StepProgram::enqueue(hipStream_t st, long long step, int n, const hipEvent_t *ev) const
{
    for (int done = 0; done < n;) {
        const int blk = std::min(n - done, (int)AMP_ROWS);
        for (int k = 0; k < blk; ++k) {
            const int q = done + k;
            fdtd_status s = enqueue_step(st, step + q, k, ev ? ev[2 * q] : nullptr, ev ? ev[2 * q + 1] : nullptr);
            if (s != FDTD_OK) return s;
        }
        FDTD_HIP_TRY(launch_advance(d_scenes, nscenes, blk, st));
        done += blk;
    }
    return FDTD_OK;
}

}  // namespace fdtd
