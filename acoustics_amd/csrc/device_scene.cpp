// device_scene.cpp -- arena allocation, table upload and per-step launch
// sequence of one scene.
#include <cstring>

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
    float *p[3], *pdir[3], *v[2][3];
    uint8_t *cpml[3], *fpml[3];
    float *f1[3], *f2[3], *f3[3], *f5h[3];
    float *gtab;
    PSrcDev *psrc;
    ImpDev *isrc;
    SpecialBox *boxes;
    int32_t *cls;
    SpecialRec *recs;
    GhostDev *ghosts;
    int32_t *gk_ptr, *gk_idx, *gu_ptr, *gu_idx;
    double *gk_w, *gu_w, *gval0, *gval1, *gtmp;
    int32_t *lptr, *lkind, *lindex;
    double *lbeta, *series;
    long long *clock;
};

void carve(const SceneSetup &S, int cap, Carver &c, Layout &L)
{
    const size_t N = (size_t)S.N;
    for (int r = 0; r < 3; ++r) L.p[r] = c.take<float>(N);
    for (int d = 0; d < 3; ++d) L.pdir[d] = c.take<float>(N);
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d) L.v[b][d] = c.take<float>(N);
    for (int d = 0; d < 3; ++d) {
        L.cpml[d] = c.take<uint8_t>(S.n[d]);
        L.fpml[d] = c.take<uint8_t>(S.n[d] + 1);
        L.f1[d] = c.take<float>(S.n[d] + 1);
        L.f2[d] = c.take<float>(S.n[d] + 1);
        L.f3[d] = c.take<float>(S.n[d]);
        L.f5h[d] = c.take<float>(S.n[d]);
    }
    L.gtab = c.take<float>(S.gtab.size());
    L.psrc = c.take<PSrcDev>(S.psrc.size());
    L.isrc = c.take<ImpDev>(S.isrc.size());
    L.boxes = c.take<SpecialBox>(S.boxes.size());
    L.cls = c.take<int32_t>(S.cls.size());
    L.recs = c.take<SpecialRec>(S.special.size());
    L.ghosts = c.take<GhostDev>(S.ghosts.size());
    L.gk_ptr = c.take<int32_t>(S.gk_ptr.size());
    L.gk_idx = c.take<int32_t>(S.gk_idx.size());
    L.gk_w = c.take<double>(S.gk_w.size());
    L.gu_ptr = c.take<int32_t>(S.gu_ptr.size());
    L.gu_idx = c.take<int32_t>(S.gu_idx.size());
    L.gu_w = c.take<double>(S.gu_w.size());
    L.gval0 = c.take<double>(S.ghosts.size());
    L.gval1 = c.take<double>(S.ghosts.size());
    L.gtmp = c.take<double>(2 * S.ghosts.size());
    L.lptr = c.take<int32_t>(S.lst_ptr.size());
    L.lkind = c.take<int32_t>(S.lst_terms.size());
    L.lindex = c.take<int32_t>(S.lst_terms.size());
    L.lbeta = c.take<double>(S.lst_terms.size());
    L.series = c.take<double>((size_t)cap * (size_t)(S.L ? S.L : 1));
    L.clock = c.take<long long>(1);
}

}  // namespace

size_t DeviceScene::bytes_needed(const SceneSetup &S, int series_cap)
{
    Carver c;
    Layout L;
    carve(S, series_cap, c, L);
    return c.off;
}

fdtd_status DeviceScene::allocate_and_upload(hipStream_t st, int cap)
{
    series_cap = cap;
    Carver c;
    Layout L;
    carve(S, cap, c, L);
    arena_bytes = c.off;
    FDTD_HIP_TRY(hipMalloc(&arena, arena_bytes));
    c.base = arena;
    c.off = 0;
    carve(S, cap, c, L);
    for (int r = 0; r < 3; ++r) p[r] = L.p[r];
    for (int d = 0; d < 3; ++d) pdir[d] = L.pdir[d];
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d) v[b][d] = L.v[b][d];
    // fields start at rest: p, pDir and both face-velocity buffers
    const size_t field_bytes = (size_t)((char *)L.cpml[0] - (char *)L.p[0]);
    FDTD_HIP_TRY(hipMemsetAsync(L.p[0], 0, field_bytes, st));
    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        if (!bytes) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    for (int d = 0; d < 3; ++d) {
        cpml[d] = L.cpml[d]; fpml[d] = L.fpml[d];
        f1[d] = L.f1[d]; f2[d] = L.f2[d]; f3[d] = L.f3[d]; f5h[d] = L.f5h[d];
        FDTD_HIP_TRY(up(cpml[d], S.cpml[d].data(), S.cpml[d].size()));
        FDTD_HIP_TRY(up(fpml[d], S.fpml[d].data(), S.fpml[d].size()));
        FDTD_HIP_TRY(up(f1[d], S.f1[d].data(), S.f1[d].size() * 4));
        FDTD_HIP_TRY(up(f2[d], S.f2[d].data(), S.f2[d].size() * 4));
        FDTD_HIP_TRY(up(f3[d], S.f3[d].data(), S.f3[d].size() * 4));
        FDTD_HIP_TRY(up(f5h[d], S.f5h[d].data(), S.f5h[d].size() * 4));
    }
    gtab = L.gtab;
    FDTD_HIP_TRY(up(gtab, S.gtab.data(), S.gtab.size() * 4));
    std::vector<PSrcDev> ps;
    for (const PSrcHost &h : S.psrc) {
        PSrcDev d{};
        d.t0 = h.t0; d.wt = h.wt; d.amp = h.amp;
        for (int e = 0; e < 3; ++e) { d.lo[e] = h.lo[e]; d.hi[e] = h.hi[e]; d.goff[e] = h.goff[e]; }
        ps.push_back(d);
    }
    psrc = L.psrc;
    FDTD_HIP_TRY(up(psrc, ps.data(), ps.size() * sizeof(PSrcDev)));
    std::vector<ImpDev> is;
    for (const ImpulseHost &h : S.isrc) is.push_back({h.obj, 0, h.A, h.w, h.ton});
    isrc = L.isrc;
    FDTD_HIP_TRY(up(isrc, is.data(), is.size() * sizeof(ImpDev)));
    boxes = L.boxes; cls = L.cls; recs = L.recs;
    FDTD_HIP_TRY(up(boxes, S.boxes.data(), S.boxes.size() * sizeof(SpecialBox)));
    FDTD_HIP_TRY(up(cls, S.cls.data(), S.cls.size() * 4));
    FDTD_HIP_TRY(up(recs, S.special.data(), S.special.size() * sizeof(SpecialRec)));
    std::vector<GhostDev> gd;
    for (const GhostEntry &g : S.ghosts) gd.push_back({g.nbr, g.obj, g.coupled, 0, g.k_raster, g.c_acc});
    ghosts = L.ghosts;
    FDTD_HIP_TRY(up(ghosts, gd.data(), gd.size() * sizeof(GhostDev)));
    gk_ptr = L.gk_ptr; gk_idx = L.gk_idx; gk_w = L.gk_w;
    gu_ptr = L.gu_ptr; gu_idx = L.gu_idx; gu_w = L.gu_w;
    FDTD_HIP_TRY(up(gk_ptr, S.gk_ptr.data(), S.gk_ptr.size() * 4));
    FDTD_HIP_TRY(up(gk_idx, S.gk_idx.data(), S.gk_idx.size() * 4));
    FDTD_HIP_TRY(up(gk_w, S.gk_w.data(), S.gk_w.size() * 8));
    FDTD_HIP_TRY(up(gu_ptr, S.gu_ptr.data(), S.gu_ptr.size() * 4));
    FDTD_HIP_TRY(up(gu_idx, S.gu_idx.data(), S.gu_idx.size() * 4));
    FDTD_HIP_TRY(up(gu_w, S.gu_w.data(), S.gu_w.size() * 8));
    gval[0] = L.gval0; gval[1] = L.gval1; gtmp = L.gtmp;
    if (!S.ghosts.empty()) {
        FDTD_HIP_TRY(hipMemsetAsync(gval[0], 0, S.ghosts.size() * 8, st));
        FDTD_HIP_TRY(hipMemsetAsync(gval[1], 0, S.ghosts.size() * 8, st));
    }
    std::vector<int32_t> kind, index;
    std::vector<double> beta;
    for (const ListenerTerm &t : S.lst_terms) { kind.push_back(t.kind); index.push_back(t.index); beta.push_back(t.beta); }
    lptr = L.lptr; lkind = L.lkind; lindex = L.lindex; lbeta = L.lbeta;
    FDTD_HIP_TRY(up(lptr, S.lst_ptr.data(), S.lst_ptr.size() * 4));
    FDTD_HIP_TRY(up(lkind, kind.data(), kind.size() * 4));
    FDTD_HIP_TRY(up(lindex, index.data(), index.size() * 4));
    FDTD_HIP_TRY(up(lbeta, beta.data(), beta.size() * 8));
    series = L.series;
    clock = L.clock;
    FDTD_HIP_TRY(hipMemsetAsync(clock, 0, sizeof(long long), st));
    // the host vectors above are pageable: make the copies complete before
    // they go out of scope
    FDTD_HIP_TRY(hipStreamSynchronize(st));
    return FDTD_OK;
}

void DeviceScene::release()
{
    if (arena) {
        (void)hipFree(arena);
        arena = nullptr;
    }
}

StepParams DeviceScene::step_params(long long s, const long long *clk, long long offset) const
{
    StepParams P{};
    for (int d = 0; d < 3; ++d) P.n[d] = S.n[d];
    P.sy = S.n[2];
    P.sx = S.n[1] * S.n[2];
    for (int d = 0; d < 3; ++d) { P.ilo[d] = S.inner_lo[d]; P.ihi[d] = S.inner_hi[d]; }
    const int ind = (int)((1 + s) % 3);  // reference ring starts at _pCollocatedInd = 1
    P.pcur = p[ind];
    P.pprev = p[(ind + 2) % 3];
    P.pnext = p[(ind + 1) % 3];
    const int vb = (int)(s & 1);
    for (int d = 0; d < 3; ++d) {
        P.pdir[d] = pdir[d];
        P.vin[d] = v[vb][d];
        P.vout[d] = v[vb ^ 1][d];
        P.f1[d] = f1[d]; P.f2[d] = f2[d]; P.f3[d] = f3[d]; P.f5h[d] = f5h[d];
        P.cpml[d] = cpml[d]; P.fpml[d] = fpml[d];
    }
    P.lam2f = (float)S.lam2;
    P.lam2 = S.lam2;
    P.clock = clk;
    P.step_offset = offset;
    P.t_start = S.t_start;
    P.dt = S.dt;
    P.n_psrc = (int)S.psrc.size();
    P.psrc = psrc;
    P.gtab = gtab;
    P.n_isrc = (int)S.isrc.size();
    P.isrc = isrc;
    P.n_boxes = (int)S.boxes.size();
    P.raster = S.boundary == FDTD_BOUNDARY_RASTERIZE ? 1 : 0;
    P.boxes = boxes;
    P.cls = cls;
    P.recs = recs;
    P.ghosts = ghosts;
    P.gval = gval[s & 1];
    return P;
}

GhostParams DeviceScene::ghost_params(long long s, const long long *clk, long long offset) const
{
    GhostParams G{};
    G.G = (int)S.ghosts.size();
    G.iters = S.coupled_iters;
    G.pcur = p[(int)((1 + s) % 3)];
    G.gval = gval[s & 1];
    G.gtmp = gtmp;
    G.ghosts = ghosts;
    G.gk_ptr = gk_ptr; G.gk_idx = gk_idx; G.gk_w = gk_w;
    G.gu_ptr = gu_ptr; G.gu_idx = gu_idx; G.gu_w = gu_w;
    G.isrc = isrc;
    G.n_isrc = (int)S.isrc.size();
    G.clock = clk;
    G.step_offset = offset;
    G.t_start = S.t_start;
    G.dt = S.dt;
    return G;
}

ListenParams DeviceScene::listen_params(long long s, const long long *clk, long long offset) const
{
    ListenParams P{};
    P.L = S.L;
    P.raster = S.boundary == FDTD_BOUNDARY_RASTERIZE ? 1 : 0;
    const int ind = (int)((1 + s) % 3);
    P.pold = p[ind];
    P.pnew = p[(ind + 1) % 3];
    P.gval = gval[s & 1];
    P.ghosts = ghosts;
    P.ptr = lptr; P.kind = lkind; P.index = lindex; P.beta = lbeta;
    P.isrc = isrc;
    P.n_isrc = (int)S.isrc.size();
    P.series = series;
    P.cap = series_cap;
    P.clock = clk;
    P.step_offset = offset;
    P.t_start = S.t_start;
    P.dt = S.dt;
    return P;
}

fdtd_status DeviceScene::enqueue_step(long long s, const long long *clk, long long offset, int tile, hipStream_t st,
                                      hipEvent_t ev_begin, hipEvent_t ev_end) const
{
    if (needs_ghost_kernel()) FDTD_HIP_TRY(launch_ghost(ghost_params(s, clk, offset), st));
    if (ev_begin) FDTD_HIP_TRY(hipEventRecord(ev_begin, st));
    FDTD_HIP_TRY(launch_step(step_params(s, clk, offset), tile, st));
    if (ev_end) FDTD_HIP_TRY(hipEventRecord(ev_end, st));
    FDTD_HIP_TRY(launch_listen(listen_params(s, clk, offset), st));
    return FDTD_OK;
}

}  // namespace fdtd
