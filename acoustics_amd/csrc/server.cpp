// server.cpp -- continuous batching of impact-sound solve requests on one
// GPU.  A worker thread owns the device: between step blocks it admits
// queued requests while their device state fits the HBM budget (and a slot
// cap), uploads them, rebuilds the batched work list, runs one step block
// for every resident request (requests admitted at different times are at
// different steps -- each carries its own device step counter), and retires
// finished requests (copy the FP64 listener series, check it is finite, free
// the memory); a request whose signal stops being finite is evicted at the
// next periodic check, and the others run on unaffected.  Requests arriving while others run join at the next block:
// their allocation and upload run on a second stream while the resident
// set's block runs, and the work list is rebuilt after that block.
// Each request's tile list is built once, when it is admitted; a rebuild
// concatenates the resident lists.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <system_error>
#include <thread>

#include "device_scene.h"

using namespace fdtd;

namespace {

using Clock = std::chrono::steady_clock;

// blocks between two finiteness checks of the resident requests (one
// listener row each, a small synchronous copy per request)
constexpr int64_t kEvictCheckBlocks = 16;

struct Request {
    int64_t id = 0;
    SceneSetup setup;             // built on the submitting thread
    int steps = 0;
    size_t bytes = 0;             // device bytes when admitted
    Clock::time_point submitted, finished;
    fdtd_status status = FDTD_ERR_BUSY;
    bool cancel = false;
    std::vector<double> result;   // [steps][L]
    std::unique_ptr<DeviceScene> dev;
    long long launched = 0;       // steps launched since admission
    // the request's tiles and listener terms, built once at admission with
    // scene id 0 and term indices from 0; a rebuild only renumbers them
    std::vector<TileDesc> tiles;
    std::vector<TermRef> terms;
};

}  // namespace

struct fdtd_server {
    fdtd_server_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t upload = nullptr;   // admissions upload here while a block runs on `stream`
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<int64_t> queue;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<Request *> active;
    int64_t next_id = 1;
    bool stop = false;
    fdtd_server_stats stats{};
    int64_t resident = 0;
    // step program of the resident set; a block's launches are enqueued
    // directly (each step kernel runs for hundreds of microseconds, so the
    // host enqueues a block far faster than the device runs it, and a
    // rebuild needs no graph capture)
    StepProgram prog;
    bool dirty = true;
    int block = 64;

    void run();
    fdtd_status rebuild();
    void finish(Request *r, fdtd_status st);
};

void fdtd_server::finish(Request *r, fdtd_status st)
{
    r->status = st;
    r->finished = Clock::now();
    if (r->dev) { r->dev->release(); r->dev.reset(); }
    resident -= (int64_t)r->bytes;
    if (st == FDTD_OK) ++stats.completed;
    else if (st == FDTD_ERR_CANCELLED) ++stats.cancelled;
    else ++stats.failed;
}

fdtd_status fdtd_server::rebuild()
{
    prog.release();
    if (active.empty()) return FDTD_OK;
    std::vector<SceneDev> devs;
    std::vector<TileDesc> tiles;
    std::vector<TermRef> terms;
    std::vector<MotionDesc> motion;
    for (Request *r : active) {
        const int id = (int)devs.size();
        devs.push_back(r->dev->dev);
        const int tb = (int)terms.size();
        for (TermRef q : r->terms) { q.scene = id; terms.push_back(q); }
        for (TileDesc t : r->tiles) { t.scene = id; t.tb += tb; t.te += tb; tiles.push_back(t); }
        if (r->dev->needs_motion_kernel()) motion.push_back({id});
    }
    return prog.upload(devs, tiles, terms, stream, tile_plan(0).fused, motion);
}

void fdtd_server::run()
{
    (void)hipSetDevice(device);
    for (;;) {
        std::vector<Request *> admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty())
                cv.wait(lk, [&] { return stop || !queue.empty(); });
            if (stop) break;
            // cancellations of resident requests take effect at the block boundary
            for (Request *r : active)
                if (r->cancel) { finish(r, FDTD_ERR_CANCELLED); dirty = true; }
            active.erase(std::remove_if(active.begin(), active.end(), [](Request *r) { return r->status != FDTD_ERR_BUSY; }),
                         active.end());
            // admission by HBM budget and slot cap, first come first served
            while (!queue.empty()) {
                Request *r = reqs.at(queue.front()).get();
                if (r->cancel) {
                    queue.pop_front();
                    resident += (int64_t)r->bytes;  // finish() returns the (never used) reservation
                    finish(r, FDTD_ERR_CANCELLED);
                    continue;
                }
                if (opt.max_active > 0 && (int)(active.size() + admit.size()) >= opt.max_active) break;
                if (resident + (int64_t)r->bytes > opt.hbm_budget_bytes) break;
                queue.pop_front();
                resident += (int64_t)r->bytes;
                admit.push_back(r);
            }
            stats.queued = (int64_t)queue.size();
            stats.peak_resident_bytes = std::max(stats.peak_resident_bytes, resident);
        }
        // the block of the current resident set starts first; the admitted
        // requests upload on their own stream while it runs, and join the
        // work list at the next rebuild
        const bool run_block = !active.empty() && !dirty;
        if (run_block && prog.enqueue(stream, 0, block) != FDTD_OK) {
            (void)hipStreamSynchronize(stream);
            std::lock_guard<std::mutex> lk(mu);
            set_last_error("step block failed on the device");
            for (Request *r : active) finish(r, FDTD_ERR_DEVICE);
            for (Request *r : admit) finish(r, FDTD_ERR_DEVICE);
            active.clear();
            done_cv.notify_all();
            continue;
        }
        std::vector<Request *> fresh;
        for (Request *r : admit) {
            r->dev = std::make_unique<DeviceScene>();
            r->dev->S = std::move(r->setup);
            fdtd_status s = r->dev->allocate_and_upload(upload, std::max(1, r->steps), r->steps);
            if (s == FDTD_OK) r->dev->append_tiles(0, tile_plan(0), r->tiles, r->terms);
            std::lock_guard<std::mutex> lk(mu);
            if (s != FDTD_OK) { finish(r, s); done_cv.notify_all(); continue; }
            fresh.push_back(r);
        }
        if (run_block && hipStreamSynchronize(stream) != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            set_last_error("step block failed on the device");
            for (Request *r : active) finish(r, FDTD_ERR_DEVICE);
            for (Request *r : fresh) finish(r, FDTD_ERR_DEVICE);
            active.clear();
            done_cv.notify_all();
            continue;
        }
        {
            // one critical section from finish() to the removal from `active`:
            // a client that sees a request finished may take (and free) it at
            // once, so no worker list may still hold it when the lock drops
            std::lock_guard<std::mutex> lk(mu);
            if (run_block) {
                ++stats.blocks_run;
                // every kEvictCheckBlocks blocks, a request whose latest
                // listener row is no longer finite (a blown-up scene: NaN
                // and infinity never leave the field once in it) is evicted
                // at once instead of holding its slot to the last step
                const bool check = stats.blocks_run % kEvictCheckBlocks == 0;
                std::vector<double> row;
                for (Request *r : active) {
                    r->launched += block;
                    if (r->launched < r->steps) {
                        const int L = r->dev->S.L;
                        if (!check || !L) continue;
                        row.resize(L);
                        fdtd_status s = r->dev->read_series(r->launched - 1, 1, row.data());
                        for (double v : row)
                            if (!std::isfinite(v)) { s = FDTD_ERR_NONFINITE; break; }
                        if (s != FDTD_OK) {
                            set_last_error("request " + std::to_string(r->id) + " evicted at step " +
                                           std::to_string(r->launched) + ": listener signal is not finite");
                            finish(r, s);
                            dirty = true;
                        }
                        continue;
                    }
                    const int L = r->dev->S.L;
                    r->result.resize((size_t)r->steps * L);
                    fdtd_status s = FDTD_OK;
                    if (r->steps && L) s = r->dev->read_series(0, r->steps, r->result.data());
                    for (double v : r->result)
                        if (!std::isfinite(v)) { s = FDTD_ERR_NONFINITE; break; }
                    finish(r, s);
                    dirty = true;
                }
            }
            if (!fresh.empty()) dirty = true;
            active.insert(active.end(), fresh.begin(), fresh.end());
            active.erase(std::remove_if(active.begin(), active.end(), [](Request *r) { return r->status != FDTD_ERR_BUSY; }),
                         active.end());
            stats.active = (int64_t)active.size();
            done_cv.notify_all();
        }
        if (dirty && !active.empty()) {
            fdtd_status s = rebuild();
            dirty = false;
            if (s != FDTD_OK) {
                std::lock_guard<std::mutex> lk(mu);
                for (Request *r : active) finish(r, s);
                active.clear();
                done_cv.notify_all();
            }
        }
    }
    prog.release();
    std::lock_guard<std::mutex> lk(mu);
    for (Request *r : active) finish(r, FDTD_ERR_CANCELLED);
    active.clear();
    done_cv.notify_all();
}

extern "C" {

void fdtd_server_options_default(fdtd_server_options *o)
{
    if (!o) return;
    o->hbm_budget_bytes = (int64_t)64 << 30;
    o->max_active = 0;
    o->steps_per_block = 64;
    o->device = -1;
}

int64_t fdtd_scene_device_bytes(const fdtd_scene *scene)
{
    if (!scene) return -1;
    SceneSetup S;
    std::string err;
    if (build_setup(*scene, S, err) != FDTD_OK) { set_last_error(err); return -1; }
    return (int64_t)DeviceScene::bytes_needed(S, 1);
}

// cell index of a special record or of a bare cell index (plan check lookups)
static inline int32_t cell_of(const SpecialRec &r) { return r.cell; }
static inline int32_t cell_of(int32_t c) { return c; }

static fdtd_status fail_early(int plan, const std::string &why)
{
    set_last_error("plan " + std::to_string(plan) + ": " + why);
    return FDTD_ERR_INVALID;
}

fdtd_status fdtd_plan_check(const fdtd_scene *scene, int tile_config, int64_t *n_tiles)
{
    if (!scene) return FDTD_ERR_INVALID;
    DeviceScene D;
    std::string err;
    fdtd_status st = build_setup(*scene, D.S, err);
    if (st != FDTD_OK) { set_last_error(err); return st; }
    const int plan = (tile_config >= 0 && tile_config < kNumTilePlans) ? tile_config : default_tile_plan(D.S);
    const TilePlan tp = tile_plan(plan);
    // the scene block as the kernels see it, without device arrays
    D.sy_dev = row_stride(D.S);
    SceneDev d{};
    for (int a = 0; a < 3; ++a) d.n[a] = D.S.n[a];
    d.sy = D.sy_dev;
    d.sx = D.S.n[1] * D.sy_dev;
    d.n_special = (int)D.S.special.size();
    d.scheme = D.S.scheme;
    d.mv = D.S.moving ? 1 : 0;
    for (int a = 0; a < 3; ++a) { d.u_lo[a] = D.S.mv_lo[a]; d.u_n[a] = D.S.mv_hi[a] - D.S.mv_lo[a]; }
    std::vector<TileDesc> tiles;
    std::vector<TermRef> terms;
    D.append_tiles(0, tp, tiles, terms);
    std::vector<TileDesc> lists[NUM_KINDS];
    build_launch_lists(tiles, tp.fused, lists);
    const std::string e = check_launch_lists(lists, tp.fused, {d}, terms.size());
    if (!e.empty()) { set_last_error("plan " + std::to_string(plan) + ": " + e); return FDTD_ERR_INVALID; }
    // the launches carry every tile once (the lists only reorder and pad):
    // the same count and the same sum of per-tile hashes
    {
        auto sig = [](const std::vector<TileDesc> &v, uint64_t *sum) {
            size_t n = 0;
            for (const TileDesc &t : v) {
                if (t.type == TILE_EMPTY) continue;
                uint64_t h = 1469598103934665603ull;
                for (int x : {t.type, t.i0, t.i1, t.j0, t.j1, t.k0, t.k1, t.slab, t.tb, t.te}) h = (h ^ (uint32_t)x) * 1099511628211ull;
                *sum += h;
                ++n;
            }
            return n;
        };
        uint64_t a = 0, b = 0;
        size_t na = 0;
        for (int k = 0; k < NUM_KINDS; ++k) na += sig(lists[k], &a);
        if (na != sig(tiles, &b) || a != b) return fail_early(plan, "the launch lists do not carry every tile exactly once");
    }
    // every cell lies in exactly one box tile; object cells (solid or next to
    // an object) only in boxes that skip them, and each record in one
    // object-cell tile.  A moving scene's reclassified box is written by its
    // moving-object tiles only: the interior tiles over it skip it.
    std::vector<uint8_t> hit((size_t)D.S.N, 0), skips((size_t)D.S.N, 0);
    std::vector<int> rec_hits(D.S.special.size(), 0);
    std::vector<uint8_t> in_box((size_t)D.S.N, 0);
    if (D.S.moving)
        for (int i = D.S.mv_lo[0]; i < D.S.mv_hi[0]; ++i)
            for (int j = D.S.mv_lo[1]; j < D.S.mv_hi[1]; ++j)
                for (int k = D.S.mv_lo[2]; k < D.S.mv_hi[2]; ++k) in_box[((int64_t)i * D.S.n[1] + j) * D.S.n[2] + k] = 1;
    for (const TileDesc &t : tiles) {
        if (t.type == TILE_SPECIAL) {
            for (int r = t.i0; r < t.i1; ++r) ++rec_hits[r];
            continue;
        }
        const int kfirst = t.k0 + ((t.type == TILE_INTERIOR_VEC || t.type == TILE_MAC_VEC) ? t.lead : 0);   // cells the tile writes
        for (int i = t.i0; i < t.i1; ++i)
            for (int j = t.j0; j < t.j1; ++j)
                for (int k = kfirst; k < t.k1; ++k) {
                    const int64_t c = ((int64_t)i * D.S.n[1] + j) * D.S.n[2] + k;
                    if (in_box[c] && t.type != TILE_MOVING) {
                        if (t.type != TILE_INTERIOR_SKIP) return fail_early(plan, "a moving scene's box lies in a tile that would overwrite it");
                        continue;
                    }
                    if (!in_box[c] && t.type == TILE_MOVING) return fail_early(plan, "a moving-object tile leaves its box");
                    ++hit[c];
                    skips[c] = t.type == TILE_INTERIOR_SKIP || t.type == TILE_MOVING;
                }
    }
    std::vector<uint8_t> object_cell((size_t)D.S.N, 0);
    if (D.S.has_solid)
        for (int64_t c = 0; c < D.S.N; ++c) object_cell[c] = D.S.bulk[c] ? 0 : 1;
    for (const SpecialRec &r : D.S.special) object_cell[r.cell] = 1;
    auto fail = [&](const std::string &why) {
        set_last_error("plan " + std::to_string(plan) + ": " + why);
        return FDTD_ERR_INVALID;
    };
    // (with the Engquist boundary the outermost layer is in no tile; with
    // plane constraints, the interior layers cut off the tile box -- solid
    // cells and the fluid cells next to them -- are in none either, and the
    // PML tiles hold their solid cells at zero)
    const int *n = D.S.n;
    const SceneSetup &S = D.S;
    for (int64_t c = 0; c < D.S.N; ++c) {
        const int i = (int)(c / ((int64_t)n[1] * n[2])), j = (int)((c / n[2]) % n[1]), k = (int)(c % n[2]);
        const bool outer = i == 0 || j == 0 || k == 0 || i == n[0] - 1 || j == n[1] - 1 || k == n[2] - 1;
        const bool pml = !S.abc && (S.cpml[0][i] || S.cpml[1][j] || S.cpml[2][k]);
        const bool cut = !pml && !S.abc && (i < S.tile_lo[0] || i >= S.tile_hi[0] || j < S.tile_lo[1] || j >= S.tile_hi[1] ||
                                             k < S.tile_lo[2] || k >= S.tile_hi[2]);
        if (cut && S.bulk.size() && S.bulk[c] && !std::binary_search(S.special.begin(), S.special.end(), (int32_t)c,
                [](const auto &a, const auto &b) {
                    return cell_of(a) < cell_of(b);
                }))
            return fail("a fluid cell off the interior tiles is no object-adjacent record");
        const int want = ((D.S.abc && outer) || cut) ? 0 : 1;
        if (hit[c] != want) return fail("a cell is covered " + std::to_string(hit[c]) + " times");
        if (object_cell[c] && hit[c] && !pml && !skips[c]) return fail("an object cell lies in a tile that would overwrite it");
    }
    for (int h : rec_hits)
        if (h != 1) return fail("an object-adjacent record is covered " + std::to_string(h) + " times");
    if (n_tiles) {
        int64_t n = 0;
        for (int k = 0; k < NUM_KINDS; ++k) n += (int64_t)lists[k].size();
        *n_tiles = n;
    }
    return FDTD_OK;
}

fdtd_status fdtd_server_create(const fdtd_server_options *opt_in, fdtd_server **out)
{
    if (!out) return FDTD_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_last_error("no HIP device"); return FDTD_ERR_NO_DEVICE; }
    std::unique_ptr<fdtd_server> s(new (std::nothrow) fdtd_server());
    if (!s) return FDTD_ERR_OOM;
    fdtd_server_options_default(&s->opt);
    if (opt_in) s->opt = *opt_in;
    if (s->opt.hbm_budget_bytes <= 0) { set_last_error("hbm_budget_bytes must be positive"); return FDTD_ERR_INVALID; }
    s->block = s->opt.steps_per_block > 0 ? s->opt.steps_per_block : 64;
    if (s->opt.device >= 0) FDTD_HIP_TRY(hipSetDevice(s->opt.device));
    FDTD_HIP_TRY(hipGetDevice(&s->device));
    FDTD_HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    const hipError_t e = hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s->stream);
        set_last_error(std::string("upload stream: ") + hipGetErrorString(e));
        return FDTD_ERR_DEVICE;
    }
    fdtd_server *raw = s.get();
    try {
        raw->worker = std::thread([raw] { raw->run(); });
    } catch (const std::system_error &) {
        (void)hipStreamDestroy(raw->stream);
        (void)hipStreamDestroy(raw->upload);
        set_last_error("could not start the server's worker thread");
        return FDTD_ERR_DEVICE;
    }
    *out = s.release();
    return FDTD_OK;
}

fdtd_status fdtd_server_submit(fdtd_server *s, const fdtd_scene *scene, int num_steps, int64_t *id)
{
    if (!s || !scene || num_steps <= 0 || !id) { set_last_error("bad submit arguments"); return FDTD_ERR_INVALID; }
    auto r = std::make_unique<Request>();
    std::string err;
    fdtd_status st = build_setup(*scene, r->setup, err);
    if (st != FDTD_OK) { set_last_error(err); return st; }
    if (r->setup.moving && num_steps > r->setup.mv_steps) {
        set_last_error("moving objects are planned up to stop_time");
        return FDTD_ERR_INVALID;
    }
    r->steps = num_steps;
    r->bytes = DeviceScene::bytes_needed(r->setup, num_steps);
    r->submitted = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if ((int64_t)r->bytes > s->opt.hbm_budget_bytes) {
        set_last_error("request needs " + std::to_string(r->bytes) + " device bytes, over the whole HBM budget");
        return FDTD_ERR_OOM;
    }
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    ++s->stats.submitted;
    s->cv.notify_all();
    return FDTD_OK;
}

fdtd_status fdtd_server_poll(fdtd_server *s, int64_t id)
{
    if (!s) return FDTD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return FDTD_ERR_NOT_FOUND;
    return it->second->status;
}

fdtd_status fdtd_server_wait(fdtd_server *s, int64_t id, int timeout_ms)
{
    if (!s) return FDTD_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return FDTD_ERR_NOT_FOUND;
    Request *r = it->second.get();
    auto pred = [&] { return r->status != FDTD_ERR_BUSY; };
    if (timeout_ms < 0) s->done_cv.wait(lk, pred);
    else s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    return r->status;
}

fdtd_status fdtd_server_take_result(fdtd_server *s, int64_t id, double *out, double *elapsed_ms)
{
    if (!s) return FDTD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return FDTD_ERR_NOT_FOUND;
    Request *r = it->second.get();
    if (r->status == FDTD_ERR_BUSY) return FDTD_ERR_BUSY;
    const fdtd_status st = r->status;
    if (elapsed_ms) *elapsed_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
    if (st == FDTD_OK && out && !r->result.empty()) std::memcpy(out, r->result.data(), sizeof(double) * r->result.size());
    s->reqs.erase(it);
    return st;
}

fdtd_status fdtd_server_cancel(fdtd_server *s, int64_t id)
{
    if (!s) return FDTD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return FDTD_ERR_NOT_FOUND;
    it->second->cancel = true;
    s->cv.notify_all();
    return FDTD_OK;
}

fdtd_status fdtd_server_get_stats(fdtd_server *s, fdtd_server_stats *out)
{
    if (!s || !out) return FDTD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->active = (int64_t)s->active.size();
    out->queued = (int64_t)s->queue.size();
    return FDTD_OK;
}

fdtd_status fdtd_server_destroy(fdtd_server *s)
{
    if (!s) return FDTD_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamDestroy(s->stream);
    (void)hipStreamDestroy(s->upload);
    delete s;
    return FDTD_OK;
}

}  // extern "C"
