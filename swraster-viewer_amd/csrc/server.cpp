// server.cpp -- render server: continuous batching over resident scenes.
//
// Requests (scene, camera, resolution) are queued by sr_server_submit and
// served by one worker thread that owns `streams` pipeline slots, each with
// its own HIP stream.  Whenever a slot is free the worker forms a batch from
// the head of the queue (FIFO) -- as many requests as fit both the view limit
// and the admission budget -- and launches it asynchronously, so requests
// arriving while other batches run join the next batch and batches on
// different slots overlap on the device.  Each batch's outputs (triangle ids,
// depth, RGBA8, tile luminance) live in one region of a framebuffer arena
// allocated once; a region returns to the arena when every request in it has
// been released.  Admission counts the arena bytes in use plus an estimate
// of the batch's bin and setup memory against `admission_bytes`.
// The numerics are the batch path's: one request renders the same bits
// alone or in any mix.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "runtime.h"
#include "swraster.h"

using namespace sr;

#define HIP_CHECK_RET(expr)                                                                     \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

const float DEFAULT_EXPOSURE = 2.0f;

// First-fit allocator over [0, size) of the framebuffer arena (host-side
// bookkeeping only).
struct Arena {
    std::map<uint64_t, uint64_t> free_;   // offset -> size
    uint64_t used = 0;
    void init(uint64_t size) {
        free_.clear();
        free_[0] = size;
        used = 0;
    }
    bool alloc(uint64_t size, uint64_t &off) {
        size = (size + 255) & ~uint64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < size) continue;
            off = it->first;
            uint64_t rest = it->second - size;
            free_.erase(it);
            if (rest) free_[off + size] = rest;
            used += size;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t size) {
        size = (size + 255) & ~uint64_t(255);
        used -= size;
        auto it = free_.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Region {
    uint64_t off = 0, size = 0;
    int refs = 0;
};

struct Request {
    uint64_t id = 0;
    sr_request r;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = SR_OK;
    bool done = false, released = false;
    int region = -1;
    uint64_t px_off = 0;          // element offset of this view in its region's planes
    uint64_t plane = 0;           // pixels per plane of its region
};

struct Slot {
    Pipeline pipe;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<uint64_t> reqs;
    // launch arguments kept for a synchronous re-run after a capacity overflow
    std::vector<DevView> hv;
    std::vector<DevDraw> draws;
    uint32_t xv = 0, in_tri = 0, items = 0;
    bool tangents = false;
    uint32_t *ids = nullptr, *rgba = nullptr;
    float *depth = nullptr, *lum = nullptr;
    float *color = nullptr;         // colour planes when the batch has translucent primitives
    uint64_t plane = 0;             // pixels of the batch (colour plane stride)
    std::vector<uint64_t> px_off;
    std::vector<int> width, height;
};

}  // namespace

struct sr_server {
    int device = 0;
    int tile = 16;
    int max_views = 64;
    // raster launch: one workgroup per (view, tile) item.  Served batches
    // hold few large views, where the persistent raster's shared work
    // counter costs more than its balancing gains (1.24x the plain path's
    // frames/s at equal p99 on the mixed 720p-2160p stream, against 0.92x
    // with 2048 persistent workgroups).
    int persistent = 0;
    uint64_t budget = 0;
    std::vector<const sr_scene *> scenes;
    DevBuf arena_mem;
    Arena arena;
    std::vector<Region> regions;
    std::vector<int> free_regions;
    std::vector<Slot> slots;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> pending;
    std::deque<sr_completion> completions;
    uint64_t next_id = 1;
    uint64_t gen = 0;             // bumped by submissions and releases
    int inflight = 0;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::thread worker;
    std::atomic<bool> stop{false};
    std::chrono::steady_clock::time_point t0;
    std::string worker_error;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void run();
    // both are called with mu held and release it around host-heavy work
    bool launch_batch(Slot &s, std::unique_lock<std::mutex> &lk);
    void retire(Slot &s, std::unique_lock<std::mutex> &lk);
    void complete(const std::vector<uint64_t> &ids, int rc);
};

void sr_server::complete(const std::vector<uint64_t> &ids, int rc) {
    double t = now();
    for (uint64_t id : ids) {
        Request &q = reqs[id];
        if (q.status == SR_OK) q.status = rc;
        q.done = true;
        q.t_done = t;
        sr_completion c;
        memset(&c, 0, sizeof(c));
        c.request_id = q.id;
        c.user_tag = q.r.user_tag;
        c.status = q.status;
        c.width = q.r.width;
        c.height = q.r.height;
        c.submit_time_s = q.t_submit;
        c.start_time_s = q.t_start;
        c.complete_time_s = q.t_done;
        completions.push_back(c);
    }
}

// Visibility buffer, RGBA8 and tile luminance of a request; a batch with a
// translucent scene also holds colour planes for all its views.
static uint64_t fb_bytes(const sr_request &r) {
    uint64_t px = (uint64_t)r.width * r.height;
    uint64_t ref = (uint64_t)((r.width + 63) / 64) * ((r.height + 63) / 64);
    return px * 12 + ref * 4;
}
static uint64_t color_bytes(const sr_request &r) { return (uint64_t)r.width * r.height * 12; }

// Blend results to RGBA8: one blit per view from the batch's colour planes.
static int blit_views(Slot &s) {
    for (size_t i = 0; i < s.hv.size(); i++) {
        hipError_t e = launch_blit(s.color + s.px_off[i], s.plane, s.width[i], s.height[i], s.hv[i].exposure,
                                   s.rgba + s.px_off[i], s.pipe.stream);
        if (e != hipSuccess) return fail(SR_ERR_HIP, "blit: %s", hipGetErrorString(e));
    }
    return SR_OK;
}

static uint64_t bin_estimate(const sr_scene *s, const sr_request &r) {
    // setup records (80 B) and clip/scan state per input triangle, one 4-byte
    // bin entry per triangle and about one per pixel, transformed vertices
    // (clip position, attributes and snap: 64 B)
    return s->triangles * 96 + (uint64_t)r.width * r.height * 4 + s->vertices * 64;
}

bool sr_server::launch_batch(Slot &s, std::unique_lock<std::mutex> &lk) {
    // (locked) take requests FIFO while they fit the view limit and the budget
    std::vector<uint64_t> take;
    uint64_t bytes = 0, bins = 0, cbytes = 0;
    bool tr = false;
    for (uint64_t id : pending) {
        const Request &q = reqs[id];
        uint64_t b = fb_bytes(q.r), e = bin_estimate(scenes[q.r.scene], q.r);
        bool qtr = tr || scenes[q.r.scene]->any_translucent;
        uint64_t c = qtr ? cbytes + color_bytes(q.r) : 0;
        if ((int)take.size() >= max_views) break;
        if (!take.empty() && arena.used + bytes + b + c + bins + e > budget) break;
        take.push_back(id);
        bytes += b;
        bins += e;
        cbytes += color_bytes(q.r);
        tr = qtr;
    }
    if (take.empty()) return false;
    bool with_color = false;       // the region holds colour planes
    auto batch_bytes = [&]() {
        uint64_t b = 0, c = 0;
        with_color = false;
        for (uint64_t id : take) {
            b += fb_bytes(reqs[id].r);
            c += color_bytes(reqs[id].r);
            with_color = with_color || scenes[reqs[id].r.scene]->any_translucent;
        }
        return b + (with_color ? c : 0);
    };
    bytes = batch_bytes();
    uint64_t off = 0;
    while (!arena.alloc(bytes, off)) {
        if (take.size() == 1) return false;     // wait for releases
        take.pop_back();
        bytes = batch_bytes();
    }
    for (size_t i = 0; i < take.size(); i++) pending.pop_front();
    int rg;
    if (!free_regions.empty()) {
        rg = free_regions.back();
        free_regions.pop_back();
    } else {
        rg = (int)regions.size();
        regions.push_back(Region());
    }
    regions[rg].off = off;
    regions[rg].size = bytes;
    regions[rg].refs = (int)take.size();
    int n = (int)take.size();
    std::vector<sr_request> rq(n);
    std::vector<const sr_scene *> rs(n);
    for (int i = 0; i < n; i++) {
        rq[i] = reqs[take[i]].r;
        rs[i] = scenes[rq[i].scene];
    }
    s.reqs = take;
    s.busy = true;                  // the slot and these requests now belong to this launch
    inflight++;
    char *base = (char *)arena_mem.ptr + off;
    double t = now();

    // (unlocked) host set-up and the asynchronous launch
    lk.unlock();
    std::vector<int> st(n, SR_OK);
    std::vector<uint64_t> px_off(n, 0);
    s.hv.assign(n, DevView());
    s.draws.clear();
    s.xv = s.in_tri = s.items = 0;
    s.tangents = false;
    uint64_t px_total = 0;
    uint32_t lum_total = 0;
    for (int i = 0; i < n; i++) {
        uint32_t tri0 = s.in_tri, xv0 = s.xv;
        size_t nd = s.draws.size();
        int rc = build_draws(rs[i], rq[i].camera, i, s.in_tri, s.xv, s.draws, true);
        if (rc) {                      // the view renders nothing; the request reports rc
            st[i] = rc;
            s.draws.resize(nd);
            s.in_tri = tri0;
            s.xv = xv0;
        }
        fill_view(s.hv[i], rs[i], rq[i].camera, rq[i].width, rq[i].height, tile, s.items, tri0, px_total, lum_total,
                  rq[i].exposure > 0.0f ? rq[i].exposure : DEFAULT_EXPOSURE, 1u);
        px_off[i] = px_total;
        s.items += (uint32_t)(s.hv[i].tiles_x * s.hv[i].tiles_y);
        lum_total += (uint32_t)(s.hv[i].ref_tiles_x * s.hv[i].ref_tiles_y);
        px_total += (uint64_t)rq[i].width * rq[i].height;
        s.tangents = s.tangents || rs[i]->any_tangents;
    }
    // translucent primitives blend over the opaque colour before the
    // tonemap: such a batch writes colour planes and blits each view after
    bool has_tr = false;
    for (const DevDraw &d : s.draws) has_tr = has_tr || !(d.flags & SR_DRAW_OPAQUE);
    if (has_tr)
        for (DevView &v : s.hv) v.flags &= ~1u;
    s.ids = (uint32_t *)base;
    s.depth = (float *)(base + px_total * 4);
    s.rgba = (uint32_t *)(base + px_total * 8);
    s.lum = (float *)(base + px_total * 12);
    s.color = has_tr ? (float *)(base + px_total * 12 + (uint64_t)lum_total * 4) : nullptr;
    s.plane = px_total;
    s.px_off = px_off;
    s.width.resize(n);
    s.height.resize(n);
    for (int i = 0; i < n; i++) {
        s.width[i] = rq[i].width;
        s.height[i] = rq[i].height;
    }
    int rc = has_tr && !with_color ? fail(SR_ERR_UNSUPPORTED, "translucent draws in a batch without colour planes")
                                   : s.pipe.run(s.hv, s.draws, s.xv, s.in_tri, s.items, s.tangents, s.ids, s.depth,
                                                s.color, has_tr ? px_total : 0, s.rgba, s.lum, persistent, false);
    if (rc == SR_OK && has_tr) rc = blit_views(s);
    if (rc == SR_OK && hipEventRecord(s.done, s.pipe.stream) != hipSuccess) rc = fail(SR_ERR_HIP, "event record");
    if (rc != SR_OK) hipStreamSynchronize(s.pipe.stream);
    lk.lock();

    // (locked) publish
    for (int i = 0; i < n; i++) {
        Request &q = reqs[take[i]];
        q.region = rg;
        q.px_off = px_off[i];
        q.plane = px_total;
        q.t_start = t;
        q.status = st[i];
    }
    if (rc != SR_OK) {
        complete(take, rc);
        s.reqs.clear();
        s.busy = false;
        inflight--;
    }
    return true;
}

void sr_server::retire(Slot &s, std::unique_lock<std::mutex> &lk) {
    // settle reads the batch's status, records its counts (which choose the
    // depth-rejecting raster for the slot's next batch) and, on a capacity
    // overflow, grows the capacities first; the batch then re-runs
    // synchronously with the grown buffers
    bool again = false;
    int rc = s.pipe.settle(s.color != nullptr, again);
    if (rc == SR_OK && again) {
        lk.unlock();
        rc = s.pipe.run(s.hv, s.draws, s.xv, s.in_tri, s.items, s.tangents, s.ids, s.depth, s.color,
                        s.color ? s.plane : 0, s.rgba, s.lum, persistent, true);
        if (rc == SR_OK && s.color) {
            rc = blit_views(s);
            if (rc == SR_OK && hipStreamSynchronize(s.pipe.stream) != hipSuccess) rc = fail(SR_ERR_HIP, "sync");
        }
        lk.lock();
    }
    complete(s.reqs, rc);
    s.reqs.clear();
    s.busy = false;
    inflight--;
}

void sr_server::run() {
    hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    uint64_t seen = gen;
    while (!stop) {
        bool progressed = false;
        for (Slot &s : slots) {
            if (s.busy && !s.reqs.empty() && hipEventQuery(s.done) == hipSuccess) {
                retire(s, lk);
                progressed = true;
            }
        }
        for (Slot &s : slots) {
            if (!s.busy && !pending.empty() && launch_batch(s, lk)) progressed = true;
        }
        if (progressed) {
            cv_done.notify_all();
            continue;
        }
        if (inflight == 0) {
            // nothing on the device: sleep until a submission or a release
            // (a release can make room for a request that did not fit)
            cv.wait(lk, [&] { return stop.load() || gen != seen; });
            seen = gen;
        } else {
            // batches in flight: poll their events, yielding the core between
            // polls (a timed sleep overshoots by tens of microseconds, a
            // sizeable part of a frame)
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }
}

extern "C" {

SR_API int sr_server_create(const sr_server_opts *opts, sr_server **out) {
    if (!out) return fail(SR_ERR_INVALID_ARGUMENT, "null output");
    int32_t ndev = 0;
    sr_device_count(&ndev);
    int dev = opts ? opts->device : 0;
    if (dev < 0 || dev >= ndev) return fail(SR_ERR_INVALID_ARGUMENT, "no HIP device %d (have %d)", dev, ndev);
    sr_server *s = new sr_server();
    s->device = dev;
    s->budget = opts && opts->admission_bytes ? opts->admission_bytes : (uint64_t)64 << 30;
    s->max_views = opts && opts->max_batch_views > 0 ? opts->max_batch_views : 16;
    int nslots = opts && opts->streams > 0 ? opts->streams : 2;
    int tile = opts ? opts->tile_size : 0;
    s->tile = (tile == 8 || tile == 16 || tile == 32) ? tile : 16;
    hipSetDevice(dev);
    // framebuffers take at most half the budget; bins and setup the rest
    int rc = s->arena_mem.ensure(s->budget / 2);
    if (rc) {
        delete s;
        return rc;
    }
    s->arena.init(s->budget / 2);
    s->slots.resize(nslots);
    for (Slot &sl : s->slots) {
        if ((rc = sl.pipe.init(dev, s->tile)) ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            if (!rc) rc = fail(SR_ERR_HIP, "event create");
            for (Slot &x : s->slots) {
                x.pipe.destroy();
                if (x.done) hipEventDestroy(x.done);
            }
            s->arena_mem.release();
            delete s;
            return rc;
        }
        sl.pipe.timing = SR_STAGE_TIMING_OFF;    // request latency is taken on the host
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return SR_OK;
}

SR_API int sr_server_destroy(sr_server *s) {
    if (!s) return SR_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    hipSetDevice(s->device);
    for (Slot &sl : s->slots) {
        sl.pipe.destroy();
        if (sl.done) hipEventDestroy(sl.done);
    }
    s->arena_mem.release();
    delete s;
    return SR_OK;
}

SR_API int sr_server_add_scene(sr_server *s, const sr_scene *scene, int32_t *index) {
    if (!s || !scene || !index) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->device != s->device) return fail(SR_ERR_INVALID_ARGUMENT, "scene on another device");
    std::lock_guard<std::mutex> g(s->mu);
    *index = (int32_t)s->scenes.size();
    s->scenes.push_back(scene);
    return SR_OK;
}

SR_API int sr_server_submit(sr_server *s, const sr_request *r, uint64_t *id) {
    if (!s || !r) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (r->width <= 0 || r->height <= 0 || r->width > 32000 || r->height > 32000)
        return fail(SR_ERR_INVALID_ARGUMENT, "bad resolution %dx%d", r->width, r->height);
    std::lock_guard<std::mutex> g(s->mu);
    if (r->scene < 0 || r->scene >= (int)s->scenes.size()) return fail(SR_ERR_NOT_FOUND, "no scene %d", r->scene);
    if (fb_bytes(*r) > s->budget / 2 || fb_bytes(*r) + bin_estimate(s->scenes[r->scene], *r) > s->budget)
        return fail(SR_ERR_CAPACITY, "request exceeds the admission budget on its own");
    Request q;
    q.id = s->next_id++;
    q.r = *r;
    q.t_submit = s->now();
    s->reqs[q.id] = q;
    s->pending.push_back(q.id);
    s->gen++;
    if (id) *id = q.id;
    s->cv.notify_all();
    return SR_OK;
}

SR_API int sr_server_poll(sr_server *s, sr_completion *out, int32_t max, int32_t *count) {
    if (!s || !count || (max > 0 && !out)) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    int n = 0;
    while (n < max && !s->completions.empty()) {
        out[n++] = s->completions.front();
        s->completions.pop_front();
    }
    *count = n;
    return SR_OK;
}

static int finished_request(sr_server *s, uint64_t id, Request **out) {
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.released) return fail(SR_ERR_NOT_FOUND, "no request %llu", (unsigned long long)id);
    if (!it->second.done) return fail(SR_ERR_BUSY, "request %llu not complete", (unsigned long long)id);
    if (it->second.status != SR_OK) return fail(it->second.status, "request %llu failed", (unsigned long long)id);
    *out = &it->second;
    return SR_OK;
}

SR_API int sr_server_read_pixels(sr_server *s, uint64_t id, uint32_t *pixels) {
    if (!s || !pixels) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    Request *q;
    const char *src;
    size_t bytes;
    {
        std::lock_guard<std::mutex> g(s->mu);
        int rc = finished_request(s, id, &q);
        if (rc) return rc;
        src = (const char *)s->arena_mem.ptr + s->regions[q->region].off + q->plane * 8 + q->px_off * 4;
        bytes = (size_t)(q->r.height & ~1) * q->r.width * 4;
    }
    hipSetDevice(s->device);
    HIP_CHECK_RET(hipMemcpy(pixels, src, bytes, hipMemcpyDeviceToHost));
    return SR_OK;
}

SR_API int sr_server_read_vbuffer(sr_server *s, uint64_t id, uint32_t *tri_id, float *depth) {
    if (!s) return fail(SR_ERR_INVALID_ARGUMENT, "null server");
    Request *q;
    const char *base;
    size_t n;
    uint64_t plane, off;
    {
        std::lock_guard<std::mutex> g(s->mu);
        int rc = finished_request(s, id, &q);
        if (rc) return rc;
        base = (const char *)s->arena_mem.ptr + s->regions[q->region].off;
        n = (size_t)q->r.width * q->r.height;
        plane = q->plane;
        off = q->px_off;
    }
    hipSetDevice(s->device);
    if (tri_id) HIP_CHECK_RET(hipMemcpy(tri_id, base + off * 4, n * 4, hipMemcpyDeviceToHost));
    if (depth) HIP_CHECK_RET(hipMemcpy(depth, base + (plane + off) * 4, n * 4, hipMemcpyDeviceToHost));
    return SR_OK;
}

SR_API int sr_server_release(sr_server *s, uint64_t id) {
    if (!s) return fail(SR_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.released) return fail(SR_ERR_NOT_FOUND, "no request %llu", (unsigned long long)id);
    if (!it->second.done) return fail(SR_ERR_BUSY, "request %llu not complete", (unsigned long long)id);
    int rg = it->second.region;
    s->reqs.erase(it);
    if (rg >= 0 && --s->regions[rg].refs == 0) {
        s->arena.release(s->regions[rg].off, s->regions[rg].size);
        s->free_regions.push_back(rg);
    }
    s->gen++;
    s->cv.notify_all();
    return SR_OK;
}

SR_API int sr_server_drain(sr_server *s) {
    if (!s) return fail(SR_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->inflight == 0; });
    return SR_OK;
}

SR_API double sr_server_clock(const sr_server *s) { return s ? s->now() : 0.0; }

}  // extern "C"
