// renderer.cpp -- the drop-in per-frame renderer (Renderer, src/renderer.rs:145-355)
// and the batched multi-view launch.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "runtime.h"
#include "swraster.h"

using namespace sr;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

static const float DEFAULT_EXPOSURE = 2.0f;

struct sr_renderer {
    int W = 0, H = 0;
    int device = 0;
    Pipeline pipe;
    DevBuf ids, depth, color, rgba, lum;
    DevBuf expo;                 // {ev, target, exposure} of the device exposure step
    bool expo_on_device = false; // the device state is newer than the host fields
    int ref_tiles = 0;
    float auto_exposure = DEFAULT_EXPOSURE, auto_exposure_target = DEFAULT_EXPOSURE, auto_exposure_ev = 1.0f;
    sr_frame_stats stats;
    bool times_pending = false;  // the checked frame's stage times are read on request
    bool has_frame = false;
    // The last frame is queued and not yet checked for overflow; what a rerun
    // needs is kept until then (the scene counts it in pending_frames).
    bool pending = false;
    const sr_scene *p_scene = nullptr;
    std::vector<DevView> p_views;
    std::vector<DevDraw> p_draws;
    uint32_t p_xv = 0, p_in = 0, p_items = 0;
    bool p_tangents = false, p_translucent = false;
    // a device exposure step queued behind the pending frame: its arguments,
    // and its state before the step saved at expo + 4 floats
    bool p_expo = false, pe_use_host = false;
    float pe_dt = 0.0f, pe_ev = 0.0f, pe_target = 0.0f;
};

struct sr_batch {
    int device = 0;
    Pipeline pipe;
    DevBuf ids, depth, rgba, lum;
    DevBuf color;                // colour planes, only for batches with translucent primitives
    std::vector<uint64_t> fb_offset;
    std::vector<int> width, height;
    sr_batch_stats stats;
    int persistent = 0;
    bool use_graph = false;
};

static float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();    // an unrecorded stage event is not an error of a later launch
        return 0.0f;
    }
    return ms;
}

extern "C" {

// ---- renderer ---------------------------------------------------------------
SR_API int sr_renderer_create(int32_t width, int32_t height, const sr_renderer_opts *opts, sr_renderer **out) {
    if (!out || width <= 0 || height <= 0 || width > 32000 || height > 32000)
        return fail(SR_ERR_INVALID_ARGUMENT, "renderer size must be in 1..32000");
    int32_t ndev = 0;
    sr_device_count(&ndev);
    int dev = opts ? opts->device : 0;
    if (dev < 0 || dev >= ndev) return fail(SR_ERR_INVALID_ARGUMENT, "no HIP device %d (have %d)", dev, ndev);
    sr_renderer *r = new sr_renderer();
    r->W = width;
    r->H = height;
    r->device = dev;
    memset(&r->stats, 0, sizeof(r->stats));
    r->auto_exposure_ev = std::log2(DEFAULT_EXPOSURE);
    int rc = r->pipe.init(dev, opts ? opts->tile_size : 0);
    size_t np = (size_t)width * height;
    int rtx = (width + SR_REF_TILE - 1) / SR_REF_TILE, rty = (height + SR_REF_TILE - 1) / SR_REF_TILE;
    r->ref_tiles = rtx * rty;
    if (!rc) rc = r->ids.ensure(np * 4);
    if (!rc) rc = r->depth.ensure(np * 4);
    if (!rc) rc = r->color.ensure(np * 12);
    if (!rc) rc = r->rgba.ensure(np * 4);
    if (!rc) rc = r->lum.ensure((size_t)r->ref_tiles * 4);
    if (!rc) rc = r->expo.ensure(32);
    if (rc) {
        r->pipe.destroy();
        delete r;
        return rc;
    }
    // the first frame's tile luminance is 1.0 as in the reference (src/renderer.rs:119)
    std::vector<float> ones(r->ref_tiles, 1.0f);
    hipMemcpy(r->lum.ptr, ones.data(), ones.size() * 4, hipMemcpyHostToDevice);
    *out = r;
    return SR_OK;
}

// Stage times and counts of the frame just checked.
static void fill_stats(sr_renderer *r) {
    r->stats.input_triangles = r->p_in;
    r->stats.setup_triangles = r->pipe.last_records;
    r->stats.binned_pairs = r->pipe.last_pairs;
    r->stats.tiles = r->p_items;
    r->stats.draws = (uint32_t)r->p_draws.size();
    r->stats.depth_rejection = r->pipe.last_hiz ? 1u : 0u;
    r->times_pending = true;
}

// The checked frame's stage times, from its events or its wall-clock
// stamps (which stay in the pipeline's clock buffer until the next frame).
static void stage_times(sr_renderer *r) {
    if (!r->times_pending) return;
    r->times_pending = false;
    sr_frame_stats &s = r->stats;
    s.ms_setup = s.ms_bin = s.ms_raster_shade = s.ms_raster = s.ms_shade = s.ms_total = 0.0f;
    const Pipeline &p = r->pipe;
    if (p.timing == SR_STAGE_TIMING_EVENTS) {
        s.ms_setup = ev_ms(p.ev[0], p.ev[1]);
        s.ms_bin = ev_ms(p.ev[1], p.ev[2]);
        s.ms_raster_shade = ev_ms(p.ev[2], p.ev[3]);
        s.ms_raster = ev_ms(p.ev[2], p.ev[4]);
        s.ms_shade = ev_ms(p.ev[4], p.ev[3]);
        s.ms_total = ev_ms(p.ev[0], p.ev[3]);
        return;
    }
    if (p.timing != SR_STAGE_TIMING_CLOCK || !p.clock_dev || p.clock_ticks_per_ms <= 0.0) return;
    StageClock c;
    hipSetDevice(r->device);
    if (hipMemcpyAsync(&c, p.clock_dev, sizeof(c), hipMemcpyDeviceToHost, p.stream) != hipSuccess ||
        hipStreamSynchronize(p.stream) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    auto end_of = [&](int k) {
        uint64_t m = 0;
        for (int i = 0; i < CLOCK_SPREAD; i++) m = std::max<uint64_t>(m, c.end[k][i]);
        return m;
    };
    // a frame without triangles runs no setup: its setup ends where it starts
    uint64_t t0 = c.start[CLOCK_FRONT], t1 = c.start[CLOCK_SETUP_END], t2 = c.start[CLOCK_RASTER];
    if (!t1) t1 = t0;
    uint64_t te = std::max(end_of(CLOCK_SHADE_END), end_of(CLOCK_TR_END));
    auto ms = [&](uint64_t a, uint64_t b) {
        return a && b && b >= a ? (float)((double)(b - a) / p.clock_ticks_per_ms) : 0.0f;
    };
    // the stamps split the frame as the reference's timers do (clip/bin,
    // then raster with shading); the raster / shading split takes events
    s.ms_setup = ms(t0, t1);
    s.ms_bin = ms(t1, t2);
    s.ms_raster_shade = ms(t2, te);
    s.ms_total = ms(t0, te);
}

static hipError_t queue_exposure(sr_renderer *r) {
    return launch_exposure((const float *)r->lum.ptr, (uint32_t)r->ref_tiles, r->pe_dt, r->pe_use_host, r->pe_ev,
                           r->pe_target, (float *)r->expo.ptr, r->pipe.stream);
}

// Completes the queued frame: waits for the stream, checks the frame's
// status and, when a capacity overflowed, runs the frame again with grown
// buffers (replaying a device exposure step queued behind it).  reran tells
// the caller to queue again what it queued behind the frame.
static int finish(sr_renderer *r, bool *reran = nullptr) {
    if (reran) *reran = false;
    if (!r->pending) return SR_OK;
    r->pending = false;
    hipSetDevice(r->device);
    int rc = SR_OK;
    bool again = false;
    if (hipStreamSynchronize(r->pipe.stream) != hipSuccess) rc = fail(SR_ERR_HIP, "frame: %s", hipGetErrorString(hipGetLastError()));
    if (!rc) rc = r->pipe.settle(r->p_translucent, again);
    if (!rc && again) {
        if (r->p_expo && hipMemcpyAsync(r->expo.ptr, (float *)r->expo.ptr + 4, 12, hipMemcpyDeviceToDevice,
                                        r->pipe.stream) != hipSuccess)
            rc = fail(SR_ERR_HIP, "exposure state restore");
        if (!rc)
            rc = r->pipe.run(r->p_views, r->p_draws, r->p_xv, r->p_in, r->p_items, r->p_tangents,
                             (uint32_t *)r->ids.ptr, (float *)r->depth.ptr, (float *)r->color.ptr,
                             (uint64_t)r->W * r->H, nullptr, (float *)r->lum.ptr, 0, true);
        if (!rc && r->p_expo && queue_exposure(r) != hipSuccess) rc = fail(SR_ERR_HIP, "exposure replay");
        if (reran) *reran = true;
    }
    r->p_expo = false;
    const_cast<sr_scene *>(r->p_scene)->pending_frames--;
    r->p_scene = nullptr;
    if (rc) {
        r->has_frame = false;
        return rc;
    }
    fill_stats(r);
    return SR_OK;
}

SR_API int sr_renderer_synchronize(sr_renderer *r) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    return finish(r);
}

SR_API int sr_renderer_destroy(sr_renderer *r) {
    if (!r) return SR_OK;
    (void)finish(r);
    hipSetDevice(r->device);
    r->pipe.destroy();
    r->ids.release();
    r->depth.release();
    r->color.release();
    r->rgba.release();
    r->lum.release();
    r->expo.release();
    delete r;
    return SR_OK;
}

// The frame is queued on the renderer's stream; the first call that reads
// it (blit, present, exposure from its luminance, vbuffer/colour reads,
// stats, synchronize) waits for it, checks it and reruns it if a capacity
// overflowed, so a frame costs one host round trip with its blit.
SR_API int sr_renderer_render_scene(sr_renderer *r, const sr_scene *scene, const sr_camera *camera) {
    if (!r || !scene || !camera) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (scene->device != r->device) return fail(SR_ERR_INVALID_ARGUMENT, "scene and renderer are on different devices");
    int rc = finish(r);        // the previous frame's overflow check sizes this one
    if (rc) return rc;
    r->has_frame = false;
    r->p_draws.clear();
    r->p_in = r->p_xv = 0;
    rc = build_draws(scene, *camera, 0, r->p_in, r->p_xv, r->p_draws, true);
    if (rc) return rc;
    r->p_views.assign(1, DevView());
    int T = r->pipe.tile_size;
    fill_view(r->p_views[0], scene, *camera, r->W, r->H, T, 0, 0, 0, 0, r->auto_exposure, 0);
    r->p_items = (uint32_t)(r->p_views[0].tiles_x * r->p_views[0].tiles_y);
    r->p_tangents = scene->any_tangents;
    r->p_translucent = false;
    for (const DevDraw &d : r->p_draws) r->p_translucent = r->p_translucent || !(d.flags & SR_DRAW_OPAQUE);
    rc = r->pipe.run(r->p_views, r->p_draws, r->p_xv, r->p_in, r->p_items, r->p_tangents, (uint32_t *)r->ids.ptr,
                     (float *)r->depth.ptr, (float *)r->color.ptr, (uint64_t)r->W * r->H, nullptr,
                     (float *)r->lum.ptr, 0, false);
    if (rc) {
        hipStreamSynchronize(r->pipe.stream);
        return rc;
    }
    r->has_frame = true;
    r->pending = true;
    r->p_expo = false;
    r->p_scene = scene;
    const_cast<sr_scene *>(scene)->pending_frames++;
    return SR_OK;
}

// Brings the host's exposure fields up to date with a device exposure step.
static int sync_exposure(sr_renderer *r) {
    if (!r->expo_on_device) return SR_OK;
    int rc = finish(r);
    if (rc) return rc;
    hipSetDevice(r->device);
    float st[4];
    HIP_TRY(hipMemcpyAsync(st, r->expo.ptr, 12, hipMemcpyDeviceToHost, r->pipe.stream));
    HIP_TRY(hipStreamSynchronize(r->pipe.stream));
    r->auto_exposure_ev = st[0];
    r->auto_exposure_target = st[1];
    r->auto_exposure = st[2];
    r->expo_on_device = false;
    return SR_OK;
}

// Renderer::update_auto_exposure (src/renderer.rs:258-290)
SR_API int sr_renderer_update_auto_exposure(sr_renderer *r, float dt) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    int src = sync_exposure(r);
    if (src) return src;
    size_t n = (size_t)r->ref_tiles;
    if (n == 0) return SR_OK;
    float dtc = std::fmax(dt, 0.0f);
    float tau = std::fmax(1.0f, 1e-4f);
    float alpha = 1.0f - std::exp(-(dtc / tau));
    if (alpha == 0.0f) {
        // the adaptation step is zero: exposure is 2^ev whatever the target
        r->auto_exposure = std::pow(2.0f, r->auto_exposure_ev);
        return SR_OK;
    }
    if ((src = finish(r))) return src;
    hipSetDevice(r->device);
    std::vector<float> s(n);
    HIP_TRY(hipMemcpy(s.data(), r->lum.ptr, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) s[i] = std::log2(std::fmax(s[i], 1e-4f));
    auto key = [](float f) {
        int32_t b;
        memcpy(&b, &f, 4);
        return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
    };
    std::sort(s.begin(), s.end(), [&](float a, float b) { return key(a) < key(b); });
    size_t trim = (size_t)std::floor((float)n * 0.10f);
    trim = std::min(trim, (n - 1) / 2);
    float sum = 0.0f;
    for (size_t i = trim; i < n - trim; i++) sum += s[i];
    float mean = sum / (float)(n - 2 * trim);
    float y = std::fmin(std::fmax(0.45f, 0.0f), 1.0f);
    float k = 0.2f;
    float denom = std::fmax(1.0f + k - y, 1e-6f);
    float meter = std::fmax((y * k) / denom, 1e-4f);
    float target_ev = std::log2(meter) - mean;
    float target = std::fmin(std::fmax(std::pow(2.0f, target_ev), 0.05f), 32.0f);
    r->auto_exposure_target = target;
    float tev = std::log2(r->auto_exposure_target);
    r->auto_exposure_ev += (tev - r->auto_exposure_ev) * alpha;
    r->auto_exposure = std::pow(2.0f, r->auto_exposure_ev);
    return SR_OK;
}

// The same step on the device, queued on the renderer's stream without a
// host round trip; later blits read its exposure there.
SR_API int sr_renderer_update_auto_exposure_device(sr_renderer *r, float dt) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    if (r->ref_tiles > SR_EXPOSURE_MAX) return sr_renderer_update_auto_exposure(r, dt);
    if (r->p_expo) {               // one step replays after a rerun; a second one waits for the frame
        int rc = finish(r);
        if (rc) return rc;
    }
    hipSetDevice(r->device);
    r->pe_dt = dt;
    r->pe_use_host = !r->expo_on_device;
    r->pe_ev = r->auto_exposure_ev;
    r->pe_target = r->auto_exposure_target;
    if (r->pending) {
        HIP_TRY(hipMemcpyAsync((float *)r->expo.ptr + 4, r->expo.ptr, 12, hipMemcpyDeviceToDevice, r->pipe.stream));
        r->p_expo = true;
    }
    HIP_TRY(queue_exposure(r));
    r->expo_on_device = true;
    return SR_OK;
}

SR_API float sr_renderer_exposure(const sr_renderer *r) {
    if (!r) return 0.0f;
    if (sync_exposure(const_cast<sr_renderer *>(r))) return 0.0f;
    return r->auto_exposure;
}

SR_API int sr_renderer_set_exposure(sr_renderer *r, float exposure) {
    if (!r || !(exposure > 0.0f)) return fail(SR_ERR_INVALID_ARGUMENT, "exposure must be positive");
    r->auto_exposure = exposure;
    r->auto_exposure_ev = std::log2(exposure);
    r->expo_on_device = false;
    return SR_OK;
}

// The blit of the current frame into dst, queued behind the frame and, if
// checking the frame reran it, queued again.
static int blit_frame(sr_renderer *r, uint32_t *dst, bool frame_bytes) {
    hipSetDevice(r->device);
    for (int k = 0; k < 2; k++) {
        HIP_TRY(launch_blit((const float *)r->color.ptr, (uint64_t)r->W * r->H, r->W, r->H, r->auto_exposure, dst,
                            r->pipe.stream, r->expo_on_device ? (const float *)r->expo.ptr : nullptr, frame_bytes));
        bool reran = false;
        int rc = finish(r, &reran);
        if (rc) return rc;
        if (!reran) break;
    }
    return SR_OK;
}

SR_API int sr_renderer_blit_to_device(sr_renderer *r, void *device_pixels) {
    if (!r || !device_pixels) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (!r->has_frame) return fail(SR_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    int rc = blit_frame(r, (uint32_t *)device_pixels, false);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(r->pipe.stream));
    return SR_OK;
}

SR_API int sr_renderer_blit_to_buffer(sr_renderer *r, uint32_t *pixels, size_t width, size_t height) {
    if (!r || !pixels) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if ((int)width != r->W || (int)height != r->H)
        return fail(SR_ERR_INVALID_ARGUMENT, "buffer is %zux%zu, renderer is %dx%d", width, height, r->W, r->H);
    if (!r->has_frame) return fail(SR_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    int rc = blit_frame(r, (uint32_t *)r->rgba.ptr, false);
    if (rc) return rc;
    size_t rows = (size_t)(r->H & ~1);
    HIP_TRY(hipMemcpyAsync(pixels, r->rgba.ptr, rows * (size_t)r->W * 4, hipMemcpyDeviceToHost, r->pipe.stream));
    HIP_TRY(hipStreamSynchronize(r->pipe.stream));
    return SR_OK;
}

// The presented frame: the blit with each pixel stored as the bytes r, g, b,
// a (App::render_frame's copy into the pixels frame, src/main.rs:574-590).
static int present(sr_renderer *r, uint32_t *dst) { return blit_frame(r, dst, true); }

SR_API int sr_renderer_present(sr_renderer *r, uint8_t *frame, size_t width, size_t height) {
    if (!r || !frame) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if ((int)width != r->W || (int)height != r->H)
        return fail(SR_ERR_INVALID_ARGUMENT, "frame is %zux%zu, renderer is %dx%d", width, height, r->W, r->H);
    if (!r->has_frame) return fail(SR_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    hipSetDevice(r->device);
    int rc = present(r, (uint32_t *)r->rgba.ptr);
    if (rc) return rc;
    size_t rows = (size_t)(r->H & ~1);
    HIP_TRY(hipMemcpyAsync(frame, r->rgba.ptr, rows * (size_t)r->W * 4, hipMemcpyDeviceToHost, r->pipe.stream));
    HIP_TRY(hipStreamSynchronize(r->pipe.stream));
    return SR_OK;
}

SR_API int sr_renderer_present_to_device(sr_renderer *r, void *device_frame) {
    if (!r || !device_frame) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (!r->has_frame) return fail(SR_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    hipSetDevice(r->device);
    int rc = present(r, (uint32_t *)device_frame);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(r->pipe.stream));
    return SR_OK;
}

SR_API int sr_renderer_read_vbuffer(sr_renderer *r, uint32_t *tri_id, float *depth) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    if (int rc = finish(r)) return rc;
    if (!r->has_frame) return fail(SR_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    hipSetDevice(r->device);
    size_t n = (size_t)r->W * r->H;
    if (tri_id) HIP_TRY(hipMemcpy(tri_id, r->ids.ptr, n * 4, hipMemcpyDeviceToHost));
    if (depth) HIP_TRY(hipMemcpy(depth, r->depth.ptr, n * 4, hipMemcpyDeviceToHost));
    return SR_OK;
}

SR_API int sr_renderer_read_color(sr_renderer *r, float *rgb) {
    if (!r || !rgb) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = finish(r)) return rc;
    if (!r->has_frame) return fail(SR_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    hipSetDevice(r->device);
    size_t n = (size_t)r->W * r->H;
    std::vector<float> planes(n * 3);
    HIP_TRY(hipMemcpy(planes.data(), r->color.ptr, n * 12, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        rgb[3 * i] = planes[i];
        rgb[3 * i + 1] = planes[n + i];
        rgb[3 * i + 2] = planes[2 * n + i];
    }
    return SR_OK;
}

SR_API int sr_renderer_read_tile_luminance(sr_renderer *r, float *lum, int32_t *count) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    if (int rc = finish(r)) return rc;
    if (count) *count = r->ref_tiles;
    if (lum) {
        hipSetDevice(r->device);
        HIP_TRY(hipMemcpy(lum, r->lum.ptr, (size_t)r->ref_tiles * 4, hipMemcpyDeviceToHost));
    }
    return SR_OK;
}

SR_API int sr_renderer_get_stats(const sr_renderer *r, sr_frame_stats *stats) {
    if (!r || !stats) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = finish(const_cast<sr_renderer *>(r))) return rc;
    stage_times(const_cast<sr_renderer *>(r));
    *stats = r->stats;
    return SR_OK;
}

SR_API int sr_renderer_set_stage_timing(sr_renderer *r, int32_t mode) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    if (mode != SR_STAGE_TIMING_CLOCK && mode != SR_STAGE_TIMING_EVENTS && mode != SR_STAGE_TIMING_OFF)
        return fail(SR_ERR_INVALID_ARGUMENT, "stage timing mode %d (0 clock, 1 events, 2 off)", mode);
    // the queued frame keeps the mode it was launched with
    if (int rc = finish(r)) return rc;
    stage_times(r);
    r->pipe.timing = mode;
    return SR_OK;
}

SR_API int sr_renderer_read_bins(sr_renderer *r, uint32_t *start, uint32_t *end, uint32_t *list, uint64_t list_len,
                                 uint32_t *items) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    if (int rc = finish(r)) return rc;
    if (!r->has_frame) return fail(SR_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    hipSetDevice(r->device);
    size_t n = r->p_items;
    if (items) *items = (uint32_t)n;
    if (start || end) {
        // the device marks the bounds of tiles with entries and leaves the
        // others (0, 0); the tiles' runs follow each other in tile order, so
        // an empty tile starts and ends where the previous one ends
        std::vector<uint32_t> s(n), e(n);
        HIP_TRY(hipMemcpy(s.data(), r->pipe.bin_start.ptr, n * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(e.data(), r->pipe.bin_end.ptr, n * 4, hipMemcpyDeviceToHost));
        uint32_t at = 0;
        for (size_t i = 0; i < n; i++) {
            if (e[i] == s[i]) s[i] = e[i] = at;
            at = e[i];
        }
        if (start) memcpy(start, s.data(), n * 4);
        if (end) memcpy(end, e.data(), n * 4);
    }
    if (list && list_len) {
        uint64_t total = std::min<uint64_t>(r->pipe.last_pairs, r->pipe.pair_cap);
        HIP_TRY(hipMemcpy(list, r->pipe.list.ptr, std::min<uint64_t>(total, list_len) * 4, hipMemcpyDeviceToHost));
    }
    return SR_OK;
}

SR_API int sr_renderer_device_buffers(sr_renderer *r, void **tri_id, void **depth, void **color_rgb) {
    if (!r) return fail(SR_ERR_INVALID_ARGUMENT, "null renderer");
    if (int rc = finish(r)) return rc;
    if (tri_id) *tri_id = r->ids.ptr;
    if (depth) *depth = r->depth.ptr;
    if (color_rgb) *color_rgb = r->color.ptr;
    return SR_OK;
}

// ---- batch -------------------------------------------------------------------
SR_API int sr_batch_create(const sr_batch_opts *opts, sr_batch **out) {
    if (!out) return fail(SR_ERR_INVALID_ARGUMENT, "null output");
    int32_t ndev = 0;
    sr_device_count(&ndev);
    int dev = opts ? opts->device : 0;
    if (dev < 0 || dev >= ndev) return fail(SR_ERR_INVALID_ARGUMENT, "no HIP device %d (have %d)", dev, ndev);
    sr_batch *b = new sr_batch();
    b->device = dev;
    memset(&b->stats, 0, sizeof(b->stats));
    int rc = b->pipe.init(dev, opts ? opts->tile_size : 0);
    if (rc) {
        delete b;
        return rc;
    }
    // a batch reports its whole time (events around the batch, or carried by
    // its first and last launches when it runs without a graph)
    b->pipe.timing = SR_STAGE_TIMING_EVENTS;
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
    b->persistent = opts && opts->persistent_blocks > 0 ? opts->persistent_blocks : cus * 8;
    b->use_graph = opts && opts->use_graph;
    *out = b;
    return SR_OK;
}

SR_API int sr_batch_configure(sr_batch *b, int32_t persistent_blocks, int32_t use_graph) {
    if (!b) return fail(SR_ERR_INVALID_ARGUMENT, "null batch");
    if (persistent_blocks < -1) return fail(SR_ERR_INVALID_ARGUMENT, "workgroup count below -1");
    if (persistent_blocks == -1) b->persistent = 0;     // one raster workgroup per (view, tile)
    if (persistent_blocks > 0) b->persistent = persistent_blocks;
    b->use_graph = use_graph != 0;
    return SR_OK;
}

SR_API int sr_batch_destroy(sr_batch *b) {
    if (!b) return SR_OK;
    hipSetDevice(b->device);
    b->pipe.destroy();
    b->ids.release();
    b->depth.release();
    b->rgba.release();
    b->lum.release();
    b->color.release();
    delete b;
    return SR_OK;
}

SR_API int sr_batch_render(sr_batch *b, const sr_view *views, int32_t n) {
    if (!b || !views || n <= 0) return fail(SR_ERR_INVALID_ARGUMENT, "empty batch");
    std::vector<DevView> hv(n);
    std::vector<DevDraw> draws;
    uint32_t in_tri = 0, xv = 0, items = 0, lum = 0;
    uint64_t fb = 0;
    bool tangents = false;
    int T = b->pipe.tile_size;
    b->fb_offset.assign(n, 0);
    b->width.assign(n, 0);
    b->height.assign(n, 0);
    for (int i = 0; i < n; i++) {
        const sr_view &v = views[i];
        if (!v.scene || v.width <= 0 || v.height <= 0 || v.width > 32000 || v.height > 32000)
            return fail(SR_ERR_INVALID_ARGUMENT, "view %d is invalid", i);
        if (v.scene->device != b->device) return fail(SR_ERR_INVALID_ARGUMENT, "view %d scene on another device", i);
        uint32_t tri0 = in_tri;
        int rc = build_draws(v.scene, v.camera, i, in_tri, xv, draws, true);
        if (rc) return rc;
        fill_view(hv[i], v.scene, v.camera, v.width, v.height, T, items, tri0, fb, lum,
                  v.exposure > 0.0f ? v.exposure : DEFAULT_EXPOSURE, 1u);
        b->fb_offset[i] = fb;
        b->width[i] = v.width;
        b->height[i] = v.height;
        items += (uint32_t)(hv[i].tiles_x * hv[i].tiles_y);
        lum += (uint32_t)(hv[i].ref_tiles_x * hv[i].ref_tiles_y);
        fb += (uint64_t)v.width * v.height;
        tangents = tangents || v.scene->any_tangents;
    }
    // Translucent primitives blend over the opaque colour before the
    // tonemap, so such a batch writes colour planes and blits each view
    // afterwards instead of fusing the tonemap into the shading.
    bool has_tr = false;
    for (const DevDraw &d : draws) has_tr = has_tr || !(d.flags & SR_DRAW_OPAQUE);
    if (has_tr)
        for (DevView &v : hv) v.flags &= ~1u;
    hipSetDevice(b->device);
    int rc;
    if ((rc = b->ids.ensure(fb * 4))) return rc;
    if ((rc = b->depth.ensure(fb * 4))) return rc;
    if ((rc = b->rgba.ensure(fb * 4))) return rc;
    if ((rc = b->lum.ensure((size_t)lum * 4))) return rc;
    if (has_tr && (rc = b->color.ensure(fb * 12))) return rc;
    rc = b->pipe.run(hv, draws, xv, in_tri, items, tangents, (uint32_t *)b->ids.ptr, (float *)b->depth.ptr,
                     has_tr ? (float *)b->color.ptr : nullptr, has_tr ? fb : 0, (uint32_t *)b->rgba.ptr,
                     (float *)b->lum.ptr, b->persistent, true, b->use_graph);
    if (rc) return rc;
    if (has_tr) {
        for (int i = 0; i < n; i++)
            HIP_TRY(launch_blit((const float *)b->color.ptr + b->fb_offset[i], fb, b->width[i], b->height[i],
                                hv[i].exposure, (uint32_t *)b->rgba.ptr + b->fb_offset[i], b->pipe.stream));
        HIP_TRY(hipStreamSynchronize(b->pipe.stream));
    }
    b->stats.input_triangles = in_tri;
    b->stats.setup_triangles = b->pipe.last_records;
    b->stats.binned_pairs = b->pipe.last_pairs;
    b->stats.pixels = fb;
    b->stats.views = (uint32_t)n;
    b->stats.work_items = items;
    b->stats.ms_total = ev_ms(b->pipe.ev[0], b->pipe.ev[3]);
    return SR_OK;
}

SR_API int sr_batch_read_vbuffer(sr_batch *b, int32_t view, uint32_t *tri_id, float *depth) {
    if (!b || view < 0 || view >= (int)b->fb_offset.size()) return fail(SR_ERR_INVALID_ARGUMENT, "bad view index");
    hipSetDevice(b->device);
    size_t n = (size_t)b->width[view] * b->height[view];
    if (tri_id)
        HIP_TRY(hipMemcpy(tri_id, (uint32_t *)b->ids.ptr + b->fb_offset[view], n * 4, hipMemcpyDeviceToHost));
    if (depth) HIP_TRY(hipMemcpy(depth, (float *)b->depth.ptr + b->fb_offset[view], n * 4, hipMemcpyDeviceToHost));
    return SR_OK;
}

SR_API int sr_batch_read_pixels(sr_batch *b, int32_t view, uint32_t *pixels) {
    if (!b || !pixels || view < 0 || view >= (int)b->fb_offset.size())
        return fail(SR_ERR_INVALID_ARGUMENT, "bad view index");
    hipSetDevice(b->device);
    size_t rows = (size_t)(b->height[view] & ~1);
    HIP_TRY(hipMemcpy(pixels, (uint32_t *)b->rgba.ptr + b->fb_offset[view], rows * b->width[view] * 4,
                      hipMemcpyDeviceToHost));
    return SR_OK;
}

SR_API int sr_batch_get_stats(const sr_batch *b, sr_batch_stats *stats) {
    if (!b || !stats) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    *stats = b->stats;
    return SR_OK;
}

}  // extern "C"
