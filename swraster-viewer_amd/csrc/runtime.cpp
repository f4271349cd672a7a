// runtime.cpp -- C ABI of the MI355X rasterization pipeline (host side).
//
// Owns device memory and streams, uploads scenes to HBM, builds the per-frame
// draw list (node order, primitive sphere culling, src/renderer.rs:201-468)
// and launches the kernel sequence of kernels.hip.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "prep.h"
#include "runtime.h"
#include "swraster.h"

namespace sr {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
int DevBuf::ensure(size_t need) {
    if (need <= bytes && ptr) return SR_OK;
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    size_t alloc = std::max<size_t>(need, 256);
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(SR_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", alloc, hipGetErrorString(e));
    }
    bytes = alloc;
    return SR_OK;
}
void DevBuf::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}
int PinBuf::ensure(size_t need, unsigned flags) {
    if (need <= bytes && ptr) return SR_OK;
    if (ptr) hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    size_t alloc = std::max<size_t>(need, 256);
    hipError_t e = hipHostMalloc(&ptr, alloc, flags);
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(SR_ERR_OUT_OF_MEMORY, "hipHostMalloc(%zu) failed: %s", alloc, hipGetErrorString(e));
    }
    bytes = alloc;
    return SR_OK;
}
void PinBuf::release() {
    if (ptr) hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// ---------------------------------------------------------------------------
// host-side geometry helpers (restated from src/scene.rs and src/renderer.rs)
// ---------------------------------------------------------------------------
static inline float dot4sse(const float a[4], const float b[4]) {
    float xx = a[0] * b[0], yy = a[1] * b[1], zz = a[2] * b[2], ww = a[3] * b[3];
    return (xx + zz) + (yy + ww);
}
static inline void matvec(const float *m, const float v[4], float o[4]) {
    for (int i = 0; i < 4; i++) {
        float acc = m[i] * v[0];
        acc = acc + m[4 + i] * v[1];
        acc = acc + m[8 + i] * v[2];
        acc = acc + m[12 + i] * v[3];
        o[i] = acc;
    }
}
static inline void matmul(const float *a, const float *b, float *o) {
    for (int c = 0; c < 4; c++) matvec(a, b + 4 * c, o + 4 * c);
}
static inline float len3(float x, float y, float z) { return std::sqrt((x * x + y * y) + z * z); }

// Mat4 * BoundingSphere (src/scene.rs:53-63)
static Sphere transform_sphere(const float *m, const Sphere &s) {
    float c0[4] = {m[0], m[1], m[2], m[3]}, c1[4] = {m[4], m[5], m[6], m[7]}, c2[4] = {m[8], m[9], m[10], m[11]};
    float l0 = std::sqrt(dot4sse(c0, c0)), l1 = std::sqrt(dot4sse(c1, c1)), l2 = std::sqrt(dot4sse(c2, c2));
    float max_scale = ((l0 + l1) + l2) / 3.0f;
    float p[4] = {s.c[0], s.c[1], s.c[2], 1.0f}, o[4];
    matvec(m, p, o);
    return {{o[0], o[1], o[2]}, s.r * max_scale};
}

enum { FT_INSIDE, FT_OUTSIDE, FT_INTERSECTING };
// test_sphere_frustum (src/renderer.rs:130-142)
static int sphere_frustum(const Sphere &s, const sr_camera &c) {
    float p[4] = {s.c[0], s.c[1], s.c[2], 1.0f}, cv[4];
    matvec(c.view_matrix, p, cv);
    int result = FT_INSIDE;
    for (int i = 0; i < 6; i++) {
        float d = dot4sse(c.view_clip_planes[i], cv);
        if (d < -s.r) return FT_OUTSIDE;
        else if (d < s.r) result = FT_INTERSECTING;
    }
    return result;
}

// Draw list of one view in canonical order (src/renderer.rs:201-468).
int build_draws(const sr_scene *scene, const sr_camera &cam, int view, uint32_t &in_tri, uint32_t &xv,
                std::vector<DevDraw> &out, bool allow_translucent) {
    const size_t nn = scene->nodes.size();
    std::vector<int> order(nn);
    std::vector<float> key(nn);
    for (size_t i = 0; i < nn; i++) {
        order[i] = (int)i;
        const Sphere &b = scene->nodes[i].bs_world;
        float dx = cam.position[0] - b.c[0], dy = cam.position[1] - b.c[1], dz = cam.position[2] - b.c[2];
        key[i] = (dx * dx + dy * dy) + dz * dz;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        float ka = key[a], kb = key[b];
        if (std::isnan(ka)) return false;
        if (std::isnan(kb)) return true;
        return ka < kb;
    });
    for (int ni : order) {
        const HostNode &n = scene->nodes[ni];
        if (n.mesh < 0) continue;
        float mvp[16];
        matmul(cam.view_project_matrix, n.transform, mvp);
        const HostMesh &m = scene->meshes[n.mesh];
        for (int pass = 0; pass < 2; pass++) {
            const std::vector<int> &list = pass == 0 ? m.opaque : m.translucent;
            for (int pl : list) {
                int pi = m.prims[pl];
                const HostPrim &p = scene->prims[pi];
                Sphere bw = transform_sphere(n.transform, p.bs);
                int ft = sphere_frustum(bw, cam);
                if (ft == FT_OUTSIDE) continue;
                if (pass == 1 && !allow_translucent)
                    return fail(SR_ERR_UNSUPPORTED, "translucent primitives are not supported on this path");
                DevDraw d;
                memset(&d, 0, sizeof(d));
                memcpy(d.mvp, mvp, 64);
                memcpy(d.model, n.transform, 64);
                d.vtx_base = p.vtx_base;
                d.num_vtx = p.num_vtx;
                d.idx_base = p.idx_base;
                d.num_tri = p.num_tri;
                d.in_tri_base = in_tri;
                d.xv_base = xv;
                d.view = view;
                d.material = p.material;
                d.flags = (ft == FT_INTERSECTING ? SR_DRAW_CLIP : 0u) | (pass == 0 ? SR_DRAW_OPAQUE : 0u) |
                          (p.tangents ? SR_DRAW_TANGENTS : 0u);
                in_tri += p.num_tri;
                xv += p.num_vtx;
                if (p.num_tri == 0) continue;
                out.push_back(d);
            }
        }
    }
    return SR_OK;
}

// ---------------------------------------------------------------------------
// per-frame pipeline state shared by renderer and batch
// ---------------------------------------------------------------------------
int Pipeline::init(int dev, int tile) {
    device = dev;
    tile_size = (tile == 8 || tile == 16 || tile == 32) ? tile : 16;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    clock_ticks_per_ms = (double)khz;
    if (int r = clock_buf.ensure(sizeof(StageClock))) return r;
    // coherent: the shading kernel stores the frame's status here
    return status_host.ensure(sizeof(DevStatus), hipHostMallocCoherent);
}

void Pipeline::destroy() {
    hipSetDevice(device);
    if (stream) hipStreamSynchronize(stream);
    drop_graph();
    DevBuf *bufs[] = {&views, &xv, &xt, &tri_counts, &tri_chunks, &tri_base, &bin_start, &bin_end, &partials,
                      &list, &rec_geo, &rec_raster, &center, &clip_queue, &clip_buf, &chunk_pairs, &pair_keys,
                      &pair_vals, &pair_keys2, &pair_vals2, &sort_hist, &tr_count,
                      &tr_start, &tr_end, &tr_list, &tr_scratch, &center_color, &clock_buf};
    for (DevBuf *b : bufs) b->release();
    stage.release();
    status_host.release();
    for (auto &e : ev)
        if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
    stream = nullptr;
}

void Pipeline::drop_graph() {
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    memset(&graph_key, 0, sizeof(graph_key));
}

// The frame's whole launch sequence (the upload of the view and draw tables
// and the zeroed status from pinned staging, every kernel, the status
// read-back) as one executable graph, re-captured only when the launch shape
// or a buffer changes.  The sequence clears memory with the upload and fill
// kernels only (no memset nodes).
int Pipeline::launch_graph(const FrameArgs &a, size_t vbytes, size_t upload) {
    GraphKey k;
    memset(&k, 0, sizeof(k));
    k.n[0] = (uint64_t)a.num_views; k.n[1] = (uint64_t)a.num_draws; k.n[2] = a.total_xv; k.n[3] = a.total_in;
    k.n[4] = a.total_items; k.n[5] = (uint64_t)a.tile_size; k.n[6] = (uint64_t)a.any_tangents; k.n[7] = a.rec_cap;
    k.n[8] = a.pair_cap; k.n[9] = a.clip_cap; k.n[10] = (uint64_t)a.persistent_blocks; k.n[11] = vbytes;
    k.n[12] = upload; k.n[13] = a.color_plane; k.n[14] = (uint64_t)a.has_translucent;
    k.n[15] = ((uint64_t)a.tr_cap << 32) | a.tr_slots | ((uint64_t)a.hiz << 31);
    const void *ptrs[] = {a.views, a.draws, a.xv, a.xt, a.tri_counts, a.tri_chunks, a.tri_base, a.bin_start, a.bin_end,
                          a.scan_partials, a.list, a.rec_geo, a.rec_raster, a.clip_queue, a.clip_buf, a.status,
                          a.out_ids, a.out_depth, a.out_color, a.out_rgba, a.out_lum, a.center_ids, stage.ptr,
                          status_host.ptr, a.tr_count, a.tr_start, a.tr_end, a.tr_list, a.tr_scratch,
                          a.center_color, a.chunk_pairs, a.pair_keys, a.pair_vals, a.pair_keys2, a.pair_vals2, a.sort_hist,
                          a.clock};
    static_assert(sizeof(ptrs) / sizeof(ptrs[0]) <= sizeof(k.p) / sizeof(k.p[0]), "graph key too small");
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) k.p[i] = (uint64_t)(uintptr_t)ptrs[i];
    if (!graph_exec || memcmp(&k, &graph_key, sizeof(k)) != 0) {
        drop_graph();
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipMemcpyAsync((void *)a.views, stage.ptr, upload, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = launch_frame(a, stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(stream, &g);
        if (e == hipSuccess) e = ce;
        if (e == hipSuccess) e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
        if (g) hipGraphDestroy(g);
        if (e != hipSuccess) {
            graph_exec = nullptr;
            return fail(SR_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        }
        graph_key = k;
    }
    HIP_TRY(hipEventRecord(ev[0], stream));
    HIP_TRY(hipGraphLaunch(graph_exec, stream));
    HIP_TRY(hipEventRecord(ev[3], stream));
    return SR_OK;
}

int Pipeline::run(const std::vector<DevView> &hv, const std::vector<DevDraw> &hd, uint32_t total_xv, uint32_t total_in,
                  uint32_t total_items, bool any_tangents, uint32_t *out_ids, float *out_depth, float *out_color,
                  uint64_t color_plane, uint32_t *out_rgba, float *out_lum, int persistent_blocks, bool sync,
                  bool use_graph) {
    HIP_TRY(hipSetDevice(device));
    // the view table, the draw table and the frame's status (zeroed) share
    // one device buffer at 256-byte aligned offsets, so one copy uploads the
    // tables and clears the status before the first kernel; the binning
    // state follows the status (the frame writes it before reading it)
    size_t vbytes = (hv.size() * sizeof(DevView) + 255) & ~(size_t)255;
    size_t dbytes = (std::max<size_t>(hd.size(), 1) * sizeof(DevDraw) + 255) & ~(size_t)255;
    size_t upload = vbytes + dbytes + sizeof(DevStatus);
    const int sort_passes = bin_sort_passes(total_items);
    if (sort_passes < 1) return fail(SR_ERR_CAPACITY, "%u screen tiles in one frame; at most 2^24", total_items);
    int r;
    if ((r = stage.ensure(upload))) return r;
    memcpy(stage.ptr, hv.data(), hv.size() * sizeof(DevView));
    if (!hd.empty()) memcpy((char *)stage.ptr + vbytes, hd.data(), hd.size() * sizeof(DevDraw));
    if ((r = views.ensure(upload + sizeof(BinSortState)))) return r;
    if ((r = xv.ensure((size_t)std::max<uint32_t>(total_xv, 1) * 64))) return r;
    if (any_tangents && (r = xt.ensure((size_t)std::max<uint32_t>(total_xv, 1) * 16))) return r;
    if ((r = tri_counts.ensure(((size_t)total_in + 1) * 4))) return r;
    if ((r = tri_chunks.ensure(((size_t)total_in / 256 + 2) * 4))) return r;
    if ((r = tri_base.ensure(((size_t)total_in + 1) * 4))) return r;
    if ((r = chunk_pairs.ensure(((size_t)total_in / 256 + 2) * 4))) return r;
    if ((r = bin_start.ensure((size_t)std::max<uint32_t>(total_items, 1) * 4))) return r;
    if ((r = bin_end.ensure((size_t)std::max<uint32_t>(total_items, 1) * 4))) return r;
    size_t lum_slots = 1;
    for (const DevView &v : hv)
        lum_slots = std::max<size_t>(lum_slots, (size_t)v.lum_offset + (size_t)v.ref_tiles_x * v.ref_tiles_y);
    if ((r = center.ensure(lum_slots * 16))) return r;
    bool has_tr = false;
    for (const DevDraw &d : hd) has_tr = has_tr || !(d.flags & SR_DRAW_OPAQUE);
    if (has_tr) {
        if ((r = tr_count.ensure(lum_slots * 4))) return r;
        if ((r = tr_start.ensure(lum_slots * 4))) return r;
        if ((r = tr_end.ensure(lum_slots * 4))) return r;
        if ((r = center_color.ensure(lum_slots * 64))) return r;
        if (tr_cap < 4096) tr_cap = 4096;
    }
    size_t maxn = std::max<size_t>((size_t)total_in + 1, total_items);
    if ((r = partials.ensure(((maxn + 4095) / 4096 + 1) * 4))) return r;
    if (rec_cap < total_in + total_in / 8 + 1024) rec_cap = total_in + total_in / 8 + 1024;
    if (pair_cap < rec_cap * 2) pair_cap = rec_cap * 2;
    if (clip_cap < 4096) clip_cap = 4096;
    for (int attempt = 0; attempt < 4; attempt++) {
        // every attempt is a frame of its own number
        DevStatus init;
        memset(&init, 0, sizeof(init));
        init.seq = ++seq;
        memcpy((char *)stage.ptr + vbytes + dbytes, &init, sizeof(init));
        if ((r = rec_geo.ensure((size_t)rec_cap * 16))) return r;
        if ((r = rec_raster.ensure((size_t)rec_cap * 64))) return r;
        if ((r = list.ensure((size_t)pair_cap * 4))) return r;
        if ((r = clip_queue.ensure((size_t)clip_cap * 4))) return r;
        if ((r = clip_buf.ensure((size_t)clip_cap * 16 * 16 * 4))) return r;
        if ((r = pair_keys.ensure((size_t)pair_cap * 4)) || (r = pair_vals.ensure((size_t)pair_cap * 4))) return r;
        // (the last pass writes its bins to the key array it does not read)
        if ((r = pair_keys2.ensure((size_t)pair_cap * 4))) return r;
        if (sort_passes > 1 && (r = pair_vals2.ensure((size_t)pair_cap * 4))) return r;
        if ((r = sort_hist.ensure((size_t)bin_sort_tiles(pair_cap) * 256 * 4))) return r;
        if (has_tr && (r = tr_list.ensure((size_t)tr_cap * 4))) return r;
        if (has_tr && (r = tr_scratch.ensure((size_t)tr_cap * 8))) return r;
        FrameArgs a;
        memset(&a, 0, sizeof(a));
        a.views = (const DevView *)views.ptr;
        a.num_views = (int)hv.size();
        a.draws = (const DevDraw *)((char *)views.ptr + vbytes);
        a.num_draws = (int)hd.size();
        a.total_xv = total_xv;
        a.total_in = total_in;
        a.total_items = total_items;
        a.tile_size = tile_size;
        a.any_tangents = any_tangents ? 1 : 0;
        a.xv = (float4 *)xv.ptr;
        a.xt = any_tangents ? (float4 *)xt.ptr : nullptr;
        a.tri_counts = (uint32_t *)tri_counts.ptr;
        a.tri_chunks = (uint32_t *)tri_chunks.ptr;
        a.tri_base = (uint32_t *)tri_base.ptr;
        a.bin_start = (uint32_t *)bin_start.ptr;
        a.bin_end = (uint32_t *)bin_end.ptr;
        a.scan_partials = (uint32_t *)partials.ptr;
        a.list = (uint32_t *)list.ptr;
        a.rec_geo = (uint4 *)rec_geo.ptr;
        a.rec_raster = (float4 *)rec_raster.ptr;
        a.rec_cap = rec_cap;
        a.pair_cap = pair_cap;
        a.clip_queue = (uint32_t *)clip_queue.ptr;
        a.clip_buf = (float *)clip_buf.ptr;
        a.clip_cap = clip_cap;
        a.status = (DevStatus *)((char *)views.ptr + vbytes + dbytes);
        a.sort = (BinSortState *)((char *)a.status + sizeof(DevStatus));
        a.clock = timing == SR_STAGE_TIMING_CLOCK ? (StageClock *)clock_buf.ptr : nullptr;
        clock_dev = a.clock;
        a.pair_keys = (uint32_t *)pair_keys.ptr;
        a.pair_vals = (uint32_t *)pair_vals.ptr;
        a.pair_keys2 = (uint32_t *)pair_keys2.ptr;
        a.pair_vals2 = (uint32_t *)pair_vals2.ptr;
        a.chunk_pairs = (uint32_t *)chunk_pairs.ptr;
        a.sort_hist = (uint32_t *)sort_hist.ptr;
        a.status_out = (DevStatus *)status_host.ptr;
        status_dev = a.status;
        a.out_ids = out_ids;
        a.out_depth = out_depth;
        a.out_color = out_color;
        a.color_plane = color_plane;
        a.out_rgba = out_rgba;
        a.out_lum = out_lum;
        a.center_ids = (uint32_t *)center.ptr;
        if (has_tr) {
            a.has_translucent = 1;
            a.tr_slots = (uint32_t)lum_slots;
            a.tr_count = (uint32_t *)tr_count.ptr;
            a.tr_start = (uint32_t *)tr_start.ptr;
            a.tr_end = (uint32_t *)tr_end.ptr;
            a.tr_list = (uint32_t *)tr_list.ptr;
            a.tr_scratch = (unsigned long long *)tr_scratch.ptr;
            a.tr_cap = tr_cap;
            a.center_color = (float4 *)center_color.ptr;
        }
        a.persistent_blocks = persistent_blocks;
        // depth rejection pays where triangles span many tiles and tiles
        // hold many of them (high overdraw); judged from the last frame
        a.hiz = tile_size == 16 && last_pairs > 32ull * total_items && last_pairs > 4ull * last_records;
        last_hiz = a.hiz != 0;
        if (use_graph) {
            if ((r = launch_graph(a, vbytes, upload))) return r;
        } else {
            HIP_TRY(hipMemcpyAsync(views.ptr, stage.ptr, upload, hipMemcpyHostToDevice, stream));
            for (int i = 0; i < 5; i++) a.events[i] = timing == SR_STAGE_TIMING_EVENTS ? ev[i] : nullptr;
            HIP_TRY(launch_frame(a, stream));
        }
        if (!sync) return SR_OK;
        HIP_TRY(hipStreamSynchronize(stream));
        bool again = false;
        if ((r = settle(has_tr, again))) return r;
        if (!again) return SR_OK;
    }
    return fail(SR_ERR_CAPACITY, "frame buffers did not converge after reallocation");
}

const DevStatus &Pipeline::frame_status() {
    DevStatus *h = (DevStatus *)status_host.ptr;
    if (h->seq != seq && status_dev) {
        if (hipMemcpy(h, status_dev, sizeof(DevStatus), hipMemcpyDeviceToHost) != hipSuccess) (void)hipGetLastError();
    }
    return *h;
}

int Pipeline::settle(bool has_tr, bool &again) {
    const DevStatus &s = frame_status();
    last_records = s.total_records;
    last_pairs = s.total_pairs;
    again = false;
    if (s.clip_overflow)
        return fail(SR_ERR_UNSUPPORTED, "a frustum-clipped polygon exceeded 16 vertices (degenerate rounding)");
    bool tr_ok = !has_tr || (!s.overflow_tr && s.total_tr_pairs <= tr_cap);
    if (!s.overflow_records && !s.overflow_pairs && !s.overflow_clip && s.total_records <= rec_cap &&
        s.total_pairs <= pair_cap && tr_ok)
        return SR_OK;
    again = true;
    if (!tr_ok) tr_cap = s.total_tr_pairs + s.total_tr_pairs / 8 + 4096;
    if (s.overflow_clip || s.clip_queued > clip_cap) clip_cap = s.clip_queued + s.clip_queued / 4 + 1024;
    if (clip_cap > (1u << 27))      // clip-queue slots are 27-bit fields of the counts and record tags
        return fail(SR_ERR_CAPACITY, "more than 2^27 triangles cross the frustum");
    if (s.total_records > rec_cap || s.overflow_records) rec_cap = s.total_records + s.total_records / 8 + 1024;
    if (s.total_pairs > pair_cap || s.overflow_pairs) pair_cap = s.total_pairs + s.total_pairs / 8 + 4096;
    if (s.overflow_pairs && s.total_pairs <= pair_cap && !s.overflow_records) pair_cap *= 2;
    return SR_OK;
}

void fill_view(DevView &v, const sr_scene *scene, const sr_camera &cam, int W, int H, int T, uint32_t item_base,
               uint32_t in_tri_base, uint64_t fb_offset, uint32_t lum_offset, float exposure, uint32_t flags) {
    memset(&v, 0, sizeof(v));
    v.scene = (const DevScene *)scene->dscene.ptr;
    memcpy(v.cam_pos, cam.position, 12);
    v.one_over_w = cam.one_over_width;
    v.one_over_h = cam.one_over_height;
    memcpy(v.sky, cam.skybox_matrix_transposed, 64);
    v.W = W;
    v.H = H;
    v.W_sub = W * 16;
    v.H_sub = H * 16;
    v.Wf = (float)W;
    v.Hf = (float)H;
    v.ref_tiles_x = (W + SR_REF_TILE - 1) / SR_REF_TILE;
    v.ref_tiles_y = (H + SR_REF_TILE - 1) / SR_REF_TILE;
    v.tiles_x = v.ref_tiles_x * (SR_REF_TILE / T);
    v.tiles_y = v.ref_tiles_y * (SR_REF_TILE / T);
    v.item_base = item_base;
    v.in_tri_base = in_tri_base;
    v.fb_offset = fb_offset;
    v.lum_offset = lum_offset;
    v.exposure = exposure;
    v.flags = flags;
}

}  // namespace sr

using namespace sr;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

SR_API const char *sr_last_error(void) { return g_last_error.c_str(); }

SR_API int sr_version(int32_t *major, int32_t *minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return SR_OK;
}

SR_API int sr_device_count(int32_t *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return SR_OK;
}

// ---- load-time producers ---------------------------------------------------
SR_API int sr_prepare_environment(const uint32_t *faces, uint32_t size, int32_t threads, sr_environment *out) {
    if (!faces || !out || size == 0 || (size & (size - 1)))
        return fail(SR_ERR_INVALID_ARGUMENT, "cubemap faces must be a power-of-two size");
    memset(out, 0, sizeof(*out));
    size_t face = (size_t)size * size;
    uint32_t *cube = (uint32_t *)malloc(face * 6 * 4);
    memcpy(cube, faces, face * 6 * 4);
    std::vector<uint32_t> spec, lut;
    uint32_t mips = 0;
    compute_specular_cubemap(faces, size, 64, spec, mips, threads);
    compute_brdf_lut(128, lut, threads);
    compute_irradiance_sh4(faces, size, out->irradiance_sh);
    uint32_t *sp = (uint32_t *)malloc(spec.size() * 4);
    memcpy(sp, spec.data(), spec.size() * 4);
    uint32_t *lp = (uint32_t *)malloc(lut.size() * 4);
    memcpy(lp, lut.data(), lut.size() * 4);
    out->cubemap_size = size;
    out->cubemap = cube;
    out->specular_size = size;
    out->specular_mips = mips;
    out->specular = sp;
    out->brdf_lut_size = 128;
    out->brdf_lut = lp;
    return SR_OK;
}

SR_API int sr_prepare_environment_device(const uint32_t *faces, uint32_t size, int32_t device, sr_environment *out) {
    if (!faces || !out || size == 0 || (size & (size - 1)))
        return fail(SR_ERR_INVALID_ARGUMENT, "cubemap faces must be a power-of-two size");
    int32_t ndev = 0;
    sr_device_count(&ndev);
    if (device < 0 || device >= ndev) return fail(SR_ERR_INVALID_ARGUMENT, "no HIP device %d (have %d)", device, ndev);
    memset(out, 0, sizeof(*out));
    std::vector<uint32_t> spec, lut;
    uint32_t mips = 0;
    int r = prepare_environment_device(faces, size, device, spec, mips, lut, 128, out->irradiance_sh);
    if (r) return r;
    size_t face = (size_t)size * size;
    uint32_t *cube = (uint32_t *)malloc(face * 6 * 4);
    uint32_t *sp = (uint32_t *)malloc(spec.size() * 4);
    uint32_t *lp = (uint32_t *)malloc(lut.size() * 4);
    if (!cube || !sp || !lp) {
        free(cube);
        free(sp);
        free(lp);
        return fail(SR_ERR_OUT_OF_MEMORY, "environment");
    }
    memcpy(cube, faces, face * 6 * 4);
    memcpy(sp, spec.data(), spec.size() * 4);
    memcpy(lp, lut.data(), lut.size() * 4);
    out->cubemap_size = size;
    out->cubemap = cube;
    out->specular_size = size;
    out->specular_mips = mips;
    out->specular = sp;
    out->brdf_lut_size = 128;
    out->brdf_lut = lp;
    return SR_OK;
}

SR_API int sr_environment_free(sr_environment *env) {
    if (!env) return SR_OK;
    free((void *)env->cubemap);
    free((void *)env->specular);
    free((void *)env->brdf_lut);
    memset(env, 0, sizeof(*env));
    return SR_OK;
}

SR_API int sr_prepare_voxel_grid(const sr_scene_desc *scene, int32_t grid, int32_t threads, sr_voxel_grid *out) {
    if (!scene || !out || grid <= 0) return fail(SR_ERR_INVALID_ARGUMENT, "bad voxel grid arguments");
    memset(out, 0, sizeof(*out));
    std::vector<float> light;
    int r = build_voxel_grid(scene, grid, threads, light, out->world_min, out->world_max);
    if (r) return r;
    float *lp = (float *)malloc(light.size() * 4);
    memcpy(lp, light.data(), light.size() * 4);
    out->dims[0] = out->dims[1] = out->dims[2] = grid;
    out->light_intensity = lp;
    out->sh_scale = 0.25f;
    out->sky_visibility = 1.0f;
    out->sh4 = nullptr;
    return SR_OK;
}

SR_API int sr_prepare_voxel_grid_device(const sr_scene_desc *scene, int32_t grid, int32_t device,
                                        sr_voxel_grid *out) {
    if (!scene || !out || grid <= 0 || grid > 1024) return fail(SR_ERR_INVALID_ARGUMENT, "bad voxel grid arguments");
    int32_t ndev = 0;
    sr_device_count(&ndev);
    if (device < 0 || device >= ndev) return fail(SR_ERR_INVALID_ARGUMENT, "no HIP device %d (have %d)", device, ndev);
    memset(out, 0, sizeof(*out));
    std::vector<float> light;
    VoxelFrame fr;
    bool fallback = false;
    int r = bake_voxels_device(scene, grid, device, light, fr, fallback);
    if (r) return r;
    if (fallback) {
        r = build_voxel_grid(scene, grid, 0, light, out->world_min, out->world_max);
        if (r) return r;
    } else {
        memcpy(out->world_min, fr.wmin, 12);
        memcpy(out->world_max, fr.wmax, 12);
    }
    float *lp = (float *)malloc(light.size() * 4);
    if (!lp) return fail(SR_ERR_OUT_OF_MEMORY, "voxel grid");
    memcpy(lp, light.data(), light.size() * 4);
    out->dims[0] = out->dims[1] = out->dims[2] = grid;
    out->light_intensity = lp;
    out->sh_scale = 0.25f;
    out->sky_visibility = 1.0f;
    out->sh4 = nullptr;
    return SR_OK;
}

SR_API int sr_voxel_grid_free(sr_voxel_grid *g) {
    if (!g) return SR_OK;
    free((void *)g->light_intensity);
    free((void *)g->sh4);
    memset(g, 0, sizeof(*g));
    return SR_OK;
}

SR_API int sr_scene_desc_bounds(const sr_scene_desc *scene, float mn[3], float mx[3], float center[3], float *diag) {
    if (!scene) return fail(SR_ERR_INVALID_ARGUMENT, "null scene");
    float a[3], b[3], c[3], d;
    scene_bounds(scene, a, b, c, &d);
    if (mn) memcpy(mn, a, 12);
    if (mx) memcpy(mx, b, 12);
    if (center) memcpy(center, c, 12);
    if (diag) *diag = d;
    return SR_OK;
}

SR_API int sr_texture_mip_chain(const sr_texture_desc *tex, uint32_t *texels_out, uint32_t *total, uint32_t *levels,
                                uint32_t offsets[32], uint32_t widths[32], uint32_t heights[32]) {
    if (!tex || !tex->texels || tex->width == 0 || tex->height == 0)
        return fail(SR_ERR_INVALID_ARGUMENT, "empty texture");
    std::vector<uint32_t> data, off, w, h;
    build_mip_chain(tex->width, tex->height, tex->type, tex->texels, data, off, w, h);
    if (texels_out) memcpy(texels_out, data.data(), data.size() * 4);
    if (total) *total = (uint32_t)data.size();
    if (levels) *levels = (uint32_t)off.size();
    for (size_t i = 0; i < off.size() && i < 32; i++) {
        if (offsets) offsets[i] = off[i];
        if (widths) widths[i] = w[i];
        if (heights) heights[i] = h[i];
    }
    return SR_OK;
}

// ---- camera ------------------------------------------------------------------
SR_API int sr_camera_new(const float position[3], const float look_at[3], float fov, float width, float height,
                         float far_plane, sr_camera *out) {
    if (!position || !look_at || !out) return fail(SR_ERR_INVALID_ARGUMENT, "null camera argument");
    camera_new(position, look_at, fov, width, height, far_plane, out);
    return SR_OK;
}

SR_API int sr_camera_from_pose(const float position[3], float yaw, float pitch, float roll, float fov, float width,
                               float height, float far_plane, sr_camera *out) {
    if (!position || !out) return fail(SR_ERR_INVALID_ARGUMENT, "null camera argument");
    float look[3] = {position[0], position[1], position[2] - 1.0f};
    camera_new(position, look, fov, width, height, far_plane, out);
    out->yaw = yaw;
    out->pitch = pitch;
    out->roll = roll;
    return SR_OK;
}

SR_API int sr_camera_update_matrices(sr_camera *camera) {
    if (!camera) return fail(SR_ERR_INVALID_ARGUMENT, "null camera");
    camera_update(camera);
    return SR_OK;
}

}  // extern "C"
