// noalloc.hip -- one frame without allocating device memory
// (sr_render_frame_noalloc).
//
// The per-frame pipeline keeps transient buffers that the renderer grows on
// demand: the view and draw tables with the frame's status, transformed
// vertices, per-triangle counts and post-clip ids, setup records, bins and
// their lists with the binning's pair arrays and tile counts, the clip
// queue and clipped polygons, tile centres and luminance.  This
// entry point carves all of them out of one file-scope device array,
// `scratch`, so it calls no hipMalloc or hipFree; the caller passes the
// frame's outputs.  The array's host-side name is only a stand-in for the
// symbol: its device address comes from hipGetSymbolAddress, once per
// device, and only that address is handed to copies and kernels.  Every
// call on a device shares that device's array and leaves the last call's
// contents in it; the frame clears what must start at zero itself (the
// status by its upload, the tiles' list bounds in the first kernel),
// and calls that overlap on one device race on it.

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "runtime.h"
#include "swraster.h"

using namespace sr;

// SR_NOALLOC_SCRATCH_BYTES (160 MiB) of device memory for one frame's
// transient buffers.
static __device__ __align__(256) uint8_t scratch[SR_NOALLOC_SCRATCH_BYTES];

namespace {

const int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on the current device (looked up once).
int scratch_ptr(int dev, uint8_t **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(SR_ERR_INVALID_ARGUMENT, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SR_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = (uint8_t *)g_scratch[dev];
    return SR_OK;
}

// Consecutive 256-byte aligned pieces of the scratch array.
struct Carve {
    uint64_t used = 0;
    uint64_t take(uint64_t bytes) {
        uint64_t o = used;
        used += (bytes + 255) & ~(uint64_t)255;
        return o;
    }
};

}  // namespace

extern "C" {

SR_API int sr_render_frame_noalloc(const sr_scene *scene, const sr_camera *camera, int32_t width, int32_t height,
                                   float exposure, uint32_t *tri_id, float *depth, uint32_t *rgba, void *stream) {
    if (!scene || !camera || !tri_id || !depth || !rgba) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (width <= 0 || height <= 0 || width > 32000 || height > 32000)
        return fail(SR_ERR_INVALID_ARGUMENT, "frame size must be in 1..32000");
    hipStream_t st = (hipStream_t)stream;
    if (hipSetDevice(scene->device) != hipSuccess) return fail(SR_ERR_HIP, "device %d", scene->device);
    uint8_t *base = nullptr;
    if (int rc = scratch_ptr(scene->device, &base)) return rc;

    // normal-mapped materials read transformed tangents, and a translucent
    // primitive needs colour planes and per-tile lists: neither is part of
    // this layout
    if (scene->any_tangents)
        return fail(SR_ERR_UNSUPPORTED, "the no-allocation entry point renders scenes without normal maps");
    std::vector<DevDraw> draws;
    uint32_t total_in = 0, total_xv = 0;
    if (int rc = build_draws(scene, *camera, 0, total_in, total_xv, draws, false)) return rc;
    const int T = 16;
    DevView view;
    fill_view(view, scene, *camera, width, height, T, 0, 0, 0, 0, exposure > 0.0f ? exposure : 2.0f, 1u);
    uint32_t items = (uint32_t)(view.tiles_x * view.tiles_y);
    uint32_t ref_tiles = (uint32_t)(view.ref_tiles_x * view.ref_tiles_y);

    // layout: fixed parts sized by the frame, then clip room, records and
    // list entries from what is left
    size_t vbytes = 256, dbytes = (std::max<size_t>(draws.size(), 1) * sizeof(DevDraw) + 255) & ~(size_t)255;
    size_t upload = vbytes + dbytes + sizeof(DevStatus);    // the binning state follows, not uploaded
    const int sort_passes = bin_sort_passes(items);
    if (sort_passes < 1) return fail(SR_ERR_CAPACITY, "%u screen tiles; at most 2^24", items);
    Carve c;
    uint64_t o_views = c.take(upload + sizeof(BinSortState));
    uint64_t o_xv = c.take((uint64_t)std::max<uint32_t>(total_xv, 1) * 64);
    uint64_t o_counts = c.take(((uint64_t)total_in + 1) * 4);
    uint64_t o_chunks = c.take(((uint64_t)total_in / 256 + 2) * 4);
    uint64_t o_chunk_pairs = c.take(((uint64_t)total_in / 256 + 2) * 4);
    uint64_t o_base = c.take(((uint64_t)total_in + 1) * 4);
    uint64_t o_bins = c.take((uint64_t)std::max<uint32_t>(items, 1) * 8);
    uint64_t maxn = std::max<uint64_t>((uint64_t)total_in + 1, items);
    uint64_t o_part = c.take(((maxn + 4095) / 4096 + 1) * 4);
    uint64_t o_center = c.take((uint64_t)ref_tiles * 16);
    uint64_t o_lum = c.take((uint64_t)ref_tiles * 4);
    uint32_t clip_cap = 1024 + total_in / 256;
    uint64_t o_clipq = c.take((uint64_t)clip_cap * 4);
    uint64_t o_clipb = c.take((uint64_t)clip_cap * 16 * 16 * 4);
    const uint64_t cap = SR_NOALLOC_SCRATCH_BYTES;
    // records (80 B each: geometry and raster) and list entries (4 B each,
    // 8 B of expanded pairs, 4 B of the other pass's keys and 4 B more of its
    // values when the radix sort takes more than one pass, 1/4 B of tile
    // digit counts; at least one entry per record)
    const uint64_t rec_bytes = 80, pair_bytes = 4 + 8 + 4 + (sort_passes > 1 ? 4 : 0) + 1;
    uint64_t need_min = c.used + 256 * 8 + ((uint64_t)total_in + 1024) * (rec_bytes + pair_bytes);
    if (need_min > cap)
        return fail(SR_ERR_CAPACITY,
                    "a %d x %d frame of %u triangles needs at least %llu bytes of scratch; the no-allocation entry "
                    "point holds %llu",
                    width, height, total_in, (unsigned long long)need_min, (unsigned long long)cap);
    uint64_t rest = cap - c.used - 256 * 8 - 1024;
    uint64_t rec_want = (uint64_t)total_in + total_in / 8 + 1024;
    uint32_t rec_cap =
        (uint32_t)std::min<uint64_t>(std::min<uint64_t>(rec_want, rest * 4 / 5 / rec_bytes), 0xFFFFFFF0ull);
    uint32_t pair_cap = (uint32_t)std::min<uint64_t>((rest - (uint64_t)rec_cap * rec_bytes) / pair_bytes, 0xFFFFFFF0ull);
    uint64_t o_geo = c.take((uint64_t)rec_cap * 16);
    uint64_t o_ras = c.take((uint64_t)rec_cap * 64);
    uint64_t o_list = c.take((uint64_t)pair_cap * 4);
    uint64_t o_pairs[4] = {0, 0, 0, 0};
    for (int k = 0; k < (sort_passes > 1 ? 4 : 3); k++) o_pairs[k] = c.take((uint64_t)pair_cap * 4);
    uint64_t o_hist = c.take((uint64_t)bin_sort_tiles(pair_cap) * 256 * 4);
    if (c.used > cap) return fail(SR_ERR_CAPACITY, "scratch layout exceeds %llu bytes", (unsigned long long)cap);

    // view and draw tables and the zeroed status: one upload
    std::vector<uint8_t> stage(upload, 0);
    memcpy(stage.data(), &view, sizeof(DevView));
    if (!draws.empty()) memcpy(stage.data() + vbytes, draws.data(), draws.size() * sizeof(DevDraw));
    if (hipMemcpyAsync(base + o_views, stage.data(), upload, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(SR_ERR_HIP, "table upload: %s", hipGetErrorString(hipGetLastError()));

    FrameArgs a;
    memset(&a, 0, sizeof(a));
    a.views = (const DevView *)(base + o_views);
    a.num_views = 1;
    a.draws = (const DevDraw *)(base + o_views + vbytes);
    a.num_draws = (int)draws.size();
    a.total_xv = total_xv;
    a.total_in = total_in;
    a.total_items = items;
    a.tile_size = T;
    a.any_tangents = 0;
    a.xv = (float4 *)(base + o_xv);
    a.xt = nullptr;
    a.tri_counts = (uint32_t *)(base + o_counts);
    a.tri_chunks = (uint32_t *)(base + o_chunks);
    a.tri_base = (uint32_t *)(base + o_base);
    a.bin_start = (uint32_t *)(base + o_bins);
    a.bin_end = a.bin_start + items;
    a.scan_partials = (uint32_t *)(base + o_part);
    a.list = (uint32_t *)(base + o_list);
    a.pair_keys = (uint32_t *)(base + o_pairs[0]);
    a.pair_vals = (uint32_t *)(base + o_pairs[1]);
    a.pair_keys2 = (uint32_t *)(base + o_pairs[2]);
    if (sort_passes > 1) a.pair_vals2 = (uint32_t *)(base + o_pairs[3]);
    a.sort_hist = (uint32_t *)(base + o_hist);
    a.chunk_pairs = (uint32_t *)(base + o_chunk_pairs);
    a.rec_geo = (uint4 *)(base + o_geo);
    a.rec_raster = (float4 *)(base + o_ras);
    a.rec_cap = rec_cap;
    a.pair_cap = pair_cap;
    a.clip_queue = (uint32_t *)(base + o_clipq);
    a.clip_buf = (float *)(base + o_clipb);
    a.clip_cap = clip_cap;
    a.status = (DevStatus *)(base + o_views + vbytes + dbytes);
    a.sort = (BinSortState *)(base + o_views + vbytes + dbytes + sizeof(DevStatus));
    a.status_out = nullptr;
    a.out_ids = tri_id;
    a.out_depth = depth;
    a.out_color = nullptr;
    a.color_plane = 0;
    a.out_rgba = rgba;
    a.out_lum = (float *)(base + o_lum);
    a.center_ids = (uint32_t *)(base + o_center);
    a.persistent_blocks = 0;
    a.hiz = 0;
    if (launch_frame(a, st) != hipSuccess) return fail(SR_ERR_HIP, "launch: %s", hipGetErrorString(hipGetLastError()));
    DevStatus s;
    if (hipMemcpyAsync(&s, a.status, sizeof(s), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(SR_ERR_HIP, "frame: %s", hipGetErrorString(hipGetLastError()));
    if (s.clip_overflow)
        return fail(SR_ERR_UNSUPPORTED, "a frustum-clipped polygon exceeded 16 vertices (degenerate rounding)");
    if (s.overflow_records || s.overflow_pairs || s.overflow_clip || s.total_records > rec_cap ||
        s.total_pairs > pair_cap || s.clip_queued > clip_cap)
        return fail(SR_ERR_CAPACITY,
                    "the frame needs %u records, %u list entries and %u clipped triangles; the scratch holds %u, %u "
                    "and %u",
                    s.total_records, s.total_pairs, s.clip_queued, rec_cap, pair_cap, clip_cap);
    return SR_OK;
}

}  // extern "C"
