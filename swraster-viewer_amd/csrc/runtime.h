// runtime.h -- host-side objects behind the C ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "swraster.h"

namespace sr {

extern thread_local std::string g_last_error;
int fail(int code, const char *fmt, ...);

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
};

struct PinBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need, unsigned flags = hipHostMallocDefault);
    void release();
};

struct Sphere {
    float c[3];
    float r;
};

struct HostPrim {
    uint32_t vtx_base, num_vtx, idx_base, num_tri;
    int32_t material;
    bool tangents;
    Sphere bs;
};
struct HostMesh {
    std::vector<int> prims, opaque, translucent;
};
struct HostNode {
    float transform[16];
    int mesh;
    Sphere bs_world;
};

struct FrameArgs;

// Per-frame pipeline state (transient buffers, stream, events).
struct Pipeline {
    int device = 0;
    int tile_size = 16;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int timing = SR_STAGE_TIMING_CLOCK;   // how eager frames time their stages
    double clock_ticks_per_ms = 0.0;      // the device's wall clock (StageClock)
    StageClock *clock_dev = nullptr;      // the last eager frame's stamps (clock_buf)
    DevBuf clock_buf;
    DevBuf views, xv, xt, tri_counts, tri_chunks, tri_base, bin_start, bin_end, partials, list;
    DevBuf rec_geo, rec_raster, center, clip_queue, clip_buf;
    DevBuf chunk_pairs, pair_keys, pair_vals, pair_keys2, pair_vals2, sort_hist;
    DevBuf tr_count, tr_start, tr_end, tr_list, tr_scratch, center_color;
    PinBuf stage, status_host;
    uint32_t rec_cap = 0, pair_cap = 0, clip_cap = 0, tr_cap = 0;
    uint32_t last_records = 0, last_pairs = 0;
    bool last_hiz = false;            // the last frame's raster used depth rejection
    struct GraphKey {
        uint64_t n[16];
        uint64_t p[40];
    } graph_key = {};
    hipGraphExec_t graph_exec = nullptr;
    int init(int device, int tile);
    void destroy();
    void drop_graph();
    int launch_graph(const FrameArgs &a, size_t vbytes, size_t upload);
    int run(const std::vector<DevView> &views, const std::vector<DevDraw> &draws, uint32_t total_xv,
            uint32_t total_in, uint32_t total_items, bool any_tangents, uint32_t *out_ids, float *out_depth,
            float *out_color, uint64_t color_plane, uint32_t *out_rgba, float *out_lum, int persistent_blocks,
            bool sync, bool use_graph = false);
    // After the stream has drained a frame: records its counts, grows what
    // overflowed and sets again when the frame must be run once more.
    int settle(bool has_translucent, bool &again);
    // The last frame's status after its stream has drained: the copy the
    // shading kernel stored in status_host, or, if that copy's frame number
    // does not match, a copy read from the device.
    const DevStatus &frame_status();
    uint32_t seq = 0;                 // number of the last frame launched
    DevStatus *status_dev = nullptr;  // its status on the device
};

}  // namespace sr

struct sr_scene {
    int device = 0;
    std::vector<sr::DevBuf> buffers;
    sr::DevBuf dscene;            // DevScene struct in device memory
    std::vector<sr::HostNode> nodes;
    std::vector<sr::HostMesh> meshes;
    std::vector<sr::HostPrim> prims;
    bool any_tangents = false;
    bool any_translucent = false;     // a node draws a translucent primitive
    std::atomic<int> pending_frames{0};   // renderer frames queued and not yet checked
    uint64_t triangles = 0, vertices = 0, device_bytes = 0;
};

namespace sr {
// allow_translucent: translucent primitives become draws without SR_DRAW_OPAQUE
// (else they fail with SR_ERR_UNSUPPORTED).
int build_draws(const sr_scene *scene, const sr_camera &cam, int view, uint32_t &in_tri, uint32_t &xv,
                std::vector<DevDraw> &out, bool allow_translucent);
void fill_view(DevView &v, const sr_scene *scene, const sr_camera &cam, int W, int H, int T, uint32_t item_base,
               uint32_t in_tri_base, uint64_t fb_offset, uint32_t lum_offset, float exposure, uint32_t flags);
}  // namespace sr
