// kernels.h -- host-callable launchers for the gfx950 kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include "common.h"

namespace sr {

struct FrameArgs {
    // inputs
    const DevView *views;         // device
    int32_t num_views;
    const DevDraw *draws;         // device
    int32_t num_draws;
    uint32_t total_xv;            // transformed vertices (sum over draws)
    uint32_t total_in;            // input triangles (sum over draws)
    uint32_t total_items;         // (view, tile) items == bins
    int32_t tile_size;            // 8, 16 or 32
    int32_t any_tangents;
    // transient buffers
    float4 *xv;                   // total_xv clip positions, 2 * total_xv attribute float4, then
                                  // total_xv snaps (px, py, 1/w, z/w)
    float4 *xt;                   // tangent float4 per transformed vertex (or null)
    uint32_t *tri_counts;         // total_in
    uint32_t *tri_chunks;         // (total_in + 255) / 256 fan-count sums, then first ids
    uint32_t *tri_base;           // total_in (+1)
    uint32_t *bin_start;          // total_items: each bin's first list entry (zeroed at the frame's start)
    uint32_t *bin_end;            // total_items: one past its last (0 for a bin without entries)
    uint32_t *scan_partials;      // scratch for scans
    uint32_t *list;               // pair_cap record ids, each bin's in submission order
    uint32_t *chunk_pairs;              // (total_in + 255) / 256 + 1: pairs of each setup chunk, then first pair
    uint32_t *pair_keys, *pair_vals;    // pair_cap each: (bin, record) pairs in record order, then every
    uint32_t *pair_keys2, *pair_vals2;  // pair_cap each: the other pass output (vals: two passes or more)
    uint32_t *sort_hist;                // 256 * bin_sort_tiles(pair_cap): per-tile digit counts, then offsets
    BinSortState *sort;                 // right after the status
    StageClock *clock;                  // stage stamps (end slots zeroed by the transform; null: none)
    uint4 *rec_geo;               // rec_cap
    float4 *rec_raster;           // rec_cap * 4
    uint32_t rec_cap, pair_cap;
    uint32_t *clip_queue;         // clip_cap input triangles that cross a frustum plane
    float *clip_buf;              // clip_cap * 16 vertices * 16 channels
    uint32_t clip_cap;
    DevStatus *status;
    DevStatus *status_out;        // pinned host copy of the status, stored by k_shade
    // outputs (per view at DevView.fb_offset)
    uint32_t *out_ids;
    float *out_depth;
    float *out_color;             // three planes of color_plane floats
    uint64_t color_plane;
    uint32_t *out_rgba;           // fused blit target (views with flags & 1)
    float *out_lum;
    uint32_t *center_ids;         // 4 per 64px tile (sum over views)
    // translucent triangles (views without the fused blit only)
    int32_t has_translucent;
    uint32_t tr_slots;            // 64px tiles (sum over views)
    uint32_t *tr_count;           // tr_slots
    uint32_t *tr_start;           // tr_slots
    uint32_t *tr_end;             // tr_slots (cursor, then end)
    uint32_t *tr_list;            // tr_cap record ids
    unsigned long long *tr_scratch;  // tr_cap: sort room for lists longer than LDS holds
    uint32_t tr_cap;
    float4 *center_color;         // 4 (colour, depth) per 64px tile
    int32_t persistent_blocks;    // 0 = one block per item
    int32_t hiz;                  // 16-px raster with depth rejection between task rounds
    hipEvent_t events[5];         // optional stage events (the launches that carry them cost about 4 us each):
                                  // start, setup done, bins done, shade done, raster done
};

// Stable binning sizes: bits of the bin index, radix passes (8 bits each;
// 0 if more than 3 would be needed, i.e. 2^24 bins or more), and the
// 4096-pair tiles of a pass.
int bin_sort_bits(uint32_t bins);
int bin_sort_passes(uint32_t bins);
uint32_t bin_sort_tiles(uint32_t pair_cap);

// Launch the whole per-frame sequence on a stream.  Returns hipSuccess or the
// first launch error.
hipError_t launch_frame(const FrameArgs &a, hipStream_t stream);

// Exclusive scan of n u32 (out, and out2 if not null: the same, or the
// inclusive scan when flags has bit 0); the total goes to *total (device)
// and to *total2 if not null.  partials holds (n + 4095) / 4096 + 1 words.
// start_stamp (if not null) receives the wall clock at the scan's start.
hipError_t launch_scan(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *out2, uint32_t *partials,
                       uint32_t *total, hipStream_t stream, uint32_t flags = 0, uint32_t *total2 = nullptr,
                       uint64_t *start_stamp = nullptr);

// Exposure + tonemap + quantize + pack of one view's linear colour planes;
// frame_bytes stores each pixel as the bytes r, g, b, a.
hipError_t launch_blit(const float *color, uint64_t plane, int32_t W, int32_t H, float exposure, uint32_t *out,
                       hipStream_t st, const float *exposure_dev = nullptr, bool frame_bytes = false);
// Most 64-px reference tiles the device exposure step sorts (8K UHD has 8160).
#define SR_EXPOSURE_MAX 8192
// state = {ev, target, exposure} on the device (k_exposure).
hipError_t launch_exposure(const float *lum, uint32_t n, float dt, bool use_host, float ev, float target, float *state,
                           hipStream_t st);

}  // namespace sr
