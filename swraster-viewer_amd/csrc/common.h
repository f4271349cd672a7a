// common.h -- device data layout shared by the host runtime and the kernels.
//
// Everything a frame reads lives in HBM in these forms:
//   scene   : vertex attributes as SoA float arrays, u32 indices, materials,
//             one RGBA8 texel pool holding every mip chain, the skybox and
//             specular cubemaps, the BRDF LUT and the voxel grid.
//   frame   : per-draw transformed vertices (48 B AoS, one gather per vertex),
//             per-triangle setup records (geo 16 B, raster 64 B, shade 128 B,
//             tangent 48 B when normal mapped), per-tile bins (u32 ids),
//             visibility buffer (u32 id + f32 depth planes), linear colour
//             (three f32 planes) and RGBA8 pixels.
#pragma once

#include <cstdint>

// Pointers into device memory held in structs.  In the device pass they are
// global-address-space pointers, so loads through them are global (or, when
// uniform, scalar) loads rather than generic FLAT loads, which also occupy
// the LDS/scalar wait counter; everywhere else they are plain pointers of the
// same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define SR_GLOBAL __attribute__((address_space(1)))
#else
#define SR_GLOBAL
#endif

#define SR_MAX_MIPS 16
#define SR_REF_TILE 64          // reference raster tile (src/renderer.rs:14)
#define SR_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

// material flags
#define SR_MAT_ALPHA_TESTED 1u
#define SR_MAT_TRANSLUCENT 2u
#define SR_MAT_NORMAL_MAP 4u

// draw flags
#define SR_DRAW_CLIP 1u
#define SR_DRAW_OPAQUE 2u
#define SR_DRAW_TANGENTS 4u

// record geo flags (rec_geo.z)
#define SR_GEO_EMPTY 1u          // culled or empty bounding box: no bins
#define SR_GEO_ALPHA 2u          // alpha-tested material
#define SR_GEO_NORMAL_MAP 4u
#define SR_GEO_TRANSLUCENT 8u    // translucent material: 64-px reference-tile lists, forward pass

struct DevTexRef {
    int32_t tex;                  // -1 = none
    int32_t wrap_s, wrap_t;
};

struct DevTexture {
    uint64_t base;                // first texel in the scene pool
    uint32_t width, height;
    float whv[4];                 // width, width, height, height
    uint32_t max_mip;
    uint32_t num_mips;
    uint32_t mip_off[SR_MAX_MIPS];    // each level in 4x4-texel blocks, row-major over the padded level
    uint32_t mip_w[SR_MAX_MIPS];
    uint32_t mip_h[SR_MAX_MIPS];
};

struct DevMaterial {
    float base[4];
    float metallic, roughness;
    float emissive[3];
    float occlusion_strength;
    float alpha_cutoff;
    float transmission;
    uint32_t flags;
    DevTexRef base_tex, mr_tex, normal_tex, emissive_tex, occlusion_tex, transmission_tex;
};

struct DevScene {
    const SR_GLOBAL float *px, *py, *pz;    // positions
    const SR_GLOBAL float *nx, *ny, *nz;    // normals
    const SR_GLOBAL float *tu, *tv;         // texcoords
    const SR_GLOBAL float *tangent;         // float4 per vertex (may be null)
    const SR_GLOBAL uint32_t *indices;
    const SR_GLOBAL DevMaterial *materials;
    const SR_GLOBAL DevTexture *textures;
    const SR_GLOBAL uint32_t *texels;
    const SR_GLOBAL uint32_t *cubemap;
    uint32_t cube_size;
    uint32_t spec_size, spec_mips;
    const SR_GLOBAL uint32_t *specular;
    const SR_GLOBAL uint32_t *lut;
    uint32_t lut_size;
    const SR_GLOBAL float *vox_light;       // constant-SH form: per cell, w0 at its 8 trilinear corners
    const SR_GLOBAL float *vox_sh4;         // full form: 16 floats per voxel (or null)
    int32_t vox_dims[3];
    float vox_min[3];
    float vox_size[3];            // (max - min) / dims, as src/voxelgrid.rs:154-161
    float sh_const[4][4];         // constant-SH coefficients (w of [1] = sky)
    float light_dir[3];
    float light_color[3];
    float spec_max_mip_f;
};

// One view = one frame of one camera at one resolution.
struct DevView {
    const SR_GLOBAL DevScene *scene;
    float cam_pos[3];
    float one_over_w, one_over_h;
    float sky[16];                // skybox_matrix_transposed
    int32_t W, H;
    int32_t W_sub, H_sub;         // W * 16, H * 16
    float Wf, Hf;
    int32_t tiles_x, tiles_y;     // GPU tiles
    int32_t ref_tiles_x, ref_tiles_y;
    uint32_t item_base;           // first global (view, tile) item == first bin
    uint32_t in_tri_base;         // first global input triangle
    uint64_t fb_offset;           // first pixel of this view in the frame buffers
    uint32_t lum_offset;          // first 64px-tile luminance slot
    float exposure;
    uint32_t flags;               // 1 = fused blit to RGBA8
    uint32_t pad;
};

struct DevDraw {
    float mvp[16];
    float model[16];
    uint32_t vtx_base;            // first scene vertex of the primitive
    uint32_t num_vtx;
    uint32_t idx_base;            // first scene index of the primitive
    uint32_t num_tri;
    uint32_t in_tri_base;         // first global input triangle
    uint32_t xv_base;             // first transformed vertex
    int32_t view;
    int32_t material;
    uint32_t flags;
    uint32_t pad[3];
};

// Frame status written by the device, read back by the host after a frame.
struct DevStatus {
    uint32_t total_records;       // post-clip triangles
    uint32_t total_pairs;         // binned (triangle, tile) pairs
    uint32_t overflow_records;
    uint32_t overflow_pairs;
    uint32_t work_counter;        // persistent-kernel queue head
    uint32_t clip_queued;         // triangles queued for clipping
    uint32_t overflow_clip;
    uint32_t clip_overflow;       // a clipped polygon outgrew its room (frame fails)
    uint32_t total_tr_pairs;      // (translucent triangle, 64-px tile) pairs
    uint32_t overflow_tr;
    uint32_t seq;                 // the frame's number, set by the host in the upload
};

// Working state of the stable binning (radix passes over the bin index),
// right after the status (every field is written in the frame before it is
// read): each pass's element count per digit (its row scan writes them, its
// scatter scans them into the digits' first positions) and the expansion's
// total.
struct BinSortState {
    uint32_t digit_total[3][256];
    uint32_t chunk_total;         // pairs of all records (the expansion's size)
    uint32_t pad[3];
};

// Stage boundaries of one frame on the GPU's constant-rate wall clock
// (wall_clock64, which ticks hipDeviceAttributeWallClockRate times per ms).
// A kernel's first workgroup stamps its start: the transform's, the bin
// scan's (the end of setup) and the raster's.  For the frame's end every
// workgroup of the shading (and translucent) pass raises one of
// CLOCK_SPREAD end slots to the time it finished (an atomic max that
// returns nothing, spread so that few workgroups share an address), and the
// end is the largest slot; the transform zeroes the end slots.
#define CLOCK_SPREAD 256
enum { CLOCK_FRONT = 0, CLOCK_SETUP_END, CLOCK_RASTER, CLOCK_STARTS };
enum { CLOCK_SHADE_END = 0, CLOCK_TR_END, CLOCK_ENDS };
struct StageClock {
    uint64_t start[4];                          // CLOCK_STARTS used
    uint64_t end[CLOCK_ENDS][CLOCK_SPREAD];     // shading, translucent shading
};
