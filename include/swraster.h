/*
 * swraster.h -- C ABI of the MI355X (gfx950) rasterization pipeline.
 *
 * The pipeline replaces the per-frame path of the reference software
 * rasterizer: clip-space projection with frustum clipping, tile binning,
 * per-tile rasterization into a visibility buffer, visibility-buffer PBR
 * shading, auto exposure and the tonemapping blit.
 *
 * Reference interfaces replaced (paths are relative to the reference tree):
 *   Renderer::new                  src/renderer.rs:165   -> sr_renderer_create
 *   Renderer::render_scene         src/renderer.rs:201   -> sr_renderer_render_scene
 *   Renderer::update_auto_exposure src/renderer.rs:258   -> sr_renderer_update_auto_exposure
 *   Renderer::blit_to_buffer       src/renderer.rs:293   -> sr_renderer_blit_to_buffer
 *   TileRasterizer::render_tile    src/tilerasterizer.rs:72 (runs inside
 *                                  sr_renderer_render_scene, one GPU
 *                                  workgroup per screen tile)
 *   RenderCamera::new              src/rendercamera.rs:28 -> sr_camera_new
 *   RenderCamera::update_matrices  src/rendercamera.rs:78 -> sr_camera_update_matrices
 *   Scene (constructed by Scene::from_gltf, src/scene.rs:145, and
 *   load_scene, src/main.rs:100-283)   -> sr_scene_desc + sr_scene_create
 *
 * Conventions
 *   - every function returns int: SR_OK (0) or a negative SR_ERR_* code and
 *     never throws across the ABI; sr_last_error() gives thread-local text.
 *   - a renderer owns its frame buffers and one HIP stream; it is not
 *     re-entrant (the reference takes &mut self).  Distinct renderers may
 *     run concurrently.  A scene is immutable after creation and may be
 *     shared by any number of renderers, batches and servers.
 *   - pixels are packed r<<24 | g<<16 | b<<8 | a (src/util.rs:98-100).
 *   - matrices are column-major float[16] (glam Mat4 layout).
 */
#ifndef SWRASTER_H
#define SWRASTER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SR_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------ */
#define SR_OK 0
#define SR_ERR_INVALID_ARGUMENT (-1)
#define SR_ERR_HIP (-2)
#define SR_ERR_OUT_OF_MEMORY (-3)
#define SR_ERR_UNSUPPORTED (-4)
#define SR_ERR_CAPACITY (-5)
#define SR_ERR_BUSY (-6)
#define SR_ERR_NOT_FOUND (-7)

/* ---- textures and materials (src/texture.rs:19-42, src/scene.rs:104-121) */
#define SR_TEX_SRGB 0
#define SR_TEX_NORMAL 1
#define SR_TEX_METALLIC_ROUGHNESS 2
#define SR_TEX_LINEAR 3

#define SR_WRAP_REPEAT 0
#define SR_WRAP_MIRRORED_REPEAT 1
#define SR_WRAP_CLAMP_TO_EDGE 2

#define SR_ALPHA_OPAQUE 0
#define SR_ALPHA_MASK 1
#define SR_ALPHA_BLEND 2

/* Base level of an RGBA8 texture; the mip chain is generated on load exactly
 * as Texture::generate_mipmaps (src/texture.rs:45-128). */
typedef struct sr_texture_desc {
    uint32_t width, height;
    int32_t type;             /* SR_TEX_* */
    const uint32_t *texels;   /* width*height packed RGBA8 */
} sr_texture_desc;

typedef struct sr_texref {
    int32_t texture;          /* index into sr_scene_desc.textures, -1 = none */
    int32_t wrap_s, wrap_t;   /* SR_WRAP_* */
} sr_texref;

typedef struct sr_material_desc {
    float base_color_factor[4];
    float metallic_factor;
    float roughness_factor;
    float emissive_factor[3];
    float occlusion_strength;
    float alpha_cutoff;
    float transmission;       /* KHR_materials_transmission factor */
    int32_t alpha_mode;       /* SR_ALPHA_* */
    int32_t is_translucent;   /* material has KHR_materials_transmission */
    sr_texref base_color;
    sr_texref metallic_roughness;
    sr_texref normal;
    sr_texref emissive;
    sr_texref occlusion;
    sr_texref transmission_texture;
} sr_material_desc;

/* One glTF primitive (src/scene.rs:94-102).  Positions, normals and texture
 * coordinates are tightly packed float arrays (xyz, xyz, uv).  texcoords may
 * be NULL (zeros are used, src/scene.rs:455-464). */
typedef struct sr_primitive_desc {
    const float *positions;   /* num_vertices * 3 */
    const float *normals;     /* num_vertices * 3 */
    const float *texcoords;   /* num_vertices * 2 or NULL */
    const float *tangents;    /* num_vertices * 4 or NULL (computed as
                                 compute_tangents, src/scene.rs:548-635) */
    const uint32_t *indices;  /* num_indices (triangle list) */
    uint32_t num_vertices;
    uint32_t num_indices;
    int32_t material;         /* index into materials; absent -> 0 (src/scene.rs:493) */
} sr_primitive_desc;

typedef struct sr_mesh_desc {
    int32_t first_primitive;  /* primitives[first .. first+count) */
    int32_t num_primitives;
} sr_mesh_desc;

/* A flattened node: world transform and optional mesh (src/scene.rs:79-85). */
typedef struct sr_node_desc {
    float transform[16];
    int32_t mesh;             /* -1 = none */
} sr_node_desc;

/* Image-based lighting inputs derived from the skybox cubemap at load time
 * (src/scene.rs:152-228): the skybox faces (mip 0 only is ever sampled), the
 * GGX-prefiltered specular cubemap (every mip at full face resolution,
 * src/texture.rs:330-420), the 4-coefficient irradiance SH
 * (src/texture.rs:289-328) and the split-sum BRDF LUT (src/texture.rs:199-235).
 * sr_prepare_environment() fills one from the skybox faces. */
typedef struct sr_environment {
    uint32_t cubemap_size;              /* face width == height */
    const uint32_t *cubemap;            /* 6 * size^2, faces +X,-X,+Y,-Y,+Z,-Z */
    uint32_t specular_size;             /* face size of every specular mip */
    uint32_t specular_mips;             /* 1 + ilog2(size) */
    const uint32_t *specular;           /* specular_mips * 6 * size^2 */
    uint32_t brdf_lut_size;
    const uint32_t *brdf_lut;           /* size^2 (mip 0 only is sampled) */
    float irradiance_sh[4][3];
} sr_environment;

/* Voxel lighting grid (src/voxelgrid.rs:6-18).  Without a GI bake the
 * reference fills every voxel with SH = scale * irradiance_sh, w0 = sun
 * visibility and w1 = sky visibility (src/gi.rs:123-149, src/main.rs:280);
 * sh4 == NULL selects exactly that form and only the per-voxel sun
 * visibility is stored.  sr_prepare_voxel_grid() builds it. */
typedef struct sr_voxel_grid {
    int32_t dims[3];
    float world_min[3];
    float world_max[3];
    const float *light_intensity;       /* dims product: blurred sun visibility */
    float sh_scale;                     /* 0.25 (GI_FALLBACK_SCENE_SH_SCALE) */
    float sky_visibility;               /* 1.0 */
    const float *sh4;                   /* optional dims product * 16 floats, NULL */
} sr_voxel_grid;

typedef struct sr_light {
    float direction[3];       /* normalized (-0.2, 1, 0.5) by default */
    float color[3];           /* (1, 1, 0.95) * 5 by default */
} sr_light;

typedef struct sr_scene_desc {
    const sr_primitive_desc *primitives; int32_t num_primitives;
    const sr_mesh_desc *meshes;          int32_t num_meshes;
    const sr_node_desc *nodes;           int32_t num_nodes;
    const sr_material_desc *materials;   int32_t num_materials;
    const sr_texture_desc *textures;     int32_t num_textures;
    const sr_environment *environment;
    const sr_voxel_grid *voxels;
    sr_light light;
} sr_scene_desc;

/* ---- camera (src/rendercamera.rs:5-25) --------------------------------- */
typedef struct sr_camera {
    float position[3];
    float fov, width, height, one_over_width, one_over_height;
    float near_plane, far_plane;
    float yaw, pitch, roll;
    float view_matrix[16];
    float projection_matrix[16];
    float view_project_matrix[16];
    float inverse_view_project_matrix[16];
    float skybox_matrix_transposed[16];
    float view_clip_planes[6][4];
    float view_normal[3];
} sr_camera;

/* ---- renderer options and statistics ------------------------------------ */
typedef struct sr_renderer_opts {
    int32_t tile_size;        /* GPU screen tile: 8, 16 or 32 (0 = tuned default) */
    int32_t raster_mode;      /* 0 = tuned default; see DESIGN */
    int32_t device;           /* HIP device ordinal */
    int32_t reserved[5];
} sr_renderer_opts;

typedef struct sr_frame_stats {
    uint64_t input_triangles;      /* submitted, pre-cull */
    uint64_t setup_triangles;      /* after frustum clipping (fan output) */
    uint64_t binned_pairs;         /* (triangle, tile) references */
    uint32_t tiles;
    uint32_t draws;
    float ms_setup;                /* transform+clip+setup */
    float ms_bin;                  /* scans + scatter */
    float ms_raster_shade;         /* per-tile raster + vbuffer shading */
    float ms_blit;
    float ms_total;
    float ms_raster;               /* coverage, depth and visibility buffer */
    float ms_shade;                /* visibility-buffer shading */
    uint32_t depth_rejection;      /* 1: the raster rejected entries behind the
                                      tile's stored depth (chosen from the
                                      previous frame's overdraw; same result) */
} sr_frame_stats;

/* ---- batched views and render server ------------------------------------ */
typedef struct sr_view {
    const struct sr_scene *scene;
    sr_camera camera;
    int32_t width, height;
    float exposure;           /* exposure applied by the fused blit (<= 0: the
                                 reference's initial 2.0) */
} sr_view;

typedef struct sr_batch_opts {
    int32_t tile_size;        /* 0 = tuned default */
    int32_t use_graph;        /* capture the launch sequence in a hipGraph */
    int32_t persistent_blocks;/* 0 = 8 per CU */
    int32_t device;
    int32_t reserved[4];
} sr_batch_opts;

typedef struct sr_batch_stats {
    uint64_t input_triangles;
    uint64_t setup_triangles;
    uint64_t binned_pairs;
    uint64_t pixels;
    uint32_t views;
    uint32_t work_items;      /* (view, tile) items drained by the persistent kernel */
    float ms_total;
} sr_batch_stats;

typedef struct sr_server_opts {
    uint64_t admission_bytes; /* framebuffer + bin budget (0 = 64 GiB) */
    int32_t max_batch_views;  /* 0 = 16 */
    int32_t streams;          /* batches in flight (0 = 2) */
    int32_t tile_size;        /* 0 = tuned default */
    int32_t device;
    int32_t reserved[4];
} sr_server_opts;

typedef struct sr_request {
    int32_t scene;            /* index returned by sr_server_add_scene */
    sr_camera camera;
    int32_t width, height;
    float exposure;
    uint64_t user_tag;
} sr_request;

typedef struct sr_completion {
    uint64_t request_id;
    uint64_t user_tag;
    int32_t status;
    int32_t width, height;
    double submit_time_s;     /* server clock */
    double start_time_s;
    double complete_time_s;
} sr_completion;

typedef struct sr_scene sr_scene;
typedef struct sr_renderer sr_renderer;
typedef struct sr_batch sr_batch;
typedef struct sr_server sr_server;

/* ---- library ------------------------------------------------------------ */
SR_API const char *sr_last_error(void);
SR_API int sr_version(int32_t *major, int32_t *minor);
SR_API int sr_device_count(int32_t *count);

/* ---- load-time input producers (host, C++) ------------------------------- */
/* Builds the GGX specular cubemap, irradiance SH and BRDF LUT from the
 * skybox faces (src/texture.rs:199-420; sample counts 64 and 128 as in
 * src/scene.rs:179, :221).  Arrays are owned by the library until
 * sr_environment_free. */
SR_API int sr_prepare_environment(const uint32_t *cubemap_faces, uint32_t face_size,
                                  int32_t threads, sr_environment *out);
/* The same inputs computed on a HIP device (csrc/envmap.hip): the BRDF LUT
 * bit-identical, the specular cubemap within 1 LSB per channel and the SH
 * within float rounding of the host version (the sRGB decode's power is
 * correctly rounded here; the host's libm powf is not always). */
SR_API int sr_prepare_environment_device(const uint32_t *cubemap_faces, uint32_t face_size,
                                         int32_t device, sr_environment *out);
SR_API int sr_environment_free(sr_environment *env);
/* Builds the 128^3 voxel grid over the scene bounds: occupancy and active
 * mask (src/gi.rs:151-265), sun visibility by ray casting
 * (src/gi.rs:267-314, src/raytracer.rs:172-260), 3x3x3 blur
 * (src/voxelgrid.rs:371-419).  scene->voxels is ignored. */
SR_API int sr_prepare_voxel_grid(const sr_scene_desc *scene, int32_t grid_size,
                                 int32_t threads, sr_voxel_grid *out);
/* The same bake on a HIP device (csrc/bake.hip), bit-identical to
 * sr_prepare_voxel_grid: occupancy, dilation, sun rays against the
 * triangles binned on the plane normal to the light (all rays are
 * parallel), blur.  Rays meeting more translucent surfaces than the device
 * keeps make it bake on the host instead. */
SR_API int sr_prepare_voxel_grid_device(const sr_scene_desc *scene, int32_t grid_size,
                                        int32_t device, sr_voxel_grid *out);
/* Voxel grid with the GI bake (the --gi path of load_scene,
 * src/main.rs:238-282: build_active_voxel_mask, compute_sun_visibility,
 * bake_voxel_gi, src/gi.rs:151-412): per active voxel 64 primary rays with
 * a one-bounce surface estimate, projected to 4-coefficient SH, floored and
 * blurred; inactive voxels keep the scene SH.  Fills sh4 (16 floats per
 * voxel) and light_intensity.  device >= 0 bakes on that HIP device, -1 on
 * `threads` host threads (0 = all); both give identical grids.  The scene
 * needs its environment.  The reference's brotli .gi cache file is not
 * read or written.  SR_ERR_CAPACITY: a ray met more than 32 translucent
 * surfaces.  Free with sr_voxel_grid_free. */
SR_API int sr_prepare_gi(const sr_scene_desc *scene, int32_t grid_size, int32_t device,
                         int32_t threads, sr_voxel_grid *out);
/* Cache files of the reference, same layout: a 20-byte little-endian header
 * (magic, version, three u32) then a brotli stream (quality 6, window 22) of
 * the little-endian payload.  Brotli is the system's libbrotlienc.so.1 /
 * libbrotlidec.so.1, opened at run time (SR_ERR_UNSUPPORTED without them).
 * Loads return SR_ERR_NOT_FOUND for a missing file or one the reference
 * would not use (other magic, version, dimensions or payload size) and
 * SR_ERR_INVALID_ARGUMENT for a corrupt stream; either way the caller bakes
 * again, as the reference does.
 * .gi (save_gi_cache / try_load_gi_cache, src/gi.rs:30-117): "VGI0", 8,
 * dims; 16 floats per voxel (4 x rgbw, x fastest).  The load fills
 * sh4[dims product * 16]. */
SR_API int sr_gi_cache_save(const char *path, const sr_voxel_grid *grid);
SR_API int sr_gi_cache_load(const char *path, const int32_t dims[3], float *sh4);
/* .ggx (save_ / try_load_prefiltered_specular_cubemap_cache,
 * src/texture.rs:426-552): "GGX0", 1, width, height, mips; every mip of
 * the six faces at full face size.  The load fills specular[(1 +
 * floor(log2(face_size))) * 6 * face_size^2]. */
SR_API int sr_specular_cache_save(const char *path, const sr_environment *env);
SR_API int sr_specular_cache_load(const char *path, uint32_t face_size, uint32_t *specular);
SR_API int sr_voxel_grid_free(sr_voxel_grid *grid);
/* Scene bounds over all node-transformed positions (src/scene.rs:333-351). */
SR_API int sr_scene_desc_bounds(const sr_scene_desc *scene, float min[3], float max[3],
                                float center[3], float *diagonal);
/* Mip chain of a texture (src/texture.rs:45-128): writes the level offsets,
 * widths and heights (at most 32 levels) and, if texels_out != NULL, the
 * whole chain (*total_texels entries). */
SR_API int sr_texture_mip_chain(const sr_texture_desc *tex, uint32_t *texels_out,
                                uint32_t *total_texels, uint32_t *num_levels,
                                uint32_t offsets[32], uint32_t widths[32],
                                uint32_t heights[32]);

/* ---- glTF scene files (host, C++) ---------------------------------------- */
/* load_gltf + Scene::from_gltf (src/main.rs:81-97, src/scene.rs:145-815): a
 * .gltf (external or base64 data-URI buffers) or .glb file, every mesh,
 * node (world transforms from the hierarchy) and material, PNG textures by
 * URI relative to the file (the reference's TextureCache,
 * src/texture.rs:866-1015).  Missing normals are computed as
 * compute_smooth_normals (src/scene.rs:515-546), missing texture
 * coordinates are zeros.  The result is host memory owned by the handle. */
typedef struct sr_gltf sr_gltf;
SR_API int sr_gltf_load(const char *path, sr_gltf **out);
/* Views of the loaded arrays as a scene description (valid until
 * sr_gltf_destroy); environment and voxels are NULL, the light is the
 * reference's default sun. */
SR_API int sr_gltf_scene_desc(sr_gltf *gltf, sr_scene_desc *desc);
SR_API int sr_gltf_destroy(sr_gltf *gltf);
/* PNG decoding as used for glTF textures (non-interlaced; every colour type
 * and bit depth; tRNS): packed r<<24|g<<16|b<<8|a texels.  With texels NULL
 * only the size is returned.  Images over 2^27 pixels (512 MiB of RGBA8) are
 * refused. */
SR_API int sr_png_decode(const uint8_t *data, size_t size, uint32_t *width, uint32_t *height,
                         uint32_t *texels, size_t capacity);
/* JPEG decoding as used for glTF textures (the reference's image crate,
 * src/texture.rs:897-1015): baseline, extended sequential and progressive
 * Huffman, 8-bit, grey or YCbCr (RGB with an Adobe marker), any sampling
 * factors, restart intervals; packed r<<24|g<<16|b<<8|0xFF texels.  With
 * texels NULL only the size is returned.  Frames over 2^27 pixels are
 * refused before anything is allocated. */
SR_API int sr_jpeg_decode(const uint8_t *data, size_t size, uint32_t *width, uint32_t *height,
                          uint32_t *texels, size_t capacity);

/* ---- camera ------------------------------------------------------------- */
/* RenderCamera::new (src/rendercamera.rs:28-63): yaw/pitch from the look-at
 * direction, near = 0.01 * far.  Matrices are computed by update_matrices. */
SR_API int sr_camera_new(const float position[3], const float look_at[3], float fov,
                         float width, float height, float far_plane, sr_camera *out);
/* A camera given directly by its reference-convention pose. */
SR_API int sr_camera_from_pose(const float position[3], float yaw, float pitch, float roll,
                               float fov, float width, float height, float far_plane,
                               sr_camera *out);
/* RenderCamera::update_matrices (src/rendercamera.rs:78-86). */
SR_API int sr_camera_update_matrices(sr_camera *camera);

/* ---- scene (resident in HBM) -------------------------------------------- */
SR_API int sr_scene_create(const sr_scene_desc *desc, int32_t device, sr_scene **out);
SR_API int sr_scene_destroy(sr_scene *scene);
SR_API int sr_scene_info(const sr_scene *scene, uint64_t *triangles, uint64_t *vertices,
                         uint64_t *device_bytes);

/* ---- renderer: the drop-in per-frame path -------------------------------- */
SR_API int sr_renderer_create(int32_t width, int32_t height, const sr_renderer_opts *opts,
                              sr_renderer **out);
SR_API int sr_renderer_destroy(sr_renderer *renderer);
/* Waits for the queued frame and checks it (see render_scene). */
SR_API int sr_renderer_synchronize(sr_renderer *renderer);
/* Renderer::render_scene: transform, clip, bin, raster and shade one frame.
 * The frame is queued on the renderer's stream.  The first call that reads
 * it (blit, present, an exposure step that meters it, vbuffer / colour /
 * luminance reads, stats, sr_renderer_synchronize, the next render_scene)
 * waits for it, checks it, reruns it once with grown buffers if a capacity
 * overflowed, and reports any error the frame met.  Until then the scene is
 * in use: sr_scene_destroy returns SR_ERR_BUSY. */
SR_API int sr_renderer_render_scene(sr_renderer *renderer, const sr_scene *scene,
                                    const sr_camera *camera);
/* Renderer::update_auto_exposure: trimmed mean of the per-64px-tile centre
 * luminances, exponential adaptation with tau = 1 s. */
SR_API int sr_renderer_update_auto_exposure(sr_renderer *renderer, float delta_time);
/* The same step on the device (src/renderer.rs:258-290 run by one
 * workgroup), queued on the renderer's stream with no host round trip; the
 * blits that follow read the new exposure on the device, and
 * sr_renderer_exposure copies it back.  The device's log2f / expf / powf may
 * differ from the host libm's by an ulp, so the exposure agrees with
 * sr_renderer_update_auto_exposure to a few ulps (RGBA8 within 1 LSB).
 * Above 8192 reference tiles it runs the host step instead. */
SR_API int sr_renderer_update_auto_exposure_device(sr_renderer *renderer, float delta_time);
/* Renderer::blit_to_buffer: exposure, tonemap, quantize, pack into a host
 * buffer of width*height u32 owned by the caller.  Rows are written in
 * pairs; an odd last row is left untouched as in the reference. */
SR_API int sr_renderer_blit_to_buffer(sr_renderer *renderer, uint32_t *pixels,
                                      size_t width, size_t height);
/* The same blit into device memory (width*height u32); nothing crosses PCIe. */
SR_API int sr_renderer_blit_to_device(sr_renderer *renderer, void *device_pixels);
/* App::render_frame's presentation copy (src/main.rs:574-590: each packed
 * pixel unpacked into the bytes r, g, b, a of the pixels frame) fused into the
 * blit: frame receives width*height*4 bytes in that order.  As with the blit,
 * rows come in pairs, so an odd last row is not written. */
SR_API int sr_renderer_present(sr_renderer *renderer, uint8_t *frame, size_t width, size_t height);
/* The same into device memory (width*height*4 bytes). */
SR_API int sr_renderer_present_to_device(sr_renderer *renderer, void *device_frame);
SR_API float sr_renderer_exposure(const sr_renderer *renderer);
SR_API int sr_renderer_set_exposure(sr_renderer *renderer, float exposure);
/* Visibility buffer of the last frame: canonical triangle id (0xFFFFFFFF =
 * no fragment) and depth (+inf = sky), width*height each; either may be NULL. */
SR_API int sr_renderer_read_vbuffer(sr_renderer *renderer, uint32_t *tri_id, float *depth);
/* Linear shaded colour of the last frame, width*height*3 floats. */
SR_API int sr_renderer_read_color(sr_renderer *renderer, float *rgb);
/* Centre-quad luminance of every 64px tile (src/tilerasterizer.rs:103-106). */
SR_API int sr_renderer_read_tile_luminance(sr_renderer *renderer, float *lum, int32_t *count);
SR_API int sr_renderer_get_stats(const sr_renderer *renderer, sr_frame_stats *stats);
/* How frames time their stages for sr_frame_stats (the reference's clip/bin
 * and raster timers, src/renderer.rs:206-219).  SR_STAGE_TIMING_CLOCK (the
 * default): the kernels stamp the stage boundaries on the GPU's
 * constant-rate wall clock, read back only when stats are asked for; it
 * fills ms_setup, ms_bin, ms_raster_shade and ms_total (about 2 us per
 * frame).  SR_STAGE_TIMING_EVENTS: HIP events carried by the stage kernels'
 * own launches, which also split ms_raster from ms_shade (about 18 us per
 * frame).  SR_STAGE_TIMING_OFF: no stage times (the ms fields read 0).
 * Applies from the next frame. */
#define SR_STAGE_TIMING_CLOCK 0
#define SR_STAGE_TIMING_EVENTS 1
#define SR_STAGE_TIMING_OFF 2
SR_API int sr_renderer_set_stage_timing(sr_renderer *renderer, int32_t mode);

/* One frame without allocating device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Renders
 * `scene` from `camera` at width x height into the caller's device buffers
 * (width*height each): tri_id and depth as sr_renderer_read_vbuffer gives
 * them, rgba as sr_renderer_blit_to_buffer packs the frame with the given
 * exposure (<= 0: the reference's initial 2.0; every row, also an odd last
 * one), on `stream` (a hipStream_t; NULL: the null stream), and waits for
 * it.  The frame's transient buffers come from one device array of
 * SR_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from its first
 * call on that GPU until the process ends; calls that overlap on one GPU
 * race on it.  SR_ERR_CAPACITY: the frame's buffers do not fit (checked
 * before launching for the sizes known up front, after it for the records,
 * list entries and clipped triangles the frame produced).
 * SR_ERR_UNSUPPORTED: translucent primitives or normal-mapped materials. */
#define SR_NOALLOC_SCRATCH_BYTES (160ull << 20)
SR_API int sr_render_frame_noalloc(const sr_scene *scene, const sr_camera *camera, int32_t width, int32_t height,
                                   float exposure, uint32_t *tri_id, float *depth, uint32_t *rgba, void *stream);
/* Parity hook: the last frame's per-tile lists (bins) of set-up triangle
 * ids, the form of the reference's per-tile packet queues
 * (src/bumpqueue.rs:96-112).  *items receives the number of screen tiles;
 * start and end (items entries each, either may be NULL) bound tile i's ids
 * in list, which takes up to list_len of them (list may be NULL).  Each
 * tile's ids ascend: the queue order of the reference's single-threaded
 * front end. */
SR_API int sr_renderer_read_bins(sr_renderer *renderer, uint32_t *start, uint32_t *end, uint32_t *list,
                                 uint64_t list_len, uint32_t *items);
/* Device pointers of the last frame's visibility buffer and colour (for
 * in-HBM consumers); valid until the next render call. */
SR_API int sr_renderer_device_buffers(sr_renderer *renderer, void **tri_id, void **depth,
                                      void **color_rgb);

/* ---- batched views (one launch sequence for many views) ------------------ */
SR_API int sr_batch_create(const sr_batch_opts *opts, sr_batch **out);
SR_API int sr_batch_destroy(sr_batch *batch);
/* Changes the persistent raster workgroup count (> 0; 0 keeps it; -1 launches
 * one raster workgroup per (view, tile) instead) and hipGraph use of an
 * existing batch; the tile size is fixed at creation.  Used by the autotuner
 * (no reference counterpart). */
SR_API int sr_batch_configure(sr_batch *batch, int32_t persistent_blocks, int32_t use_graph);
/* Renders every view; per-view outputs are written to device memory owned
 * by the batch and read back with sr_batch_read_*. */
SR_API int sr_batch_render(sr_batch *batch, const sr_view *views, int32_t num_views);
SR_API int sr_batch_read_vbuffer(sr_batch *batch, int32_t view, uint32_t *tri_id, float *depth);
SR_API int sr_batch_read_pixels(sr_batch *batch, int32_t view, uint32_t *pixels);
SR_API int sr_batch_get_stats(const sr_batch *batch, sr_batch_stats *stats);

/* ---- render server: continuous batching over resident scenes ------------- */
SR_API int sr_server_create(const sr_server_opts *opts, sr_server **out);
SR_API int sr_server_destroy(sr_server *server);
SR_API int sr_server_add_scene(sr_server *server, const sr_scene *scene, int32_t *scene_index);
SR_API int sr_server_submit(sr_server *server, const sr_request *request, uint64_t *request_id);
/* Pops up to max completions; *count receives how many were written. */
SR_API int sr_server_poll(sr_server *server, sr_completion *out, int32_t max, int32_t *count);
SR_API int sr_server_read_pixels(sr_server *server, uint64_t request_id, uint32_t *pixels);
SR_API int sr_server_read_vbuffer(sr_server *server, uint64_t request_id, uint32_t *tri_id,
                                  float *depth);
SR_API int sr_server_release(sr_server *server, uint64_t request_id);
SR_API int sr_server_drain(sr_server *server);
SR_API double sr_server_clock(const sr_server *server);

#ifdef __cplusplus
}
#endif

#endif /* SWRASTER_H */
