// prep.h -- host-side load-time input producers and camera math.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

#include "swraster.h"

namespace sr {

// fn(i) for i in [0, n) over `threads` host threads (all cores when <= 0),
// each taking the next index from a shared counter.
template <class F> void parallel_for(size_t n, int threads, F &&fn) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (threads == 1 || n < 2) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads && (size_t)t < n; t++)
        pool.emplace_back([&]() {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= n) break;
                fn(i);
            }
        });
    for (auto &th : pool) th.join();
}

void build_mip_chain(uint32_t width, uint32_t height, int type, const uint32_t *base, std::vector<uint32_t> &data,
                     std::vector<uint32_t> &offsets, std::vector<uint32_t> &widths, std::vector<uint32_t> &heights);
void compute_brdf_lut(uint32_t size, std::vector<uint32_t> &out, int threads);
void compute_irradiance_sh4(const uint32_t *faces, uint32_t size, float sh[4][3]);
void compute_specular_cubemap(const uint32_t *faces, uint32_t size, uint32_t sample_count, std::vector<uint32_t> &out,
                              uint32_t &num_mips, int threads);
void scene_bounds(const sr_scene_desc *d, float mn[3], float mx[3], float center[3], float *diag);
// active_out (optional): the bake's active-voxel mask (occupancy dilated by 2).
int build_voxel_grid(const sr_scene_desc *d, int grid, int threads, std::vector<float> &light, float wmin[3],
                     float wmax[3], std::vector<uint8_t> *active_out = nullptr);

// Shared by the host bake and the device bake (bake.hip).
struct WorldTri {
    float p[9];          // p0, p1, p2 in world space
    int32_t material;
};
struct VoxelFrame {
    int grid;
    float wmin[3], wmax[3];
    float lo[3], vs[3];  // grid origin and voxel size
    float cmin[3];       // centre of voxel (0, 0, 0)
    float ldir[3];       // normalized sun direction
    float bias;          // ray origin offset along ldir
    float area_ref;      // occupancy sampling reference area
};
// uvs (optional): the three vertices' texture coordinates of each triangle.
void collect_world_triangles(const sr_scene_desc *d, int threads, std::vector<WorldTri> &out,
                             std::vector<float> *uvs = nullptr);
void voxel_frame(const sr_scene_desc *d, int grid, VoxelFrame &f);
// The bake on a HIP device (bake.hip).  fallback: a ray met more translucent
// surfaces than the device keeps per ray; the caller bakes on the host.
int bake_voxels_device(const sr_scene_desc *d, int grid, int device, std::vector<float> &light, VoxelFrame &fr,
                       bool &fallback, std::vector<uint8_t> *active_out = nullptr);
// IBL inputs on a HIP device (envmap.hip).
int prepare_environment_device(const uint32_t *faces, uint32_t size, int device, std::vector<uint32_t> &spec,
                               uint32_t &mips, std::vector<uint32_t> &lut, uint32_t lut_size, float sh[4][3]);
void camera_new(const float pos[3], const float look[3], float fov, float w, float h, float far_plane, sr_camera *c);
void camera_update(sr_camera *c);
void compute_tangents(const sr_primitive_desc &p, std::vector<float> &out);
// Column-major 4x4 product a * b and rotation matrix of quaternion (x, y, z, w)
// (glam Mat4::mul_mat4, Mat4::from_quat).
void mat4_mul(const float a[16], const float b[16], float out[16]);
void mat4_from_quat(const float q[4], float out[16]);

}  // namespace sr
