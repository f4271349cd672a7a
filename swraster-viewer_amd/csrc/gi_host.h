// gi_host.h -- host-side inputs of the voxel GI bake, shared by the host bake
// (gi.cpp) and the device bake (gi.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "gi_core.h"
#include "prep.h"

namespace sr {

struct GiHost {
    VoxelFrame frame;
    std::vector<GiTri> tris;
    std::vector<GiNode> nodes;       // node 0 is the root
    std::vector<int32_t> order;      // leaf ranges index this
    std::vector<GiMat> mats;
    std::vector<GiTex> texs;
    std::vector<uint32_t> texels;
    std::vector<float> dirs;         // GI_SETS * GI_SAMPLES * 3 primary directions
    const uint32_t *cubemap = nullptr;
    uint32_t cube_size = 0;
    float sh[4][3];
    float light_color[3];
    float primary_bias;
    GiRayParams params;
};

int gi_prepare(const sr_scene_desc *d, int grid, GiHost &h);
GiScene gi_scene_view(const GiHost &h);
void gi_initial(const GiHost &h, const std::vector<float> &light, std::vector<float> &sh4);
int gi_bake_host(const GiHost &h, const std::vector<uint8_t> &active, const std::vector<float> &light, int threads,
                 std::vector<float> &out);
// The same bake on a HIP device (gi.hip); identical results.
int gi_bake_device(const GiHost &h, const std::vector<uint8_t> &active, const std::vector<float> &light, int device,
                   std::vector<float> &out);

}  // namespace sr
