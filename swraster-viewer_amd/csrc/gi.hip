// gi.hip -- the voxel GI bake on the GPU (bake_voxel_gi, src/gi.rs:316-412;
// blur_gi_sh4, src/voxelgrid.rs:422-472).
//
// One wavefront per active voxel: lane s traces primary direction s of the
// voxel's direction set (GI_SAMPLES == 64 == the wave width) through the
// uploaded BVH, with its surface bounce (one sun and four sky transmittance
// rays) where it hits; the 64 radiances meet in LDS and lane 0 reduces them
// in sample order (gi_finish_voxel), so every voxel's coefficients are the
// host bake's (gi.cpp: the same gi_core.h operations in the same order).
// A 3x3x3 blur kernel over the whole grid follows.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "gi_host.h"
#include "runtime.h"

namespace sr {
namespace {

static_assert(GI_SAMPLES == 64, "one lane per primary direction");

struct GiDev {
    GiScene S;
    const float *dirs;            // GI_SETS * GI_SAMPLES * 3
    const uint32_t *voxels;       // active voxel indices
    uint32_t nactive;
    int grid;
    float cmin[3], vs[3];
    float primary_bias;
    GiRayParams P;
    const float *light;           // per voxel sun visibility
    float *coeffs;                // per voxel 16 floats
    uint32_t *over;               // set if a transmittance ray met more than GI_HIT_CAP surfaces
};

// initialize_voxel_gi_from_scene with scale 1 (src/gi.rs:123-149).
__global__ void k_gi_init(GiDev G, size_t nvox) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (size_t)gridDim.x * blockDim.x) {
        float o[16];
        for (int c = 0; c < 4; c++) {
            for (int k = 0; k < 3; k++) o[4 * c + k] = G.S.sh[c][k] * 1.0f;
            o[4 * c + 3] = 0.0f;
        }
        o[3] = G.light[i];
        o[7] = 1.0f;
        float4 *dst = reinterpret_cast<float4 *>(G.coeffs + 16 * i);
        for (int c = 0; c < 4; c++) dst[c] = make_float4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
    }
}

// Traversal stack of one lane in LDS, entries strided by the wave width so
// that the lanes of a wave touch consecutive words.
struct GiLdsStack {
    int32_t *base;
    int n;
    __device__ void push(int32_t v) { base[64 * n++] = v; }
    __device__ int32_t pop() { return base[64 * --n]; }
};

// One wave per active voxel, four per workgroup.
__global__ __launch_bounds__(256) void k_gi_voxels(GiDev G, uint32_t first) {
    __shared__ float rad[4][GI_SAMPLES * 3];
    __shared__ uint32_t sky[4][GI_SAMPLES];
    __shared__ int32_t stacks[4][GI_STACK * 64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t a = first + blockIdx.x * 4 + w;
    const bool live = a < G.nactive;
    uint32_t vi = 0;
    int x = 0, y = 0, z = 0, set = 0;
    if (live) {
        vi = G.voxels[a];
        x = (int)(vi % (uint32_t)G.grid);
        y = (int)((vi / (uint32_t)G.grid) % (uint32_t)G.grid);
        z = (int)(vi / ((uint32_t)G.grid * (uint32_t)G.grid));
        set = ((x % 3) + 3 * (y % 3) + (z % 3)) % GI_SETS;
        G3 c = g3(G.cmin[0] + (float)x * G.vs[0], G.cmin[1] + (float)y * G.vs[1], G.cmin[2] + (float)z * G.vs[2]);
        const float *d = G.dirs + ((size_t)set * GI_SAMPLES + lane) * 3;
        bool hit_sky = false, over = false;
        GiLdsStack st{&stacks[w][lane], 0};
        G3 r = gi_primary_ray(G.S, c, g3(d[0], d[1], d[2]), G.primary_bias, G.P, hit_sky, over, st);
        rad[w][3 * lane] = r.x;
        rad[w][3 * lane + 1] = r.y;
        rad[w][3 * lane + 2] = r.z;
        sky[w][lane] = hit_sky ? 1u : 0u;
        if (over) *G.over = 1u;
    }
    __syncthreads();
    if (live && lane == 0) {
        uint32_t hits = 0;
        for (int s = 0; s < GI_SAMPLES; s++) hits += sky[w][s];
        float o[16];
        gi_finish_voxel(G.S, G.dirs + (size_t)set * GI_SAMPLES * 3, rad[w], hits, G.light[vi], o);
        float4 *dst = reinterpret_cast<float4 *>(G.coeffs + 16 * (size_t)vi);
        for (int c = 0; c < 4; c++) dst[c] = make_float4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
    }
}

// blur_gi_sh4 (src/voxelgrid.rs:422-472)
__global__ void k_gi_blur(const float *src, float *dst, int grid) {
    size_t n = (size_t)grid * grid * grid;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int x = (int)(i % grid), y = (int)((i / grid) % grid), z = (int)(i / ((size_t)grid * grid));
        float o[16];
        gi_blur_voxel(src, grid, grid, grid, x, y, z, o);
        float4 *d = reinterpret_cast<float4 *>(dst + 16 * i);
        for (int c = 0; c < 4; c++) d[c] = make_float4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
    }
}

struct DevMem {
    std::vector<void *> ptrs;
    ~DevMem() {
        for (void *p : ptrs) hipFree(p);
    }
    template <class T> hipError_t get(T *&p, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(q);
        p = (T *)q;
        return e;
    }
    template <class T> hipError_t put(const T *&p, const T *host, size_t count) {
        T *q;
        hipError_t e = get(q, count);
        if (e == hipSuccess && count) e = hipMemcpy(q, host, count * sizeof(T), hipMemcpyHostToDevice);
        p = q;
        return e;
    }
};

inline uint32_t blocks(size_t n) { return (uint32_t)std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 65536); }

}  // namespace

#define GI_TRY(expr)                                                                           \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

int gi_bake_device(const GiHost &h, const std::vector<uint8_t> &active, const std::vector<float> &light, int device,
                   std::vector<float> &out) {
    const int grid = h.frame.grid;
    const size_t nvox = (size_t)grid * grid * grid;
    if (active.size() != nvox || light.size() != nvox) return fail(SR_ERR_INVALID_ARGUMENT, "GI grid size mismatch");
    std::vector<uint32_t> list;
    for (size_t i = 0; i < nvox; i++)
        if (active[i]) list.push_back((uint32_t)i);
    GI_TRY(hipSetDevice(device));
    DevMem mem;
    GiDev G;
    memset(&G, 0, sizeof(G));
    G.S = gi_scene_view(h);
    GI_TRY(mem.put(G.S.tris, h.tris.data(), h.tris.size()));
    GI_TRY(mem.put(G.S.nodes, h.nodes.data(), h.nodes.size()));
    GI_TRY(mem.put(G.S.order, h.order.data(), h.order.size()));
    GI_TRY(mem.put(G.S.mats, h.mats.data(), h.mats.size()));
    GI_TRY(mem.put(G.S.texs, h.texs.data(), h.texs.size()));
    GI_TRY(mem.put(G.S.texels, h.texels.data(), h.texels.size()));
    GI_TRY(mem.put(G.S.cubemap, h.cubemap, (size_t)6 * h.cube_size * h.cube_size));
    GI_TRY(mem.put(G.dirs, h.dirs.data(), h.dirs.size()));
    GI_TRY(mem.put(G.voxels, list.data(), list.size()));
    GI_TRY(mem.put(G.light, light.data(), nvox));
    float *blurred;
    GI_TRY(mem.get(G.coeffs, nvox * 16));
    GI_TRY(mem.get(blurred, nvox * 16));
    GI_TRY(mem.get(G.over, 1));
    GI_TRY(hipMemset(G.over, 0, 4));
    G.nactive = (uint32_t)list.size();
    G.grid = grid;
    memcpy(G.cmin, h.frame.cmin, 12);
    memcpy(G.vs, h.frame.vs, 12);
    G.primary_bias = h.primary_bias;
    G.P = h.params;
    hipLaunchKernelGGL(k_gi_init, dim3(blocks(nvox)), dim3(256), 0, nullptr, G, nvox);
    GI_TRY(hipGetLastError());
    out.resize(nvox * 16);
    if (list.empty()) {     // nothing active: the scene SH, not blurred (src/gi.rs:338-341)
        GI_TRY(hipMemcpy(out.data(), G.coeffs, nvox * 64, hipMemcpyDeviceToHost));
        return SR_OK;
    }
    // launches of at most 65536 voxels each
    for (uint32_t first = 0; first < G.nactive; first += 65536) {
        uint32_t n = std::min<uint32_t>(65536, G.nactive - first);
        hipLaunchKernelGGL(k_gi_voxels, dim3((n + 3) / 4), dim3(256), 0, nullptr, G, first);
        GI_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_gi_blur, dim3(blocks(nvox)), dim3(256), 0, nullptr, (const float *)G.coeffs, blurred, grid);
    GI_TRY(hipGetLastError());
    uint32_t over = 0;
    GI_TRY(hipMemcpy(&over, G.over, 4, hipMemcpyDeviceToHost));
    if (over) return fail(SR_ERR_CAPACITY, "a GI ray met more than %d translucent surfaces", GI_HIT_CAP);
    GI_TRY(hipMemcpy(out.data(), blurred, nvox * 64, hipMemcpyDeviceToHost));
    return SR_OK;
}

}  // namespace sr
