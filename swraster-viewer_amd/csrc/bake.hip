// bake.hip -- the voxel sun-visibility bake on the GPU (build_active_voxel_mask,
// compute_sun_visibility, src/gi.rs:151-314; ray_triangle_intersect,
// src/raytracer.rs:225-260; blur_grid, src/voxelgrid.rs:371-419).
//
// Same voxel frame, world triangles and floating-point operations as the
// host bake (prep.cpp), so the grids are identical.  Sun rays are all
// parallel, so instead of a BVH the triangles are binned on a 2-D grid over
// the plane normal to the light: a ray meets exactly the triangles whose
// projection covers its cell (bins are widened by a margin far above the
// rounding of the projection), each once.  An opaque hit ends the ray; the
// transmissions of translucent hits multiply in order of distance (equal
// distances in triangle order, as the host sorts them).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.h"
#include "prep.h"
#include "runtime.h"

namespace sr {
namespace {

struct F3d { float x, y, z; };
__device__ __forceinline__ F3d mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3d sub(F3d a, F3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dotf(F3d a, F3d b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ F3d crossf(F3d a, F3d b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Bake {
    const WorldTri *tris;
    uint32_t ntri;
    int grid;
    float lo[3], vs[3], cmin[3], ldir[3], bias, area_ref;
    uint8_t *occupied, *active;
    float *vis, *light;
    // light-plane grid
    double e1[3], e2[3];
    double u0, v0, inv_cu, inv_cv, margin;
    int G;
    uint32_t *cell_count, *cell_start, *cell_cursor, *cell_list;
    const uint8_t *translucent;
    const float *transmission;
    int nmat;
    uint32_t *status;   // [0] translucent hits beyond HIT_CAP on some ray
};

#define HIT_CAP 24

__device__ __forceinline__ size_t vox(const Bake &B, size_t x, size_t y, size_t z) {
    return (z * (size_t)B.grid + y) * (size_t)B.grid + x;
}

// Occupancy by surface sampling (src/gi.rs:151-240): one thread per triangle.
__global__ void __launch_bounds__(256) k_bake_occupy(Bake B) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.ntri; i += gridDim.x * blockDim.x) {
        const WorldTri &t = B.tris[i];
        F3d p0 = mk(t.p[0], t.p[1], t.p[2]), p1 = mk(t.p[3], t.p[4], t.p[5]), p2 = mk(t.p[6], t.p[7], t.p[8]);
        F3d c = crossf(sub(p1, p0), sub(p2, p0));
        float area = 0.5f * sqrtf(dotf(c, c));
        float ts = ceilf((area / B.area_ref) * 2.0f);
        uint32_t target = !(ts > 0.0f) ? 1u : (ts >= 4096.0f ? 4096u : (uint32_t)ts);
        if (target < 1u) target = 1u;
        uint32_t ns = (uint32_t)ceilf(sqrtf((float)target));
        for (uint32_t iu = 0; iu < ns; iu++)
            for (uint32_t iv = 0; iv < ns - iu; iv++) {
                float u = ((float)iu + 0.5f) / (float)ns, v = ((float)iv + 0.5f) / (float)ns;
                float w = (1.0f - u) - v;
                if (w < 0.0f) continue;
                float px = (p0.x * w + p1.x * u) + p2.x * v;
                float py = (p0.y * w + p1.y * u) + p2.y * v;
                float pz = (p0.z * w + p1.z * u) + p2.z * v;
                float fx = (px - B.lo[0]) / B.vs[0], fy = (py - B.lo[1]) / B.vs[1], fz = (pz - B.lo[2]) / B.vs[2];
                // written so that NaN coordinates are skipped as well
                if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f)) continue;
                double x = floor((double)fx), y = floor((double)fy), z = floor((double)fz);
                if (!(x < B.grid && y < B.grid && z < B.grid)) continue;
                B.occupied[vox(B, (size_t)x, (size_t)y, (size_t)z)] = 1;
            }
    }
}

// Active mask: occupied voxels dilated by 2 in every axis (src/gi.rs:241-265).
__global__ void __launch_bounds__(256) k_bake_dilate(Bake B) {
    size_t n = (size_t)B.grid * B.grid * B.grid;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int x = (int)(i % B.grid), y = (int)((i / B.grid) % B.grid), z = (int)(i / ((size_t)B.grid * B.grid));
        uint8_t a = 0;
        for (int nz = max(z - 2, 0); nz <= min(z + 2, B.grid - 1) && !a; nz++)
            for (int ny = max(y - 2, 0); ny <= min(y + 2, B.grid - 1) && !a; ny++)
                for (int nx = max(x - 2, 0); nx <= min(x + 2, B.grid - 1); nx++)
                    if (B.occupied[vox(B, nx, ny, nz)]) { a = 1; break; }
        B.active[i] = a;
    }
}

__device__ __forceinline__ void project(const Bake &B, double x, double y, double z, double &u, double &v) {
    u = (x * B.e1[0] + y * B.e1[1]) + z * B.e1[2];
    v = (x * B.e2[0] + y * B.e2[1]) + z * B.e2[2];
}

// Cell index along one axis, clamped (NaN -> 0).
__device__ __forceinline__ int cell_of(double c, int G) {
    double f = floor(c);
    if (!(f >= 0.0)) return 0;
    return f > (double)(G - 1) ? G - 1 : (int)f;
}

// Triangles into the light-plane cells their widened projection covers:
// count (list null) or scatter.
__global__ void __launch_bounds__(256) k_bake_bin(Bake B, uint32_t *counts, uint32_t *list) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.ntri; i += gridDim.x * blockDim.x) {
        const WorldTri &t = B.tris[i];
        double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
        for (int k = 0; k < 3; k++) {
            double u, v;
            project(B, t.p[3 * k], t.p[3 * k + 1], t.p[3 * k + 2], u, v);
            umin = fmin(umin, u); umax = fmax(umax, u);
            vmin = fmin(vmin, v); vmax = fmax(vmax, v);
        }
        int cu0 = cell_of((umin - B.margin - B.u0) * B.inv_cu, B.G), cu1 = cell_of((umax + B.margin - B.u0) * B.inv_cu, B.G);
        int cv0 = cell_of((vmin - B.margin - B.v0) * B.inv_cv, B.G), cv1 = cell_of((vmax + B.margin - B.v0) * B.inv_cv, B.G);
        for (int cv = cv0; cv <= cv1; cv++)
            for (int cu = cu0; cu <= cu1; cu++) {
                uint32_t pos = atomicAdd(&counts[(size_t)cv * B.G + cu], 1u);
                if (list) list[pos] = i;
            }
    }
}

// ray_triangle_intersect (src/raytracer.rs:225-260), the host's operations.
__device__ __forceinline__ bool ray_tri(F3d o, F3d d, const WorldTri &t, float tmin, float tmax, float &tout) {
    F3d p0 = mk(t.p[0], t.p[1], t.p[2]);
    F3d e1 = sub(mk(t.p[3], t.p[4], t.p[5]), p0), e2 = sub(mk(t.p[6], t.p[7], t.p[8]), p0);
    F3d pv = crossf(d, e2);
    float det = dotf(e1, pv);
    if (fabsf(det) <= 1.0e-8f) return false;
    float inv = 1.0f / det;
    F3d tv = sub(o, p0);
    float u = dotf(tv, pv) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    F3d qv = crossf(tv, e1);
    float v = dotf(d, qv) * inv;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    float tt = dotf(e2, qv) * inv;
    if (tt < tmin || tt > tmax) return false;
    tout = tt;
    return true;
}

// Sun visibility of each active voxel (src/gi.rs:267-314).
__global__ void __launch_bounds__(256) k_bake_rays(Bake B) {
    size_t n = (size_t)B.grid * B.grid * B.grid;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float vis = 1.0f;
        if (B.active[i]) {
            int x = (int)(i % B.grid), y = (int)((i / B.grid) % B.grid), z = (int)(i / ((size_t)B.grid * B.grid));
            F3d c = mk(B.cmin[0] + (float)x * B.vs[0], B.cmin[1] + (float)y * B.vs[1], B.cmin[2] + (float)z * B.vs[2]);
            F3d d = mk(B.ldir[0], B.ldir[1], B.ldir[2]);
            F3d o = mk(c.x + d.x * B.bias, c.y + d.y * B.bias, c.z + d.z * B.bias);
            double u, v;
            project(B, o.x, o.y, o.z, u, v);
            size_t cell = (size_t)cell_of((v - B.v0) * B.inv_cv, B.G) * B.G + cell_of((u - B.u0) * B.inv_cu, B.G);
            uint32_t s = B.cell_start[cell], e = s + B.cell_count[cell];
            float ht[HIT_CAP];
            uint32_t hi[HIT_CAP];
            int nh = 0;
            bool opaque = false, over = false;
            for (uint32_t k = s; k < e; k++) {
                uint32_t ti = B.cell_list[k];
                const WorldTri &t = B.tris[ti];
                float tt;
                if (!ray_tri(o, d, t, 1.0e-4f, INFINITY, tt)) continue;
                int m = (t.material >= 0 && t.material < B.nmat) ? t.material : 0;
                if (!B.translucent[m]) {
                    opaque = true;
                    break;
                }
                if (nh == HIT_CAP) {
                    over = true;
                    continue;
                }
                // insertion by (distance, triangle index)
                int j = nh++;
                while (j > 0 && (ht[j - 1] > tt || (ht[j - 1] == tt && hi[j - 1] > ti))) {
                    ht[j] = ht[j - 1];
                    hi[j] = hi[j - 1];
                    j--;
                }
                ht[j] = tt;
                hi[j] = ti;
            }
            if (opaque) {
                vis = 0.0f;
            } else {
                if (over) atomicOr(&B.status[0], 1u);
                float tr = 1.0f;
                for (int j = 0; j < nh; j++) {
                    int m = B.tris[hi[j]].material;
                    m = (m >= 0 && m < B.nmat) ? m : 0;
                    tr *= B.transmission[m];
                    if (tr <= 0.0001f) {
                        tr = 0.0f;
                        break;
                    }
                }
                vis = tr;
            }
        }
        B.vis[i] = vis;
    }
}

// blur_grid (src/voxelgrid.rs:371-419): 3x3x3 mean in the host's summation
// order, then squared.
__global__ void __launch_bounds__(256) k_bake_blur(Bake B) {
    size_t n = (size_t)B.grid * B.grid * B.grid;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int x = (int)(i % B.grid), y = (int)((i / B.grid) % B.grid), z = (int)(i / ((size_t)B.grid * B.grid));
        float sum = 0.0f;
        int count = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int nx = x + dx, ny = y + dy, nz = z + dz;
                    if (nx >= 0 && nx < B.grid && ny >= 0 && ny < B.grid && nz >= 0 && nz < B.grid) {
                        sum += B.vis[vox(B, nx, ny, nz)];
                        count++;
                    }
                }
        float m = count > 0 ? sum / (float)count : 0.0f;
        B.light[i] = count > 0 ? m * m : 0.0f;
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
};

inline uint32_t blocks(size_t n) { return (uint32_t)std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 65536); }

}  // namespace

#define BAKE_TRY(expr)                                                                         \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

int bake_voxels_device(const sr_scene_desc *d, int grid, int device, std::vector<float> &light, VoxelFrame &fr,
                       bool &fallback, std::vector<uint8_t> *active_out) {
    fallback = false;
    voxel_frame(d, grid, fr);
    std::vector<WorldTri> tris;
    collect_world_triangles(d, 0, tris);
    int nmat = std::max(d->num_materials, 1);
    std::vector<uint8_t> translucent(nmat, 0);
    std::vector<float> transmission(nmat, 0.0f);
    for (int i = 0; i < d->num_materials; i++) {
        translucent[i] = d->materials[i].is_translucent != 0;
        transmission[i] = d->materials[i].transmission;
    }
    size_t nvox = (size_t)grid * grid * grid;
    BAKE_TRY(hipSetDevice(device));
    hipStream_t st = nullptr;
    DevMem mem;
    Bake B;
    memset(&B, 0, sizeof(B));
    B.ntri = (uint32_t)tris.size();
    B.grid = grid;
    memcpy(B.lo, fr.lo, 12);
    memcpy(B.vs, fr.vs, 12);
    memcpy(B.cmin, fr.cmin, 12);
    memcpy(B.ldir, fr.ldir, 12);
    B.bias = fr.bias;
    B.area_ref = fr.area_ref;
    B.nmat = nmat;
    WorldTri *dt;
    uint8_t *dtr;
    float *dtm;
    BAKE_TRY(mem.get(dt, tris.size()));
    BAKE_TRY(mem.get(B.occupied, nvox));
    BAKE_TRY(mem.get(B.active, nvox));
    BAKE_TRY(mem.get(B.vis, nvox));
    BAKE_TRY(mem.get(B.light, nvox));
    BAKE_TRY(mem.get(dtr, (size_t)nmat));
    BAKE_TRY(mem.get(dtm, (size_t)nmat));
    BAKE_TRY(mem.get(B.status, 4));
    if (!tris.empty()) BAKE_TRY(hipMemcpy(dt, tris.data(), tris.size() * sizeof(WorldTri), hipMemcpyHostToDevice));
    BAKE_TRY(hipMemcpy(dtr, translucent.data(), (size_t)nmat, hipMemcpyHostToDevice));
    BAKE_TRY(hipMemcpy(dtm, transmission.data(), (size_t)nmat * 4, hipMemcpyHostToDevice));
    BAKE_TRY(hipMemset(B.occupied, 0, nvox));
    BAKE_TRY(hipMemset(B.status, 0, 16));
    B.tris = dt;
    B.translucent = dtr;
    B.transmission = dtm;

    // light-plane frame: e1, e2 orthonormal to the sun direction
    double L[3] = {fr.ldir[0], fr.ldir[1], fr.ldir[2]};
    double ln = std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
    for (double &c : L) c /= ln;
    double a[3] = {1, 0, 0};
    if (std::fabs(L[0]) > 0.9) { a[0] = 0; a[1] = 1; }
    double e1[3] = {L[1] * a[2] - L[2] * a[1], L[2] * a[0] - L[0] * a[2], L[0] * a[1] - L[1] * a[0]};
    double n1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (double &c : e1) c /= n1;
    double e2[3] = {L[1] * e1[2] - L[2] * e1[1], L[2] * e1[0] - L[0] * e1[2], L[0] * e1[1] - L[1] * e1[0]};
    memcpy(B.e1, e1, 24);
    memcpy(B.e2, e2, 24);
    // projected extent of the grid box (rays start inside it, triangles lie in it)
    double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
    for (int k = 0; k < 8; k++) {
        double p[3] = {(k & 1) ? fr.wmax[0] : fr.wmin[0], (k & 2) ? fr.wmax[1] : fr.wmin[1], (k & 4) ? fr.wmax[2] : fr.wmin[2]};
        double u = (p[0] * e1[0] + p[1] * e1[1]) + p[2] * e1[2], v = (p[0] * e2[0] + p[1] * e2[1]) + p[2] * e2[2];
        umin = std::min(umin, u); umax = std::max(umax, u);
        vmin = std::min(vmin, v); vmax = std::max(vmax, v);
    }
    double ext = std::max(umax - umin, vmax - vmin);
    // widen each triangle's bins far beyond the rounding of the projection
    // and of the f32 hit test (relative 2^-24 of the scene size)
    B.margin = 1.0e-4 * ext + 1.0e-6;
    int G = (int)std::sqrt((double)std::max<size_t>(tris.size(), 1));
    G = std::min(std::max(G, 16), 2048);
    B.G = G;
    B.u0 = umin - 0.01 * ext;
    B.v0 = vmin - 0.01 * ext;
    B.inv_cu = G / ((umax - umin) + 0.02 * ext + 1e-30);
    B.inv_cv = G / ((vmax - vmin) + 0.02 * ext + 1e-30);
    size_t ncell = (size_t)G * G;
    uint32_t *partials, *total;
    BAKE_TRY(mem.get(B.cell_count, ncell));
    BAKE_TRY(mem.get(B.cell_start, ncell));
    BAKE_TRY(mem.get(B.cell_cursor, ncell));
    BAKE_TRY(mem.get(partials, (ncell + 4095) / 4096 + 1));
    BAKE_TRY(mem.get(total, 1));
    BAKE_TRY(hipMemset(B.cell_count, 0, ncell * 4));

    if (B.ntri) hipLaunchKernelGGL(k_bake_occupy, dim3(blocks(B.ntri)), dim3(256), 0, st, B);
    hipLaunchKernelGGL(k_bake_dilate, dim3(blocks(nvox)), dim3(256), 0, st, B);
    if (B.ntri) hipLaunchKernelGGL(k_bake_bin, dim3(blocks(B.ntri)), dim3(256), 0, st, B, B.cell_count, nullptr);
    BAKE_TRY(hipGetLastError());
    BAKE_TRY(launch_scan(B.cell_count, (uint32_t)ncell, B.cell_start, B.cell_cursor, partials, total, st));
    uint32_t npairs = 0;
    BAKE_TRY(hipMemcpy(&npairs, total, 4, hipMemcpyDeviceToHost));
    BAKE_TRY(mem.get(B.cell_list, npairs));
    if (B.ntri) hipLaunchKernelGGL(k_bake_bin, dim3(blocks(B.ntri)), dim3(256), 0, st, B, B.cell_cursor, B.cell_list);
    hipLaunchKernelGGL(k_bake_rays, dim3(blocks(nvox)), dim3(256), 0, st, B);
    hipLaunchKernelGGL(k_bake_blur, dim3(blocks(nvox)), dim3(256), 0, st, B);
    BAKE_TRY(hipGetLastError());
    uint32_t status = 0;
    BAKE_TRY(hipMemcpy(&status, B.status, 4, hipMemcpyDeviceToHost));
    if (status) {
        fallback = true;    // some ray met more translucent surfaces than it keeps
        return SR_OK;
    }
    light.resize(nvox);
    BAKE_TRY(hipMemcpy(light.data(), B.light, nvox * 4, hipMemcpyDeviceToHost));
    if (active_out) {
        active_out->resize(nvox);
        BAKE_TRY(hipMemcpy(active_out->data(), B.active, nvox, hipMemcpyDeviceToHost));
    }
    return SR_OK;
}

}  // namespace sr
