// scene.cpp -- scene upload to HBM (sr_scene_create).
//
// Mirrors what Scene::from_gltf and load_scene leave in memory for the
// per-frame path (src/scene.rs:145-354, src/main.rs:100-283): vertex streams,
// materials, texture mip chains, image-based lighting, voxel grid, light, and
// the node / mesh / primitive hierarchy with bounding spheres.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "prep.h"
#include "runtime.h"
#include "swraster.h"

using namespace sr;

namespace {

// compute_bounding_sphere (src/scene.rs:498-511); glam Vec3A min/max are SSE.
Sphere bounding_sphere(const float *pos, uint32_t n) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = 0; v < n; v++)
        for (int k = 0; k < 3; k++) {
            float p = pos[3 * v + k];
            mn[k] = mn[k] < p ? mn[k] : p;
            mx[k] = mx[k] > p ? mx[k] : p;
        }
    Sphere s;
    for (int k = 0; k < 3; k++) s.c[k] = (mn[k] + mx[k]) * 0.5f;
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    s.r = std::sqrt((dx * dx + dy * dy) + dz * dz) / 2.0f;
    return s;
}

Sphere xform_sphere(const float *m, const Sphere &s) {
    auto l4 = [](const float *c) {
        float xx = c[0] * c[0], yy = c[1] * c[1], zz = c[2] * c[2], ww = c[3] * c[3];
        return std::sqrt((xx + zz) + (yy + ww));
    };
    float max_scale = ((l4(m) + l4(m + 4)) + l4(m + 8)) / 3.0f;
    float o[3];
    for (int i = 0; i < 3; i++) {
        float acc = m[i] * s.c[0];
        acc = acc + m[4 + i] * s.c[1];
        acc = acc + m[8 + i] * s.c[2];
        acc = acc + m[12 + i] * 1.0f;
        o[i] = acc;
    }
    return {{o[0], o[1], o[2]}, s.r * max_scale};
}

void grow(Sphere &a, const Sphere &b) {
    float dx = a.c[0] - b.c[0], dy = a.c[1] - b.c[1], dz = a.c[2] - b.c[2];
    float d = std::sqrt((dx * dx + dy * dy) + dz * dz);
    if (d + b.r > a.r) a.r = d + b.r;
}

int upload(sr_scene *s, const void *src, size_t bytes, void **dst) {
    DevBuf b;
    int r = b.ensure(std::max<size_t>(bytes, 4));
    if (r) return r;
    if (bytes) {
        hipError_t e = hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            b.release();
            return fail(SR_ERR_HIP, "scene upload: %s", hipGetErrorString(e));
        }
    }
    s->device_bytes += b.bytes;
    s->buffers.push_back(b);
    *dst = b.ptr;
    return SR_OK;
}

DevTexRef texref(const sr_texref &r) {
    DevTexRef o;
    o.tex = r.texture;
    o.wrap_s = r.wrap_s;
    o.wrap_t = r.wrap_t;
    return o;
}

}  // namespace

extern "C" {

SR_API int sr_scene_create(const sr_scene_desc *d, int32_t device, sr_scene **out) {
    if (!d || !out) return fail(SR_ERR_INVALID_ARGUMENT, "null scene descriptor");
    if (!d->environment || !d->voxels) return fail(SR_ERR_INVALID_ARGUMENT, "scene needs an environment and a voxel grid");
    if (d->num_materials <= 0) return fail(SR_ERR_INVALID_ARGUMENT, "scene needs at least one material");
    const sr_environment &env = *d->environment;
    const sr_voxel_grid &vg = *d->voxels;
    if (!env.cubemap || !env.specular || !env.brdf_lut || env.specular_mips == 0 || env.specular_mips > 16)
        return fail(SR_ERR_INVALID_ARGUMENT, "incomplete environment");
    if (vg.dims[0] <= 0 || vg.dims[1] <= 0 || vg.dims[2] <= 0 || (!vg.light_intensity && !vg.sh4))
        return fail(SR_ERR_INVALID_ARGUMENT, "incomplete voxel grid");
    for (int i = 0; i < d->num_primitives; i++) {
        const sr_primitive_desc &p = d->primitives[i];
        if (!p.positions || !p.normals || (!p.indices && p.num_indices))
            return fail(SR_ERR_INVALID_ARGUMENT, "primitive %d is missing positions, normals or indices", i);
        if (p.material < 0 || p.material >= d->num_materials)
            return fail(SR_ERR_INVALID_ARGUMENT, "primitive %d has material %d out of range", i, p.material);
        for (uint32_t k = 0; k < p.num_indices; k++)
            if (p.indices[k] >= p.num_vertices)
                return fail(SR_ERR_INVALID_ARGUMENT, "primitive %d index %u out of range", i, p.indices[k]);
    }
    for (int i = 0; i < d->num_meshes; i++) {
        const sr_mesh_desc &m = d->meshes[i];
        if (m.first_primitive < 0 || m.num_primitives < 0 || m.first_primitive + m.num_primitives > d->num_primitives)
            return fail(SR_ERR_INVALID_ARGUMENT, "mesh %d primitive range out of bounds", i);
    }
    for (int i = 0; i < d->num_nodes; i++)
        if (d->nodes[i].mesh >= d->num_meshes) return fail(SR_ERR_INVALID_ARGUMENT, "node %d mesh out of range", i);
    for (int i = 0; i < d->num_materials; i++) {
        const sr_texref *refs[6] = {&d->materials[i].base_color, &d->materials[i].metallic_roughness,
                                    &d->materials[i].normal, &d->materials[i].emissive,
                                    &d->materials[i].occlusion, &d->materials[i].transmission_texture};
        for (auto *r : refs)
            if (r->texture >= d->num_textures) return fail(SR_ERR_INVALID_ARGUMENT, "material %d texture out of range", i);
    }
    int32_t ndev = 0;
    sr_device_count(&ndev);
    if (device < 0 || device >= ndev) return fail(SR_ERR_INVALID_ARGUMENT, "no HIP device %d (have %d)", device, ndev);
    hipError_t he = hipSetDevice(device);
    if (he != hipSuccess) return fail(SR_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(he));

    sr_scene *s = new sr_scene();
    s->device = device;
    int r = SR_OK;
    auto bail = [&](int code) {
        for (auto &b : s->buffers) b.release();
        s->dscene.release();
        delete s;
        return code;
    };

    // materials
    std::vector<DevMaterial> mats(d->num_materials);
    std::vector<bool> mat_normal(d->num_materials), mat_translucent(d->num_materials);
    for (int i = 0; i < d->num_materials; i++) {
        const sr_material_desc &m = d->materials[i];
        DevMaterial &o = mats[i];
        memset(&o, 0, sizeof(o));
        memcpy(o.base, m.base_color_factor, 16);
        o.metallic = m.metallic_factor;
        o.roughness = m.roughness_factor;
        memcpy(o.emissive, m.emissive_factor, 12);
        o.occlusion_strength = m.occlusion_strength;
        o.alpha_cutoff = m.alpha_cutoff;
        o.transmission = m.transmission;
        o.flags = 0;
        if (m.alpha_mode == SR_ALPHA_MASK && m.base_color.texture >= 0) o.flags |= SR_MAT_ALPHA_TESTED;
        if (m.is_translucent) o.flags |= SR_MAT_TRANSLUCENT;
        if (m.normal.texture >= 0) o.flags |= SR_MAT_NORMAL_MAP;
        o.base_tex = texref(m.base_color);
        o.mr_tex = texref(m.metallic_roughness);
        o.normal_tex = texref(m.normal);
        o.emissive_tex = texref(m.emissive);
        o.occlusion_tex = texref(m.occlusion);
        o.transmission_tex = texref(m.transmission_texture);
        mat_normal[i] = m.normal.texture >= 0;
        mat_translucent[i] = m.is_translucent != 0;
    }

    // vertex streams (SoA) and indices
    uint64_t nv = 0, ni = 0;
    for (int i = 0; i < d->num_primitives; i++) {
        nv += d->primitives[i].num_vertices;
        ni += d->primitives[i].num_indices;
    }
    if (nv > 0xFFFFFFF0ull || ni > 0xFFFFFFF0ull) return bail(fail(SR_ERR_UNSUPPORTED, "scene exceeds 2^32 vertices"));
    std::vector<float> px(nv), py(nv), pz(nv), nx(nv), ny(nv), nz(nv), tu(nv), tv(nv);
    std::vector<uint32_t> idx(ni);
    std::vector<float> tan;
    bool any_tan = false;
    for (int i = 0; i < d->num_primitives; i++)
        if (mat_normal[d->primitives[i].material]) any_tan = true;
    if (any_tan) tan.assign(nv * 4, 0.0f);
    s->prims.resize(d->num_primitives);
    uint32_t vb = 0, ib = 0;
    for (int i = 0; i < d->num_primitives; i++) {
        const sr_primitive_desc &p = d->primitives[i];
        HostPrim &h = s->prims[i];
        h.vtx_base = vb;
        h.num_vtx = p.num_vertices;
        h.idx_base = ib;
        h.num_tri = p.num_indices / 3;
        h.material = p.material;
        h.tangents = mat_normal[p.material];
        h.bs = bounding_sphere(p.positions, p.num_vertices);
        for (uint32_t v = 0; v < p.num_vertices; v++) {
            px[vb + v] = p.positions[3 * v];
            py[vb + v] = p.positions[3 * v + 1];
            pz[vb + v] = p.positions[3 * v + 2];
            nx[vb + v] = p.normals[3 * v];
            ny[vb + v] = p.normals[3 * v + 1];
            nz[vb + v] = p.normals[3 * v + 2];
            tu[vb + v] = p.texcoords ? p.texcoords[2 * v] : 0.0f;
            tv[vb + v] = p.texcoords ? p.texcoords[2 * v + 1] : 0.0f;
        }
        if (h.tangents) {
            std::vector<float> t;
            if (p.tangents) t.assign(p.tangents, p.tangents + 4 * (size_t)p.num_vertices);
            else compute_tangents(p, t);
            memcpy(&tan[4 * (size_t)vb], t.data(), t.size() * 4);
        }
        memcpy(&idx[ib], p.indices, (size_t)p.num_indices * 4);
        vb += p.num_vertices;
        ib += p.num_indices;
        s->triangles += h.num_tri;
    }
    s->vertices = nv;
    s->any_tangents = any_tan;

    // textures: mip chains (built on the host's cores, one texture per
    // task) into one pool
    std::vector<DevTexture> texs(d->num_textures);
    std::vector<uint32_t> pool;
    for (int i = 0; i < d->num_textures; i++) {
        const sr_texture_desc &t = d->textures[i];
        if (!t.texels || t.width == 0 || t.height == 0) return bail(fail(SR_ERR_INVALID_ARGUMENT, "texture %d empty", i));
    }
    struct Chain { std::vector<uint32_t> data, off, w, h; };
    std::vector<Chain> chains(d->num_textures);
    parallel_for((size_t)d->num_textures, std::min(16, (int)std::max(1u, std::thread::hardware_concurrency())),
                 [&](size_t i) {
                     const sr_texture_desc &t = d->textures[i];
                     Chain &c = chains[i];
                     build_mip_chain(t.width, t.height, t.type, t.texels, c.data, c.off, c.w, c.h);
                 });
    size_t pool_size = 0;
    for (const Chain &c : chains)
        for (size_t k = 0; k < c.w.size(); k++) pool_size += (size_t)((c.w[k] + 3) / 4) * ((c.h[k] + 3) / 4) * 16;
    pool.reserve(pool_size);
    for (int i = 0; i < d->num_textures; i++) {
        const sr_texture_desc &t = d->textures[i];
        std::vector<uint32_t> &data = chains[i].data, &off = chains[i].off, &w = chains[i].w, &h = chains[i].h;
        if (off.size() > SR_MAX_MIPS) return bail(fail(SR_ERR_UNSUPPORTED, "texture %d has too many mips", i));
        DevTexture &o = texs[i];
        memset(&o, 0, sizeof(o));
        o.base = pool.size();
        o.width = t.width;
        o.height = t.height;
        o.whv[0] = (float)t.width; o.whv[1] = (float)t.width;
        o.whv[2] = (float)t.height; o.whv[3] = (float)t.height;
        o.num_mips = (uint32_t)off.size();
        o.max_mip = o.num_mips - 1;
        // each level stored in 4x4-texel blocks (64 B), blocks row-major
        // over the level padded to multiples of 4: a pixel's nearest texel
        // shares its cache line with the texels of the pixels around it
        uint32_t at = 0;
        for (size_t k = 0; k < off.size(); k++) {
            const uint32_t lw = w[k], lh = h[k], bw = (lw + 3) / 4, bh = (lh + 3) / 4;
            o.mip_off[k] = at;
            o.mip_w[k] = lw;
            o.mip_h[k] = lh;
            size_t first = pool.size();
            pool.resize(first + (size_t)bw * bh * 16, 0u);
            const uint32_t *src = data.data() + off[k];
            for (uint32_t y = 0; y < lh; y++)
                for (uint32_t x = 0; x < lw; x++)
                    pool[first + (((size_t)(y >> 2) * bw + (x >> 2)) << 4) + ((y & 3) << 2) + (x & 3)] =
                        src[(size_t)y * lw + x];
            at += bw * bh * 16;
        }
        std::vector<uint32_t>().swap(data);
    }

    DevScene ds;
    memset(&ds, 0, sizeof(ds));
    void *p = nullptr;
#define UP(vec, field, type)                                                           \
    if ((r = upload(s, (vec).data(), (vec).size() * sizeof((vec)[0]), &p))) return bail(r); \
    ds.field = (type)p;
    UP(px, px, const float *); UP(py, py, const float *); UP(pz, pz, const float *);
    UP(nx, nx, const float *); UP(ny, ny, const float *); UP(nz, nz, const float *);
    UP(tu, tu, const float *); UP(tv, tv, const float *);
    UP(idx, indices, const uint32_t *);
    UP(mats, materials, const DevMaterial *);
    if (!texs.empty()) { UP(texs, textures, const DevTexture *); }
    if (!pool.empty()) { UP(pool, texels, const uint32_t *); }
    if (any_tan) { UP(tan, tangent, const float *); }
#undef UP
    size_t face = (size_t)env.cubemap_size * env.cubemap_size;
    if ((r = upload(s, env.cubemap, face * 6 * 4, &p))) return bail(r);
    ds.cubemap = (const uint32_t *)p;
    ds.cube_size = env.cubemap_size;
    size_t sface = (size_t)env.specular_size * env.specular_size;
    if ((r = upload(s, env.specular, sface * 6 * env.specular_mips * 4, &p))) return bail(r);
    ds.specular = (const uint32_t *)p;
    ds.spec_size = env.specular_size;
    ds.spec_mips = env.specular_mips;
    ds.spec_max_mip_f = (float)(env.specular_mips - 1);
    if ((r = upload(s, env.brdf_lut, (size_t)env.brdf_lut_size * env.brdf_lut_size * 4, &p))) return bail(r);
    ds.lut = (const uint32_t *)p;
    ds.lut_size = env.brdf_lut_size;
    size_t nvox = (size_t)vg.dims[0] * vg.dims[1] * vg.dims[2];
    if (vg.sh4) {
        if ((r = upload(s, vg.sh4, nvox * 16 * 4, &p))) return bail(r);
        ds.vox_sh4 = (const float *)p;
    } else {
        // the eight trilinear corners of every cell, contiguous and already
        // clamped at the far faces as the lookup clamps them
        // (src/voxelgrid.rs:275-300): one 32-byte read per lookup
        const int W = vg.dims[0], H = vg.dims[1], D = vg.dims[2];
        std::vector<float> c8(nvox * 8);
        const float *L = vg.light_intensity;
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    int x1 = std::min(x + 1, W - 1), y1 = std::min(y + 1, H - 1), z1 = std::min(z + 1, D - 1);
                    auto at = [&](int xx, int yy, int zz) { return L[((size_t)zz * H + yy) * W + xx]; };
                    float *o = &c8[(((size_t)z * H + y) * W + x) * 8];
                    o[0] = at(x, y, z);   o[1] = at(x, y, z1);  o[2] = at(x, y1, z);  o[3] = at(x, y1, z1);
                    o[4] = at(x1, y, z);  o[5] = at(x1, y, z1); o[6] = at(x1, y1, z); o[7] = at(x1, y1, z1);
                }
        if ((r = upload(s, c8.data(), nvox * 8 * 4, &p))) return bail(r);
        ds.vox_light = (const float *)p;
    }
    for (int k = 0; k < 3; k++) {
        ds.vox_dims[k] = vg.dims[k];
        ds.vox_min[k] = vg.world_min[k];
        float ws = vg.world_max[k] - vg.world_min[k];
        ds.vox_size[k] = ws / (float)vg.dims[k];
    }
    for (int c = 0; c < 4; c++) {
        for (int k = 0; k < 3; k++) ds.sh_const[c][k] = env.irradiance_sh[c][k] * vg.sh_scale;
        ds.sh_const[c][3] = 0.0f;
    }
    ds.sh_const[1][3] = vg.sky_visibility;
    memcpy(ds.light_dir, d->light.direction, 12);
    memcpy(ds.light_color, d->light.color, 12);
    if ((r = s->dscene.ensure(sizeof(DevScene)))) return bail(r);
    if (hipMemcpy(s->dscene.ptr, &ds, sizeof(ds), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SR_ERR_HIP, "scene descriptor upload failed"));
    s->device_bytes += s->dscene.bytes;

    // hierarchy
    s->meshes.resize(d->num_meshes);
    for (int i = 0; i < d->num_meshes; i++) {
        HostMesh &m = s->meshes[i];
        for (int k = 0; k < d->meshes[i].num_primitives; k++) {
            int pi = d->meshes[i].first_primitive + k;
            m.prims.push_back(pi);
            if (mat_translucent[d->primitives[pi].material]) m.translucent.push_back(k);
            else m.opaque.push_back(k);
        }
    }
    s->nodes.resize(d->num_nodes);
    for (int i = 0; i < d->num_nodes; i++) {
        HostNode &n = s->nodes[i];
        memcpy(n.transform, d->nodes[i].transform, 64);
        n.mesh = d->nodes[i].mesh;
        for (int k = 0; k < 3; k++) {
            float acc = n.transform[k] * 0.0f;
            acc = acc + n.transform[4 + k] * 0.0f;
            acc = acc + n.transform[8 + k] * 0.0f;
            acc = acc + n.transform[12 + k] * 1.0f;
            n.bs_world.c[k] = acc;
        }
        n.bs_world.r = 0.0f;
        if (n.mesh >= 0)
            for (int pi : s->meshes[n.mesh].prims) grow(n.bs_world, xform_sphere(n.transform, s->prims[pi].bs));
        if (n.mesh >= 0 && !s->meshes[n.mesh].translucent.empty()) s->any_translucent = true;
    }
    *out = s;
    return SR_OK;
}

SR_API int sr_scene_destroy(sr_scene *s) {
    if (!s) return SR_OK;
    if (s->pending_frames.load() > 0)
        return fail(SR_ERR_BUSY, "a renderer's queued frame still uses this scene (read it or synchronize first)");
    hipSetDevice(s->device);
    hipDeviceSynchronize();
    for (auto &b : s->buffers) b.release();
    s->dscene.release();
    delete s;
    return SR_OK;
}

SR_API int sr_scene_info(const sr_scene *s, uint64_t *triangles, uint64_t *vertices, uint64_t *device_bytes) {
    if (!s) return fail(SR_ERR_INVALID_ARGUMENT, "null scene");
    if (triangles) *triangles = s->triangles;
    if (vertices) *vertices = s->vertices;
    if (device_bytes) *device_bytes = s->device_bytes;
    return SR_OK;
}

}  // extern "C"
