// gi.cpp -- voxel GI bake (bake_voxel_gi, src/gi.rs:316-412; the --gi path
// of load_scene, src/main.rs:238-282): host preparation shared with the
// device bake (gi.hip), the host bake itself, and the C ABI.
//
// Per active voxel: GI_SAMPLES primary rays over one of GI_SETS Hammersley
// sphere sets; a ray that hits a surface returns its one-bounce radiance
// (sun through a transmittance ray, sky through four hemisphere
// transmittance rays, albedo and emission from the hit's material), a ray
// that escapes returns the skybox; the radiances project onto 4-coefficient
// SH, become irradiance, are clamped and floored; every voxel then takes the
// 3x3x3 mean.  Inactive voxels keep the scene SH (initialize_voxel_gi_from_
// scene with scale 1).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "gi_core.h"
#include "gi_host.h"
#include "prep.h"
#include "runtime.h"
#include "swraster.h"

namespace sr {
namespace {

template <class F> void par_for(size_t n, int threads, F &&fn) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (threads == 1 || n < 2) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= n) break;
                fn(i);
            }
        });
    for (auto &th : pool) th.join();
}

float radical_inverse(uint32_t bits) {
    bits = (bits << 16) | (bits >> 16);
    bits = ((bits & 0x55555555u) << 1) | ((bits & 0xAAAAAAAAu) >> 1);
    bits = ((bits & 0x33333333u) << 2) | ((bits & 0xCCCCCCCCu) >> 2);
    bits = ((bits & 0x0F0F0F0Fu) << 4) | ((bits & 0xF0F0F0F0u) >> 4);
    bits = ((bits & 0x00FF00FFu) << 8) | ((bits & 0xFF00FF00u) >> 8);
    return (float)bits * 2.3283064e-10f;
}

// Median-split BVH over the triangles' boxes (the same construction as the
// sun-visibility bake's).
struct Builder {
    const std::vector<GiTri> &tris;
    std::vector<float> box;          // 6 per triangle
    std::vector<int32_t> &order;
    std::vector<GiNode> &nodes;
    Builder(const std::vector<GiTri> &t, std::vector<int32_t> &o, std::vector<GiNode> &n) : tris(t), order(o), nodes(n) {}
    int build(int first, int count, int depth) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = first; i < first + count; i++) {
            const float *b = &box[6 * (size_t)order[i]];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::fmin(lo[a], b[a]);
                hi[a] = std::fmax(hi[a], b[3 + a]);
                float c = (b[a] + b[3 + a]) * 0.5f;
                clo[a] = std::fmin(clo[a], c);
                chi[a] = std::fmax(chi[a], c);
            }
        }
        int id = (int)nodes.size();
        GiNode nd;
        memcpy(nd.lo, lo, 12);
        memcpy(nd.hi, hi, 12);
        nd.left = nd.right = -1;
        nd.first = first;
        nd.count = count;
        nodes.push_back(nd);
        if (count <= 4 || depth >= GI_STACK - 4) return id;   // depth-first stacks hold depth + 1 entries
        float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
        int axis = ext[0] > ext[1] ? (ext[0] > ext[2] ? 0 : 2) : (ext[1] > ext[2] ? 1 : 2);
        auto key = [&](int t) { return box[6 * (size_t)t + axis] + box[6 * (size_t)t + 3 + axis]; };
        int mid = first + count / 2;
        std::nth_element(order.begin() + first, order.begin() + mid, order.begin() + first + count,
                         [&](int a, int c) { return key(a) < key(c); });
        int l = build(first, mid - first, depth + 1);
        int r = build(mid, first + count - mid, depth + 1);
        nodes[id].left = l;
        nodes[id].right = r;
        nodes[id].count = 0;
        return id;
    }
};

}  // namespace

int gi_prepare(const sr_scene_desc *d, int grid, GiHost &h) {
    if (!d->environment || !d->environment->cubemap || d->environment->cubemap_size == 0)
        return fail(SR_ERR_INVALID_ARGUMENT, "the GI bake needs the scene's environment (skybox and SH)");
    voxel_frame(d, grid, h.frame);
    // RayTracer::new (src/raytracer.rs:76-133): world triangles with face
    // normal and texture coordinates, degenerate ones skipped
    std::vector<WorldTri> wt;
    std::vector<float> uvs;
    collect_world_triangles(d, 0, wt, &uvs);
    h.tris.resize(wt.size());
    for (size_t i = 0; i < wt.size(); i++) {
        GiTri &g = h.tris[i];
        memcpy(g.p, wt[i].p, 36);
        G3 p0 = g3(g.p[0], g.p[1], g.p[2]);
        G3 nn = gnormalize(gcross(gsub(g3(g.p[3], g.p[4], g.p[5]), p0), gsub(g3(g.p[6], g.p[7], g.p[8]), p0)));
        g.n[0] = nn.x;
        g.n[1] = nn.y;
        g.n[2] = nn.z;
        memcpy(g.uv, &uvs[6 * i], 24);
        g.material = wt[i].material;
    }
    // BVH over the triangles' boxes widened by 1e-5 (src/raytracer.rs:108-109)
    h.order.resize(h.tris.size());
    h.nodes.clear();
    Builder B(h.tris, h.order, h.nodes);
    B.box.resize(h.tris.size() * 6);
    for (size_t i = 0; i < h.tris.size(); i++) {
        const float *p = h.tris[i].p;
        for (int a = 0; a < 3; a++) {
            B.box[6 * i + a] = std::fmin(std::fmin(p[a], p[3 + a]), p[6 + a]) - 1.0e-5f;
            B.box[6 * i + 3 + a] = std::fmax(std::fmax(p[a], p[3 + a]), p[6 + a]) + 1.0e-5f;
        }
        h.order[i] = (int32_t)i;
    }
    if (!h.tris.empty()) B.build(0, (int)h.tris.size(), 0);
    // materials and their base-colour / emissive textures (base level)
    h.mats.assign(std::max(d->num_materials, 1), GiMat{});
    std::map<int, int> texmap;
    auto tex_of = [&](const sr_texref &r) -> int {
        if (r.texture < 0 || r.texture >= d->num_textures) return -1;
        auto it = texmap.find(r.texture);
        if (it != texmap.end()) return it->second;
        const sr_texture_desc &t = d->textures[r.texture];
        if (!t.texels || t.width == 0 || t.height == 0) return -1;
        GiTex g = {(uint32_t)h.texels.size(), t.width, t.height};
        h.texels.insert(h.texels.end(), t.texels, t.texels + (size_t)t.width * t.height);
        h.texs.push_back(g);
        return texmap[r.texture] = (int)h.texs.size() - 1;
    };
    for (int i = 0; i < d->num_materials; i++) {
        const sr_material_desc &m = d->materials[i];
        GiMat &g = h.mats[i];
        for (int k = 0; k < 3; k++) {
            g.base[k] = m.base_color_factor[k];
            g.emissive[k] = m.emissive_factor[k];
        }
        g.base_tex = tex_of(m.base_color);
        g.emis_tex = tex_of(m.emissive);
        g.base_ws = m.base_color.wrap_s; g.base_wt = m.base_color.wrap_t;
        g.emis_ws = m.emissive.wrap_s; g.emis_wt = m.emissive.wrap_t;
        g.translucent = m.is_translucent != 0;
        g.transmission = m.transmission;
    }
    if (d->num_materials == 0) {
        GiMat &g = h.mats[0];
        g.base[0] = g.base[1] = g.base[2] = 1.0f;
        g.base_tex = g.emis_tex = -1;
    }
    // build_direction_sets (src/gi.rs:415-445): Hammersley points with a
    // Cranley-Patterson rotation per set, mapped to the sphere
    const float PI_F = 3.14159265358979323846f;
    h.dirs.resize((size_t)GI_SETS * GI_SAMPLES * 3);
    for (int s = 0; s < GI_SETS; s++) {
        float ox = radical_inverse((uint32_t)s + 1), oy = ((float)s + 0.5f) / (float)GI_SETS;
        for (int i = 0; i < GI_SAMPLES; i++) {
            float xx = (float)i / (float)GI_SAMPLES + ox, xy = radical_inverse((uint32_t)i) + oy;
            xx = xx - std::trunc(xx);
            xy = xy - std::trunc(xy);
            float z = 1.0f - 2.0f * xx;
            float r = std::sqrt(std::fmax(1.0f - z * z, 0.0f));
            float phi = (2.0f * PI_F) * xy;
            G3 v = gnormalize(g3(r * std::cos(phi), r * std::sin(phi), z));
            float *o = &h.dirs[((size_t)s * GI_SAMPLES + i) * 3];
            o[0] = v.x; o[1] = v.y; o[2] = v.z;
        }
    }
    const sr_environment &env = *d->environment;
    h.cube_size = env.cubemap_size;
    h.cubemap = env.cubemap;
    memcpy(h.sh, env.irradiance_sh, sizeof(h.sh));
    memcpy(h.light_color, d->light.color, 12);
    // bake constants (src/gi.rs:322-333)
    const float *vs = h.frame.vs;
    float vlen = std::sqrt((vs[0] * vs[0] + vs[1] * vs[1]) + vs[2] * vs[2]);
    float mve = std::fmax(std::fmin(std::fmin(vs[0], vs[1]), vs[2]), 1.0e-6f);
    h.primary_bias = std::fmax(vlen * 0.35f, 1.0e-4f * 4.0f);
    h.params.hit_bias = std::fmax(vlen * 0.05f, 1.0e-4f * 8.0f);
    h.params.tmin = std::fmax(mve * 0.15f, 1.0e-4f);
    return SR_OK;
}

GiScene gi_scene_view(const GiHost &h) {
    GiScene S;
    memset(&S, 0, sizeof(S));
    S.tris = h.tris.data();
    S.nodes = h.nodes.data();
    S.order = h.order.data();
    S.nnodes = (int32_t)h.nodes.size();
    S.mats = h.mats.data();
    S.nmat = (int32_t)h.mats.size();
    S.texs = h.texs.data();
    S.texels = h.texels.data();
    S.cubemap = h.cubemap;
    S.cube_size = h.cube_size;
    memcpy(S.sh, h.sh, sizeof(S.sh));
    memcpy(S.light_dir, h.frame.ldir, 12);
    memcpy(S.light_color, h.light_color, 12);
    return S;
}

// Initial coefficients of every voxel (initialize_voxel_gi_from_scene with
// scale 1, src/gi.rs:123-149).
void gi_initial(const GiHost &h, const std::vector<float> &light, std::vector<float> &sh4) {
    size_t n = light.size();
    sh4.assign(n * 16, 0.0f);
    for (size_t i = 0; i < n; i++) {
        float *o = &sh4[16 * i];
        for (int c = 0; c < 4; c++)
            for (int k = 0; k < 3; k++) o[4 * c + k] = h.sh[c][k] * 1.0f;
        o[3] = light[i];
        o[7] = 1.0f;
    }
}

int gi_bake_host(const GiHost &h, const std::vector<uint8_t> &active, const std::vector<float> &light, int threads,
                 std::vector<float> &out) {
    int G = h.frame.grid;
    size_t nvox = (size_t)G * G * G;
    std::vector<float> coeffs;
    gi_initial(h, light, coeffs);
    out.resize(nvox * 16);
    if (std::find(active.begin(), active.end(), (uint8_t)1) == active.end()) {
        out = coeffs;       // nothing active: the scene SH, not blurred (src/gi.rs:338-341)
        return SR_OK;
    }
    GiScene S = gi_scene_view(h);
    std::atomic<bool> over{false};
    par_for(nvox, threads, [&](size_t i) {
        if (!active[i]) return;
        int x = (int)(i % G), y = (int)((i / G) % G), z = (int)(i / ((size_t)G * G));
        G3 c = g3(h.frame.cmin[0] + (float)x * h.frame.vs[0], h.frame.cmin[1] + (float)y * h.frame.vs[1],
                  h.frame.cmin[2] + (float)z * h.frame.vs[2]);
        int set = ((x % 3) + 3 * (y % 3) + (z % 3)) % GI_SETS;
        const float *dirs = &h.dirs[(size_t)set * GI_SAMPLES * 3];
        float rad[GI_SAMPLES * 3];
        uint32_t sky_hits = 0;
        bool ov = false;
        GiLocalStack st;
        for (int s = 0; s < GI_SAMPLES; s++) {
            bool sky = false;
            G3 r = gi_primary_ray(S, c, g3(dirs[3 * s], dirs[3 * s + 1], dirs[3 * s + 2]), h.primary_bias, h.params, sky,
                                  ov, st);
            rad[3 * s] = r.x; rad[3 * s + 1] = r.y; rad[3 * s + 2] = r.z;
            sky_hits += sky ? 1u : 0u;
        }
        if (ov) over = true;
        gi_finish_voxel(S, dirs, rad, sky_hits, light[i], &coeffs[16 * i]);
    });
    if (over) return fail(SR_ERR_CAPACITY, "a GI ray met more than %d translucent surfaces", GI_HIT_CAP);
    par_for((size_t)G, threads, [&](size_t zz) {
        for (int y = 0; y < G; y++)
            for (int x = 0; x < G; x++)
                gi_blur_voxel(coeffs.data(), G, G, G, x, y, (int)zz, &out[16 * (((size_t)zz * G + y) * G + x)]);
    });
    return SR_OK;
}

}  // namespace sr

using namespace sr;

extern "C" {

SR_API int sr_prepare_gi(const sr_scene_desc *scene, int32_t grid, int32_t device, int32_t threads,
                         sr_voxel_grid *out) {
    if (!scene || !out || grid <= 0 || grid > 512) return fail(SR_ERR_INVALID_ARGUMENT, "bad GI arguments");
    if (device >= 0) {
        int32_t ndev = 0;
        sr_device_count(&ndev);
        if (device >= ndev) return fail(SR_ERR_INVALID_ARGUMENT, "no HIP device %d (have %d)", device, ndev);
    }
    memset(out, 0, sizeof(*out));
    int rc;
    // sun visibility and the active mask first (as the reference, src/main.rs:240-246)
    std::vector<float> light;
    std::vector<uint8_t> active;
    VoxelFrame fr;
    if (device >= 0) {
        bool fallback = false;
        rc = bake_voxels_device(scene, grid, device, light, fr, fallback, &active);
        if (!rc && fallback) rc = build_voxel_grid(scene, grid, threads, light, fr.wmin, fr.wmax, &active);
    } else {
        rc = build_voxel_grid(scene, grid, threads, light, fr.wmin, fr.wmax, &active);
    }
    if (rc) return rc;
    GiHost h;
    if ((rc = gi_prepare(scene, grid, h))) return rc;
    std::vector<float> sh4;
    rc = device >= 0 ? gi_bake_device(h, active, light, device, sh4) : gi_bake_host(h, active, light, threads, sh4);
    if (rc) return rc;
    size_t nvox = (size_t)grid * grid * grid;
    float *lp = (float *)malloc(nvox * 4);
    float *sp = (float *)malloc(nvox * 64);
    if (!lp || !sp) {
        free(lp);
        free(sp);
        return fail(SR_ERR_OUT_OF_MEMORY, "GI grid");
    }
    memcpy(lp, light.data(), nvox * 4);
    memcpy(sp, sh4.data(), nvox * 64);
    out->dims[0] = out->dims[1] = out->dims[2] = grid;
    memcpy(out->world_min, h.frame.wmin, 12);
    memcpy(out->world_max, h.frame.wmax, 12);
    out->light_intensity = lp;
    out->sh_scale = 1.0f;
    out->sky_visibility = 1.0f;
    out->sh4 = sp;
    return SR_OK;
}

}  // extern "C"
