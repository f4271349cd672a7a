// gi_core.h -- the voxel GI bake's per-ray and per-voxel operations
// (bake_voxel_gi, src/gi.rs:316-412; estimate_surface_bounce_radiance and
// helpers, src/gi.rs:450-665; RayTracer, src/raytracer.rs:135-260;
// blur_gi_sh4, src/voxelgrid.rs:422-472), compiled for the host bake
// (gi.cpp, g++) and the device bake (gi.hip) from this one source so both
// evaluate the same operations in the same order.
//
// Defined where the reference depends on its BVH crate's traversal order:
// the nearest hit is the smallest distance, ties to the lowest triangle
// index; transmittance multiplies translucent hits by distance, ties by
// triangle index.  The sRGB decode's power is computed in double and rounded
// once (correctly rounded over its input range).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define GI_HD __host__ __device__ inline
#else
#define GI_HD inline
#endif

namespace sr {

struct GiTri {
    float p[9];          // world-space vertices
    float n[3];          // normalized face normal
    float uv[6];
    int32_t material;
};
struct GiNode {
    float lo[3], hi[3];
    int32_t left, right, first, count;   // leaf: left < 0, order[first .. first+count)
};
struct GiMat {
    float base[3];
    float emissive[3];
    int32_t base_tex, emis_tex;          // -1 = none
    int32_t base_ws, base_wt, emis_ws, emis_wt;
    int32_t translucent;
    float transmission;
};
struct GiTex {
    uint32_t offset, width, height;      // base level in the texel pool
};
struct GiScene {
    const GiTri *tris;
    const GiNode *nodes;
    const int32_t *order;
    int32_t nnodes;
    const GiMat *mats;
    int32_t nmat;
    const GiTex *texs;
    const uint32_t *texels;
    const uint32_t *cubemap;             // 6 faces of cube_size^2, mip 0
    uint32_t cube_size;
    float sh[4][3];                      // scene irradiance SH
    float light_dir[3], light_color[3];
};

#define GI_SAMPLES 64
#define GI_SETS 9
#define GI_STACK 48          // traversal stack entries; the BVH is at most 44 levels deep
#define GI_HIT_CAP 32

// Traversal stack in a private array (host).  The device keeps its stacks
// in LDS with the same interface and capacity, so both walk identically.
struct GiLocalStack {
    int32_t s[GI_STACK];
    int n = 0;
    GI_HD void push(int32_t v) { s[n++] = v; }
    GI_HD int32_t pop() { return s[--n]; }
};

struct G3 { float x, y, z; };
GI_HD G3 g3(float x, float y, float z) { return {x, y, z}; }
GI_HD G3 gadd(G3 a, G3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
GI_HD G3 gsub(G3 a, G3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
GI_HD G3 gmul(G3 a, G3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
GI_HD G3 gscale(G3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
GI_HD float gdot(G3 a, G3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
GI_HD G3 gcross(G3 a, G3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
GI_HD G3 gnormalize(G3 a) { return gscale(a, 1.0f / sqrtf(gdot(a, a))); }
GI_HD float gmaxf(float a, float b) { return a > b ? a : b; }   // SSE max: NaN in a gives b
GI_HD float gminf(float a, float b) { return a < b ? a : b; }
GI_HD G3 gmax0(G3 a) { return {gmaxf(a.x, 0.0f), gmaxf(a.y, 0.0f), gmaxf(a.z, 0.0f)}; }

GI_HD float gi_srgb_to_linear(float s) {
    if (s <= 0.04045f) return s / 12.92f;
    return (float)pow((double)((s + 0.055f) / 1.055f), (double)2.4f);
}
GI_HD float gi_unpack(uint32_t c, int shift) { return (float)((c >> shift) & 0xFFu) / 255.0f; }
GI_HD uint32_t gi_sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

// ray_triangle_intersect (src/raytracer.rs:225-260)
GI_HD bool gi_ray_tri(G3 o, G3 d, const GiTri &t, float tmin, float tmax, float &tt, float &u, float &v) {
    G3 p0 = g3(t.p[0], t.p[1], t.p[2]);
    G3 e1 = gsub(g3(t.p[3], t.p[4], t.p[5]), p0), e2 = gsub(g3(t.p[6], t.p[7], t.p[8]), p0);
    G3 pv = gcross(d, e2);
    float det = gdot(e1, pv);
    if (fabsf(det) <= 1.0e-8f) return false;
    float inv = 1.0f / det;
    G3 tv = gsub(o, p0);
    u = gdot(tv, pv) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    G3 qv = gcross(tv, e1);
    v = gdot(d, qv) * inv;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    tt = gdot(e2, qv) * inv;
    if (tt < tmin || tt > tmax) return false;
    return true;
}

// Slab test of a node box for t in [0, tmax], with a small tolerance.
GI_HD bool gi_hit_box(const GiNode &n, G3 o, G3 inv, float tmax) {
    float t0 = 0.0f, t1 = tmax;
    const float op[3] = {o.x, o.y, o.z}, ip[3] = {inv.x, inv.y, inv.z};
    for (int a = 0; a < 3; a++) {
        float ta = (n.lo[a] - op[a]) * ip[a], tb = (n.hi[a] - op[a]) * ip[a];
        if (ta > tb) {
            float s = ta;
            ta = tb;
            tb = s;
        }
        if (ta != ta) ta = -INFINITY;   // 0 * inf
        if (tb != tb) tb = INFINITY;
        t0 = t0 < ta ? ta : t0;
        t1 = tb < t1 ? tb : t1;
        if (t0 > t1 * 1.0000001f + 1e-7f) return false;
    }
    return true;
}

struct GiHit {
    int32_t tri;
    float t, u, v;
};

// RayTracer::trace_nearest (src/raytracer.rs:135-172): depth-first, right
// child first; boxes beyond the current nearest distance are skipped.
template <class Stack>
GI_HD bool gi_trace_nearest(const GiScene &S, G3 o, G3 d, float tmin, GiHit &hit, Stack &st) {
    if (S.nnodes == 0) return false;
    G3 inv = g3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    st.n = 0;
    st.push(0);
    float best = INFINITY;
    int32_t best_tri = -1;
    float bu = 0.0f, bv = 0.0f;
    while (st.n > 0) {
        const GiNode &n = S.nodes[st.pop()];
        if (!gi_hit_box(n, o, inv, best)) continue;
        if (n.left < 0) {
            for (int32_t k = n.first; k < n.first + n.count; k++) {
                int32_t ti = S.order[k];
                float t, u, v;
                if (!gi_ray_tri(o, d, S.tris[ti], tmin, best, t, u, v)) continue;
                if (t < best || (t == best && ti < best_tri)) {
                    best = t;
                    best_tri = ti;
                    bu = u;
                    bv = v;
                }
            }
        } else if (st.n + 2 <= GI_STACK) {
            st.push(n.left);
            st.push(n.right);
        }
    }
    if (best_tri < 0) return false;
    hit.tri = best_tri;
    hit.t = best;
    hit.u = bu;
    hit.v = bv;
    return true;
}

// RayTracer::trace_transmittance (src/raytracer.rs:174-207).  over is set if
// more than GI_HIT_CAP translucent surfaces lie on the ray.
template <class Stack>
GI_HD float gi_trace_transmittance(const GiScene &S, G3 o, G3 d, float tmin, bool &over, Stack &st) {
    if (S.nnodes == 0) return 1.0f;
    G3 inv = g3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    st.n = 0;
    st.push(0);
    float ht[GI_HIT_CAP];
    int32_t hi[GI_HIT_CAP];
    int nh = 0;
    while (st.n > 0) {
        const GiNode &n = S.nodes[st.pop()];
        if (!gi_hit_box(n, o, inv, INFINITY)) continue;
        if (n.left < 0) {
            for (int32_t k = n.first; k < n.first + n.count; k++) {
                int32_t ti = S.order[k];
                float t, u, v;
                if (!gi_ray_tri(o, d, S.tris[ti], tmin, INFINITY, t, u, v)) continue;
                int32_t m = S.tris[ti].material;
                m = (m >= 0 && m < S.nmat) ? m : 0;
                if (!S.mats[m].translucent) return 0.0f;     // any opaque hit
                if (nh == GI_HIT_CAP) {
                    over = true;
                    continue;
                }
                int j = nh++;
                while (j > 0 && (ht[j - 1] > t || (ht[j - 1] == t && hi[j - 1] > ti))) {
                    ht[j] = ht[j - 1];
                    hi[j] = hi[j - 1];
                    j--;
                }
                ht[j] = t;
                hi[j] = ti;
            }
        } else if (st.n + 2 <= GI_STACK) {
            st.push(n.left);
            st.push(n.right);
        }
    }
    float tr = 1.0f;
    for (int j = 0; j < nh; j++) {
        int32_t m = S.tris[hi[j]].material;
        m = (m >= 0 && m < S.nmat) ? m : 0;
        tr *= S.mats[m].transmission;
        if (tr <= 0.0001f) return 0.0f;
    }
    return tr;
}

// Texture::apply_wrap_mode (src/texture.rs:578-589)
GI_HD float gi_wrap(float texel, float dim, int mode) {
    float t = texel;
    if (mode == 0) {              // repeat
        t = texel - floorf(texel / dim) * dim;
    } else if (mode == 1) {       // mirrored repeat
        float two = dim * 2.0f;
        float m = texel - floorf(texel / two) * two;
        t = gminf(m, two - m);
    }
    return gminf(t, dim - 1.0f);
}

// sample_bilinear_rgb at mip 0 (src/texture.rs:730-790), lane 0, then the
// sRGB decode per channel.
GI_HD G3 gi_sample_bilinear_srgb(const GiScene &S, int32_t tex, int ws, int wt, float u, float v) {
    const GiTex &T = S.texs[tex];
    float wf = (float)T.width, hf = (float)T.height;
    float xf = u * wf - 0.5f, yf = v * hf - 0.5f;
    float x0 = floorf(xf), y0 = floorf(yf);
    float x1 = x0 + 1.0f, y1 = y0 + 1.0f;
    float fx = xf - x0, fy = yf - y0;
    float omfx = 1.0f - fx, omfy = 1.0f - fy;
    uint32_t x0i = gi_sat_u32(gi_wrap(x0, wf, ws)), y0i = gi_sat_u32(gi_wrap(y0, hf, wt));
    uint32_t x1i = gi_sat_u32(gi_wrap(x1, wf, ws)), y1i = gi_sat_u32(gi_wrap(y1, hf, wt));
    const uint32_t *px = S.texels + T.offset;
    uint32_t c00 = px[y0i * T.width + x0i], c10 = px[y0i * T.width + x1i];
    uint32_t c01 = px[y1i * T.width + x0i], c11 = px[y1i * T.width + x1i];
    float w00 = omfx * omfy, w10 = fx * omfy, w01 = omfx * fy, w11 = fx * fy;
    float c[3];
    for (int k = 0; k < 3; k++) {
        int sh = 24 - 8 * k;
        c[k] = ((gi_unpack(c00, sh) * w00 + gi_unpack(c10, sh) * w10) + gi_unpack(c01, sh) * w01) +
               gi_unpack(c11, sh) * w11;
    }
    return g3(gi_srgb_to_linear(c[0]), gi_srgb_to_linear(c[1]), gi_srgb_to_linear(c[2]));
}

// sample_sky_radiance (src/gi.rs:584-597): nearest skybox texel, decoded.
GI_HD G3 gi_sky_radiance(const GiScene &S, G3 n) {
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    bool mx = (ax >= ay) && (ax >= az);
    bool my = (ay > ax) && (ay >= az);
    float sx = n.x >= 0.0f ? 1.0f : -1.0f, sy = n.y >= 0.0f ? 1.0f : -1.0f, sz = n.z >= 0.0f ? 1.0f : -1.0f;
    float u = mx ? -n.z * sx : (my ? n.x : n.x * sz);
    float v = mx ? -n.y : (my ? n.z * sy : -n.y);
    float den = mx ? ax : (my ? ay : az);
    den = gmaxf(den, 1.0e-19f);
    float uf = 0.5f * (u / den + 1.0f), vf = 0.5f * (v / den + 1.0f);
    float slicef = mx ? (n.x >= 0.0f ? 0.0f : 1.0f) : (my ? (n.y >= 0.0f ? 2.0f : 3.0f) : (n.z >= 0.0f ? 4.0f : 5.0f));
    uint32_t slice = gi_sat_u32(slicef);
    float wf = (float)S.cube_size;
    float x = roundf(uf * (wf - 1.0f)), y = roundf(vf * (wf - 1.0f));
    uint32_t xi = gi_sat_u32(gminf(gmaxf(x, 0.0f), wf - 1.0f)), yi = gi_sat_u32(gminf(gmaxf(y, 0.0f), wf - 1.0f));
    uint32_t c = S.cubemap[(size_t)slice * S.cube_size * S.cube_size + (size_t)yi * S.cube_size + xi];
    return g3(gi_srgb_to_linear(gi_unpack(c, 24)), gi_srgb_to_linear(gi_unpack(c, 16)), gi_srgb_to_linear(gi_unpack(c, 8)));
}

struct GiRayParams {
    float hit_bias, tmin;
};

// sky_visibility_hemisphere4 (src/gi.rs:549-570)
template <class Stack>
GI_HD float gi_sky_visibility(const GiScene &S, G3 origin, G3 normal, float dir_bias, float tmin, bool &over,
                              Stack &st) {
    G3 up = fabsf(normal.z) < 0.999f ? g3(0.0f, 0.0f, 1.0f) : g3(1.0f, 0.0f, 0.0f);
    G3 tangent = gnormalize(gcross(normal, up));
    G3 bitangent = gcross(normal, tangent);
    const float dl[4][3] = {{0.0f, 0.0f, 1.0f}, {0.8660254f, 0.0f, 0.5f}, {-0.4330127f, 0.75f, 0.5f},
                            {-0.4330127f, -0.75f, 0.5f}};
    float vis = 0.0f;
    for (int k = 0; k < 4; k++) {
        G3 dir = gnormalize(gadd(gadd(gscale(tangent, dl[k][0]), gscale(bitangent, dl[k][1])), gscale(normal, dl[k][2])));
        G3 o = gadd(origin, gscale(dir, dir_bias));
        vis += gi_trace_transmittance(S, o, dir, tmin, over, st);
    }
    return vis * 0.25f;
}

// estimate_surface_bounce_radiance (src/gi.rs:450-547)
template <class Stack>
GI_HD G3 gi_bounce(const GiScene &S, const GiHit &hit, G3 o, G3 incoming, const GiRayParams &P, bool &over,
                   Stack &st) {
    const GiTri &tri = S.tris[hit.tri];
    int32_t mi = (tri.material >= 0 && tri.material < S.nmat) ? tri.material : 0;
    const GiMat &M = S.mats[mi];
    G3 normal = g3(tri.n[0], tri.n[1], tri.n[2]);
    if (gdot(normal, gscale(incoming, -1.0f)) < 0.0f) normal = gscale(normal, -1.0f);
    float w = (1.0f - hit.u) - hit.v;
    float uvx = (tri.uv[0] * w + tri.uv[2] * hit.u) + tri.uv[4] * hit.v;
    float uvy = (tri.uv[1] * w + tri.uv[3] * hit.u) + tri.uv[5] * hit.v;
    G3 emissive = g3(M.emissive[0], M.emissive[1], M.emissive[2]);
    if (M.emis_tex >= 0) emissive = gmul(emissive, gi_sample_bilinear_srgb(S, M.emis_tex, M.emis_ws, M.emis_wt, uvx, uvy));
    emissive = gmax0(emissive);
    G3 albedo = g3(M.base[0], M.base[1], M.base[2]);
    if (M.base_tex >= 0) albedo = gmul(albedo, gi_sample_bilinear_srgb(S, M.base_tex, M.base_ws, M.base_wt, uvx, uvy));
    albedo = g3(gminf(gmaxf(albedo.x, 0.0f), 1.0f), gminf(gmaxf(albedo.y, 0.0f), 1.0f), gminf(gmaxf(albedo.z, 0.0f), 1.0f));
    albedo = gscale(albedo, 1.0f);
    if (gdot(albedo, albedo) <= 0.0f && gdot(emissive, emissive) <= 0.0f) return g3(0.0f, 0.0f, 0.0f);
    float dir_bias = P.hit_bias * 0.5f;
    G3 position = gadd(o, gscale(incoming, hit.t));
    G3 shading_origin = gadd(position, gscale(normal, P.hit_bias));
    G3 L = g3(S.light_dir[0], S.light_dir[1], S.light_dir[2]);
    float ndl = gmaxf(gdot(normal, L), 0.0f);
    G3 e_sun = g3(0.0f, 0.0f, 0.0f);
    if (ndl > 0.0f) {
        G3 so = gadd(shading_origin, gscale(L, dir_bias));
        float vis = gi_trace_transmittance(S, so, L, P.tmin, over, st);
        e_sun = gscale(g3(S.light_color[0], S.light_color[1], S.light_color[2]), ndl * vis);
    }
    float sky = gi_sky_visibility(S, shading_origin, normal, dir_bias, P.tmin, over, st);
    G3 irr;
    irr.x = ((S.sh[0][0] + S.sh[1][0] * normal.y) + S.sh[2][0] * normal.z) + S.sh[3][0] * normal.x;
    irr.y = ((S.sh[0][1] + S.sh[1][1] * normal.y) + S.sh[2][1] * normal.z) + S.sh[3][1] * normal.x;
    irr.z = ((S.sh[0][2] + S.sh[1][2] * normal.y) + S.sh[2][2] * normal.z) + S.sh[3][2] * normal.x;
    G3 e_sky = gscale(gmax0(irr), sky);
    G3 e_total = gmax0(gadd(e_sun, e_sky));
    const float PI_F = 3.14159265358979323846f;
    G3 bounce = g3(0.0f, 0.0f, 0.0f);
    if (gdot(e_total, e_total) > 0.0f) bounce = gmul(albedo, g3(e_total.x / PI_F, e_total.y / PI_F, e_total.z / PI_F));
    return gadd(bounce, emissive);
}

// Incident radiance along primary direction dir from the voxel centre
// (one iteration of the ray loop of bake_voxel_gi, src/gi.rs:372-388).
template <class Stack>
GI_HD G3 gi_primary_ray(const GiScene &S, G3 center, G3 dir, float primary_bias, const GiRayParams &P, bool &sky,
                        bool &over, Stack &st) {
    G3 o = gadd(center, gscale(dir, primary_bias));
    GiHit hit;
    if (gi_trace_nearest(S, o, dir, P.tmin, hit, st)) {
        sky = false;
        return gi_bounce(S, hit, o, dir, P, over, st);
    }
    sky = true;
    return gi_sky_radiance(S, dir);
}

// The per-voxel reduction of bake_voxel_gi (src/gi.rs:389-404) over the
// incident radiances of its GI_SAMPLES directions, in order: SH projection,
// irradiance, clamp, minimum brightness floor, sun and sky visibility.
GI_HD void gi_finish_voxel(const GiScene &S, const float *dirs, const float *radiance, uint32_t sky_hits,
                           float sun_visibility, float out[16]) {
    const float PI_F = 3.14159265358979323846f;
    const float weight = (4.0f * PI_F) / (float)GI_SAMPLES;
    G3 sh[4] = {g3(0, 0, 0), g3(0, 0, 0), g3(0, 0, 0), g3(0, 0, 0)};
    for (int s = 0; s < GI_SAMPLES; s++) {
        G3 r = g3(radiance[3 * s], radiance[3 * s + 1], radiance[3 * s + 2]);
        if (gdot(r, r) <= 0.0f) continue;
        G3 d = g3(dirs[3 * s], dirs[3 * s + 1], dirs[3 * s + 2]);
        const float basis[4] = {0.282095f, 0.488603f * d.y, 0.488603f * d.z, 0.488603f * d.x};
        for (int i = 0; i < 4; i++) sh[i] = gadd(sh[i], gscale(r, basis[i] * weight));
    }
    const float k0 = PI_F * 0.282095f, k1 = ((2.0f * PI_F) / 3.0f) * 0.488603f;
    sh[0] = gscale(sh[0], k0);
    for (int i = 1; i < 4; i++) sh[i] = gscale(sh[i], k1);
    for (int i = 0; i < 4; i++) {
        G3 c = gmax0(sh[i]);
        out[4 * i] = c.x;
        out[4 * i + 1] = c.y;
        out[4 * i + 2] = c.z;
        out[4 * i + 3] = 0.0f;
    }
    // apply_min_brightness_floor (src/gi.rs:630-665)
    const float MINB = 0.005f;
    G3 l0 = gscale(g3(S.sh[0][0], S.sh[0][1], S.sh[0][2]), MINB);
    float floor_lum = (0.2126f * l0.x + 0.7152f * l0.y) + 0.0722f * l0.z;
    if (floor_lum > 0.0f) {
        float cur = (0.2126f * out[0] + 0.7152f * out[1]) + 0.0722f * out[2];
        if (cur < floor_lum) {
            float t = (floor_lum - cur) / floor_lum;
            t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
            for (int i = 0; i < 4; i++) {
                G3 c = g3(out[4 * i], out[4 * i + 1], out[4 * i + 2]);
                G3 target = gscale(g3(S.sh[i][0], S.sh[i][1], S.sh[i][2]), MINB);
                G3 b = gadd(gscale(c, 1.0f - t), gscale(target, t));   // glam 0.30 Vec3A::lerp
                out[4 * i] = b.x;
                out[4 * i + 1] = b.y;
                out[4 * i + 2] = b.z;
            }
        }
    }
    out[3] = sun_visibility;
    out[7] = (float)sky_hits / (float)GI_SAMPLES;
}

// blur_gi_sh4 (src/voxelgrid.rs:422-472) for voxel (x, y, z).
GI_HD void gi_blur_voxel(const float *src, int W, int H, int D, int x, int y, int z, float out[16]) {
    float sum[16];
    for (int k = 0; k < 16; k++) sum[k] = 0.0f;
    float count = 0.0f;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int nx = x + dx, ny = y + dy, nz = z + dz;
                if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) continue;
                const float *s = src + 16 * (((size_t)nz * H + ny) * W + nx);
                for (int k = 0; k < 16; k++) sum[k] += s[k];
                count += 1.0f;
            }
    const float *self = src + 16 * (((size_t)z * H + y) * W + x);
    for (int k = 0; k < 16; k++) out[k] = count > 0.0f ? sum[k] / count : self[k];
}

}  // namespace sr
