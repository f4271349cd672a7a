// prep.cpp -- load-time input producers and camera (host C++).
//
// These run once per scene (or per camera update) on the host and produce
// the inputs the per-frame GPU path consumes:
//   - texture mip chains            (src/texture.rs:45-128)
//   - image-based lighting: BRDF LUT, irradiance SH4, GGX-prefiltered
//     specular cubemap              (src/texture.rs:135-420)
//   - voxel sun-visibility grid     (src/gi.rs:123-314, src/raytracer.rs,
//                                    src/voxelgrid.rs:371-419)
//   - scene bounds                  (src/scene.rs:333-351)
//   - RenderCamera                  (src/rendercamera.rs)
// Arithmetic follows the reference expressions in f32 with no contraction;
// transcendental functions come from the host libm.

#include "prep.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace sr {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static inline uint32_t ilog2u(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }
static inline uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 4294967296.0f) return UINT32_MAX;
    return (uint32_t)f;
}
static inline void unpack(uint32_t c, float o[4]) {
    o[0] = (float)((c >> 24) & 0xFF) / 255.0f;
    o[1] = (float)((c >> 16) & 0xFF) / 255.0f;
    o[2] = (float)((c >> 8) & 0xFF) / 255.0f;
    o[3] = (float)(c & 0xFF) / 255.0f;
}
static inline uint32_t pack(const float c[4]) {
    return (sat_u32(c[0] * 255.0f) << 24) | (sat_u32(c[1] * 255.0f) << 16) | (sat_u32(c[2] * 255.0f) << 8) |
           sat_u32(c[3] * 255.0f);
}
static inline float srgb_to_linear(float s) {
    if (s <= 0.04045f) return s / 12.92f;
    return std::pow((s + 0.055f) / 1.055f, 2.4f);
}
static inline float linear_to_srgb(float s) {
    if (s <= 0.0031308f) return s * 12.92f;
    return std::pow(s, 1.0f / 2.4f) * 1.055f - 0.055f;
}

struct F3 { float x, y, z; };
static inline F3 f3(float x, float y, float z) { return {x, y, z}; }
static inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline F3 mulv(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline float len(F3 a) { return std::sqrt(dot(a, a)); }
static inline F3 normalize(F3 a) { float r = 1.0f / len(a); return a * r; }
static inline F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline F3 fmin3(F3 a, F3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
static inline F3 fmax3(F3 a, F3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }

// Mat4 (column-major) * (x, y, z, 1), sequential column sums
static inline void mat_point(const float *m, const float p[3], float out[4]) {
    for (int i = 0; i < 4; i++) {
        float acc = m[i] * p[0];
        acc = acc + m[4 + i] * p[1];
        acc = acc + m[8 + i] * p[2];
        acc = acc + m[12 + i] * 1.0f;
        out[i] = acc;
    }
}

// ---------------------------------------------------------------------------
// Mip chains (src/texture.rs:45-128)
// ---------------------------------------------------------------------------
void build_mip_chain(uint32_t width, uint32_t height, int type, const uint32_t *base, std::vector<uint32_t> &data,
                     std::vector<uint32_t> &offsets, std::vector<uint32_t> &widths, std::vector<uint32_t> &heights) {
    data.assign(base, base + (size_t)width * height);
    offsets = {0};
    widths = {width};
    heights = {height};
    uint32_t num_mips = 1 + ilog2u(std::max(width, height));
    for (uint32_t mip = 1; mip < num_mips; mip++) {
        offsets.push_back((uint32_t)data.size());
        uint32_t mw = std::max(width >> mip, 1u), mh = std::max(height >> mip, 1u);
        uint32_t po = offsets[mip - 1], pw = widths[mip - 1], ph = heights[mip - 1];
        size_t start = data.size();
        data.resize(start + (size_t)mw * mh);
        for (uint32_t y = 0; y < mh; y++)
            for (uint32_t x = 0; x < mw; x++) {
                uint32_t x0 = x * 2, y0 = y * 2;
                uint32_t x1 = std::min(x0 + 1, pw - 1), y1 = std::min(y0 + 1, ph - 1);
                float p[4][4];
                unpack(data[po + y0 * pw + x0], p[0]);
                unpack(data[po + y0 * pw + x1], p[1]);
                unpack(data[po + y1 * pw + x0], p[2]);
                unpack(data[po + y1 * pw + x1], p[3]);
                float a[4];
                if (type == SR_TEX_SRGB) {
                    float l[4][4];
                    for (int k = 0; k < 4; k++) {
                        for (int c = 0; c < 3; c++) l[k][c] = srgb_to_linear(p[k][c]);
                        l[k][3] = p[k][3];
                    }
                    for (int c = 0; c < 4; c++) a[c] = (((l[0][c] + l[1][c]) + l[2][c]) + l[3][c]) / 4.0f;
                    for (int c = 0; c < 3; c++) a[c] = linear_to_srgb(a[c]);
                } else if (type == SR_TEX_METALLIC_ROUGHNESS) {
                    float r = (((p[0][1] * p[0][1] + p[1][1] * p[1][1]) + p[2][1] * p[2][1]) + p[3][1] * p[3][1]) / 4.0f;
                    float m = (((p[0][2] + p[1][2]) + p[2][2]) + p[3][2]) / 4.0f;
                    a[0] = p[0][0]; a[1] = std::sqrt(r); a[2] = m; a[3] = p[0][3];
                } else if (type == SR_TEX_NORMAL) {
                    for (int c = 0; c < 4; c++) {
                        float n0 = p[0][c] * 2.0f - 1.0f, n1 = p[1][c] * 2.0f - 1.0f;
                        float n2 = p[2][c] * 2.0f - 1.0f, n3 = p[3][c] * 2.0f - 1.0f;
                        float avg = (((n0 + n1) + n2) + n3) / 4.0f;
                        a[c] = (avg + 1.0f) / 2.0f;
                    }
                } else {
                    for (int c = 0; c < 4; c++) a[c] = (((p[0][c] + p[1][c]) + p[2][c]) + p[3][c]) / 4.0f;
                }
                data[start + (size_t)y * mw + x] = pack(a);
            }
        widths.push_back(mw);
        heights.push_back(mh);
    }
}

// ---------------------------------------------------------------------------
// Image-based lighting (src/texture.rs:135-420)
// ---------------------------------------------------------------------------
static const float PI_F = 3.14159265358979323846f;

static float radical_inverse_vdc(uint32_t bits) {
    bits = (bits << 16) | (bits >> 16);
    bits = ((bits & 0x55555555u) << 1) | ((bits & 0xAAAAAAAAu) >> 1);
    bits = ((bits & 0x33333333u) << 2) | ((bits & 0xCCCCCCCCu) >> 2);
    bits = ((bits & 0x0F0F0F0Fu) << 4) | ((bits & 0xF0F0F0F0u) >> 4);
    bits = ((bits & 0x00FF00FFu) << 8) | ((bits & 0xFF00FF00u) >> 8);
    return (float)bits * 2.3283064e-10f;
}
static inline void hammersley(uint32_t i, uint32_t n, float &x, float &y) {
    x = (float)i / (float)n;
    y = radical_inverse_vdc(i);
}
static F3 importance_sample_ggx(float xi_x, float xi_y, F3 n, float roughness) {
    float a = roughness * roughness;
    float phi = (2.0f * PI_F) * xi_x;
    float cos_theta = std::sqrt((1.0f - xi_y) / (1.0f + (a * a - 1.0f) * xi_y));
    float sin_theta = std::sqrt(std::fmax(1.0f - cos_theta * cos_theta, 0.0f));
    F3 h = f3(std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, cos_theta);
    F3 up = std::fabs(n.z) < 0.999f ? f3(0, 0, 1) : f3(1, 0, 0);
    F3 t = normalize(cross(n, up));
    F3 b = cross(n, t);
    return normalize((t * h.x + b * h.y) + n * h.z);
}
static inline float powi5(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    return x * x4;
}
static void integrate_brdf(float ndotv, float roughness, float &ra, float &rb) {
    F3 v = f3(std::sqrt(std::fmax(1.0f - ndotv * ndotv, 0.0f)), 0.0f, ndotv);
    F3 n = f3(0, 0, 1);
    const uint32_t count = 128;
    float a = 0.0f, b = 0.0f;
    for (uint32_t i = 0; i < count; i++) {
        float xx, xy;
        hammersley(i, count, xx, xy);
        F3 h = importance_sample_ggx(xx, xy, n, roughness);
        F3 l = normalize(h * (2.0f * dot(v, h)) - v);
        float ndotl = std::fmax(l.z, 0.0f);
        float ndoth = std::fmax(h.z, 0.0f);
        float vdoth = std::fmax(dot(v, h), 0.0f);
        if (ndotl > 0.0f) {
            float alpha = roughness * roughness;
            float k = ((alpha + 1.0f) * (alpha + 1.0f)) * 0.125f;
            float gv = ndotv / (ndotv * (1.0f - k) + k);
            float gl = ndotl / (ndotl * (1.0f - k) + k);
            float gvis = std::fmax(((gv * gl) * vdoth) / (ndoth * std::fmax(ndotv, 1.0e-5f)), 0.0f);
            float fc = powi5(1.0f - vdoth);
            a += (1.0f - fc) * gvis;
            b += fc * gvis;
        }
    }
    ra = a / (float)count;
    rb = b / (float)count;
}

static F3 face_uv_to_direction(uint32_t face, float u, float v) {
    F3 d;
    switch (face) {
        case 0: d = f3(1.0f, -v, -u); break;
        case 1: d = f3(-1.0f, -v, u); break;
        case 2: d = f3(u, 1.0f, v); break;
        case 3: d = f3(u, -1.0f, -v); break;
        case 4: d = f3(u, -v, 1.0f); break;
        default: d = f3(-u, -v, -1.0f); break;
    }
    return normalize(d);
}
static void direction_to_face_uv(F3 n, uint32_t &face, float &u, float &v) {
    float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
    if (ax >= ay && ax >= az) {
        if (n.x >= 0.0f) { face = 0; u = (-n.z / ax) * 0.5f + 0.5f; v = (-n.y / ax) * 0.5f + 0.5f; }
        else { face = 1; u = (n.z / ax) * 0.5f + 0.5f; v = (-n.y / ax) * 0.5f + 0.5f; }
    } else if (ay > ax && ay >= az) {
        if (n.y >= 0.0f) { face = 2; u = (n.x / ay) * 0.5f + 0.5f; v = (n.z / ay) * 0.5f + 0.5f; }
        else { face = 3; u = (n.x / ay) * 0.5f + 0.5f; v = (-n.z / ay) * 0.5f + 0.5f; }
    } else if (n.z >= 0.0f) {
        face = 4; u = (n.x / az) * 0.5f + 0.5f; v = (-n.y / az) * 0.5f + 0.5f;
    } else {
        face = 5; u = (-n.x / az) * 0.5f + 0.5f; v = (-n.y / az) * 0.5f + 0.5f;
    }
}
static inline float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// sample_bilinear_rgb at mip 0 with ClampToEdge on one cubemap face, lane 0
// (src/texture.rs:730-790), then srgb_to_linear_scalar per channel.
static F3 sample_cubemap_direction_linear(const uint32_t *faces, uint32_t size, F3 dir) {
    uint32_t face;
    float u, v;
    direction_to_face_uv(dir, face, u, v);
    u = clamp01(u);
    v = clamp01(v);
    float wf = (float)size, hf = (float)size;
    float xf = u * wf - 0.5f, yf = v * hf - 0.5f;
    float x0 = std::floor(xf), y0 = std::floor(yf);
    float x1 = x0 + 1.0f, y1 = y0 + 1.0f;
    float fx = xf - x0, fy = yf - y0;
    float omfx = 1.0f - fx, omfy = 1.0f - fy;
    auto wrap = [](float t, float dim) { float lim = dim - 1.0f; return t < lim ? t : lim; };
    uint32_t x0i = sat_u32(wrap(x0, wf)), y0i = sat_u32(wrap(y0, hf));
    uint32_t x1i = sat_u32(wrap(x1, wf)), y1i = sat_u32(wrap(y1, hf));
    uint32_t off = face * size * size;
    float p00[4], p10[4], p01[4], p11[4];
    unpack(faces[off + y0i * size + x0i], p00);
    unpack(faces[off + y0i * size + x1i], p10);
    unpack(faces[off + y1i * size + x0i], p01);
    unpack(faces[off + y1i * size + x1i], p11);
    float w00 = omfx * omfy, w10 = fx * omfy, w01 = omfx * fy, w11 = fx * fy;
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = ((p00[k] * w00 + p10[k] * w10) + p01[k] * w01) + p11[k] * w11;
    return f3(srgb_to_linear(c[0]), srgb_to_linear(c[1]), srgb_to_linear(c[2]));
}

void compute_brdf_lut(uint32_t size, std::vector<uint32_t> &out, int threads) {
    out.assign((size_t)size * size, 0);
    float sf = (float)size;
    parallel_for(size, threads, [&](size_t y) {
        float roughness = clamp01(((float)y + 0.5f) / sf);
        for (uint32_t x = 0; x < size; x++) {
            float ndotv = clamp01(((float)x + 0.5f) / sf);
            float a, b;
            integrate_brdf(std::fmax(ndotv, 1.0e-4f), std::fmax(roughness, 1.0e-4f), a, b);
            float t[4] = {clamp01(a), clamp01(b), 0.0f, 1.0f};
            out[y * size + x] = pack(t);
        }
    });
}

void compute_irradiance_sh4(const uint32_t *faces, uint32_t size, float sh[4][3]) {
    float acc[4][3] = {{0}};
    float w = (float)size, h = (float)size;
    float texel_omega = (2.0f / w) * (2.0f / h);
    for (uint32_t face = 0; face < 6; face++)
        for (uint32_t y = 0; y < size; y++) {
            float v = (((float)y + 0.5f) / h) * 2.0f - 1.0f;
            for (uint32_t x = 0; x < size; x++) {
                float u = (((float)x + 0.5f) / w) * 2.0f - 1.0f;
                F3 dir = face_uv_to_direction(face, u, v);
                float weight = texel_omega / std::pow((1.0f + u * u) + v * v, 1.5f);
                F3 col = sample_cubemap_direction_linear(faces, size, dir);
                float basis[4] = {0.282095f, 0.488603f * dir.y, 0.488603f * dir.z, 0.488603f * dir.x};
                for (int i = 0; i < 4; i++) {
                    float s = basis[i] * weight;
                    acc[i][0] += col.x * s;
                    acc[i][1] += col.y * s;
                    acc[i][2] += col.z * s;
                }
            }
        }
    const float c0 = PI_F, c1 = (2.0f * PI_F) / 3.0f;
    for (int k = 0; k < 3; k++) acc[0][k] *= c0;
    for (int i = 1; i < 4; i++)
        for (int k = 0; k < 3; k++) acc[i][k] *= c1;
    const float fold[4] = {0.282095f, 0.488603f, 0.488603f, 0.488603f};
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) sh[i][k] = acc[i][k] * fold[i];
}

void compute_specular_cubemap(const uint32_t *faces, uint32_t size, uint32_t sample_count, std::vector<uint32_t> &out,
                              uint32_t &num_mips, int threads) {
    num_mips = 1 + ilog2u(size);
    uint32_t max_mip = num_mips - 1;
    size_t face_texels = (size_t)size * size;
    out.assign(face_texels * 6 * num_mips, 0);
    size_t rows = (size_t)num_mips * 6 * size;
    parallel_for(rows, threads, [&](size_t row) {
        uint32_t mip = (uint32_t)(row / (6 * size));
        uint32_t face = (uint32_t)((row / size) % 6);
        uint32_t y = (uint32_t)(row % size);
        float roughness = max_mip > 0 ? (float)mip / (float)max_mip : 0.0f;
        float v = (((float)y + 0.5f) / (float)size) * 2.0f - 1.0f;
        for (uint32_t x = 0; x < size; x++) {
            float u = (((float)x + 0.5f) / (float)size) * 2.0f - 1.0f;
            F3 r = face_uv_to_direction(face, u, v);
            F3 color;
            if (mip == 0) {
                color = sample_cubemap_direction_linear(faces, size, r);
            } else {
                F3 accum = f3(0, 0, 0);
                float total = 0.0f;
                for (uint32_t i = 0; i < sample_count; i++) {
                    float xx, xy;
                    hammersley(i, sample_count, xx, xy);
                    F3 hv = importance_sample_ggx(xx, xy, r, std::fmax(roughness, 0.045f));
                    F3 l = normalize(hv * (2.0f * dot(r, hv)) - r);
                    float ndotl = std::fmax(dot(r, l), 0.0f);
                    if (ndotl > 0.0f) {
                        accum = accum + sample_cubemap_direction_linear(faces, size, l) * ndotl;
                        total += ndotl;
                    }
                }
                if (total > 0.0f) color = f3(accum.x / total, accum.y / total, accum.z / total);
                else color = sample_cubemap_direction_linear(faces, size, r);
            }
            float t[4] = {color.x, color.y, color.z, 1.0f};
            out[(size_t)mip * 6 * face_texels + (size_t)face * face_texels + (size_t)y * size + x] = pack(t);
        }
    });
}

// ---------------------------------------------------------------------------
// Scene bounds (src/scene.rs:333-351)
// ---------------------------------------------------------------------------
void scene_bounds(const sr_scene_desc *d, float mn[3], float mx[3], float center[3], float *diag) {
    F3 lo = f3(INFINITY, INFINITY, INFINITY), hi = f3(-INFINITY, -INFINITY, -INFINITY);
    for (int ni = 0; ni < d->num_nodes; ni++) {
        const sr_node_desc &n = d->nodes[ni];
        if (n.mesh < 0) continue;
        const sr_mesh_desc &m = d->meshes[n.mesh];
        for (int k = 0; k < m.num_primitives; k++) {
            const sr_primitive_desc &p = d->primitives[m.first_primitive + k];
            for (uint32_t v = 0; v < p.num_vertices; v++) {
                float o[4];
                mat_point(n.transform, p.positions + 3 * v, o);
                lo = fmin3(lo, f3(o[0], o[1], o[2]));
                hi = fmax3(hi, f3(o[0], o[1], o[2]));
            }
        }
    }
    mn[0] = lo.x; mn[1] = lo.y; mn[2] = lo.z;
    mx[0] = hi.x; mx[1] = hi.y; mx[2] = hi.z;
    F3 c = (lo + hi) * 0.5f;
    center[0] = c.x; center[1] = c.y; center[2] = c.z;
    *diag = len(hi - lo);
}

// ---------------------------------------------------------------------------
// Voxel sun visibility (src/gi.rs:151-314, src/raytracer.rs)
// ---------------------------------------------------------------------------
namespace {
struct Tri { F3 p0, p1, p2; int material; };
struct Box { F3 lo, hi; };
struct BNode { Box box; int left, right, first, count; };

struct Bvh {
    std::vector<Tri> tris;
    std::vector<Box> tbox;
    std::vector<int> order;
    std::vector<BNode> nodes;

    int build(int first, int count, int depth) {
        Box b = {f3(INFINITY, INFINITY, INFINITY), f3(-INFINITY, -INFINITY, -INFINITY)};
        Box cb = b;
        for (int i = first; i < first + count; i++) {
            const Box &t = tbox[order[i]];
            b.lo = fmin3(b.lo, t.lo); b.hi = fmax3(b.hi, t.hi);
            F3 c = (t.lo + t.hi) * 0.5f;
            cb.lo = fmin3(cb.lo, c); cb.hi = fmax3(cb.hi, c);
        }
        int id = (int)nodes.size();
        nodes.push_back({b, -1, -1, first, count});
        if (count <= 4 || depth > 60) return id;
        F3 ext = cb.hi - cb.lo;
        int axis = ext.x > ext.y ? (ext.x > ext.z ? 0 : 2) : (ext.y > ext.z ? 1 : 2);
        auto key = [&](int t) {
            const Box &bb = tbox[t];
            return axis == 0 ? bb.lo.x + bb.hi.x : axis == 1 ? bb.lo.y + bb.hi.y : bb.lo.z + bb.hi.z;
        };
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

    static bool hit_box(const Box &b, F3 o, F3 inv, float tmax) {
        float t0 = 0.0f, t1 = tmax;
        const float *lo = &b.lo.x, *hi = &b.hi.x, *op = &o.x, *ip = &inv.x;
        for (int a = 0; a < 3; a++) {
            float ta = (lo[a] - op[a]) * ip[a], tb = (hi[a] - op[a]) * ip[a];
            if (ta > tb) std::swap(ta, tb);
            if (ta != ta) ta = -INFINITY;  // 0 * inf
            if (tb != tb) tb = INFINITY;
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
            if (t0 > t1 * 1.0000001f + 1e-7f) return false;
        }
        return true;
    }
};

// ray_triangle_intersect (src/raytracer.rs:225-260)
static bool ray_tri(F3 o, F3 d, const Tri &t, float tmin, float tmax, float &tout) {
    F3 e1 = t.p1 - t.p0, e2 = t.p2 - t.p0;
    F3 pv = cross(d, e2);
    float det = dot(e1, pv);
    if (std::fabs(det) <= 1.0e-8f) return false;
    float inv = 1.0f / det;
    F3 tv = o - t.p0;
    float u = dot(tv, pv) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    F3 qv = cross(tv, e1);
    float v = dot(d, qv) * inv;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    float tt = dot(e2, qv) * inv;
    if (tt < tmin || tt > tmax) return false;
    tout = tt;
    return true;
}
}  // namespace

// World-space triangles of every node in node order, primitive order and
// triangle order, skipping degenerate ones (RayTracer::new,
// src/raytracer.rs:172-223).  Parallel over chunks; the order is kept.
void collect_world_triangles(const sr_scene_desc *d, int threads, std::vector<WorldTri> &out,
                             std::vector<float> *uvs) {
    struct Job { int node, prim; uint32_t first, count; };
    std::vector<Job> jobs;
    const uint32_t CHUNK = 1u << 16;
    for (int ni = 0; ni < d->num_nodes; ni++) {
        const sr_node_desc &n = d->nodes[ni];
        if (n.mesh < 0) continue;
        const sr_mesh_desc &m = d->meshes[n.mesh];
        for (int k = 0; k < m.num_primitives; k++) {
            int pi = m.first_primitive + k;
            uint32_t nt = d->primitives[pi].num_indices / 3;
            for (uint32_t f = 0; f < nt; f += CHUNK) jobs.push_back({ni, pi, f, std::min(CHUNK, nt - f)});
        }
    }
    std::vector<std::vector<WorldTri>> parts(jobs.size());
    std::vector<std::vector<float>> uparts(uvs ? jobs.size() : 0);
    parallel_for(jobs.size(), threads, [&](size_t j) {
        const Job &jb = jobs[j];
        const sr_node_desc &n = d->nodes[jb.node];
        const sr_primitive_desc &p = d->primitives[jb.prim];
        std::vector<WorldTri> &o = parts[j];
        o.reserve(jb.count);
        for (uint32_t t = jb.first; t < jb.first + jb.count; t++) {
            float a[4], b[4], c[4];
            mat_point(n.transform, p.positions + 3 * p.indices[3 * t], a);
            mat_point(n.transform, p.positions + 3 * p.indices[3 * t + 1], b);
            mat_point(n.transform, p.positions + 3 * p.indices[3 * t + 2], c);
            F3 p0 = f3(a[0], a[1], a[2]), p1 = f3(b[0], b[1], b[2]), p2 = f3(c[0], c[1], c[2]);
            F3 fn = cross(p1 - p0, p2 - p0);
            if (dot(fn, fn) <= 1.0e-12f) continue;
            o.push_back({{p0.x, p0.y, p0.z, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z}, p.material});
            if (uvs)
                for (int k = 0; k < 3; k++) {
                    uint32_t v = p.indices[3 * t + k];
                    uparts[j].push_back(p.texcoords ? p.texcoords[2 * v] : 0.0f);
                    uparts[j].push_back(p.texcoords ? p.texcoords[2 * v + 1] : 0.0f);
                }
        }
    });
    out.clear();
    for (auto &pt : parts) out.insert(out.end(), pt.begin(), pt.end());
    if (uvs) {
        uvs->clear();
        for (auto &pt : uparts) uvs->insert(uvs->end(), pt.begin(), pt.end());
    }
}

// Grid placement and sun ray constants of the voxel bake (src/gi.rs:267-290,
// src/voxelgrid.rs:154-161).
void voxel_frame(const sr_scene_desc *d, int grid, VoxelFrame &f) {
    float center[3], diag;
    scene_bounds(d, f.wmin, f.wmax, center, &diag);
    F3 lo = f3(f.wmin[0], f.wmin[1], f.wmin[2]), hi = f3(f.wmax[0], f.wmax[1], f.wmax[2]);
    F3 ws = hi - lo;
    F3 vs = f3(ws.x / (float)grid, ws.y / (float)grid, ws.z / (float)grid);
    F3 ldir = normalize(f3(d->light.direction[0], d->light.direction[1], d->light.direction[2]));
    F3 cmin = lo + vs * 0.5f;
    f.grid = grid;
    f.lo[0] = lo.x; f.lo[1] = lo.y; f.lo[2] = lo.z;
    f.vs[0] = vs.x; f.vs[1] = vs.y; f.vs[2] = vs.z;
    f.ldir[0] = ldir.x; f.ldir[1] = ldir.y; f.ldir[2] = ldir.z;
    f.cmin[0] = cmin.x; f.cmin[1] = cmin.y; f.cmin[2] = cmin.z;
    f.bias = len(vs) * 3.0f;
    f.area_ref = std::fmax(std::fmin(vs.x, std::fmin(vs.y, vs.z)), 1.0e-6f);
    f.area_ref = f.area_ref * f.area_ref;
}

int build_voxel_grid(const sr_scene_desc *d, int grid, int threads, std::vector<float> &light, float wmin[3],
                     float wmax[3], std::vector<uint8_t> *active_out) {
    VoxelFrame fr;
    voxel_frame(d, grid, fr);
    memcpy(wmin, fr.wmin, 12);
    memcpy(wmax, fr.wmax, 12);
    F3 lo = f3(fr.lo[0], fr.lo[1], fr.lo[2]);
    F3 vs = f3(fr.vs[0], fr.vs[1], fr.vs[2]);
    size_t nvox = (size_t)grid * grid * grid;

    Bvh bvh;
    {
        std::vector<WorldTri> wt;
        collect_world_triangles(d, threads, wt);
        bvh.tris.reserve(wt.size());
        bvh.tbox.reserve(wt.size());
        for (const WorldTri &w : wt) {
            Tri tr = {f3(w.p[0], w.p[1], w.p[2]), f3(w.p[3], w.p[4], w.p[5]), f3(w.p[6], w.p[7], w.p[8]), w.material};
            bvh.tris.push_back(tr);
            const float eps = 1.0e-5f;
            Box bx = {fmin3(fmin3(tr.p0, tr.p1), tr.p2) - f3(eps, eps, eps),
                      fmax3(fmax3(tr.p0, tr.p1), tr.p2) + f3(eps, eps, eps)};
            bvh.tbox.push_back(bx);
        }
    }

    // build_active_voxel_mask (src/gi.rs:151-265): occupancy by surface sampling
    std::vector<uint8_t> occupied(nvox, 0);
    float area_ref = fr.area_ref;
    for (const Tri &t : bvh.tris) {
        F3 e1 = t.p1 - t.p0, e2 = t.p2 - t.p0;
        float area = 0.5f * len(cross(e1, e2));
        float ts = std::ceil((area / area_ref) * 2.0f);
        size_t target = (size_t)(ts > 0 ? (ts >= 1.8e19f ? UINT64_MAX : (uint64_t)ts) : 0);
        target = std::min<size_t>(std::max<size_t>(target, 1), 4096);
        size_t ns = (size_t)std::ceil(std::sqrt((float)target));
        for (size_t iu = 0; iu < ns; iu++)
            for (size_t iv = 0; iv < ns - iu; iv++) {
                float u = ((float)iu + 0.5f) / (float)ns, v = ((float)iv + 0.5f) / (float)ns;
                float w = (1.0f - u) - v;
                if (w < 0.0f) continue;
                F3 p = (t.p0 * w + t.p1 * u) + t.p2 * v;
                float fx = (p.x - lo.x) / vs.x, fy = (p.y - lo.y) / vs.y, fz = (p.z - lo.z) / vs.z;
                if (fx < 0.0f || fy < 0.0f || fz < 0.0f) continue;
                double x = std::floor(fx), y = std::floor(fy), z = std::floor(fz);
                if (x >= grid || y >= grid || z >= grid) continue;
                occupied[((size_t)z * grid + (size_t)y) * grid + (size_t)x] = 1;
            }
    }
    std::vector<uint8_t> active(occupied);
    const int R = 2;
    for (int z = 0; z < grid; z++)
        for (int y = 0; y < grid; y++)
            for (int x = 0; x < grid; x++) {
                if (!occupied[((size_t)z * grid + y) * grid + x]) continue;
                for (int nz = std::max(z - R, 0); nz <= std::min(z + R, grid - 1); nz++)
                    for (int ny = std::max(y - R, 0); ny <= std::min(y + R, grid - 1); ny++)
                        for (int nx = std::max(x - R, 0); nx <= std::min(x + R, grid - 1); nx++)
                            active[((size_t)nz * grid + ny) * grid + nx] = 1;
            }

    // BVH over the triangle soup
    bvh.order.resize(bvh.tris.size());
    for (size_t i = 0; i < bvh.order.size(); i++) bvh.order[i] = (int)i;
    if (!bvh.tris.empty()) bvh.build(0, (int)bvh.tris.size(), 0);

    // materials: translucency and transmission for trace_transmittance
    std::vector<uint8_t> translucent(d->num_materials > 0 ? d->num_materials : 1, 0);
    std::vector<float> transmission(translucent.size(), 0.0f);
    for (int i = 0; i < d->num_materials; i++) {
        translucent[i] = d->materials[i].is_translucent != 0;
        transmission[i] = d->materials[i].transmission;
    }
    if (active_out) *active_out = active;

    // compute_sun_visibility (src/gi.rs:267-314)
    F3 ldir = f3(fr.ldir[0], fr.ldir[1], fr.ldir[2]);
    F3 cmin = f3(fr.cmin[0], fr.cmin[1], fr.cmin[2]);
    float bias = fr.bias;
    std::vector<float> vis(nvox, 1.0f);
    F3 inv = f3(1.0f / ldir.x, 1.0f / ldir.y, 1.0f / ldir.z);
    parallel_for((size_t)grid * grid, threads, [&](size_t zy) {
        int z = (int)(zy / grid), y = (int)(zy % grid);
        std::vector<int> stack;
        struct Hit { float t; int tri, material; };
        std::vector<Hit> hits;
        for (int x = 0; x < grid; x++) {
            size_t idx = ((size_t)z * grid + y) * grid + x;
            if (!active[idx]) continue;
            F3 c = cmin + mulv(f3((float)x, (float)y, (float)z), vs);
            F3 o = c + ldir * bias;
            hits.clear();
            if (!bvh.nodes.empty()) {
                stack.clear();
                stack.push_back(0);
                while (!stack.empty()) {
                    int ni = stack.back();
                    stack.pop_back();
                    const BNode &nd = bvh.nodes[ni];
                    if (!Bvh::hit_box(nd.box, o, inv, INFINITY)) continue;
                    if (nd.left < 0) {
                        for (int k = nd.first; k < nd.first + nd.count; k++) {
                            const Tri &t = bvh.tris[bvh.order[k]];
                            float tt;
                            if (ray_tri(o, ldir, t, 1.0e-4f, INFINITY, tt)) hits.push_back({tt, bvh.order[k], t.material});
                        }
                    } else {
                        stack.push_back(nd.left);
                        stack.push_back(nd.right);
                    }
                }
            }
            // by distance; equal distances by triangle order, so the result
            // does not depend on the traversal order
            std::sort(hits.begin(), hits.end(),
                      [](const Hit &a, const Hit &b) { return a.t < b.t || (a.t == b.t && a.tri < b.tri); });
            float tr = 1.0f;
            for (auto &h : hits) {
                int m = (h.material >= 0 && h.material < (int)translucent.size()) ? h.material : 0;
                if (!translucent[m]) { tr = 0.0f; break; }
                tr *= transmission[m];
                if (tr <= 0.0001f) { tr = 0.0f; break; }
            }
            vis[idx] = tr;
        }
    });

    // blur_grid (src/voxelgrid.rs:371-419): 3x3x3 box, then squared
    light.assign(nvox, 0.0f);
    parallel_for((size_t)grid, threads, [&](size_t zz) {
        int z = (int)zz;
        for (int y = 0; y < grid; y++)
            for (int x = 0; x < grid; x++) {
                float sum = 0.0f;
                int count = 0;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int nx = x + dx, ny = y + dy, nz = z + dz;
                            if (nx >= 0 && nx < grid && ny >= 0 && ny < grid && nz >= 0 && nz < grid) {
                                sum += vis[((size_t)nz * grid + ny) * grid + nx];
                                count++;
                            }
                        }
                light[((size_t)z * grid + y) * grid + x] = count > 0 ? std::pow(sum / (float)count, 2.0f) : 0.0f;
            }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// Camera (src/rendercamera.rs)
// ---------------------------------------------------------------------------
namespace {
struct Q { float x, y, z, w; };
static inline Q qmul(Q a, Q b) {
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
struct M { float m[16]; };
static inline M mmul(const M &a, const M &b) {
    M r;
    for (int c = 0; c < 4; c++)
        for (int i = 0; i < 4; i++) {
            float acc = a.m[i] * b.m[4 * c];
            acc = acc + a.m[4 + i] * b.m[4 * c + 1];
            acc = acc + a.m[8 + i] * b.m[4 * c + 2];
            acc = acc + a.m[12 + i] * b.m[4 * c + 3];
            r.m[4 * c + i] = acc;
        }
    return r;
}
static inline M from_quat(Q q) {
    float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2, yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    return {{1.0f - (yy + zz), xy + wz, xz - wy, 0, xy - wz, 1.0f - (xx + zz), yz + wx, 0, xz + wy, yz - wx,
             1.0f - (xx + yy), 0, 0, 0, 0, 1}};
}
}  // namespace

void mat4_mul(const float a[16], const float b[16], float out[16]) {
    M A, B;
    memcpy(A.m, a, 64);
    memcpy(B.m, b, 64);
    M r = mmul(A, B);
    memcpy(out, r.m, 64);
}

void mat4_from_quat(const float q[4], float out[16]) {
    M r = from_quat(Q{q[0], q[1], q[2], q[3]});
    memcpy(out, r.m, 64);
}

namespace {
static M inverse(const M &A) {
    const float *m = A.m;
    float m00 = m[0], m01 = m[1], m02 = m[2], m03 = m[3], m10 = m[4], m11 = m[5], m12 = m[6], m13 = m[7];
    float m20 = m[8], m21 = m[9], m22 = m[10], m23 = m[11], m30 = m[12], m31 = m[13], m32 = m[14], m33 = m[15];
    float c00 = m22 * m33 - m32 * m23, c02 = m12 * m33 - m32 * m13, c03 = m12 * m23 - m22 * m13;
    float c04 = m21 * m33 - m31 * m23, c06 = m11 * m33 - m31 * m13, c07 = m11 * m23 - m21 * m13;
    float c08 = m21 * m32 - m31 * m22, c10 = m11 * m32 - m31 * m12, c11 = m11 * m22 - m21 * m12;
    float c12 = m20 * m33 - m30 * m23, c14 = m10 * m33 - m30 * m13, c15 = m10 * m23 - m20 * m13;
    float c16 = m20 * m32 - m30 * m22, c18 = m10 * m32 - m30 * m12, c19 = m10 * m22 - m20 * m12;
    float c20 = m20 * m31 - m30 * m21, c22 = m10 * m31 - m30 * m11, c23 = m10 * m21 - m20 * m11;
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3_[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float v0[4] = {m10, m00, m00, m00}, v1[4] = {m11, m01, m01, m01};
    const float v2[4] = {m12, m02, m02, m02}, v3[4] = {m13, m03, m03, m03};
    const float sa[4] = {1, -1, 1, -1}, sb[4] = {-1, 1, -1, 1};
    M r;
    for (int i = 0; i < 4; i++) {
        r.m[i] = ((v1[i] * f0[i] - v2[i] * f1[i]) + v3[i] * f2[i]) * sa[i];
        r.m[4 + i] = ((v0[i] * f0[i] - v2[i] * f3_[i]) + v3[i] * f4[i]) * sb[i];
        r.m[8 + i] = ((v0[i] * f1[i] - v1[i] * f3_[i]) + v3[i] * f5[i]) * sa[i];
        r.m[12 + i] = ((v0[i] * f2[i] - v1[i] * f4[i]) + v2[i] * f5[i]) * sb[i];
    }
    float d0 = m00 * r.m[0], d1 = m01 * r.m[4], d2 = m02 * r.m[8], d3 = m03 * r.m[12];
    float det = ((d0 + d1) + d2) + d3;
    float rcp = 1.0f / det;
    for (int i = 0; i < 16; i++) r.m[i] = r.m[i] * rcp;
    return r;
}
static Q rotation_of(const sr_camera &c) {
    Q qy = {0, std::sin(c.yaw * 0.5f), 0, std::cos(c.yaw * 0.5f)};
    Q qp = {std::sin(c.pitch * 0.5f), 0, 0, std::cos(c.pitch * 0.5f)};
    Q qr = {0, 0, std::sin(c.roll * 0.5f), std::cos(c.roll * 0.5f)};
    return qmul(qmul(qr, qp), qy);
}
}  // namespace

void camera_new(const float pos[3], const float look[3], float fov, float w, float h, float far_plane, sr_camera *c) {
    memset(c, 0, sizeof(*c));
    F3 fwd = normalize(f3(look[0] - pos[0], look[1] - pos[1], look[2] - pos[2]));
    memcpy(c->position, pos, 12);
    c->fov = fov;
    c->width = w;
    c->height = h;
    c->one_over_width = 1.0f / w;
    c->one_over_height = 1.0f / h;
    c->near_plane = far_plane * 0.01f;
    c->far_plane = far_plane;
    c->yaw = std::atan2(fwd.x, -fwd.z);
    c->pitch = std::asin(fwd.y);
    c->roll = 0.0f;
}

void camera_update(sr_camera *c) {
    Q q = rotation_of(*c);
    M rot = from_quat(q);
    M tr = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -c->position[0], -c->position[1], -c->position[2], 1}};
    M view = mmul(rot, tr);
    float sn = std::sin(0.5f * c->fov), cs = std::cos(0.5f * c->fov);
    float hh = cs / sn;
    float ww = hh / (c->width / c->height);
    float rr = c->far_plane / (c->near_plane - c->far_plane);
    M proj = {{ww, 0, 0, 0, 0, hh, 0, 0, 0, 0, rr, -1.0f, 0, 0, rr * c->near_plane, 0}};
    M vp = mmul(proj, view);
    M ivp = inverse(vp);
    M sky = inverse(mmul(proj, rot));
    memcpy(c->view_matrix, view.m, 64);
    memcpy(c->projection_matrix, proj.m, 64);
    memcpy(c->view_project_matrix, vp.m, 64);
    memcpy(c->inverse_view_project_matrix, ivp.m, 64);
    for (int col = 0; col < 4; col++)
        for (int i = 0; i < 4; i++) c->skybox_matrix_transposed[4 * i + col] = sky.m[4 * col + i];
    float aspect = c->width / c->height;
    float ty = std::tan(c->fov * 0.5f);
    float tx = ty * aspect;
    F3 r = normalize(f3(1.0f, 0.0f, -tx)), l = normalize(f3(-1.0f, 0.0f, -tx));
    F3 t = normalize(f3(0.0f, 1.0f, -ty)), b = normalize(f3(0.0f, -1.0f, -ty));
    float planes[6][4] = {{0, 0, -1, c->near_plane}, {0, 0, 1, c->far_plane}, {r.x, r.y, r.z, 0},
                          {l.x, l.y, l.z, 0},         {t.x, t.y, t.z, 0},        {b.x, b.y, b.z, 0}};
    memcpy(c->view_clip_planes, planes, sizeof(planes));
    Q qc = {-q.x, -q.y, -q.z, q.w};
    F3 u = f3(qc.x, qc.y, qc.z), v = f3(0, 0, 1);
    F3 tt = cross(u, v) * 2.0f;
    F3 vn = (v + tt * qc.w) + cross(u, tt);
    c->view_normal[0] = vn.x;
    c->view_normal[1] = vn.y;
    c->view_normal[2] = vn.z;
}

// compute_tangents (src/scene.rs:548-635)
void compute_tangents(const sr_primitive_desc &p, std::vector<float> &out) {
    size_t n = p.num_vertices;
    std::vector<F3> tan(n, f3(0, 0, 0)), bit(n, f3(0, 0, 0));
    out.assign(n * 4, 0.0f);
    auto P = [&](uint32_t i) { return f3(p.positions[3 * i], p.positions[3 * i + 1], p.positions[3 * i + 2]); };
    auto UV = [&](uint32_t i, int k) { return p.texcoords ? p.texcoords[2 * i + k] : 0.0f; };
    for (uint32_t t = 0; t + 2 < p.num_indices; t += 3) {
        uint32_t i0 = p.indices[t], i1 = p.indices[t + 1], i2 = p.indices[t + 2];
        F3 e1 = P(i1) - P(i0), e2 = P(i2) - P(i0);
        float d1x = UV(i1, 0) - UV(i0, 0), d1y = UV(i1, 1) - UV(i0, 1);
        float d2x = UV(i2, 0) - UV(i0, 0), d2y = UV(i2, 1) - UV(i0, 1);
        float det = d1x * d2y - d2x * d1y;
        if (std::fabs(det) < 1e-6f) continue;
        F3 tg = e1 * d2y - e2 * d1y;
        tg = f3(tg.x / det, tg.y / det, tg.z / det);
        F3 bt = e2 * d1x - e1 * d2x;
        bt = f3(bt.x / det, bt.y / det, bt.z / det);
        tan[i0] = tan[i0] + tg; tan[i1] = tan[i1] + tg; tan[i2] = tan[i2] + tg;
        bit[i0] = bit[i0] + bt; bit[i1] = bit[i1] + bt; bit[i2] = bit[i2] + bt;
    }
    for (size_t i = 0; i < n; i++) {
        F3 nn = f3(p.normals[3 * i], p.normals[3 * i + 1], p.normals[3 * i + 2]);
        F3 t = tan[i];
        F3 tg;
        float hand;
        if (dot(t, t) < 1e-6f) {
            F3 helper = std::fabs(nn.x) > 0.9f ? f3(0, 1, 0) : f3(1, 0, 0);
            tg = normalize(cross(nn, helper));
            hand = 1.0f;
        } else {
            tg = normalize(t - nn * dot(nn, t));
            hand = dot(cross(nn, tg), bit[i]) < 0.0f ? -1.0f : 1.0f;
        }
        out[4 * i] = tg.x; out[4 * i + 1] = tg.y; out[4 * i + 2] = tg.z; out[4 * i + 3] = hand;
    }
}

}  // namespace sr
