// envmap.hip -- image-based-lighting inputs on the GPU: GGX-prefiltered
// specular cubemap (generate_prefiltered_specular_cubemap,
// src/texture.rs:330-420), split-sum BRDF LUT (generate_brdf_lut,
// src/texture.rs:199-235) and 4-coefficient irradiance SH
// (compute_irradiance_sh4, src/texture.rs:289-328).
//
// The host restatement (prep.cpp) is followed operation for operation.  Its
// libm calls are handled so: the sample angles' cosines and sines (64 or 128
// distinct values) and the solid-angle weights' powers (one per texel
// position) are tabulated on the host with the same libm, so they are
// identical; the sRGB decode's x^2.4 is evaluated in double and rounded once
// (the host's powf is not always correctly rounded, so a texel can differ in
// the last bit, which the RGBA8 output mostly absorbs).  The SH sums run in
// the host's serial order.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "prep.h"
#include "runtime.h"

namespace sr {
namespace {

struct V3 { float x, y, z; };
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize(V3 a) {
    float r = 1.0f / sqrtf(dot(a, a));
    return scale(a, r);
}
__device__ __forceinline__ float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
__device__ __forceinline__ uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}
__device__ __forceinline__ uint32_t pack4(float r, float g, float b, float a) {
    return (sat_u32(r * 255.0f) << 24) | (sat_u32(g * 255.0f) << 16) | (sat_u32(b * 255.0f) << 8) | sat_u32(a * 255.0f);
}
__device__ __forceinline__ float srgb_to_linear(float s) {
    if (s <= 0.04045f) return s / 12.92f;
    return (float)pow((double)((s + 0.055f) / 1.055f), (double)2.4f);
}

struct Env {
    const uint32_t *faces;
    uint32_t size;
    const float *cs64, *cs128;   // cos, sin of 2 pi i / n (n = 64, 128), interleaved
    const float *w15;            // ((1 + u^2) + v^2)^1.5 per texel position
    uint32_t *spec;
    uint32_t mips;
    uint32_t *lut;
    uint32_t lut_size;
    float *sh_terms;             // 12 per texel, face-major
};

__device__ V3 face_uv_to_direction(uint32_t face, float u, float v) {
    V3 d;
    switch (face) {
        case 0: d = v3(1.0f, -v, -u); break;
        case 1: d = v3(-1.0f, -v, u); break;
        case 2: d = v3(u, 1.0f, v); break;
        case 3: d = v3(u, -1.0f, -v); break;
        case 4: d = v3(u, -v, 1.0f); break;
        default: d = v3(-u, -v, -1.0f); break;
    }
    return normalize(d);
}

__device__ void direction_to_face_uv(V3 n, uint32_t &face, float &u, float &v) {
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
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

// Bilinear ClampToEdge sample of the skybox at mip 0, decoded to linear.
__device__ V3 sample_linear(const Env &E, V3 dir) {
    uint32_t face;
    float u, v;
    direction_to_face_uv(dir, face, u, v);
    u = clamp01(u);
    v = clamp01(v);
    float wf = (float)E.size, hf = (float)E.size;
    float xf = u * wf - 0.5f, yf = v * hf - 0.5f;
    float x0 = floorf(xf), y0 = floorf(yf);
    float x1 = x0 + 1.0f, y1 = y0 + 1.0f;
    float fx = xf - x0, fy = yf - y0;
    float omfx = 1.0f - fx, omfy = 1.0f - fy;
    float limw = wf - 1.0f, limh = hf - 1.0f;
    uint32_t x0i = sat_u32(x0 < limw ? x0 : limw), y0i = sat_u32(y0 < limh ? y0 : limh);
    uint32_t x1i = sat_u32(x1 < limw ? x1 : limw), y1i = sat_u32(y1 < limh ? y1 : limh);
    size_t off = (size_t)face * E.size * E.size;
    uint32_t t00 = E.faces[off + (size_t)y0i * E.size + x0i], t10 = E.faces[off + (size_t)y0i * E.size + x1i];
    uint32_t t01 = E.faces[off + (size_t)y1i * E.size + x0i], t11 = E.faces[off + (size_t)y1i * E.size + x1i];
    float w00 = omfx * omfy, w10 = fx * omfy, w01 = omfx * fy, w11 = fx * fy;
    float c[3];
    for (int k = 0; k < 3; k++) {
        int sh = 24 - 8 * k;
        float p00 = (float)((t00 >> sh) & 0xFFu) / 255.0f, p10 = (float)((t10 >> sh) & 0xFFu) / 255.0f;
        float p01 = (float)((t01 >> sh) & 0xFFu) / 255.0f, p11 = (float)((t11 >> sh) & 0xFFu) / 255.0f;
        c[k] = ((p00 * w00 + p10 * w10) + p01 * w01) + p11 * w11;
    }
    return v3(srgb_to_linear(c[0]), srgb_to_linear(c[1]), srgb_to_linear(c[2]));
}

__device__ __forceinline__ float radical_inverse_vdc(uint32_t bits) {
    bits = (bits << 16) | (bits >> 16);
    bits = ((bits & 0x55555555u) << 1) | ((bits & 0xAAAAAAAAu) >> 1);
    bits = ((bits & 0x33333333u) << 2) | ((bits & 0xCCCCCCCCu) >> 2);
    bits = ((bits & 0x0F0F0F0Fu) << 4) | ((bits & 0xF0F0F0F0u) >> 4);
    bits = ((bits & 0x00FF00FFu) << 8) | ((bits & 0xFF00FF00u) >> 8);
    return (float)bits * 2.3283064e-10f;
}

// importance_sample_ggx for Hammersley sample i of n (src/texture.rs:237-262);
// cs holds cos and sin of the sample's azimuth.
__device__ V3 importance_sample_ggx(uint32_t i, const float *cs, V3 n, float roughness) {
    float xi_y = radical_inverse_vdc(i);
    float a = roughness * roughness;
    float cos_theta = sqrtf((1.0f - xi_y) / (1.0f + (a * a - 1.0f) * xi_y));
    float sin_theta = sqrtf(fmaxf(1.0f - cos_theta * cos_theta, 0.0f));
    V3 h = v3(cs[2 * i] * sin_theta, cs[2 * i + 1] * sin_theta, cos_theta);
    V3 up = fabsf(n.z) < 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
    V3 t = normalize(cross(n, up));
    V3 b = cross(n, t);
    return normalize(add(add(scale(t, h.x), scale(b, h.y)), scale(n, h.z)));
}

// One thread per output texel of every mip (64 samples above mip 0).
__global__ void __launch_bounds__(256) k_env_specular(Env E) {
    size_t face_texels = (size_t)E.size * E.size;
    size_t total = face_texels * 6 * E.mips;
    uint32_t max_mip = E.mips - 1;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        uint32_t mip = (uint32_t)(idx / (6 * face_texels));
        uint32_t face = (uint32_t)((idx / face_texels) % 6);
        uint32_t y = (uint32_t)((idx / E.size) % E.size), x = (uint32_t)(idx % E.size);
        float roughness = max_mip > 0 ? (float)mip / (float)max_mip : 0.0f;
        float v = (((float)y + 0.5f) / (float)E.size) * 2.0f - 1.0f;
        float u = (((float)x + 0.5f) / (float)E.size) * 2.0f - 1.0f;
        V3 r = face_uv_to_direction(face, u, v);
        V3 color;
        if (mip == 0) {
            color = sample_linear(E, r);
        } else {
            V3 accum = v3(0.0f, 0.0f, 0.0f);
            float tot = 0.0f;
            for (uint32_t i = 0; i < 64; i++) {
                V3 hv = importance_sample_ggx(i, E.cs64, r, fmaxf(roughness, 0.045f));
                V3 l = normalize(sub(scale(hv, 2.0f * dot(r, hv)), r));
                float ndotl = fmaxf(dot(r, l), 0.0f);
                if (ndotl > 0.0f) {
                    accum = add(accum, scale(sample_linear(E, l), ndotl));
                    tot += ndotl;
                }
            }
            if (tot > 0.0f) color = v3(accum.x / tot, accum.y / tot, accum.z / tot);
            else color = sample_linear(E, r);
        }
        E.spec[idx] = pack4(color.x, color.y, color.z, 1.0f);
    }
}

// integrate_brdf (src/texture.rs:199-235), one thread per LUT texel.
__global__ void __launch_bounds__(256) k_env_lut(Env E) {
    uint32_t n = E.lut_size * E.lut_size;
    float sf = (float)E.lut_size;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        uint32_t y = idx / E.lut_size, x = idx % E.lut_size;
        float roughness = fmaxf(clamp01(((float)y + 0.5f) / sf), 1.0e-4f);
        float ndotv = fmaxf(clamp01(((float)x + 0.5f) / sf), 1.0e-4f);
        V3 v = v3(sqrtf(fmaxf(1.0f - ndotv * ndotv, 0.0f)), 0.0f, ndotv);
        V3 nn = v3(0.0f, 0.0f, 1.0f);
        float a = 0.0f, b = 0.0f;
        for (uint32_t i = 0; i < 128; i++) {
            V3 h = importance_sample_ggx(i, E.cs128, nn, roughness);
            V3 l = normalize(sub(scale(h, 2.0f * dot(v, h)), v));
            float ndotl = fmaxf(l.z, 0.0f);
            float ndoth = fmaxf(h.z, 0.0f);
            float vdoth = fmaxf(dot(v, h), 0.0f);
            if (ndotl > 0.0f) {
                float alpha = roughness * roughness;
                float k = ((alpha + 1.0f) * (alpha + 1.0f)) * 0.125f;
                float gv = ndotv / (ndotv * (1.0f - k) + k);
                float gl = ndotl / (ndotl * (1.0f - k) + k);
                float gvis = fmaxf(((gv * gl) * vdoth) / (ndoth * fmaxf(ndotv, 1.0e-5f)), 0.0f);
                float om = 1.0f - vdoth;
                float x2 = om * om;
                float fc = om * (x2 * x2);
                a += (1.0f - fc) * gvis;
                b += fc * gvis;
            }
        }
        E.lut[idx] = pack4(clamp01(a / 128.0f), clamp01(b / 128.0f), 0.0f, 1.0f);
    }
}

// Per-texel SH terms (colour times basis times solid angle), face-major.
__global__ void __launch_bounds__(256) k_env_sh_terms(Env E) {
    size_t face_texels = (size_t)E.size * E.size;
    float w = (float)E.size, h = (float)E.size;
    float texel_omega = (2.0f / w) * (2.0f / h);
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < 6 * face_texels;
         idx += (size_t)gridDim.x * blockDim.x) {
        uint32_t face = (uint32_t)(idx / face_texels);
        uint32_t y = (uint32_t)((idx / E.size) % E.size), x = (uint32_t)(idx % E.size);
        float v = (((float)y + 0.5f) / h) * 2.0f - 1.0f;
        float u = (((float)x + 0.5f) / w) * 2.0f - 1.0f;
        V3 dir = face_uv_to_direction(face, u, v);
        float weight = texel_omega / E.w15[(size_t)y * E.size + x];
        V3 col = sample_linear(E, dir);
        float basis[4] = {0.282095f, 0.488603f * dir.y, 0.488603f * dir.z, 0.488603f * dir.x};
        float *o = E.sh_terms + 12 * idx;
        for (int i = 0; i < 4; i++) {
            float s = basis[i] * weight;
            o[3 * i] = col.x * s;
            o[3 * i + 1] = col.y * s;
            o[3 * i + 2] = col.z * s;
        }
    }
}

// The host's serial accumulation order: twelve independent running sums.
__global__ void __launch_bounds__(64) k_env_sh_sum(const float *terms, size_t texels, float *acc) {
    int k = threadIdx.x;
    if (k >= 12) return;
    float s = 0.0f;
    for (size_t t = 0; t < texels; t++) s += terms[12 * t + (size_t)k];
    acc[k] = s;
}

inline uint32_t blocks(size_t n) { return (uint32_t)std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 65536); }

}  // namespace

#define ENV_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) { rc = fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); goto done; } \
    } while (0)

int prepare_environment_device(const uint32_t *faces, uint32_t size, int device, std::vector<uint32_t> &spec,
                               uint32_t &mips, std::vector<uint32_t> &lut, uint32_t lut_size, float sh[4][3]) {
    const float PI_F = 3.14159265358979323846f;
    mips = 1;
    while ((1u << mips) <= size) mips++;           // 1 + ilog2(size)
    size_t face_texels = (size_t)size * size;
    // host tables with the host path's libm
    std::vector<float> cs64(128), cs128(256), w15(face_texels);
    for (uint32_t i = 0; i < 64; i++) {
        float phi = (2.0f * PI_F) * ((float)i / 64.0f);
        cs64[2 * i] = std::cos(phi);
        cs64[2 * i + 1] = std::sin(phi);
    }
    for (uint32_t i = 0; i < 128; i++) {
        float phi = (2.0f * PI_F) * ((float)i / 128.0f);
        cs128[2 * i] = std::cos(phi);
        cs128[2 * i + 1] = std::sin(phi);
    }
    for (uint32_t y = 0; y < size; y++) {
        float v = (((float)y + 0.5f) / (float)size) * 2.0f - 1.0f;
        for (uint32_t x = 0; x < size; x++) {
            float u = (((float)x + 0.5f) / (float)size) * 2.0f - 1.0f;
            w15[(size_t)y * size + x] = std::pow((1.0f + u * u) + v * v, 1.5f);
        }
    }
    int rc = SR_OK;
    Env E;
    memset(&E, 0, sizeof(E));
    void *d_faces = nullptr, *d_cs64 = nullptr, *d_cs128 = nullptr, *d_w15 = nullptr, *d_spec = nullptr,
         *d_lut = nullptr, *d_terms = nullptr, *d_acc = nullptr;
    float acc[12];
    spec.assign(face_texels * 6 * mips, 0);
    lut.assign((size_t)lut_size * lut_size, 0);
    ENV_TRY(hipSetDevice(device));
    ENV_TRY(hipMalloc(&d_faces, face_texels * 6 * 4));
    ENV_TRY(hipMalloc(&d_cs64, cs64.size() * 4));
    ENV_TRY(hipMalloc(&d_cs128, cs128.size() * 4));
    ENV_TRY(hipMalloc(&d_w15, w15.size() * 4));
    ENV_TRY(hipMalloc(&d_spec, spec.size() * 4));
    ENV_TRY(hipMalloc(&d_lut, lut.size() * 4));
    ENV_TRY(hipMalloc(&d_terms, face_texels * 6 * 12 * 4));
    ENV_TRY(hipMalloc(&d_acc, 12 * 4));
    ENV_TRY(hipMemcpy(d_faces, faces, face_texels * 6 * 4, hipMemcpyHostToDevice));
    ENV_TRY(hipMemcpy(d_cs64, cs64.data(), cs64.size() * 4, hipMemcpyHostToDevice));
    ENV_TRY(hipMemcpy(d_cs128, cs128.data(), cs128.size() * 4, hipMemcpyHostToDevice));
    ENV_TRY(hipMemcpy(d_w15, w15.data(), w15.size() * 4, hipMemcpyHostToDevice));
    E.faces = (const uint32_t *)d_faces;
    E.size = size;
    E.cs64 = (const float *)d_cs64;
    E.cs128 = (const float *)d_cs128;
    E.w15 = (const float *)d_w15;
    E.spec = (uint32_t *)d_spec;
    E.mips = mips;
    E.lut = (uint32_t *)d_lut;
    E.lut_size = lut_size;
    E.sh_terms = (float *)d_terms;
    hipLaunchKernelGGL(k_env_specular, dim3(blocks(spec.size())), dim3(256), 0, 0, E);
    hipLaunchKernelGGL(k_env_lut, dim3(blocks(lut.size())), dim3(256), 0, 0, E);
    hipLaunchKernelGGL(k_env_sh_terms, dim3(blocks(face_texels * 6)), dim3(256), 0, 0, E);
    hipLaunchKernelGGL(k_env_sh_sum, dim3(1), dim3(64), 0, 0, (const float *)d_terms, face_texels * 6, (float *)d_acc);
    ENV_TRY(hipGetLastError());
    ENV_TRY(hipMemcpy(spec.data(), d_spec, spec.size() * 4, hipMemcpyDeviceToHost));
    ENV_TRY(hipMemcpy(lut.data(), d_lut, lut.size() * 4, hipMemcpyDeviceToHost));
    ENV_TRY(hipMemcpy(acc, d_acc, 12 * 4, hipMemcpyDeviceToHost));
    {
        // compute_irradiance_sh4's final scaling (src/texture.rs:318-327)
        const float c0 = PI_F, c1 = (2.0f * PI_F) / 3.0f;
        const float fold[4] = {0.282095f, 0.488603f, 0.488603f, 0.488603f};
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) sh[i][k] = (acc[3 * i + k] * (i == 0 ? c0 : c1)) * fold[i];
    }
done:
    for (void *p : {d_faces, d_cs64, d_cs128, d_w15, d_spec, d_lut, d_terms, d_acc})
        if (p) hipFree(p);
    return rc;
}

}  // namespace sr
