// oracle.cpp -- CPU restatement of the reference rasterizer's per-frame path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
// MI355X pipeline: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or the CPU timing
// baseline, never as a code path of the product.
//
// It follows the reference source op for op (paths relative to the reference
// tree): src/renderer.rs (front end, binning, exposure, blit),
// src/tilerasterizer.rs (per-tile raster and vbuffer shading), src/shader.rs
// (vbuffer writer, PBR), src/texture.rs (samplers, mip generation),
// src/voxelgrid.rs (trilinear SH lookup), src/rendercamera.rs (camera),
// src/scene.rs (bounds, spheres, tangents), src/util.rs and src/math.rs.
//
// Arithmetic contract (shared with the GPU): no FMA contraction
// (-ffp-contract=off), IEEE division and square root, Rust `as` casts are
// saturating with NaN -> 0, i32 arithmetic wraps, SSE min/max semantics
// (a > b ? a : b) for vector min/max/clamp.  Two deliberate definitions
// differ from the reference binary: normalize uses 1/sqrt instead of the
// approximate _mm_rsqrt_ps (src/math.rs:34-39), and every frame starts from a
// freshly cleared vbuffer (barycentrics and packet indices are not carried
// over from earlier frames, src/tilerasterizer.rs:74 clears only depth).
// Submission order is canonical (serial): nodes by distance, opaque then
// translucent primitives, triangles, clip-fan triangles, bins row-major.
// The canonical triangle id of a packet is its position in that post-clip
// stream; the vbuffer reports that id.
//
// Parity status: the oracle is pinned by the CPU tests in tests/ (exact
// integer rasterizer on small viewports, watertightness, clipper
// invariants, sampler closed forms, tonemap/quantize closed forms; the
// shader's direct, SH, split-sum and environment terms, the skybox
// direction and the exposure step against the published models written
// independently in float64; the translucent sort and blend) and by the FP64
// shading restatement below.  Agreement with the Rust binary itself
// is unpinned where it depends on glam's SIMD operation order, rayon's
// scheduling or platform libm (no Rust toolchain exists here).

#include "swraster.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

// ----------------------------------------------------------------------------
// Scalar semantics
// ----------------------------------------------------------------------------
template <class T> static inline T maxps(T a, T b) { return a > b ? a : b; }  // _mm_max_ps
template <class T> static inline T minps(T a, T b) { return a < b ? a : b; }  // _mm_min_ps
template <class T> static inline T clampv(T x, T lo, T hi) { return minps(maxps(x, lo), hi); }

static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
static inline int32_t wabs(int32_t a) { return a < 0 ? (int32_t)(0u - (uint32_t)a) : a; }

template <class T> static inline int32_t as_i32(T f) {
    if (f != f) return 0;
    if (f >= (T)2147483648.0) return INT32_MAX;
    if (f <= (T)-2147483648.0) return INT32_MIN;
    return (int32_t)f;
}
template <class T> static inline uint32_t as_u32(T f) {
    if (!(f > (T)0)) return 0;
    if (f >= (T)4294967296.0) return UINT32_MAX;
    return (uint32_t)f;
}
template <class T> static inline uint8_t as_u8(T f) {
    if (!(f > (T)0)) return 0;
    if (f >= (T)255.0) return 255;
    return (uint8_t)f;
}
template <class T> static inline uint64_t as_usize(T f) {
    if (!(f > (T)0)) return 0;
    if (f >= (T)18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)f;
}
static inline uint32_t ilog2(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }

// ----------------------------------------------------------------------------
// Small vector types (glam layout)
// ----------------------------------------------------------------------------
struct Vec2 { float x, y; };
struct Vec3 { float x, y, z; };
struct Vec4s { float x, y, z, w; };
struct IVec2 { int32_t x, y; };

static inline Vec3 v3(float x, float y, float z) { return {x, y, z}; }
static inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline float dot3(Vec3 a, Vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline float length3(Vec3 a) { return std::sqrt(dot3(a, a)); }
static inline Vec3 normalize3(Vec3 a) { float r = 1.0f / length3(a); return a * r; }
static inline Vec3 cross3(Vec3 a, Vec3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// glam Vec4::dot on SSE2: (x*x' + z*z') + (y*y' + w*w')
static inline float dot4(Vec4s a, Vec4s b) {
    float xx = a.x * b.x, yy = a.y * b.y, zz = a.z * b.z, ww = a.w * b.w;
    return (xx + zz) + (yy + ww);
}

struct Mat4 {
    float m[16];  // column-major
    Vec4s col(int c) const { return {m[4 * c], m[4 * c + 1], m[4 * c + 2], m[4 * c + 3]}; }
};
static inline Vec4s mul(const Mat4 &a, Vec4s v) {
    Vec4s r;
    float *rp = &r.x;
    for (int i = 0; i < 4; i++) {
        float acc = a.m[i] * v.x;
        acc = acc + a.m[4 + i] * v.y;
        acc = acc + a.m[8 + i] * v.z;
        acc = acc + a.m[12 + i] * v.w;
        rp[i] = acc;
    }
    return r;
}
static inline Mat4 mul(const Mat4 &a, const Mat4 &b) {
    Mat4 r;
    for (int c = 0; c < 4; c++) {
        Vec4s v = mul(a, b.col(c));
        r.m[4 * c] = v.x; r.m[4 * c + 1] = v.y; r.m[4 * c + 2] = v.z; r.m[4 * c + 3] = v.w;
    }
    return r;
}
// Mat3A::from_mat4 * Vec3A
static inline Vec3 mul3(const Mat4 &a, Vec3 v) {
    float r[3];
    for (int i = 0; i < 3; i++) {
        float acc = a.m[i] * v.x;
        acc = acc + a.m[4 + i] * v.y;
        acc = acc + a.m[8 + i] * v.z;
        r[i] = acc;
    }
    return {r[0], r[1], r[2]};
}

// ----------------------------------------------------------------------------
// Camera (src/rendercamera.rs)
// ----------------------------------------------------------------------------
struct Quat { float x, y, z, w; };
static inline Quat qmul(Quat a, Quat b) {
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
static inline Mat4 mat_from_quat(Quat q) {
    float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    Mat4 r = {{1.0f - (yy + zz), xy + wz, xz - wy, 0.0f,
               xy - wz, 1.0f - (xx + zz), yz + wx, 0.0f,
               xz + wy, yz - wx, 1.0f - (xx + yy), 0.0f,
               0.0f, 0.0f, 0.0f, 1.0f}};
    return r;
}
static inline Mat4 mat_translation(Vec3 t) {
    Mat4 r = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, t.x, t.y, t.z, 1}};
    return r;
}
static inline Mat4 perspective_rh(float fov, float aspect, float zn, float zf) {
    float s = std::sin(0.5f * fov), c = std::cos(0.5f * fov);
    float h = c / s;
    float w = h / aspect;
    float r = zf / (zn - zf);
    Mat4 m = {{w, 0, 0, 0, 0, h, 0, 0, 0, 0, r, -1.0f, 0, 0, r * zn, 0}};
    return m;
}
static inline Mat4 transpose(const Mat4 &a) {
    Mat4 r;
    for (int c = 0; c < 4; c++)
        for (int i = 0; i < 4; i++) r.m[4 * i + c] = a.m[4 * c + i];
    return r;
}
static Mat4 inverse(const Mat4 &M) {
    const float *m = M.m;
    float m00 = m[0], m01 = m[1], m02 = m[2], m03 = m[3];
    float m10 = m[4], m11 = m[5], m12 = m[6], m13 = m[7];
    float m20 = m[8], m21 = m[9], m22 = m[10], m23 = m[11];
    float m30 = m[12], m31 = m[13], m32 = m[14], m33 = m[15];
    float coef00 = m22 * m33 - m32 * m23, coef02 = m12 * m33 - m32 * m13, coef03 = m12 * m23 - m22 * m13;
    float coef04 = m21 * m33 - m31 * m23, coef06 = m11 * m33 - m31 * m13, coef07 = m11 * m23 - m21 * m13;
    float coef08 = m21 * m32 - m31 * m22, coef10 = m11 * m32 - m31 * m12, coef11 = m11 * m22 - m21 * m12;
    float coef12 = m20 * m33 - m30 * m23, coef14 = m10 * m33 - m30 * m13, coef15 = m10 * m23 - m20 * m13;
    float coef16 = m20 * m32 - m30 * m22, coef18 = m10 * m32 - m30 * m12, coef19 = m10 * m22 - m20 * m12;
    float coef20 = m20 * m31 - m30 * m21, coef22 = m10 * m31 - m30 * m11, coef23 = m10 * m21 - m20 * m11;
    float fac0[4] = {coef00, coef00, coef02, coef03}, fac1[4] = {coef04, coef04, coef06, coef07};
    float fac2[4] = {coef08, coef08, coef10, coef11}, fac3[4] = {coef12, coef12, coef14, coef15};
    float fac4[4] = {coef16, coef16, coef18, coef19}, fac5[4] = {coef20, coef20, coef22, coef23};
    float vec0[4] = {m10, m00, m00, m00}, vec1[4] = {m11, m01, m01, m01};
    float vec2[4] = {m12, m02, m02, m02}, vec3[4] = {m13, m03, m03, m03};
    float inv0[4], inv1[4], inv2[4], inv3[4];
    for (int i = 0; i < 4; i++) {
        inv0[i] = (vec1[i] * fac0[i] - vec2[i] * fac1[i]) + vec3[i] * fac2[i];
        inv1[i] = (vec0[i] * fac0[i] - vec2[i] * fac3[i]) + vec3[i] * fac4[i];
        inv2[i] = (vec0[i] * fac1[i] - vec1[i] * fac3[i]) + vec3[i] * fac5[i];
        inv3[i] = (vec0[i] * fac2[i] - vec1[i] * fac4[i]) + vec2[i] * fac5[i];
    }
    const float sa[4] = {1, -1, 1, -1}, sb[4] = {-1, 1, -1, 1};
    Mat4 r;
    for (int i = 0; i < 4; i++) {
        r.m[i] = inv0[i] * sa[i];
        r.m[4 + i] = inv1[i] * sb[i];
        r.m[8 + i] = inv2[i] * sa[i];
        r.m[12 + i] = inv3[i] * sb[i];
    }
    Vec4s col0 = {r.m[0], r.m[4], r.m[8], r.m[12]};
    Vec4s d0 = {m00 * col0.x, m01 * col0.y, m02 * col0.z, m03 * col0.w};
    float dot1 = ((d0.x + d0.y) + d0.z) + d0.w;
    float rcp = 1.0f / dot1;
    for (int i = 0; i < 16; i++) r.m[i] = r.m[i] * rcp;
    return r;
}

static Quat camera_rotation(const sr_camera &c) {
    float sy = std::sin(c.yaw * 0.5f), cy = std::cos(c.yaw * 0.5f);
    float sp = std::sin(c.pitch * 0.5f), cp = std::cos(c.pitch * 0.5f);
    float sr = std::sin(c.roll * 0.5f), cr = std::cos(c.roll * 0.5f);
    Quat qy = {0, sy, 0, cy}, qp = {sp, 0, 0, cp}, qr = {0, 0, sr, cr};
    return qmul(qmul(qr, qp), qy);
}

static void camera_update(sr_camera &c) {
    Mat4 rot = mat_from_quat(camera_rotation(c));
    Mat4 tr = mat_translation(v3(-c.position[0], -c.position[1], -c.position[2]));
    Mat4 view = mul(rot, tr);
    Mat4 proj = perspective_rh(c.fov, c.width / c.height, c.near_plane, c.far_plane);
    Mat4 vp = mul(proj, view);
    Mat4 ivp = inverse(vp);
    Mat4 sky = transpose(inverse(mul(proj, rot)));
    memcpy(c.view_matrix, view.m, 64);
    memcpy(c.projection_matrix, proj.m, 64);
    memcpy(c.view_project_matrix, vp.m, 64);
    memcpy(c.inverse_view_project_matrix, ivp.m, 64);
    memcpy(c.skybox_matrix_transposed, sky.m, 64);
    float aspect = c.width / c.height;
    float tan_y = std::tan(c.fov * 0.5f);
    float tan_x = tan_y * aspect;
    Vec3 rp = normalize3(v3(1.0f, 0.0f, -tan_x)), lp = normalize3(v3(-1.0f, 0.0f, -tan_x));
    Vec3 tp = normalize3(v3(0.0f, 1.0f, -tan_y)), bp = normalize3(v3(0.0f, -1.0f, -tan_y));
    float planes[6][4] = {{0, 0, -1, c.near_plane}, {0, 0, 1, c.far_plane},
                          {rp.x, rp.y, rp.z, 0},     {lp.x, lp.y, lp.z, 0},
                          {tp.x, tp.y, tp.z, 0},     {bp.x, bp.y, bp.z, 0}};
    memcpy(c.view_clip_planes, planes, sizeof(planes));
    // view_normal = conjugate(rotation) * (0,0,1)  (glam Quat * Vec3: v + 2*w*(q x v) + 2*(q x (q x v)))
    Quat q = camera_rotation(c);
    Quat qc = {-q.x, -q.y, -q.z, q.w};
    Vec3 u = v3(qc.x, qc.y, qc.z), v = v3(0, 0, 1);
    Vec3 t = cross3(u, v) * 2.0f;
    Vec3 r = (v + t * qc.w) + cross3(u, t);
    c.view_normal[0] = r.x; c.view_normal[1] = r.y; c.view_normal[2] = r.z;
}

static void camera_new(const float pos[3], const float look[3], float fov, float w, float h,
                       float far_plane, sr_camera &c) {
    memset(&c, 0, sizeof(c));
    Vec3 fwd = normalize3(v3(look[0] - pos[0], look[1] - pos[1], look[2] - pos[2]));
    c.position[0] = pos[0]; c.position[1] = pos[1]; c.position[2] = pos[2];
    c.fov = fov; c.width = w; c.height = h;
    c.one_over_width = 1.0f / w; c.one_over_height = 1.0f / h;
    c.near_plane = far_plane * 0.01f; c.far_plane = far_plane;
    c.yaw = std::atan2(fwd.x, -fwd.z);
    c.pitch = std::asin(fwd.y);
    c.roll = 0.0f;
}

// ----------------------------------------------------------------------------
// Lane vectors (one 2x2 quad)
// ----------------------------------------------------------------------------
template <class T> struct V4 {
    T v[4];
    T &operator[](int i) { return v[i]; }
    const T &operator[](int i) const { return v[i]; }
    static V4 splat(T s) { return {{s, s, s, s}}; }
};
template <class T> static inline V4<T> operator+(V4<T> a, V4<T> b) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = a[i] + b[i]; return r; }
template <class T> static inline V4<T> operator-(V4<T> a, V4<T> b) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = a[i] - b[i]; return r; }
template <class T> static inline V4<T> operator*(V4<T> a, V4<T> b) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = a[i] * b[i]; return r; }
template <class T> static inline V4<T> operator/(V4<T> a, V4<T> b) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = a[i] / b[i]; return r; }
template <class T> static inline V4<T> operator+(V4<T> a, T s) { return a + V4<T>::splat(s); }
template <class T> static inline V4<T> operator-(V4<T> a, T s) { return a - V4<T>::splat(s); }
template <class T> static inline V4<T> operator*(V4<T> a, T s) { return a * V4<T>::splat(s); }
template <class T> static inline V4<T> operator+(T s, V4<T> a) { return V4<T>::splat(s) + a; }
template <class T> static inline V4<T> operator-(T s, V4<T> a) { return V4<T>::splat(s) - a; }
template <class T> static inline V4<T> operator*(T s, V4<T> a) { return V4<T>::splat(s) * a; }
template <class T> static inline V4<T> operator/(T s, V4<T> a) { return V4<T>::splat(s) / a; }
template <class T> static inline V4<T> vmax(V4<T> a, V4<T> b) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = maxps(a[i], b[i]); return r; }
template <class T> static inline V4<T> vmin(V4<T> a, V4<T> b) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = minps(a[i], b[i]); return r; }
template <class T> static inline V4<T> vclamp(V4<T> a, V4<T> lo, V4<T> hi) { return vmin(vmax(a, lo), hi); }
template <class T> static inline V4<T> vfloor(V4<T> a) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = std::floor(a[i]); return r; }
template <class T> static inline V4<T> vround(V4<T> a) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = std::round(a[i]); return r; }
template <class T> static inline V4<T> vabs(V4<T> a) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = std::fabs(a[i]); return r; }
template <class T> static inline V4<T> vneg(V4<T> a) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = -a[i]; return r; }
template <class T> static inline V4<T> vsqrt(V4<T> a) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = std::sqrt(a[i]); return r; }
template <class T> static inline V4<T> vrsqrt(V4<T> a) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = (T)1 / std::sqrt(a[i]); return r; }
struct M4 { bool v[4]; bool &operator[](int i) { return v[i]; } bool operator[](int i) const { return v[i]; } };
template <class T> static inline M4 cmpge(V4<T> a, V4<T> b) { M4 r; for (int i = 0; i < 4; i++) r[i] = a[i] >= b[i]; return r; }
template <class T> static inline M4 cmpgt(V4<T> a, V4<T> b) { M4 r; for (int i = 0; i < 4; i++) r[i] = a[i] > b[i]; return r; }
template <class T> static inline M4 cmple(V4<T> a, V4<T> b) { M4 r; for (int i = 0; i < 4; i++) r[i] = a[i] <= b[i]; return r; }
template <class T> static inline M4 cmpne(V4<T> a, V4<T> b) { M4 r; for (int i = 0; i < 4; i++) r[i] = a[i] != b[i]; return r; }
static inline M4 operator&(M4 a, M4 b) { M4 r; for (int i = 0; i < 4; i++) r[i] = a[i] && b[i]; return r; }
static inline M4 operator~(M4 a) { M4 r; for (int i = 0; i < 4; i++) r[i] = !a[i]; return r; }
static inline bool any(M4 a) { return a[0] || a[1] || a[2] || a[3]; }
static inline bool all(M4 a) { return a[0] && a[1] && a[2] && a[3]; }
template <class T> static inline V4<T> select(M4 m, V4<T> a, V4<T> b) { V4<T> r; for (int i = 0; i < 4; i++) r[i] = m[i] ? a[i] : b[i]; return r; }
struct U4 { uint32_t v[4]; uint32_t &operator[](int i) { return v[i]; } uint32_t operator[](int i) const { return v[i]; } };
static inline U4 usplat(uint32_t s) { return {{s, s, s, s}}; }
static inline U4 operator+(U4 a, U4 b) { U4 r; for (int i = 0; i < 4; i++) r[i] = a[i] + b[i]; return r; }
static inline U4 operator*(U4 a, U4 b) { U4 r; for (int i = 0; i < 4; i++) r[i] = a[i] * b[i]; return r; }
template <class T> static inline U4 as_uvec4(V4<T> a) { U4 r; for (int i = 0; i < 4; i++) r[i] = as_u32(a[i]); return r; }

template <class T> struct V3x4 {
    V4<T> x, y, z;
    static V3x4 splat(T s) { return {V4<T>::splat(s), V4<T>::splat(s), V4<T>::splat(s)}; }
    static V3x4 from3(T a, T b, T c) { return {V4<T>::splat(a), V4<T>::splat(b), V4<T>::splat(c)}; }
};
template <class T> static inline V3x4<T> operator+(V3x4<T> a, V3x4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> static inline V3x4<T> operator-(V3x4<T> a, V3x4<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> static inline V3x4<T> operator*(V3x4<T> a, V3x4<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> static inline V3x4<T> operator*(V3x4<T> a, V4<T> s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> static inline V3x4<T> operator*(V3x4<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> static inline V3x4<T> operator+(V3x4<T> a, V4<T> s) { return {a.x + s, a.y + s, a.z + s}; }
template <class T> static inline V3x4<T> operator+(V3x4<T> a, T s) { return {a.x + s, a.y + s, a.z + s}; }
template <class T> static inline V3x4<T> operator-(V3x4<T> a, T s) { return {a.x - s, a.y - s, a.z - s}; }
template <class T> static inline V3x4<T> operator/(V3x4<T> a, V3x4<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> static inline V4<T> dot(V3x4<T> a, V3x4<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> static inline V3x4<T> normalize(V3x4<T> a) {
    V4<T> r = vrsqrt(dot(a, a));
    return {a.x * r, a.y * r, a.z * r};
}
template <class T> static inline V3x4<T> cross(V3x4<T> a, V3x4<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> static inline V3x4<T> reflect(V3x4<T> i, V3x4<T> n) { return i - (n * dot(i, n)) * (T)2; }
template <class T> static inline V3x4<T> vmax3(V3x4<T> a, V4<T> m) { return {vmax(a.x, m), vmax(a.y, m), vmax(a.z, m)}; }
template <class T> static inline V3x4<T> select3(M4 m, V3x4<T> a, V3x4<T> b) { return {select(m, a.x, b.x), select(m, a.y, b.y), select(m, a.z, b.z)}; }
template <class T> static inline V3x4<T> lerp3(V3x4<T> a, V3x4<T> b, V4<T> t) {
    return {a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t};
}
// srgb_to_linear_fast (src/util.rs:111-113)
template <class T> static inline V3x4<T> srgb_to_linear_fast(V3x4<T> x) {
    return x * (x * (x * (T)0.305306011f + (T)0.682171111f) + (T)0.012522878f);
}

// util.rs scalar conversions
static inline float srgb_to_linear_scalar(float s) {
    if (s <= 0.04045f) return s / 12.92f;
    return std::pow((s + 0.055f) / 1.055f, 2.4f);
}
static inline float linear_to_srgb_scalar(float s) {
    if (s <= 0.0031308f) return s * 12.92f;
    return std::pow(s, 1.0f / 2.4f) * 1.055f - 0.055f;
}
static inline Vec4s rgba8_unpack(uint32_t c) {
    return {(float)((c >> 24) & 0xFF) / 255.0f, (float)((c >> 16) & 0xFF) / 255.0f,
            (float)((c >> 8) & 0xFF) / 255.0f, (float)(c & 0xFF) / 255.0f};
}
static inline uint32_t rgba8_pack_vec4(Vec4s c) {
    return (as_u32(c.x * 255.0f) << 24) | (as_u32(c.y * 255.0f) << 16) |
           (as_u32(c.z * 255.0f) << 8) | as_u32(c.w * 255.0f);
}
static inline uint32_t rgba8_pack_u8(uint8_t r, uint8_t g, uint8_t b, uint8_t a) {
    return ((uint32_t)r << 24) | ((uint32_t)g << 16) | ((uint32_t)b << 8) | a;
}

// ----------------------------------------------------------------------------
// Textures and samplers (src/texture.rs)
// ----------------------------------------------------------------------------
struct Texture {
    uint32_t width = 0, height = 0;
    float whv[4] = {0, 0, 0, 0};
    int type = SR_TEX_LINEAR;
    std::vector<uint32_t> data;
    uint32_t max_mip_level = 0;
    float max_mip_level_f = 0.0f;
    std::vector<uint32_t> mip_offsets, mip_widths, mip_heights, array_stride;
    std::vector<float> mip_widths_f, mip_heights_f;
    Vec4s load(size_t idx) const { return rgba8_unpack(data[idx]); }
};

static void generate_mipmaps(Texture &t) {
    uint32_t num_mips = 1 + ilog2(std::max(t.width, t.height));
    for (uint32_t mip = 1; mip < num_mips; mip++) {
        t.mip_offsets.push_back((uint32_t)t.data.size());
        uint32_t mw = std::max(t.width >> mip, 1u), mh = std::max(t.height >> mip, 1u);
        t.array_stride.push_back(mw * mh);
        uint32_t prev_off = t.mip_offsets[mip - 1], pw = t.mip_widths[mip - 1], ph = t.mip_heights[mip - 1];
        for (uint32_t y = 0; y < mh; y++)
            for (uint32_t x = 0; x < mw; x++) {
                uint32_t x0 = x * 2, y0 = y * 2;
                uint32_t x1 = std::min(x0 + 1, pw - 1), y1 = std::min(y0 + 1, ph - 1);
                Vec4s p00 = t.load(prev_off + (y0 * pw + x0));
                Vec4s p10 = t.load(prev_off + (y0 * pw + x1));
                Vec4s p01 = t.load(prev_off + (y1 * pw + x0));
                Vec4s p11 = t.load(prev_off + (y1 * pw + x1));
                Vec4s avg;
                if (t.type == SR_TEX_SRGB) {
                    auto lin = [](Vec4s p) {
                        return Vec4s{srgb_to_linear_scalar(p.x), srgb_to_linear_scalar(p.y),
                                     srgb_to_linear_scalar(p.z), p.w};
                    };
                    Vec4s a = lin(p00), b = lin(p10), c = lin(p01), d = lin(p11);
                    Vec4s s = {((a.x + b.x) + c.x) + d.x, ((a.y + b.y) + c.y) + d.y,
                               ((a.z + b.z) + c.z) + d.z, ((a.w + b.w) + c.w) + d.w};
                    Vec4s q = {s.x / 4.0f, s.y / 4.0f, s.z / 4.0f, s.w / 4.0f};
                    avg = {linear_to_srgb_scalar(q.x), linear_to_srgb_scalar(q.y),
                           linear_to_srgb_scalar(q.z), q.w};
                } else if (t.type == SR_TEX_METALLIC_ROUGHNESS) {
                    float r = (((p00.y * p00.y + p10.y * p10.y) + p01.y * p01.y) + p11.y * p11.y) / 4.0f;
                    float m = (((p00.z + p10.z) + p01.z) + p11.z) / 4.0f;
                    avg = {p00.x, std::sqrt(r), m, p00.w};
                } else if (t.type == SR_TEX_NORMAL) {
                    auto n = [](Vec4s p) {
                        return Vec4s{p.x * 2.0f - 1.0f, p.y * 2.0f - 1.0f, p.z * 2.0f - 1.0f, p.w * 2.0f - 1.0f};
                    };
                    Vec4s a = n(p00), b = n(p10), c = n(p01), d = n(p11);
                    Vec4s s = {((a.x + b.x) + c.x) + d.x, ((a.y + b.y) + c.y) + d.y,
                               ((a.z + b.z) + c.z) + d.z, ((a.w + b.w) + c.w) + d.w};
                    Vec4s q = {s.x / 4.0f, s.y / 4.0f, s.z / 4.0f, s.w / 4.0f};
                    avg = {(q.x + 1.0f) / 2.0f, (q.y + 1.0f) / 2.0f, (q.z + 1.0f) / 2.0f, (q.w + 1.0f) / 2.0f};
                } else {
                    Vec4s s = {((p00.x + p10.x) + p01.x) + p11.x, ((p00.y + p10.y) + p01.y) + p11.y,
                               ((p00.z + p10.z) + p01.z) + p11.z, ((p00.w + p10.w) + p01.w) + p11.w};
                    avg = {s.x / 4.0f, s.y / 4.0f, s.z / 4.0f, s.w / 4.0f};
                }
                t.data.push_back(rgba8_pack_vec4(avg));
            }
        t.mip_widths.push_back(mw);
        t.mip_heights.push_back(mh);
        t.mip_widths_f.push_back((float)mw);
        t.mip_heights_f.push_back((float)mh);
        t.max_mip_level += 1;
        t.max_mip_level_f += 1.0f;
    }
}

static void texture_from_base(Texture &t, uint32_t w, uint32_t h, int type, const uint32_t *texels,
                              bool mips) {
    t.width = w; t.height = h; t.type = type;
    t.whv[0] = (float)w; t.whv[1] = (float)w; t.whv[2] = (float)h; t.whv[3] = (float)h;
    t.data.assign(texels, texels + (size_t)w * h);
    t.mip_offsets = {0}; t.mip_widths = {w}; t.mip_heights = {h};
    t.mip_widths_f = {(float)w}; t.mip_heights_f = {(float)h}; t.array_stride = {0};
    if (mips) generate_mipmaps(t);
}

struct Sampled {
    const Texture *tex = nullptr;
    int wrap_s = SR_WRAP_REPEAT, wrap_t = SR_WRAP_REPEAT;
};

template <class T> static inline V4<T> apply_wrap_mode(V4<T> texel, V4<T> dim, int mode) {
    V4<T> t2;
    if (mode == SR_WRAP_CLAMP_TO_EDGE) t2 = texel;
    else if (mode == SR_WRAP_REPEAT) t2 = texel - vfloor(texel / dim) * dim;
    else {
        V4<T> two = dim * V4<T>::splat((T)2);
        V4<T> t = texel - vfloor(texel / two) * two;
        t2 = vmin(t, two - t);
    }
    return vmin(t2, dim - V4<T>::splat((T)1));
}

template <class T> static V3x4<T> gather_rgb(const Texture &t, U4 idx) {
    V3x4<T> r;
    for (int i = 0; i < 4; i++) {
        Vec4s c = t.load(idx[i]);
        r.x[i] = (T)c.x; r.y[i] = (T)c.y; r.z[i] = (T)c.z;
    }
    return r;
}
template <class T> static V4<T> gather_alpha(const Texture &t, U4 idx) {
    V4<T> r;
    for (int i = 0; i < 4; i++) r[i] = (T)t.load(idx[i]).w;
    return r;
}

template <class T> static inline void cubemap_uv_from_normal(V3x4<T> n, V4<T> &uf, V4<T> &vf, U4 &slice) {
    V4<T> ax = vabs(n.x), ay = vabs(n.y), az = vabs(n.z);
    M4 mx = cmpge(ax, ay) & cmpge(ax, az);
    M4 my = cmpgt(ay, ax) & cmpge(ay, az);
    V4<T> one = V4<T>::splat(1), neg = V4<T>::splat(-1), zero = V4<T>::splat(0);
    V4<T> sx = select(cmpge(n.x, zero), one, neg);
    V4<T> sy = select(cmpge(n.y, zero), one, neg);
    V4<T> sz = select(cmpge(n.z, zero), one, neg);
    V4<T> u_x = vneg(n.z) * sx, v_x = vneg(n.y), d_x = ax;
    V4<T> u_y = n.x, v_y = n.z * sy, d_y = ay;
    V4<T> u_z = n.x * sz, v_z = vneg(n.y), d_z = az;
    V4<T> u = select(mx, u_x, select(my, u_y, u_z));
    V4<T> v = select(mx, v_x, select(my, v_y, v_z));
    V4<T> den = select(mx, d_x, select(my, d_y, d_z));
    den = vmax(den, V4<T>::splat((T)1.0e-19f));
    uf = (T)0.5 * (u / den + (T)1);
    vf = (T)0.5 * (v / den + (T)1);
    V4<T> ix = select(cmpge(n.x, zero), V4<T>::splat(0), V4<T>::splat(1));
    V4<T> iy = select(cmpge(n.y, zero), V4<T>::splat(2), V4<T>::splat(3));
    V4<T> iz = select(cmpge(n.z, zero), V4<T>::splat(4), V4<T>::splat(5));
    slice = as_uvec4(select(mx, ix, select(my, iy, iz)));
}

template <class T> static V3x4<T> sample_cubemap_rgb(const Texture &t, V3x4<T> n, U4 mip) {
    V4<T> u, v; U4 slice;
    cubemap_uv_from_normal(n, u, v, slice);
    U4 wi, off; V4<T> wf, hf;
    for (int i = 0; i < 4; i++) {
        uint32_t m = std::min(mip[i], t.max_mip_level);
        wi[i] = t.mip_widths[m];
        wf[i] = (T)t.mip_widths_f[m];
        hf[i] = (T)t.mip_heights_f[m];
        off[i] = t.mip_offsets[m] + slice[i] * t.array_stride[m];
    }
    V4<T> one = V4<T>::splat(1);
    V4<T> x = vround(u * (wf - one)), y = vround(v * (hf - one));
    U4 xi = as_uvec4(vclamp(x, V4<T>::splat(0), wf - one));
    U4 yi = as_uvec4(vclamp(y, V4<T>::splat(0), hf - one));
    return gather_rgb<T>(t, off + yi * wi + xi);
}

template <class T> static V3x4<T> sample_cubemap_trilinear_rgb(const Texture &t, V3x4<T> n, V4<T> mip_level) {
    V4<T> maxv = V4<T>::splat((T)t.max_mip_level_f);
    V4<T> mip = vclamp(mip_level, V4<T>::splat(0), maxv);
    V4<T> m0 = vfloor(mip);
    V4<T> m1 = vmin(m0 + (T)1, maxv);
    V4<T> tt = mip - m0;
    V3x4<T> c0 = sample_cubemap_rgb(t, n, as_uvec4(m0));
    V3x4<T> c1 = sample_cubemap_rgb(t, n, as_uvec4(m1));
    return {c0.x + (c1.x - c0.x) * tt, c0.y + (c1.y - c0.y) * tt, c0.z + (c1.z - c0.z) * tt};
}

template <class T> static uint32_t compute_mip_level(const Texture &t, V4<T> du_dv) {
    V4<T> d = du_dv * V4<T>{{(T)t.whv[0], (T)t.whv[1], (T)t.whv[2], (T)t.whv[3]}};
    T dx2 = d[0] * d[0] + d[2] * d[2];
    T dy2 = d[1] * d[1] + d[3] * d[3];
    T footprint = (dx2 + dy2) * (T)0.5;
    footprint = std::fmax(footprint, (T)1);
    uint32_t mip = ilog2(as_u32(footprint)) >> 1;
    return std::min(mip, t.max_mip_level);
}

template <class T> static V3x4<T> sample4_rgb(const Sampled &s, V4<T> u, V4<T> v, V4<T> du_dv) {
    const Texture &t = *s.tex;
    uint32_t mip = compute_mip_level(t, du_dv);
    U4 wi = usplat(t.mip_widths[mip]);
    V4<T> wf = V4<T>::splat((T)t.mip_widths_f[mip]), hf = V4<T>::splat((T)t.mip_heights_f[mip]);
    U4 off = usplat(t.mip_offsets[mip]);
    U4 x = as_uvec4(apply_wrap_mode(vfloor(u * wf), wf, s.wrap_s));
    U4 y = as_uvec4(apply_wrap_mode(vfloor(v * hf), hf, s.wrap_t));
    return gather_rgb<T>(t, off + y * wi + x);
}
template <class T> static V4<T> sample4_alpha(const Sampled &s, V4<T> u, V4<T> v, V4<T> du_dv) {
    const Texture &t = *s.tex;
    uint32_t mip = compute_mip_level(t, du_dv);
    U4 wi = usplat(t.mip_widths[mip]);
    V4<T> wf = V4<T>::splat((T)t.mip_widths_f[mip]), hf = V4<T>::splat((T)t.mip_heights_f[mip]);
    U4 off = usplat(t.mip_offsets[mip]);
    U4 x = as_uvec4(apply_wrap_mode(vfloor(u * wf), wf, s.wrap_s));
    U4 y = as_uvec4(apply_wrap_mode(vfloor(v * hf), hf, s.wrap_t));
    return gather_alpha<T>(t, off + y * wi + x);
}

template <class T> static V3x4<T> sample_bilinear_rgb(const Texture &t, int wrap_s, int wrap_t, V4<T> u, V4<T> v,
                                                      uint32_t mip, U4 slice) {
    U4 wi = usplat(t.mip_widths[mip]);
    V4<T> wf = V4<T>::splat((T)t.mip_widths_f[mip]), hf = V4<T>::splat((T)t.mip_heights_f[mip]);
    U4 off = usplat(t.mip_offsets[mip]) + slice * usplat(t.array_stride[mip]);
    V4<T> xf = u * wf - V4<T>::splat((T)0.5), yf = v * hf - V4<T>::splat((T)0.5);
    V4<T> x0 = vfloor(xf), y0 = vfloor(yf);
    V4<T> x1 = x0 + V4<T>::splat(1), y1 = y0 + V4<T>::splat(1);
    V4<T> fx = xf - x0, fy = yf - y0;
    V4<T> omfx = V4<T>::splat(1) - fx, omfy = V4<T>::splat(1) - fy;
    x0 = apply_wrap_mode(x0, wf, wrap_s);
    y0 = apply_wrap_mode(y0, hf, wrap_t);
    x1 = apply_wrap_mode(x1, wf, wrap_s);
    y1 = apply_wrap_mode(y1, hf, wrap_t);
    U4 x0i = as_uvec4(x0), y0i = as_uvec4(y0), x1i = as_uvec4(x1), y1i = as_uvec4(y1);
    V3x4<T> p00 = gather_rgb<T>(t, off + y0i * wi + x0i);
    V3x4<T> p10 = gather_rgb<T>(t, off + y0i * wi + x1i);
    V3x4<T> p01 = gather_rgb<T>(t, off + y1i * wi + x0i);
    V3x4<T> p11 = gather_rgb<T>(t, off + y1i * wi + x1i);
    V4<T> w00 = omfx * omfy, w10 = fx * omfy, w01 = omfx * fy, w11 = fx * fy;
    return {((p00.x * w00 + p10.x * w10) + p01.x * w01) + p11.x * w11,
            ((p00.y * w00 + p10.y * w10) + p01.y * w01) + p11.y * w11,
            ((p00.z * w00 + p10.z * w10) + p01.z * w01) + p11.z * w11};
}

// ----------------------------------------------------------------------------
// Voxel grid lookup (src/voxelgrid.rs:264-368)
// ----------------------------------------------------------------------------
struct VoxelGrid {
    int w = 1, h = 1, d = 1;
    Vec3 wmin{0, 0, 0}, wmax{1, 1, 1};
    std::vector<float> light;           // w0 (constant-SH form)
    std::vector<float> sh4;             // full form: 16 floats per voxel, or empty
    float sh_const[4][4];               // constant form coefficients (rgb, w=0 except [0].w per voxel, [1].w = sky)
    Vec3 voxel_size() const {
        Vec3 s = wmax - wmin;
        return v3(s.x / (float)w, s.y / (float)h, s.z / (float)d);
    }
    // coefficient c of voxel index i
    Vec4s coeff(size_t i, int c) const {
        if (!sh4.empty()) { const float *p = &sh4[i * 16 + c * 4]; return {p[0], p[1], p[2], p[3]}; }
        Vec4s r = {sh_const[c][0], sh_const[c][1], sh_const[c][2], sh_const[c][3]};
        if (c == 0) r.w = light[i];
        return r;
    }
};

template <class T>
static void voxel_lookup(const VoxelGrid &g, V3x4<T> pos, V3x4<T> rgb[4], V4<T> wout[4]) {
    Vec3 vs = g.voxel_size();
    V4<T> vx = (pos.x - (T)g.wmin.x) / V4<T>::splat((T)vs.x);
    V4<T> vy = (pos.y - (T)g.wmin.y) / V4<T>::splat((T)vs.y);
    V4<T> vz = (pos.z - (T)g.wmin.z) / V4<T>::splat((T)vs.z);
    V4<T> x0f = vfloor(vx), y0f = vfloor(vy), z0f = vfloor(vz);
    uint64_t x0[4], y0[4], z0[4], x1[4], y1[4], z1[4];
    for (int i = 0; i < 4; i++) {
        x0[i] = std::min<uint64_t>(as_usize(x0f[i]), (uint64_t)(g.w - 1));
        y0[i] = std::min<uint64_t>(as_usize(y0f[i]), (uint64_t)(g.h - 1));
        z0[i] = std::min<uint64_t>(as_usize(z0f[i]), (uint64_t)(g.d - 1));
        x1[i] = std::min<uint64_t>(x0[i] + 1, (uint64_t)(g.w - 1));
        y1[i] = std::min<uint64_t>(y0[i] + 1, (uint64_t)(g.h - 1));
        z1[i] = std::min<uint64_t>(z0[i] + 1, (uint64_t)(g.d - 1));
    }
    V4<T> fx = vx - x0f, fy = vy - y0f, fz = vz - z0f;
    V4<T> one = V4<T>::splat(1);
    auto idx = [&](const uint64_t *X, const uint64_t *Y, const uint64_t *Z, int i) {
        return (size_t)(Z[i] * (uint64_t)g.w * (uint64_t)g.h + Y[i] * (uint64_t)g.w + X[i]);
    };
    for (int c = 0; c < 4; c++) {
        V4<T> s[8][4];  // [corner][channel]
        const uint64_t *xs[2] = {x0, x1}, *ys[2] = {y0, y1}, *zs[2] = {z0, z1};
        for (int cx = 0; cx < 2; cx++)
            for (int cy = 0; cy < 2; cy++)
                for (int cz = 0; cz < 2; cz++) {
                    int k = cx * 4 + cy * 2 + cz;  // v{x}{y}{z}
                    for (int i = 0; i < 4; i++) {
                        Vec4s q = g.coeff(idx(xs[cx], ys[cy], zs[cz], i), c);
                        s[k][0][i] = (T)q.x; s[k][1][i] = (T)q.y; s[k][2][i] = (T)q.z; s[k][3][i] = (T)q.w;
                    }
                }
        V4<T> out[4];
        for (int ch = 0; ch < 4; ch++) {
            V4<T> v000 = s[0][ch], v001 = s[1][ch], v010 = s[2][ch], v011 = s[3][ch];
            V4<T> v100 = s[4][ch], v101 = s[5][ch], v110 = s[6][ch], v111 = s[7][ch];
            V4<T> v00 = v000 * (one - fx) + v100 * fx;
            V4<T> v01 = v001 * (one - fx) + v101 * fx;
            V4<T> v10 = v010 * (one - fx) + v110 * fx;
            V4<T> v11 = v011 * (one - fx) + v111 * fx;
            V4<T> v0 = v00 * (one - fy) + v10 * fy;
            V4<T> v1 = v01 * (one - fy) + v11 * fy;
            out[ch] = v0 * (one - fz) + v1 * fz;
        }
        rgb[c] = {out[0], out[1], out[2]};
        wout[c] = out[3];
    }
}

// ----------------------------------------------------------------------------
// Scene
// ----------------------------------------------------------------------------
struct BoundingSphere { Vec3 center{0, 0, 0}; float radius = 0.0f; };

struct Material {
    float base_color_factor[4];
    float metallic, roughness;
    float emissive[3];
    float occlusion_strength;
    float alpha_cutoff;
    float transmission;
    bool is_alpha_tested, is_translucent;
    Sampled base, mr, normal, emissive_tex, occlusion, transmission_tex;
    bool has_base, has_mr, has_normal, has_emissive, has_occlusion, has_transmission_tex;
};

struct Primitive {
    std::vector<Vec4s> positions;
    std::vector<Vec3> normals;
    std::vector<Vec4s> tangents;
    std::vector<Vec2> texcoords;
    std::vector<uint32_t> indices;
    int material = 0;
    BoundingSphere bs;
};
struct Mesh { std::vector<int> prims, opaque, translucent; };
struct Node { Mat4 transform; int mesh = -1; BoundingSphere bs_world; };

struct Scene {
    std::vector<Primitive> prims;
    std::vector<Mesh> meshes;
    std::vector<Node> nodes;
    std::vector<Material> materials;
    std::vector<Texture> textures;
    Texture cubemap, specular, lut;
    VoxelGrid voxels;
    Vec3 light_dir, light_color;
};

static BoundingSphere compute_bounding_sphere(const std::vector<Vec4s> &pos) {
    Vec3 mn = v3(INFINITY, INFINITY, INFINITY), mx = v3(-INFINITY, -INFINITY, -INFINITY);
    for (auto &p : pos) {
        mn = v3(minps(mn.x, p.x), minps(mn.y, p.y), minps(mn.z, p.z));
        mx = v3(maxps(mx.x, p.x), maxps(mx.y, p.y), maxps(mx.z, p.z));
    }
    BoundingSphere b;
    b.center = (mn + mx) * 0.5f;
    b.radius = length3(mx - mn) / 2.0f;
    return b;
}
static inline float length4(Vec4s a) { return std::sqrt(dot4(a, a)); }
static BoundingSphere transform_sphere(const Mat4 &m, const BoundingSphere &s) {
    float max_scale = ((length4(m.col(0)) + length4(m.col(1))) + length4(m.col(2))) / 3.0f;
    Vec4s c = mul(m, Vec4s{s.center.x, s.center.y, s.center.z, 1.0f});
    BoundingSphere r;
    r.center = v3(c.x, c.y, c.z);
    r.radius = s.radius * max_scale;
    return r;
}
static void grow_by_sphere(BoundingSphere &a, const BoundingSphere &b) {
    float d = length3(a.center - b.center);
    if (d + b.radius > a.radius) a.radius = d + b.radius;
}

// compute_tangents (src/scene.rs:548-635)
static std::vector<Vec4s> compute_tangents(const std::vector<Vec4s> &pos, const std::vector<Vec2> &uv,
                                           const std::vector<Vec3> &nrm, const std::vector<uint32_t> &idx) {
    size_t n = pos.size();
    std::vector<Vec3> tan(n, v3(0, 0, 0)), bit(n, v3(0, 0, 0));
    std::vector<Vec4s> out(n, Vec4s{0, 0, 0, 0});
    for (size_t t = 0; t + 2 < idx.size(); t += 3) {
        uint32_t i0 = idx[t], i1 = idx[t + 1], i2 = idx[t + 2];
        Vec3 p0 = v3(pos[i0].x, pos[i0].y, pos[i0].z), p1 = v3(pos[i1].x, pos[i1].y, pos[i1].z),
             p2 = v3(pos[i2].x, pos[i2].y, pos[i2].z);
        Vec2 uv0 = uv[i0], uv1 = uv[i1], uv2 = uv[i2];
        Vec3 e1 = p1 - p0, e2 = p2 - p0;
        Vec2 d1 = {uv1.x - uv0.x, uv1.y - uv0.y}, d2 = {uv2.x - uv0.x, uv2.y - uv0.y};
        float det = d1.x * d2.y - d2.x * d1.y;
        if (std::fabs(det) < 1e-6f) continue;
        Vec3 tg = (e1 * d2.y - e2 * d1.y);
        tg = v3(tg.x / det, tg.y / det, tg.z / det);
        Vec3 bt = (e2 * d1.x - e1 * d2.x);
        bt = v3(bt.x / det, bt.y / det, bt.z / det);
        tan[i0] = tan[i0] + tg; tan[i1] = tan[i1] + tg; tan[i2] = tan[i2] + tg;
        bit[i0] = bit[i0] + bt; bit[i1] = bit[i1] + bt; bit[i2] = bit[i2] + bt;
    }
    for (size_t i = 0; i < n; i++) {
        Vec3 nn = nrm[i], t = tan[i];
        if (dot3(t, t) < 1e-6f) {
            Vec3 helper = std::fabs(nn.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
            Vec3 tg = normalize3(cross3(nn, helper));
            out[i] = {tg.x, tg.y, tg.z, 1.0f};
            continue;
        }
        Vec3 tg = normalize3(t - nn * dot3(nn, t));
        Vec3 b = bit[i];
        float hand = dot3(cross3(nn, tg), b) < 0.0f ? -1.0f : 1.0f;
        out[i] = {tg.x, tg.y, tg.z, hand};
    }
    return out;
}

static Sampled make_sampled(const Scene &s, const sr_texref &r) {
    Sampled o;
    o.tex = &s.textures[r.texture];
    o.wrap_s = r.wrap_s; o.wrap_t = r.wrap_t;
    return o;
}

static Scene *build_scene(const sr_scene_desc *d) {
    Scene *s = new Scene();
    s->textures.resize(d->num_textures);
    for (int i = 0; i < d->num_textures; i++) {
        const sr_texture_desc &t = d->textures[i];
        texture_from_base(s->textures[i], t.width, t.height, t.type, t.texels, true);
    }
    s->materials.resize(d->num_materials);
    for (int i = 0; i < d->num_materials; i++) {
        const sr_material_desc &m = d->materials[i];
        Material &o = s->materials[i];
        memcpy(o.base_color_factor, m.base_color_factor, 16);
        o.metallic = m.metallic_factor; o.roughness = m.roughness_factor;
        memcpy(o.emissive, m.emissive_factor, 12);
        o.occlusion_strength = m.occlusion_strength;
        o.alpha_cutoff = m.alpha_cutoff;
        o.transmission = m.transmission;
        o.is_translucent = m.is_translucent != 0;
        o.has_base = m.base_color.texture >= 0;
        o.has_mr = m.metallic_roughness.texture >= 0;
        o.has_normal = m.normal.texture >= 0;
        o.has_emissive = m.emissive.texture >= 0;
        o.has_occlusion = m.occlusion.texture >= 0;
        o.has_transmission_tex = m.transmission_texture.texture >= 0;
        if (o.has_base) o.base = make_sampled(*s, m.base_color);
        if (o.has_mr) o.mr = make_sampled(*s, m.metallic_roughness);
        if (o.has_normal) o.normal = make_sampled(*s, m.normal);
        if (o.has_emissive) o.emissive_tex = make_sampled(*s, m.emissive);
        if (o.has_occlusion) o.occlusion = make_sampled(*s, m.occlusion);
        if (o.has_transmission_tex) o.transmission_tex = make_sampled(*s, m.transmission_texture);
        o.is_alpha_tested = (m.alpha_mode == SR_ALPHA_MASK) && o.has_base;
    }
    s->prims.resize(d->num_primitives);
    for (int i = 0; i < d->num_primitives; i++) {
        const sr_primitive_desc &p = d->primitives[i];
        Primitive &o = s->prims[i];
        o.positions.resize(p.num_vertices);
        o.normals.resize(p.num_vertices);
        o.texcoords.resize(p.num_vertices);
        for (uint32_t v = 0; v < p.num_vertices; v++) {
            o.positions[v] = {p.positions[3 * v], p.positions[3 * v + 1], p.positions[3 * v + 2], 1.0f};
            o.normals[v] = v3(p.normals[3 * v], p.normals[3 * v + 1], p.normals[3 * v + 2]);
            o.texcoords[v] = p.texcoords ? Vec2{p.texcoords[2 * v], p.texcoords[2 * v + 1]} : Vec2{0, 0};
        }
        o.indices.assign(p.indices, p.indices + p.num_indices);
        if (p.tangents) {
            o.tangents.resize(p.num_vertices);
            for (uint32_t v = 0; v < p.num_vertices; v++)
                o.tangents[v] = {p.tangents[4 * v], p.tangents[4 * v + 1], p.tangents[4 * v + 2], p.tangents[4 * v + 3]};
        } else {
            o.tangents = compute_tangents(o.positions, o.texcoords, o.normals, o.indices);
        }
        o.material = p.material;
        o.bs = compute_bounding_sphere(o.positions);
    }
    s->meshes.resize(d->num_meshes);
    for (int i = 0; i < d->num_meshes; i++) {
        Mesh &m = s->meshes[i];
        for (int k = 0; k < d->meshes[i].num_primitives; k++) {
            int pi = d->meshes[i].first_primitive + k;
            m.prims.push_back(pi);
            if (s->materials[s->prims[pi].material].is_translucent) m.translucent.push_back(k);
            else m.opaque.push_back(k);
        }
    }
    s->nodes.resize(d->num_nodes);
    for (int i = 0; i < d->num_nodes; i++) {
        Node &n = s->nodes[i];
        memcpy(n.transform.m, d->nodes[i].transform, 64);
        n.mesh = d->nodes[i].mesh;
        Vec4s c = mul(n.transform, Vec4s{0, 0, 0, 1});
        n.bs_world.center = v3(c.x, c.y, c.z);
        n.bs_world.radius = 0.0f;
        if (n.mesh >= 0)
            for (int pi : s->meshes[n.mesh].prims) grow_by_sphere(n.bs_world, transform_sphere(n.transform, s->prims[pi].bs));
    }
    const sr_environment *e = d->environment;
    texture_from_base(s->cubemap, e->cubemap_size, e->cubemap_size, SR_TEX_LINEAR, e->cubemap, false);
    s->cubemap.data.assign(e->cubemap, e->cubemap + (size_t)6 * e->cubemap_size * e->cubemap_size);
    s->cubemap.array_stride = {e->cubemap_size * e->cubemap_size};
    {
        Texture &t = s->specular;
        uint32_t sz = e->specular_size, nm = e->specular_mips;
        t.width = sz; t.height = sz; t.type = SR_TEX_LINEAR;
        t.whv[0] = t.whv[1] = t.whv[2] = t.whv[3] = (float)sz;
        t.data.assign(e->specular, e->specular + (size_t)nm * 6 * sz * sz);
        t.max_mip_level = nm - 1; t.max_mip_level_f = (float)(nm - 1);
        for (uint32_t m = 0; m < nm; m++) {
            t.mip_offsets.push_back(m * 6 * sz * sz);
            t.mip_widths.push_back(sz); t.mip_heights.push_back(sz);
            t.mip_widths_f.push_back((float)sz); t.mip_heights_f.push_back((float)sz);
            t.array_stride.push_back(sz * sz);
        }
    }
    texture_from_base(s->lut, e->brdf_lut_size, e->brdf_lut_size, SR_TEX_LINEAR, e->brdf_lut, false);
    const sr_voxel_grid *vg = d->voxels;
    VoxelGrid &g = s->voxels;
    g.w = vg->dims[0]; g.h = vg->dims[1]; g.d = vg->dims[2];
    g.wmin = v3(vg->world_min[0], vg->world_min[1], vg->world_min[2]);
    g.wmax = v3(vg->world_max[0], vg->world_max[1], vg->world_max[2]);
    size_t nv = (size_t)g.w * g.h * g.d;
    if (vg->sh4) {
        g.sh4.assign(vg->sh4, vg->sh4 + nv * 16);
    } else {
        g.light.assign(vg->light_intensity, vg->light_intensity + nv);
        for (int c = 0; c < 4; c++) {
            for (int k = 0; k < 3; k++) g.sh_const[c][k] = e->irradiance_sh[c][k] * vg->sh_scale;
            g.sh_const[c][3] = 0.0f;
        }
        g.sh_const[1][3] = vg->sky_visibility;
    }
    s->light_dir = v3(d->light.direction[0], d->light.direction[1], d->light.direction[2]);
    s->light_color = v3(d->light.color[0], d->light.color[1], d->light.color[2]);
    return s;
}

// ----------------------------------------------------------------------------
// Renderer (src/renderer.rs) and tile rasterizer (src/tilerasterizer.rs)
// ----------------------------------------------------------------------------
static const int TILE_SIZE = 64;
static const int SUBPIXEL_SCALE = 16;
static const float SUBPIXEL_SCALE_F = 16.0f;
static const int SUBPIXEL_SHIFT = 4;
static const int STEP_X_SIZE = 2 * SUBPIXEL_SCALE, STEP_Y_SIZE = 2 * SUBPIXEL_SCALE;
static const int HALF_PIXEL = SUBPIXEL_SCALE / 2;
static const int ONE_HALF_PIXEL = SUBPIXEL_SCALE + HALF_PIXEL;
static const int COARSE_BLOCK_SIZE_PIXELS = 16;
static const int COARSE_BLOCK_SIZE_SUBPIXELS = COARSE_BLOCK_SIZE_PIXELS * SUBPIXEL_SCALE;
static const float DEFAULT_EXPOSURE = 2.0f;

struct Vertex { Vec4s pos_clip; Vec3 pos_world; Vec3 normal; Vec4s tangent; Vec2 uv; };
struct Triangle { Vertex v[3]; int mesh_index, primitive_index; };

struct SInterp { float a, da, db; };
struct VInterp { Vec3 a, da, db; };
static inline SInterp sinterp(float a, float b, float c) { return {a, b - a, c - a}; }
static inline VInterp vinterp(Vec3 a, Vec3 b, Vec3 c) { return {a, b - a, c - a}; }
template <class T> static inline V4<T> interp(const SInterp &s, V4<T> b1, V4<T> b2) {
    return ((T)s.a + b1 * (T)s.da) + b2 * (T)s.db;
}
template <class T> static inline V3x4<T> interp(const VInterp &s, V4<T> b1, V4<T> b2) {
    return {((T)s.a.x + b1 * (T)s.da.x) + b2 * (T)s.db.x, ((T)s.a.y + b1 * (T)s.da.y) + b2 * (T)s.db.y,
            ((T)s.a.z + b1 * (T)s.da.z) + b2 * (T)s.db.z};
}

struct RasterPacket {
    IVec2 smin, smax;
    IVec2 p[3];
    SInterp z_over_w, one_over_w;
    VInterp normals, tangents;
    SInterp tangent_sign, u_over_w, v_over_w;
    VInterp pos_world_over_w;
    float one_over_area;
    uint32_t mesh_index, primitive_index;
    float avg_z;
    float du_dv[4];
    uint32_t tri_id;  // canonical post-clip triangle id
};

struct Tile {
    IVec2 smin, smax;
    std::vector<RasterPacket> opaque, translucent;
    std::mutex lock;
    V3x4<float> color[1024];
    V4<float> depth[1024];
    U4 packet_index[1024];
    V4<float> bary1[1024], bary2[1024];
    V3x4<double> color64[1024];
    M4 written[1024];  // some fragment passed the depth test this frame
    float center_luminance = 1.0f;
};

// Renderer::update_auto_exposure (src/renderer.rs:258-290) on the tiles'
// centre luminances: trimmed mean of their log2 (10% cut at each end of the
// f32::total_cmp order), the target exposure that maps it to the metered
// grey, and the exponential step of ev towards it (tau = 1 s).  Updates ev
// and target; returns the exposure 2^ev.
static float exposure_step(const float *lum, size_t n, float dt, float &ev, float &target_out) {
    std::vector<float> s(n);
    for (size_t i = 0; i < n; i++) s[i] = std::log2(std::fmax(lum[i], 1e-4f));
    // f32::total_cmp order
    auto key = [](float f) {
        int32_t b; memcpy(&b, &f, 4);
        return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
    };
    std::sort(s.begin(), s.end(), [&](float a, float b) { return key(a) < key(b); });
    size_t trim = (size_t)std::floor((float)n * 0.10f);
    trim = std::min(trim, (n - 1) / 2);
    float sum = 0.0f;
    for (size_t i = trim; i < n - trim; i++) sum += s[i];
    float mean = sum / (float)(n - 2 * trim);
    // tonemap_inverse_scalar(0.45) (src/util.rs:43-47)
    float y = std::fmin(std::fmax(0.45f, 0.0f), 1.0f);
    float k = 0.2f;
    float denom = std::fmax(1.0f + k - y, 1e-6f);
    float meter = std::fmax((y * k) / denom, 1e-4f);
    float target_ev = std::log2(meter) - mean;
    float target = std::fmin(std::fmax(std::pow(2.0f, target_ev), 0.05f), 32.0f);
    target_out = target;
    float tev = std::log2(target);
    float tau = std::fmax(1.0f, 1e-4f);
    float alpha = 1.0f - std::exp(-(std::fmax(dt, 0.0f) / tau));
    ev += (tev - ev) * alpha;
    return std::pow(2.0f, ev);
}

struct Renderer;
static void render_tile(Renderer &r, Tile &t, const Scene &s, const sr_camera &c, bool fp64);

struct Renderer {
    int width, height, width_sub, height_sub;
    float width_f, height_f;
    int tiles_x, tiles_y;
    std::vector<Tile *> tiles;
    float auto_exposure = DEFAULT_EXPOSURE, auto_exposure_target = DEFAULT_EXPOSURE;
    float auto_exposure_ev;
    bool threaded = false;
    uint64_t setup_count = 0;
    double ms_clipbin = 0, ms_raster = 0;
    Renderer(int w, int h) : width(w), height(h) {
        width_sub = w * SUBPIXEL_SCALE; height_sub = h * SUBPIXEL_SCALE;
        width_f = (float)w; height_f = (float)h;
        tiles_x = (w + TILE_SIZE - 1) / TILE_SIZE;
        tiles_y = (h + TILE_SIZE - 1) / TILE_SIZE;
        for (int y = 0; y < tiles_y; y++)
            for (int x = 0; x < tiles_x; x++) {
                Tile *t = new Tile();
                t->smin = {x * TILE_SIZE, y * TILE_SIZE};
                t->smax = {(x + 1) * TILE_SIZE, (y + 1) * TILE_SIZE};
                for (int q = 0; q < 1024; q++) {
                    t->depth[q] = V4<float>::splat(INFINITY);
                    t->packet_index[q] = usplat(0);
                    t->bary1[q] = t->bary2[q] = V4<float>::splat(0);
                    t->color[q] = V3x4<float>::splat(0);
                }
                tiles.push_back(t);
            }
        auto_exposure_ev = std::log2(DEFAULT_EXPOSURE);
    }
    ~Renderer() { for (auto *t : tiles) delete t; }

    IVec2 clip_to_screen_subpixels(Vec4s v) const {
        float nx = v.x / v.w, ny = v.y / v.w;
        float sx = (nx + 1.0f) * width_f / 2.0f;
        float sy = (1.0f - ny) * height_f / 2.0f;
        return {as_i32(std::round(sx * SUBPIXEL_SCALE_F)), as_i32(std::round(sy * SUBPIXEL_SCALE_F))};
    }

    void bin_triangle(const Triangle &tri, bool opaque, uint32_t tri_id) {
        IVec2 p0 = clip_to_screen_subpixels(tri.v[0].pos_clip);
        IVec2 p1 = clip_to_screen_subpixels(tri.v[1].pos_clip);
        IVec2 p2 = clip_to_screen_subpixels(tri.v[2].pos_clip);
        int32_t signed_area = wsub(wmul(wsub(p1.x, p0.x), wsub(p2.y, p0.y)), wmul(wsub(p2.x, p0.x), wsub(p1.y, p0.y)));
        if (signed_area > 0) return;
        IVec2 mn_sub = {std::max(std::min(std::min(p0.x, p1.x), p2.x), 0), std::max(std::min(std::min(p0.y, p1.y), p2.y), 0)};
        IVec2 mx_sub = {std::min(std::max(std::max(p0.x, p1.x), p2.x), width_sub),
                        std::min(std::max(std::max(p0.y, p1.y), p2.y), height_sub)};
        IVec2 mn = {mn_sub.x >> SUBPIXEL_SHIFT, mn_sub.y >> SUBPIXEL_SHIFT};
        IVec2 mx = {wadd(mx_sub.x, SUBPIXEL_SCALE) >> SUBPIXEL_SHIFT, wadd(mx_sub.y, SUBPIXEL_SCALE) >> SUBPIXEL_SHIFT};
        float one_over_area = 1.0f / (float)wabs(signed_area);
        float oow[3], zow[3];
        for (int i = 0; i < 3; i++) oow[i] = 1.0f / tri.v[i].pos_clip.w;
        for (int i = 0; i < 3; i++) zow[i] = tri.v[i].pos_clip.z * oow[i];
        Vec2 uvw[3];
        Vec3 pww[3];
        for (int i = 0; i < 3; i++) {
            uvw[i] = {tri.v[i].uv.x * oow[i], tri.v[i].uv.y * oow[i]};
            pww[i] = tri.v[i].pos_world * oow[i];
        }
        float dx1 = (float)wsub(p1.x, p0.x), dx2 = (float)wsub(p2.x, p0.x);
        float dy1 = (float)wsub(p1.y, p0.y), dy2 = (float)wsub(p2.y, p0.y);
        float du1 = uvw[1].x - uvw[0].x, du2 = uvw[2].x - uvw[0].x;
        float dv1 = uvw[1].y - uvw[0].y, dv2 = uvw[2].y - uvw[0].y;
        float du_dx = du1 * dy2 - du2 * dy1;
        float du_dy = du2 * dx1 - du1 * dx2;
        float dv_dx = dv1 * dy2 - dv2 * dy1;
        float dv_dy = dv2 * dx1 - dv1 * dx2;
        float scale = one_over_area * SUBPIXEL_SCALE_F;
        int min_bin_x = mn.x / TILE_SIZE, min_bin_y = mn.y / TILE_SIZE;
        int max_bin_x = wadd(mx.x, TILE_SIZE - 1) / TILE_SIZE, max_bin_y = wadd(mx.y, TILE_SIZE - 1) / TILE_SIZE;
        float avg_z = opaque ? 0.0f : ((tri.v[0].pos_clip.z + tri.v[1].pos_clip.z) + tri.v[2].pos_clip.z) / 3.0f;
        RasterPacket pk;
        pk.p[0] = p0; pk.p[1] = p1; pk.p[2] = p2;
        pk.z_over_w = sinterp(zow[0], zow[1], zow[2]);
        pk.one_over_w = sinterp(oow[0], oow[1], oow[2]);
        pk.normals = vinterp(tri.v[0].normal, tri.v[1].normal, tri.v[2].normal);
        pk.tangents = vinterp(v3(tri.v[0].tangent.x, tri.v[0].tangent.y, tri.v[0].tangent.z),
                              v3(tri.v[1].tangent.x, tri.v[1].tangent.y, tri.v[1].tangent.z),
                              v3(tri.v[2].tangent.x, tri.v[2].tangent.y, tri.v[2].tangent.z));
        pk.tangent_sign = sinterp(tri.v[0].tangent.w, tri.v[1].tangent.w, tri.v[2].tangent.w);
        pk.u_over_w = sinterp(uvw[0].x, uvw[1].x, uvw[2].x);
        pk.v_over_w = sinterp(uvw[0].y, uvw[1].y, uvw[2].y);
        pk.pos_world_over_w = vinterp(pww[0], pww[1], pww[2]);
        pk.one_over_area = one_over_area;
        pk.mesh_index = (uint32_t)tri.mesh_index;
        pk.primitive_index = (uint32_t)tri.primitive_index;
        pk.avg_z = avg_z;
        pk.du_dv[0] = du_dx * scale; pk.du_dv[1] = du_dy * scale;
        pk.du_dv[2] = dv_dx * scale; pk.du_dv[3] = dv_dy * scale;
        pk.tri_id = tri_id;
        for (int y = min_bin_y; y < max_bin_y; y++)
            for (int x = min_bin_x; x < max_bin_x; x++) {
                int bin_index = wadd(wmul(y, tiles_x), x);
                if (bin_index >= 0 && bin_index < (int)tiles.size()) {
                    Tile &t = *tiles[bin_index];
                    IVec2 cmn = {std::max(mn.x, t.smin.x), std::max(mn.y, t.smin.y)};
                    IVec2 cmx = {std::min(mx.x, t.smax.x), std::min(mx.y, t.smax.y)};
                    if (cmx.x - cmn.x < 1 || cmx.y - cmn.y < 1) continue;
                    pk.smin = cmn; pk.smax = cmx;
                    if (threaded) {
                        std::lock_guard<std::mutex> g(t.lock);
                        (opaque ? t.opaque : t.translucent).push_back(pk);
                    } else {
                        (opaque ? t.opaque : t.translucent).push_back(pk);
                    }
                }
            }
    }

    // Returns the number of fan triangles emitted (bin_triangle calls).
    uint32_t clip_against_frustum(const Triangle &tri, bool opaque, uint32_t id_base) {
        static const Vec4s PLANES[6] = {{0, 0, 1, 1}, {0, 0, -1, 1}, {1, 0, 0, 1},
                                        {-1, 0, 0, 1}, {0, 1, 0, 1}, {0, -1, 0, 1}};
        Vertex poly[16], np[16];
        int len = 3;
        poly[0] = tri.v[0]; poly[1] = tri.v[1]; poly[2] = tri.v[2];
        for (int pi = 0; pi < 6; pi++) {
            const Vec4s &pl = PLANES[pi];
            if (len == 0) return 0;
            int nl = 0;
            for (int i = 0; i < len; i++) {
                const Vertex &cur = poly[i];
                const Vertex &prev = poly[(i + len - 1) % len];
                bool ci = dot4(pl, cur.pos_clip) >= 0.0f;
                bool pin = dot4(pl, prev.pos_clip) >= 0.0f;
                auto isect = [&](const Vertex &v0, const Vertex &v1) {
                    float d0 = dot4(pl, v0.pos_clip), d1 = dot4(pl, v1.pos_clip);
                    float t = d0 / (d0 - d1);
                    Vertex o;
                    o.pos_clip = {v0.pos_clip.x + (v1.pos_clip.x - v0.pos_clip.x) * t,
                                  v0.pos_clip.y + (v1.pos_clip.y - v0.pos_clip.y) * t,
                                  v0.pos_clip.z + (v1.pos_clip.z - v0.pos_clip.z) * t,
                                  v0.pos_clip.w + (v1.pos_clip.w - v0.pos_clip.w) * t};
                    o.pos_world = v0.pos_world + (v1.pos_world - v0.pos_world) * t;
                    o.normal = v0.normal + (v1.normal - v0.normal) * t;
                    o.tangent = {v0.tangent.x + (v1.tangent.x - v0.tangent.x) * t,
                                 v0.tangent.y + (v1.tangent.y - v0.tangent.y) * t,
                                 v0.tangent.z + (v1.tangent.z - v0.tangent.z) * t,
                                 v0.tangent.w + (v1.tangent.w - v0.tangent.w) * t};
                    o.uv = {v0.uv.x + (v1.uv.x - v0.uv.x) * t, v0.uv.y + (v1.uv.y - v0.uv.y) * t};
                    return o;
                };
                if (ci) {
                    if (!pin) np[nl++] = isect(prev, cur);
                    np[nl++] = cur;
                } else if (pin) {
                    np[nl++] = isect(prev, cur);
                }
            }
            for (int i = 0; i < nl; i++) poly[i] = np[i];
            len = nl;
        }
        if (len < 3) return 0;
        for (int i = 1; i < len - 1; i++) {
            Triangle t;
            t.v[0] = poly[0]; t.v[1] = poly[i]; t.v[2] = poly[i + 1];
            t.mesh_index = tri.mesh_index; t.primitive_index = tri.primitive_index;
            bin_triangle(t, opaque, id_base + (uint32_t)(i - 1));
        }
        return (uint32_t)(len - 2);
    }

    Triangle assemble(const Primitive &p, const Mat4 &model, const Mat4 &mvp, size_t t, int mi, int pi) const {
        Triangle tri;
        tri.mesh_index = mi; tri.primitive_index = pi;
        for (int k = 0; k < 3; k++) {
            uint32_t i = p.indices[t * 3 + k];
            Vec4s local = p.positions[i];
            Vec4s w = mul(model, local);
            tri.v[k].pos_world = v3(w.x, w.y, w.z);
            tri.v[k].pos_clip = mul(mvp, local);
            tri.v[k].normal = mul3(model, p.normals[i]);
            Vec3 tw = mul3(model, v3(p.tangents[i].x, p.tangents[i].y, p.tangents[i].z));
            tri.v[k].tangent = {tw.x, tw.y, tw.z, p.tangents[i].w};
            tri.v[k].uv = p.texcoords[i];
        }
        return tri;
    }

    enum { INSIDE, OUTSIDE, INTERSECTING };
    static int test_sphere_frustum(const BoundingSphere &s, const sr_camera &c) {
        Mat4 view;
        memcpy(view.m, c.view_matrix, 64);
        Vec4s cv = mul(view, Vec4s{s.center.x, s.center.y, s.center.z, 1.0f});
        int result = INSIDE;
        for (int i = 0; i < 6; i++) {
            Vec4s pl = {c.view_clip_planes[i][0], c.view_clip_planes[i][1], c.view_clip_planes[i][2], c.view_clip_planes[i][3]};
            float dist = dot4(pl, cv);
            if (dist < -s.radius) return OUTSIDE;
            else if (dist < s.radius) result = INTERSECTING;
        }
        return result;
    }

    struct Draw { int node, mesh, prim_local, prim; bool opaque, clip; Mat4 model, mvp; };

    std::vector<Draw> build_draws(const Scene &s, const sr_camera &c) const {
        std::vector<int> order(s.nodes.size());
        std::vector<float> key(s.nodes.size());
        for (size_t i = 0; i < s.nodes.size(); i++) {
            order[i] = (int)i;
            const BoundingSphere &b = s.nodes[i].bs_world;
            Vec3 d = v3(c.position[0], c.position[1], c.position[2]) - b.center;
            key[i] = dot3(d, d);
        }
        // OrderedFloat: NaN is greater than everything and equal to itself.
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            float ka = key[a], kb = key[b];
            if (std::isnan(ka)) return false;
            if (std::isnan(kb)) return true;
            return ka < kb;
        });
        Mat4 vp;
        memcpy(vp.m, c.view_project_matrix, 64);
        std::vector<Draw> draws;
        for (int ni : order) {
            const Node &n = s.nodes[ni];
            if (n.mesh < 0) continue;
            Mat4 mvp = mul(vp, n.transform);
            const Mesh &m = s.meshes[n.mesh];
            for (int pass = 0; pass < 2; pass++) {
                const std::vector<int> &list = pass == 0 ? m.opaque : m.translucent;
                for (int pl : list) {
                    int pi = m.prims[pl];
                    BoundingSphere bw = transform_sphere(n.transform, s.prims[pi].bs);
                    int ft = test_sphere_frustum(bw, c);
                    if (ft == OUTSIDE) continue;
                    Draw d;
                    d.node = ni; d.mesh = n.mesh; d.prim_local = pl; d.prim = pi;
                    d.opaque = pass == 0; d.clip = ft == INTERSECTING;
                    d.model = n.transform; d.mvp = mvp;
                    draws.push_back(d);
                }
            }
        }
        return draws;
    }

    // canonical: serial binning in canonical submission order (deterministic
    // ids and per-tile order); the per-tile stage is independent per tile and
    // runs on `threads` threads either way.  !canonical mirrors the
    // reference's parallel front end (128-triangle jobs, shared queues).
    void render_scene(const Scene &s, const sr_camera &c, int threads, bool fp64, bool canonical) {
        for (auto *t : tiles) { t->opaque.clear(); t->translucent.clear(); }
        auto t0 = std::chrono::steady_clock::now();
        std::vector<Draw> draws = build_draws(s, c);
        threaded = !canonical && threads > 1;
        uint32_t next_id = 0;
        if (!threaded) {
            for (const Draw &d : draws) {
                const Primitive &p = s.prims[d.prim];
                size_t ntri = p.indices.size() / 3;
                for (size_t t = 0; t < ntri; t++) {
                    Triangle tri = assemble(p, d.model, d.mvp, t, d.mesh, d.prim_local);
                    if (d.clip) next_id += clip_against_frustum(tri, d.opaque, next_id);
                    else { bin_triangle(tri, d.opaque, next_id); next_id += 1; }
                }
            }
            setup_count = next_id;
        } else {
            // Reference-like scheduling: 128-triangle jobs pushing into shared
            // per-tile queues (src/renderer.rs:207, :488-493).  Ids here are
            // (input triangle << 3 | fan index), not canonical.
            struct Job { int draw; size_t first, count; };
            std::vector<Job> jobs;
            for (size_t di = 0; di < draws.size(); di++) {
                size_t ntri = s.prims[draws[di].prim].indices.size() / 3;
                for (size_t f = 0; f < ntri; f += 128) jobs.push_back({(int)di, f, std::min<size_t>(128, ntri - f)});
            }
            std::atomic<size_t> next{0};
            std::atomic<uint64_t> emitted{0};
            auto worker = [&]() {
                uint64_t local = 0;
                for (;;) {
                    size_t j = next.fetch_add(1);
                    if (j >= jobs.size()) break;
                    const Draw &d = draws[jobs[j].draw];
                    const Primitive &p = s.prims[d.prim];
                    for (size_t t = jobs[j].first; t < jobs[j].first + jobs[j].count; t++) {
                        Triangle tri = assemble(p, d.model, d.mvp, t, d.mesh, d.prim_local);
                        uint32_t id = (uint32_t)(t << 3);
                        if (d.clip) local += clip_against_frustum(tri, d.opaque, id);
                        else { bin_triangle(tri, d.opaque, id); local += 1; }
                    }
                }
                emitted += local;
            };
            std::vector<std::thread> pool;
            for (int i = 0; i < threads; i++) pool.emplace_back(worker);
            for (auto &th : pool) th.join();
            setup_count = emitted.load();
        }
        auto t1 = std::chrono::steady_clock::now();
        if (threads <= 1) {
            for (auto *t : tiles) render_tile(*this, *t, s, c, fp64);
        } else {
            std::atomic<size_t> next{0};
            auto worker = [&]() {
                for (;;) {
                    size_t i = next.fetch_add(1);
                    if (i >= tiles.size()) break;
                    render_tile(*this, *tiles[i], s, c, fp64);
                }
            };
            std::vector<std::thread> pool;
            for (int i = 0; i < threads; i++) pool.emplace_back(worker);
            for (auto &th : pool) th.join();
        }
        auto t2 = std::chrono::steady_clock::now();
        ms_clipbin = std::chrono::duration<double, std::milli>(t1 - t0).count();
        ms_raster = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }

    void update_auto_exposure(float dt) {
        size_t n = tiles.size();
        if (n == 0) return;
        std::vector<float> lum(n);
        for (size_t i = 0; i < n; i++) lum[i] = tiles[i]->center_luminance;
        auto_exposure = exposure_step(lum.data(), n, dt, auto_exposure_ev, auto_exposure_target);
    }

    void blit(uint32_t *pixels, size_t bw, size_t bh) const {
        size_t rows = (bh / 2) * 2;
        for (size_t qy = 0; qy * 2 < rows; qy++) {
            int py = (int)qy * 2;
            int ty = py / TILE_SIZE;
            uint32_t *rowp = pixels + (size_t)py * bw;
            for (int tx = 0; tx < tiles_x; tx++) {
                const Tile &t = *tiles[ty * tiles_x + tx];
                int ly = (py - t.smin.y) / 2;
                int twq = (t.smax.x - t.smin.x) / 2;
                for (int qx = 0; qx < twq; qx++) {
                    V3x4<float> c = t.color[ly * twq + qx];
                    c = c * V4<float>::splat(auto_exposure);
                    float k = 0.2f;
                    c = (c / (c + k)) * (1.0f + k);
                    for (int sy = 0; sy < 2; sy++)
                        for (int sx = 0; sx < 2; sx++) {
                            int px = t.smin.x + qx * 2 + sx;
                            if (px < width) {
                                int l = sy * 2 + sx;
                                uint8_t r8 = as_u8(c.x[l] * 255.0f), g8 = as_u8(c.y[l] * 255.0f), b8 = as_u8(c.z[l] * 255.0f);
                                rowp[px + sy * (int)bw] = rgba8_pack_u8(r8, g8, b8, 0xFF);
                            }
                        }
                }
            }
        }
    }
};

// ---- per-tile rasterization ------------------------------------------------
static inline int index_from_xy(const Tile &t, IVec2 p) {
    int qx = (p.x - t.smin.x) / 2, qy = (p.y - t.smin.y) / 2;
    return qy * ((t.smax.x - t.smin.x) / 2) + qx;
}

struct RasterParams {
    const RasterPacket *pk;
    uint32_t packet_index;
    const Material *mat;
    float a01, b01, c01, a12, b12, c12, a20, b20, c20;
    float sx01, sx12, sx20, sy01, sy12, sy20;
    float ooa;
};

static void vbuffer_shade(Tile &t, const RasterParams &rp, int index, V4<float> b1, V4<float> b2, V4<float> z,
                          V4<float> w, M4 mask, V4<float> depth_from_test) {
    const Material &m = *rp.mat;
    if (m.is_alpha_tested) {
        // get_alpha_test_mask (src/shader.rs:311-329): raw du_dv, uv from the raster w
        V4<float> dd = {{rp.pk->du_dv[0], rp.pk->du_dv[1], rp.pk->du_dv[2], rp.pk->du_dv[3]}};
        V4<float> u = interp(rp.pk->u_over_w, b1, b2) * w;
        V4<float> v = interp(rp.pk->v_over_w, b1, b2) * w;
        V4<float> alpha = sample4_alpha(m.base, u, v, dd);
        mask = mask & cmpge(alpha, V4<float>::splat(m.alpha_cutoff));
    }
    for (int l = 0; l < 4; l++)
        if (mask[l]) {
            t.packet_index[index][l] = rp.packet_index;
            t.bary1[index][l] = b1[l];
            t.bary2[index][l] = b2[l];
            t.written[index][l] = true;
        }
    if (!m.is_alpha_tested) t.depth[index] = depth_from_test;
    else t.depth[index] = select(mask, z, t.depth[index]);
}

static void fine_raster(Tile &t, const RasterParams &rp, int xs, int ys, int xe, int ye) {
    IVec2 p = {xs, ys};
    V4<float> x = {{(float)(p.x + HALF_PIXEL), (float)(p.x + ONE_HALF_PIXEL), (float)(p.x + HALF_PIXEL), (float)(p.x + ONE_HALF_PIXEL)}};
    V4<float> y = {{(float)(p.y + HALF_PIXEL), (float)(p.y + HALF_PIXEL), (float)(p.y + ONE_HALF_PIXEL), (float)(p.y + ONE_HALF_PIXEL)}};
    V4<float> w0r = (x * rp.a12 + y * rp.b12) + rp.c12;
    V4<float> w1r = (x * rp.a20 + y * rp.b20) + rp.c20;
    V4<float> w2r = (x * rp.a01 + y * rp.b01) + rp.c01;
    V4<float> zero = V4<float>::splat(0);
    while (p.y < ye) {
        V4<float> w0 = w0r, w1 = w1r, w2 = w2r;
        p.x = xs;
        while (p.x < xe) {
            M4 mask = cmpge(w0, zero) & cmpge(w1, zero) & cmpge(w2, zero);
            if (any(mask)) {
                V4<float> b1 = w1 * rp.ooa, b2 = w2 * rp.ooa;
                V4<float> w = 1.0f / interp(rp.pk->one_over_w, b1, b2);
                V4<float> z = interp(rp.pk->z_over_w, b1, b2) * w;
                IVec2 pix = {p.x >> SUBPIXEL_SHIFT, p.y >> SUBPIXEL_SHIFT};
                int index = index_from_xy(t, pix);
                V4<float> cur = t.depth[index];
                M4 fm = mask & cmple(z, cur);
                V4<float> fd = select(fm, z, cur);
                if (any(fm)) vbuffer_shade(t, rp, index, b1, b2, z, w, fm, fd);
            }
            w0 = w0 + rp.sx12; w1 = w1 + rp.sx20; w2 = w2 + rp.sx01;
            p.x += STEP_X_SIZE;
        }
        w0r = w0r + rp.sy12; w1r = w1r + rp.sy20; w2r = w2r + rp.sy01;
        p.y += STEP_Y_SIZE;
    }
}

static void coarse_raster(Tile &t, const RasterParams &rp) {
    const RasterPacket &pk = *rp.pk;
    int xs = (pk.smin.x & ~1) * SUBPIXEL_SCALE, ys = (pk.smin.y & ~1) * SUBPIXEL_SCALE;
    int xe = pk.smax.x * SUBPIXEL_SCALE, ye = pk.smax.y * SUBPIXEL_SCALE;
    for (int by = ys; by < ye; by += COARSE_BLOCK_SIZE_SUBPIXELS) {
        for (int bx = xs; bx < xe; bx += COARSE_BLOCK_SIZE_SUBPIXELS) {
            V4<float> cx = {{(float)(bx + HALF_PIXEL), (float)(bx + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL),
                             (float)(bx + HALF_PIXEL), (float)(bx + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL)}};
            V4<float> cy = {{(float)(by + HALF_PIXEL), (float)(by + HALF_PIXEL),
                             (float)(by + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL), (float)(by + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL)}};
            V4<float> w0 = (cx * rp.a12 + cy * rp.b12) + rp.c12;
            V4<float> w1 = (cx * rp.a20 + cy * rp.b20) + rp.c20;
            V4<float> w2 = (cx * rp.a01 + cy * rp.b01) + rp.c01;
            auto maxe = [](V4<float> v) { return maxps(maxps(v[0], v[1]), maxps(v[2], v[3])); };
            if (maxe(w0) < 0.0f || maxe(w1) < 0.0f || maxe(w2) < 0.0f) continue;
            fine_raster(t, rp, bx, by, std::min(bx + COARSE_BLOCK_SIZE_SUBPIXELS, xe), std::min(by + COARSE_BLOCK_SIZE_SUBPIXELS, ye));
        }
    }
}

static inline int32_t top_left_bias(int32_t a, int32_t b) { return (a < 0 || (a == 0 && b > 0)) ? 0 : -1; }

static void rasterize_packet(Tile &t, const Scene &s, uint32_t index, const RasterPacket &pk) {
    IVec2 p0 = pk.p[0], p1 = pk.p[1], p2 = pk.p[2];
    RasterParams rp;
    rp.pk = &pk;
    rp.packet_index = index;
    const Mesh &mesh = s.meshes[pk.mesh_index];
    rp.mat = &s.materials[s.prims[mesh.prims[pk.primitive_index]].material];
    int32_t a01 = wsub(p1.y, p0.y), b01 = wsub(p0.x, p1.x);
    int32_t c01 = wadd(wsub(wmul(p1.x, p0.y), wmul(p0.x, p1.y)), top_left_bias(a01, b01));
    int32_t a12 = wsub(p2.y, p1.y), b12 = wsub(p1.x, p2.x);
    int32_t c12 = wadd(wsub(wmul(p2.x, p1.y), wmul(p1.x, p2.y)), top_left_bias(a12, b12));
    int32_t a20 = wsub(p0.y, p2.y), b20 = wsub(p2.x, p0.x);
    int32_t c20 = wadd(wsub(wmul(p0.x, p2.y), wmul(p2.x, p0.y)), top_left_bias(a20, b20));
    rp.sx01 = (float)wmul(a01, STEP_X_SIZE); rp.sx12 = (float)wmul(a12, STEP_X_SIZE); rp.sx20 = (float)wmul(a20, STEP_X_SIZE);
    rp.sy01 = (float)wmul(b01, STEP_Y_SIZE); rp.sy12 = (float)wmul(b12, STEP_Y_SIZE); rp.sy20 = (float)wmul(b20, STEP_Y_SIZE);
    rp.a01 = (float)a01; rp.b01 = (float)b01; rp.c01 = (float)c01;
    rp.a12 = (float)a12; rp.b12 = (float)b12; rp.c12 = (float)c12;
    rp.a20 = (float)a20; rp.b20 = (float)b20; rp.c20 = (float)c20;
    rp.ooa = pk.one_over_area;
    int wpx = pk.smax.x - pk.smin.x, hpx = pk.smax.y - pk.smin.y;
    if (wpx > COARSE_BLOCK_SIZE_PIXELS || hpx > COARSE_BLOCK_SIZE_PIXELS) {
        coarse_raster(t, rp);
    } else {
        int xs = (pk.smin.x & ~1) * SUBPIXEL_SCALE, ys = (pk.smin.y & ~1) * SUBPIXEL_SCALE;
        fine_raster(t, rp, xs, ys, pk.smax.x * SUBPIXEL_SCALE, pk.smax.y * SUBPIXEL_SCALE);
    }
}

// ---- shading -----------------------------------------------------------------
template <class T>
static V3x4<T> pbr_shader(const Scene &s, const sr_camera &cam, const RasterPacket &pk, const Material &m,
                          V4<T> b1, V4<T> b2, bool translucent, V3x4<T> current_color) {
    const T EPS = (T)1e-6f;
    const T PI = (T)3.14159265358979323846f;
    V4<T> one = V4<T>::splat(1), zero = V4<T>::splat(0);
    V4<T> w = (T)1 / interp(pk.one_over_w, b1, b2);
    V3x4<T> input_normal = normalize(interp(pk.normals, b1, b2));
    V3x4<T> input_tangent = normalize(interp(pk.tangents, b1, b2));
    V4<T> tangent_sign = interp(pk.tangent_sign, b1, b2);
    V3x4<T> pos_world = interp(pk.pos_world_over_w, b1, b2) * w;
    V4<T> uv_x = interp(pk.u_over_w, b1, b2) * w;
    V4<T> uv_y = interp(pk.v_over_w, b1, b2) * w;
    V4<T> du_dv = V4<T>{{(T)pk.du_dv[0], (T)pk.du_dv[1], (T)pk.du_dv[2], (T)pk.du_dv[3]}} * w;
    V3x4<T> tangent_world = normalize(input_tangent - input_normal * dot(input_normal, input_tangent));
    V4<T> hand = select(cmpge(tangent_sign, zero), one, V4<T>::splat(-1));
    V3x4<T> bitangent_world = cross(input_normal, tangent_world) * hand;
    V3x4<T> normal_world = normalize(input_normal);
    if (m.has_normal) {
        V3x4<T> tsn = sample4_rgb(m.normal, uv_x, uv_y, du_dv) * (T)2 - (T)1;
        normal_world = (tangent_world * tsn.x + bitangent_world * tsn.y) + input_normal * tsn.z;
        normal_world = normalize(normal_world);
    }
    V3x4<T> light_dir = V3x4<T>::from3((T)s.light_dir.x, (T)s.light_dir.y, (T)s.light_dir.z);
    V3x4<T> light_color = V3x4<T>::from3((T)s.light_color.x, (T)s.light_color.y, (T)s.light_color.z);
    V3x4<T> view_dir = V3x4<T>::from3((T)cam.position[0], (T)cam.position[1], (T)cam.position[2]) - pos_world;
    V3x4<T> view_normal = normalize(view_dir);
    V4<T> n_dot_l = vmax(dot(normal_world, light_dir), zero);
    V3x4<T> half_vector = normalize(light_dir + view_normal);
    V4<T> n_dot_h = vmax(dot(normal_world, half_vector), zero);
    V4<T> n_dot_v = vmax(dot(normal_world, view_normal), V4<T>::splat((T)1.0e-4f));
    V4<T> v_dot_h = vmax(dot(view_normal, half_vector), zero);
    V3x4<T> gi_rgb[4];
    V4<T> gi_w[4];
    voxel_lookup(s.voxels, pos_world, gi_rgb, gi_w);
    V4<T> vli = vclamp(gi_w[0], zero, one);
    V4<T> sky_vis = vclamp(gi_w[1], zero, one);
    V3x4<T> base = V3x4<T>::from3((T)m.base_color_factor[0], (T)m.base_color_factor[1], (T)m.base_color_factor[2]);
    if (m.has_base) base = base * srgb_to_linear_fast(sample4_rgb(m.base, uv_x, uv_y, du_dv));
    V4<T> roughness = V4<T>::splat((T)m.roughness);
    V4<T> metallic = V4<T>::splat((T)m.metallic);
    if (m.has_mr) {
        V3x4<T> mr = sample4_rgb(m.mr, uv_x, uv_y, du_dv);
        roughness = roughness * mr.y;
        metallic = metallic * mr.z;
    }
    roughness = vclamp(roughness, V4<T>::splat((T)0.045f), one);
    metallic = vclamp(metallic, zero, one);
    V4<T> ao = one;
    if (m.has_occlusion) {
        ao = sample4_rgb(m.occlusion, uv_x, uv_y, du_dv).x;
        ao = one + (ao - one) * V4<T>::splat((T)m.occlusion_strength);
    }
    V3x4<T> f0 = lerp3(V3x4<T>::splat((T)0.04f), base, metallic);
    V4<T> omvh = one - v_dot_h;
    V4<T> omvh2 = omvh * omvh;
    V4<T> omvh4 = omvh2 * omvh2;
    V4<T> omvh5 = omvh4 * omvh;
    V3x4<T> brdf_f = f0 + (V3x4<T>::splat(1) - f0) * omvh5;
    V4<T> alpha = roughness * roughness;
    V4<T> alpha2 = alpha * alpha;
    V4<T> ndh2 = n_dot_h * n_dot_h;
    V4<T> denom_d = ndh2 * (alpha2 - one) + one;
    V4<T> brdf_d = alpha2 / (V4<T>::splat(PI) * (denom_d * denom_d) + V4<T>::splat(EPS));
    V4<T> k = roughness + one;
    k = (k * k) * V4<T>::splat((T)0.125f);
    V4<T> gv_den = n_dot_v * (one - k) + k;
    V4<T> gv = n_dot_v / (gv_den + V4<T>::splat(EPS));
    V4<T> gl_den = n_dot_l * (one - k) + k;
    V4<T> gl = n_dot_l / (gl_den + V4<T>::splat(EPS));
    V4<T> brdf_g = gv * gl;
    V4<T> spec_dg = (brdf_d * brdf_g) / ((V4<T>::splat((T)4) * n_dot_l) * n_dot_v + V4<T>::splat(EPS));
    V3x4<T> kd_direct = (V3x4<T>::splat(1) - brdf_f) * (one - metallic);
    const T INV_PI = (T)(1.0f / 3.14159265358979323846f);
    V3x4<T> lambert = base * V4<T>::splat(INV_PI);
    V3x4<T> c_dd = (((light_color * kd_direct) * lambert) * n_dot_l) * vli;
    V3x4<T> c_ds = (((light_color * brdf_f) * spec_dg) * n_dot_l) * vli;
    V3x4<T> kd_ind = (V3x4<T>::splat(1) - f0) * (one - metallic);
    V3x4<T> irr = (gi_rgb[0] + gi_rgb[1] * normal_world.y) + gi_rgb[2] * normal_world.z;
    irr = irr + gi_rgb[3] * normal_world.x;
    irr = vmax3(irr, zero);
    V3x4<T> c_id = ((irr * base) * kd_ind) * V4<T>::splat(INV_PI);
    V3x4<T> refl = reflect(view_normal * (T)-1, normal_world);
    V4<T> spec_mip = roughness * V4<T>::splat((T)s.specular.max_mip_level_f);
    V3x4<T> env = sample_cubemap_trilinear_rgb(s.specular, refl, spec_mip);
    V3x4<T> lut = sample_bilinear_rgb(s.lut, SR_WRAP_CLAMP_TO_EDGE, SR_WRAP_CLAMP_TO_EDGE, vclamp(n_dot_v, zero, one),
                                      vclamp(roughness, zero, one), 0, usplat(0));
    V3x4<T> bsf = f0 * lut.x;
    bsf = bsf + lut.y;
    V3x4<T> c_is = env * bsf;
    V4<T> ao_spec = one + (ao - one) * V4<T>::splat((T)0.5f);
    c_is = c_is * (ao_spec * sky_vis);
    V3x4<T> color;
    if (translucent) {
        V4<T> tr = V4<T>::splat((T)m.transmission);
        if (m.has_transmission_tex) tr = tr * sample4_rgb(m.transmission_tex, uv_x, uv_y, du_dv).x;
        tr = vclamp(tr, zero, one);
        V4<T> inv_tr = one - tr;
        color = (((current_color * base) * tr) + ((c_dd + c_id) * inv_tr)) + c_ds;
        color = color + c_is;
    } else {
        color = ((c_dd + c_ds) + c_id) + c_is;
    }
    V3x4<T> em = V3x4<T>::splat(1);
    if (m.has_emissive) em = srgb_to_linear_fast(sample4_rgb(m.emissive_tex, uv_x, uv_y, du_dv));
    color = color + em * V3x4<T>::from3((T)m.emissive[0], (T)m.emissive[1], (T)m.emissive[2]);
    return color;
}

// The four normalized view directions of the quad at pix (compute_skybox,
// src/tilerasterizer.rs:478-505, before the cubemap lookup).
template <class T>
static V3x4<T> skybox_direction(const sr_camera &cam, IVec2 pix) {
    V4<T> px = {{(T)((float)pix.x + 0.5f), (T)((float)pix.x + 1.5f), (T)((float)pix.x + 0.5f), (T)((float)pix.x + 1.5f)}};
    V4<T> py = {{(T)((float)pix.y + 0.5f), (T)((float)pix.y + 0.5f), (T)((float)pix.y + 1.5f), (T)((float)pix.y + 1.5f)}};
    V4<T> one = V4<T>::splat(1);
    V4<T> ndc_x = (px * (T)cam.one_over_width) * (T)2 - one;
    V4<T> ndc_y = (one - py * (T)cam.one_over_height) * (T)2 - one;
    const float *m = cam.skybox_matrix_transposed;
    V3x4<T> d;
    V4<T> *out[3] = {&d.x, &d.y, &d.z};
    for (int r = 0; r < 3; r++) {
        const float *row = m + 4 * r;  // column r of the transposed matrix
        *out[r] = ((ndc_x * (T)row[0] + ndc_y * (T)row[1]) + one * (T)row[2]) + (T)row[3];
    }
    return normalize(d);
}

template <class T>
static V3x4<T> compute_skybox(const Scene &s, const sr_camera &cam, IVec2 pix) {
    V3x4<T> n = skybox_direction<T>(cam, pix);
    return srgb_to_linear_fast(sample_cubemap_rgb(s.cubemap, n, usplat(0)));
}

static void shade_vbuffer(Tile &t, const Scene &s, const sr_camera &cam, bool fp64) {
    int tw = t.smax.x - t.smin.x, th = t.smax.y - t.smin.y;
    int wq = tw / 2;
    for (int y = 0; y < th / 2; y++) {
        int sy = t.smin.y + y * 2;
        for (int qx = 0; qx < wq; qx++) {
            int sx = t.smin.x + qx * 2;
            IVec2 pix = {sx, sy};
            int index = index_from_xy(t, pix);
            U4 piv = t.packet_index[index];
            V4<float> depth = t.depth[index];
            M4 dm = cmpne(depth, V4<float>::splat(INFINITY));
            V3x4<float> out = V3x4<float>::splat(0);
            V3x4<double> out64 = V3x4<double>::splat(0);
            if (any(dm)) {
                M4 remaining = dm;
                while (any(remaining)) {
                    int lane = 0;
                    while (!remaining[lane]) lane++;
                    uint32_t pi = piv[lane];
                    M4 mask;
                    for (int l = 0; l < 4; l++) mask[l] = piv[l] == pi;
                    const RasterPacket &pk = t.opaque[pi];
                    const Mesh &mesh = s.meshes[pk.mesh_index];
                    const Material &m = s.materials[s.prims[mesh.prims[pk.primitive_index]].material];
                    V3x4<float> c = pbr_shader<float>(s, cam, pk, m, t.bary1[index], t.bary2[index], false, V3x4<float>::splat(0));
                    out = select3(mask, c, out);
                    if (fp64) {
                        V4<double> b1, b2;
                        for (int l = 0; l < 4; l++) { b1[l] = t.bary1[index][l]; b2[l] = t.bary2[index][l]; }
                        V3x4<double> c64 = pbr_shader<double>(s, cam, pk, m, b1, b2, false, V3x4<double>::splat(0));
                        out64 = select3(mask, c64, out64);
                    }
                    remaining = remaining & ~mask;
                }
            }
            if (!all(dm)) {
                V3x4<float> sky = compute_skybox<float>(s, cam, pix);
                out = select3(dm, out, sky);
                if (fp64) out64 = select3(dm, out64, compute_skybox<double>(s, cam, pix));
            }
            t.color[index] = out;
            if (fp64) t.color64[index] = out64;
        }
    }
}

// BumpQueue::sort_by (src/bumpqueue.rs:155-233): recursive Lomuto quicksort,
// rightmost pivot, "less" = compare(x, pivot) == Less.  With the comparator
// b.avg_z.cmp(&a.avg_z) (src/tilerasterizer.rs:93-94) x is "less" when
// x.avg_z > pivot.avg_z under OrderedFloat (NaN greatest).
static inline bool of_less(float a, float b) {  // OrderedFloat a < b
    if (std::isnan(a)) return false;
    if (std::isnan(b)) return true;
    return a < b;
}
static size_t lomuto_partition(std::vector<RasterPacket> &v, size_t lo, size_t hi) {
    float pivot = v[hi].avg_z;
    size_t i = lo;
    for (size_t j = lo; j < hi; j++) {
        if (of_less(pivot, v[j].avg_z)) {
            if (i != j) std::swap(v[i], v[j]);
            i += 1;
        }
    }
    if (i != hi) std::swap(v[i], v[hi]);
    return i;
}
static void lomuto_quicksort(std::vector<RasterPacket> &v, size_t lo, size_t hi) {
    if (lo < hi) {
        size_t p = lomuto_partition(v, lo, hi);
        if (p > 0) lomuto_quicksort(v, lo, p - 1);
        if (p + 1 < hi) lomuto_quicksort(v, p + 1, hi);
    }
}

void translucent_packet(Tile &t, const Scene &s, const sr_camera &cam, const RasterPacket &pk);

static void render_tile(Renderer &r, Tile &t, const Scene &s, const sr_camera &c, bool fp64) {
    (void)r;
    for (int q = 0; q < 1024; q++) {
        t.depth[q] = V4<float>::splat(INFINITY);
        t.packet_index[q] = usplat(0);
        t.bary1[q] = t.bary2[q] = V4<float>::splat(0);
        for (int l = 0; l < 4; l++) t.written[q][l] = false;
    }
    for (uint32_t i = 0; i < t.opaque.size(); i++) rasterize_packet(t, s, i, t.opaque[i]);
    shade_vbuffer(t, s, c, fp64);
    if (t.translucent.size() > 1) lomuto_quicksort(t.translucent, 0, t.translucent.size() - 1);
    // Forward translucent pass (src/tilerasterizer.rs:96-101, src/shader.rs:66-76)
    for (uint32_t i = 0; i < t.translucent.size(); i++) translucent_packet(t, s, c, t.translucent[i]);
    V3x4<float> cq = t.color[1024 / 2];
    V4<float> lum = (cq.x * 0.2126f + cq.y * 0.7152f) + cq.z * 0.0722f;
    t.center_luminance = (((lum[0] + lum[1]) + lum[2]) + lum[3]) * 0.25f;
}

// Forward shading of a translucent packet: the same coverage walk as the
// opaque path with depth test against the opaque vbuffer, no depth write,
// colour = pbr_shader<TRANSLUCENT>(current colour) on passing lanes.
static void translucent_fine(Tile &t, const Scene &s, const sr_camera &cam, const RasterParams &rp,
                             int xs, int ys, int xe, int ye) {
    IVec2 p = {xs, ys};
    V4<float> x = {{(float)(p.x + HALF_PIXEL), (float)(p.x + ONE_HALF_PIXEL), (float)(p.x + HALF_PIXEL), (float)(p.x + ONE_HALF_PIXEL)}};
    V4<float> y = {{(float)(p.y + HALF_PIXEL), (float)(p.y + HALF_PIXEL), (float)(p.y + ONE_HALF_PIXEL), (float)(p.y + ONE_HALF_PIXEL)}};
    V4<float> w0r = (x * rp.a12 + y * rp.b12) + rp.c12;
    V4<float> w1r = (x * rp.a20 + y * rp.b20) + rp.c20;
    V4<float> w2r = (x * rp.a01 + y * rp.b01) + rp.c01;
    V4<float> zero = V4<float>::splat(0);
    while (p.y < ye) {
        V4<float> w0 = w0r, w1 = w1r, w2 = w2r;
        p.x = xs;
        while (p.x < xe) {
            M4 mask = cmpge(w0, zero) & cmpge(w1, zero) & cmpge(w2, zero);
            if (any(mask)) {
                V4<float> b1 = w1 * rp.ooa, b2 = w2 * rp.ooa;
                V4<float> w = 1.0f / interp(rp.pk->one_over_w, b1, b2);
                V4<float> z = interp(rp.pk->z_over_w, b1, b2) * w;
                IVec2 pix = {p.x >> SUBPIXEL_SHIFT, p.y >> SUBPIXEL_SHIFT};
                int index = index_from_xy(t, pix);
                M4 fm = mask & cmple(z, t.depth[index]);
                if (any(fm)) {
                    V3x4<float> cur = t.color[index];
                    V3x4<float> c = pbr_shader<float>(s, cam, *rp.pk, *rp.mat, b1, b2, true, cur);
                    t.color[index] = select3(fm, c, cur);
                }
            }
            w0 = w0 + rp.sx12; w1 = w1 + rp.sx20; w2 = w2 + rp.sx01;
            p.x += STEP_X_SIZE;
        }
        w0r = w0r + rp.sy12; w1r = w1r + rp.sy20; w2r = w2r + rp.sy01;
        p.y += STEP_Y_SIZE;
    }
}

void translucent_packet(Tile &t, const Scene &s, const sr_camera &cam, const RasterPacket &pk) {
    IVec2 p0 = pk.p[0], p1 = pk.p[1], p2 = pk.p[2];
    RasterParams rp;
    rp.pk = &pk;
    rp.packet_index = 0;
    const Mesh &mesh = s.meshes[pk.mesh_index];
    rp.mat = &s.materials[s.prims[mesh.prims[pk.primitive_index]].material];
    int32_t a01 = wsub(p1.y, p0.y), b01 = wsub(p0.x, p1.x);
    int32_t c01 = wadd(wsub(wmul(p1.x, p0.y), wmul(p0.x, p1.y)), top_left_bias(a01, b01));
    int32_t a12 = wsub(p2.y, p1.y), b12 = wsub(p1.x, p2.x);
    int32_t c12 = wadd(wsub(wmul(p2.x, p1.y), wmul(p1.x, p2.y)), top_left_bias(a12, b12));
    int32_t a20 = wsub(p0.y, p2.y), b20 = wsub(p2.x, p0.x);
    int32_t c20 = wadd(wsub(wmul(p0.x, p2.y), wmul(p2.x, p0.y)), top_left_bias(a20, b20));
    rp.sx01 = (float)wmul(a01, STEP_X_SIZE); rp.sx12 = (float)wmul(a12, STEP_X_SIZE); rp.sx20 = (float)wmul(a20, STEP_X_SIZE);
    rp.sy01 = (float)wmul(b01, STEP_Y_SIZE); rp.sy12 = (float)wmul(b12, STEP_Y_SIZE); rp.sy20 = (float)wmul(b20, STEP_Y_SIZE);
    rp.a01 = (float)a01; rp.b01 = (float)b01; rp.c01 = (float)c01;
    rp.a12 = (float)a12; rp.b12 = (float)b12; rp.c12 = (float)c12;
    rp.a20 = (float)a20; rp.b20 = (float)b20; rp.c20 = (float)c20;
    rp.ooa = pk.one_over_area;
    int xs = (pk.smin.x & ~1) * SUBPIXEL_SCALE, ys = (pk.smin.y & ~1) * SUBPIXEL_SCALE;
    int xe = pk.smax.x * SUBPIXEL_SCALE, ye = pk.smax.y * SUBPIXEL_SCALE;
    int wpx = pk.smax.x - pk.smin.x, hpx = pk.smax.y - pk.smin.y;
    if (wpx > COARSE_BLOCK_SIZE_PIXELS || hpx > COARSE_BLOCK_SIZE_PIXELS) {
        for (int by = ys; by < ye; by += COARSE_BLOCK_SIZE_SUBPIXELS)
            for (int bx = xs; bx < xe; bx += COARSE_BLOCK_SIZE_SUBPIXELS) {
                V4<float> cx = {{(float)(bx + HALF_PIXEL), (float)(bx + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL),
                                 (float)(bx + HALF_PIXEL), (float)(bx + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL)}};
                V4<float> cy = {{(float)(by + HALF_PIXEL), (float)(by + HALF_PIXEL),
                                 (float)(by + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL), (float)(by + COARSE_BLOCK_SIZE_SUBPIXELS - HALF_PIXEL)}};
                V4<float> w0 = (cx * rp.a12 + cy * rp.b12) + rp.c12;
                V4<float> w1 = (cx * rp.a20 + cy * rp.b20) + rp.c20;
                V4<float> w2 = (cx * rp.a01 + cy * rp.b01) + rp.c01;
                auto maxe = [](V4<float> v) { return maxps(maxps(v[0], v[1]), maxps(v[2], v[3])); };
                if (maxe(w0) < 0.0f || maxe(w1) < 0.0f || maxe(w2) < 0.0f) continue;
                translucent_fine(t, s, cam, rp, bx, by, std::min(bx + COARSE_BLOCK_SIZE_SUBPIXELS, xe),
                                 std::min(by + COARSE_BLOCK_SIZE_SUBPIXELS, ye));
            }
    } else {
        translucent_fine(t, s, cam, rp, xs, ys, xe, ye);
    }
}

}  // namespace orc

// ----------------------------------------------------------------------------
// C interface for the test harness (ctypes)
// ----------------------------------------------------------------------------
using namespace orc;

extern "C" {

__attribute__((visibility("default"))) void *orc_scene_create(const sr_scene_desc *d) { return build_scene(d); }
__attribute__((visibility("default"))) void orc_scene_destroy(void *s) { delete (Scene *)s; }
__attribute__((visibility("default"))) void *orc_renderer_create(int w, int h) { return new Renderer(w, h); }
__attribute__((visibility("default"))) void orc_renderer_destroy(void *r) { delete (Renderer *)r; }

__attribute__((visibility("default"))) int orc_render(void *rp, void *sp, const sr_camera *cam, int threads, int fp64,
                                                      int canonical) {
    Renderer &r = *(Renderer *)rp;
    r.render_scene(*(Scene *)sp, *cam, threads, fp64 != 0, canonical != 0);
    return 0;
}
__attribute__((visibility("default"))) double orc_stage_ms(void *rp, int stage) {
    Renderer &r = *(Renderer *)rp;
    return stage == 0 ? r.ms_clipbin : r.ms_raster;
}
__attribute__((visibility("default"))) uint64_t orc_setup_count(void *rp) { return ((Renderer *)rp)->setup_count; }

// Visibility buffer: canonical id of the last depth-passing fragment (or
// 0xFFFFFFFF), depth value; colour: linear shaded RGB.
__attribute__((visibility("default"))) int orc_read_frame(void *rp, uint32_t *ids, float *depth, float *rgb,
                                                          double *rgb64, float *bary) {
    Renderer &r = *(Renderer *)rp;
    for (auto *tp : r.tiles) {
        Tile &t = *tp;
        for (int q = 0; q < 1024; q++) {
            int qx = q % 32, qy = q / 32;
            for (int l = 0; l < 4; l++) {
                int x = t.smin.x + qx * 2 + (l & 1), y = t.smin.y + qy * 2 + (l >> 1);
                if (x >= r.width || y >= r.height) continue;
                size_t o = (size_t)y * r.width + x;
                float d = t.depth[q][l];
                if (ids) ids[o] = t.written[q][l] ? t.opaque[t.packet_index[q][l]].tri_id : 0xFFFFFFFFu;
                if (depth) depth[o] = d;
                if (rgb) { rgb[3 * o] = t.color[q].x[l]; rgb[3 * o + 1] = t.color[q].y[l]; rgb[3 * o + 2] = t.color[q].z[l]; }
                if (rgb64) { rgb64[3 * o] = t.color64[q].x[l]; rgb64[3 * o + 1] = t.color64[q].y[l]; rgb64[3 * o + 2] = t.color64[q].z[l]; }
                if (bary) { bary[2 * o] = t.bary1[q][l]; bary[2 * o + 1] = t.bary2[q][l]; }
            }
        }
    }
    return 0;
}
__attribute__((visibility("default"))) int orc_update_auto_exposure(void *rp, float dt) {
    ((Renderer *)rp)->update_auto_exposure(dt);
    return 0;
}
__attribute__((visibility("default"))) float orc_exposure(void *rp) { return ((Renderer *)rp)->auto_exposure; }
__attribute__((visibility("default"))) int orc_blit(void *rp, uint32_t *px, size_t w, size_t h) {
    ((Renderer *)rp)->blit(px, w, h);
    return 0;
}
__attribute__((visibility("default"))) int orc_tile_luminance(void *rp, float *out) {
    Renderer &r = *(Renderer *)rp;
    for (size_t i = 0; i < r.tiles.size(); i++) out[i] = r.tiles[i]->center_luminance;
    return (int)r.tiles.size();
}
__attribute__((visibility("default"))) int orc_camera_new(const float pos[3], const float look[3], float fov, float w,
                                                          float h, float far_plane, sr_camera *out) {
    camera_new(pos, look, fov, w, h, far_plane, *out);
    return 0;
}
__attribute__((visibility("default"))) int orc_camera_update(sr_camera *c) {
    camera_update(*c);
    return 0;
}
// Tonemap + quantize one linear value with a given exposure (closed-form tests).
// The per-tile translucent sort on its own: order[i] = the index of the
// packet that ends at position i, packets given by avg_z in queue order.
__attribute__((visibility("default"))) int orc_sort_translucent(const float *avg_z, uint32_t *order, size_t n) {
    std::vector<RasterPacket> v(n);
    for (size_t i = 0; i < n; i++) {
        v[i].avg_z = avg_z[i];
        v[i].tri_id = (uint32_t)i;
    }
    if (n > 1) lomuto_quicksort(v, 0, n - 1);
    for (size_t i = 0; i < n; i++) order[i] = v[i].tri_id;
    return 0;
}

__attribute__((visibility("default"))) uint32_t orc_quantize(float c, float exposure) {
    float x = c * exposure;
    float k = 0.2f;
    x = (x / (x + k)) * (1.0f + k);
    return as_u8(x * 255.0f);
}
__attribute__((visibility("default"))) int orc_mip_chain(const sr_texture_desc *t, uint32_t *out, uint32_t *n) {
    Texture tex;
    texture_from_base(tex, t->width, t->height, t->type, t->texels, true);
    if (out) memcpy(out, tex.data.data(), tex.data.size() * 4);
    *n = (uint32_t)tex.data.size();
    return 0;
}

// ---- unit entry points for closed-form tests ---------------------------------
// sample4_rgb with its mip selection on a texture built from `t` (mips
// generated); returns the chosen mip level.
__attribute__((visibility("default"))) int orc_sample4(const sr_texture_desc *t, int wrap_s, int wrap_t, const float u[4],
                                                       const float v[4], const float du_dv[4], float out_rgb[12]) {
    Texture tex;
    texture_from_base(tex, t->width, t->height, t->type, t->texels, true);
    Sampled s;
    s.tex = &tex;
    s.wrap_s = wrap_s;
    s.wrap_t = wrap_t;
    V4<float> U, Vv, D;
    for (int i = 0; i < 4; i++) { U[i] = u[i]; Vv[i] = v[i]; D[i] = du_dv[i]; }
    V3x4<float> c = sample4_rgb(s, U, Vv, D);
    for (int i = 0; i < 4; i++) { out_rgb[3 * i] = c.x[i]; out_rgb[3 * i + 1] = c.y[i]; out_rgb[3 * i + 2] = c.z[i]; }
    return (int)compute_mip_level(tex, D);
}

// Trilinear voxel lookup on a full SH4 grid (16 floats per voxel) at four
// world positions; writes rgb[4 coeff][3 ch][4 lane] and w[4 coeff][4 lane].
__attribute__((visibility("default"))) int orc_voxel_lookup(int w, int h, int d, const float wmin[3], const float wmax[3],
                                                            const float *sh4, const float pos[12], float *rgb_out,
                                                            float *w_out) {
    VoxelGrid g;
    g.w = w; g.h = h; g.d = d;
    g.wmin = v3(wmin[0], wmin[1], wmin[2]);
    g.wmax = v3(wmax[0], wmax[1], wmax[2]);
    g.sh4.assign(sh4, sh4 + (size_t)w * h * d * 16);
    V3x4<float> p;
    for (int i = 0; i < 4; i++) { p.x[i] = pos[3 * i]; p.y[i] = pos[3 * i + 1]; p.z[i] = pos[3 * i + 2]; }
    V3x4<float> rgb[4];
    V4<float> wo[4];
    voxel_lookup(g, p, rgb, wo);
    for (int c = 0; c < 4; c++)
        for (int l = 0; l < 4; l++) {
            rgb_out[(c * 3 + 0) * 4 + l] = rgb[c].x[l];
            rgb_out[(c * 3 + 1) * 4 + l] = rgb[c].y[l];
            rgb_out[(c * 3 + 2) * 4 + l] = rgb[c].z[l];
            w_out[c * 4 + l] = wo[c][l];
        }
    return 0;
}

// cubemap_uv_from_normal for four directions.
__attribute__((visibility("default"))) int orc_cubemap_uv(const float n[12], float uv[8], uint32_t face[4]) {
    V3x4<float> v;
    for (int i = 0; i < 4; i++) { v.x[i] = n[3 * i]; v.y[i] = n[3 * i + 1]; v.z[i] = n[3 * i + 2]; }
    V4<float> u, vv;
    U4 s;
    cubemap_uv_from_normal(v, u, vv, s);
    for (int i = 0; i < 4; i++) { uv[2 * i] = u[i]; uv[2 * i + 1] = vv[i]; face[i] = s[i]; }
    return 0;
}

// pbr_shader on one packet with constant interpolators: every vertex has
// normal n, world position p (1/w = 1) and texture coordinate uv; du_dv is
// the packet's UV derivative term.  All four lanes shade the same point;
// lane 0's colour in f32 and in the f64 instantiation.
__attribute__((visibility("default"))) int orc_shade_probe(void *sp, const sr_camera *cam, int material,
                                                           const float n[3], const float p[3], const float uv[2],
                                                           const float du_dv[4], float out32[3], double out64[3]) {
    const Scene &s = *(const Scene *)sp;
    if (material < 0 || material >= (int)s.materials.size()) return -1;
    RasterPacket pk;
    memset(&pk, 0, sizeof(pk));
    Vec3 nn = v3(n[0], n[1], n[2]), pp = v3(p[0], p[1], p[2]);
    pk.normals = vinterp(nn, nn, nn);
    pk.tangents = vinterp(v3(1, 0, 0), v3(1, 0, 0), v3(1, 0, 0));
    pk.tangent_sign = sinterp(1.0f, 1.0f, 1.0f);
    pk.one_over_w = sinterp(1.0f, 1.0f, 1.0f);
    pk.z_over_w = sinterp(0.5f, 0.5f, 0.5f);
    pk.pos_world_over_w = vinterp(pp, pp, pp);
    pk.u_over_w = sinterp(uv[0], uv[0], uv[0]);
    pk.v_over_w = sinterp(uv[1], uv[1], uv[1]);
    for (int i = 0; i < 4; i++) pk.du_dv[i] = du_dv[i];
    const Material &m = s.materials[material];
    V4<float> z = V4<float>::splat(0.0f);
    V3x4<float> c = pbr_shader<float>(s, *cam, pk, m, z, z, false, V3x4<float>::splat(0));
    V4<double> zd = V4<double>::splat(0.0);
    V3x4<double> cd = pbr_shader<double>(s, *cam, pk, m, zd, zd, false, V3x4<double>::splat(0));
    out32[0] = c.x[0]; out32[1] = c.y[0]; out32[2] = c.z[0];
    out64[0] = cd.x[0]; out64[1] = cd.y[0]; out64[2] = cd.z[0];
    return 0;
}

// compute_skybox's four view directions for the quad at (x, y) (x, y even):
// dir[lane * 3 + axis].
__attribute__((visibility("default"))) int orc_skybox_dir(const sr_camera *cam, int x, int y, float dir[12]) {
    V3x4<float> d = skybox_direction<float>(*cam, IVec2{x, y});
    for (int l = 0; l < 4; l++) { dir[3 * l] = d.x[l]; dir[3 * l + 1] = d.y[l]; dir[3 * l + 2] = d.z[l]; }
    return 0;
}

// The scene's environment samplers on one lane: which = 0 the specular
// cubemap trilinearly at direction a (mip level `level`), 1 the BRDF LUT
// bilinearly at (u, v) = (a[0], a[1]) with clamp-to-edge (as pbr_shader
// reads it), 2 the skybox cubemap (nearest, mip 0) at direction a.
__attribute__((visibility("default"))) int orc_env_sample(void *sp, int which, const float a[3], float level,
                                                          float out[3]) {
    const Scene &s = *(const Scene *)sp;
    V3x4<float> d = V3x4<float>::from3(a[0], a[1], a[2]), c;
    if (which == 0) c = sample_cubemap_trilinear_rgb(s.specular, d, V4<float>::splat(level));
    else if (which == 1)
        c = sample_bilinear_rgb(s.lut, SR_WRAP_CLAMP_TO_EDGE, SR_WRAP_CLAMP_TO_EDGE, V4<float>::splat(a[0]),
                                V4<float>::splat(a[1]), 0, usplat(0));
    else if (which == 2) c = sample_cubemap_rgb(s.cubemap, d, usplat(0));
    else return -1;
    out[0] = c.x[0]; out[1] = c.y[0]; out[2] = c.z[0];
    return 0;
}

// The auto-exposure step on given tile luminances, from (ev, target);
// state = {ev, target, exposure} after the step.
__attribute__((visibility("default"))) int orc_exposure_step(const float *lum, size_t n, float dt, float ev,
                                                             float target, float state[3]) {
    if (n == 0) return -1;
    float e = exposure_step(lum, n, dt, ev, target);
    state[0] = ev; state[1] = target; state[2] = e;
    return 0;
}

}  // extern "C"
