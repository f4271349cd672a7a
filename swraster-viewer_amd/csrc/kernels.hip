// kernels.hip -- the per-frame rasterization pipeline for gfx950 (MI355X).
//
// Stage map (reference paths are relative to the reference tree):
//   k_front         one launch of two parts: SoA vertex fetch and MVP / model
//                   / normal transform once per (draw, vertex)
//                                           src/renderer.rs:499-521
//                   and the frustum classification of every input triangle
//                   (fan counts, clip queue) src/renderer.rs:579-665
//   k_clip          clipping of the queued triangles with all attributes
//   k_scan_*        device-wide exclusive scan (post-clip ids, chunk pairs)
//   k_setup         clipping with attributes, snap, cull, bbox, setup records,
//                   pair counts per chunk   src/renderer.rs:668-846
//   k_bin_expand    stable binning: the records' pairs in record order
//   k_bin_hist, k_bin_rowscan, k_bin_scatter, k_bin_bounds
//                   stable binning: radix passes over the bin index, then
//                   each tile's list bounds src/bumpqueue.rs:96-112
//   k_raster_shade  one workgroup per screen tile: coverage with the
//                   reference's exact f32 edge-evaluation chains into a 64-bit
//                   (depth, id) LDS buffer, then visibility-buffer PBR
//                   shading, skybox and centre luminance
//                                           src/tilerasterizer.rs:72-543,
//                                           src/shader.rs:102-329
//   k_blit          exposure, tonemap, quantize, pack
//                                           src/renderer.rs:293-355
//
// Arithmetic contract: built with -ffp-contract=off, IEEE division and
// square root, denormals preserved; vector min/max/clamp use the SSE form
// (a > b ? a : b); Rust `as` casts saturate with NaN -> 0; i32 wraps.
//
// Depth resolution: the reference's per-tile sequential depth test
// (z <= current, later fragments win ties) leaves at each pixel the last
// fragment, in submission order, among those with the minimal depth; NaN
// depths never pass and +0 == -0.  The kernel computes exactly that as an
// order-independent minimum over 64-bit keys (orderable depth << 32 |
// ~triangle id), so the result does not depend on the order of a tile's
// list (the lists are in submission order anyway).  The winner's
// barycentrics and depth bits are then recomputed with the same evaluation
// chain.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "swraster.h"

namespace sr {

// ---------------------------------------------------------------------------
// arithmetic helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float maxps(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float minps(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float clampv(float x, float lo, float hi) { return minps(maxps(x, lo), hi); }
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
__device__ __forceinline__ int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
__device__ __forceinline__ int32_t wabs(int32_t a) { return a < 0 ? (int32_t)(0u - (uint32_t)a) : a; }
__device__ __forceinline__ int32_t sat_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}
__device__ __forceinline__ uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}
__device__ __forceinline__ uint32_t sat_u8(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 255.0f) return 255u;
    return (uint32_t)f;
}
// `as usize` then min(limit): saturating, NaN and negatives -> 0
__device__ __forceinline__ uint32_t sat_index(float f, uint32_t limit) {
    if (!(f > 0.0f)) return 0u;
    if (f >= (float)limit) return limit;
    uint32_t v = (uint32_t)f;
    return v < limit ? v : limit;
}
__device__ __forceinline__ float rsqrt_ieee(float x) { return 1.0f / sqrtf(x); }

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 muls(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 adds(f3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 normalize3(f3 a) {
    float r = rsqrt_ieee(dot3(a, a));
    return {a.x * r, a.y * r, a.z * r};
}
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 max3s(f3 a, float m) { return {maxps(a.x, m), maxps(a.y, m), maxps(a.z, m)}; }
// srgb_to_linear_fast (src/util.rs:111-113)
__device__ __forceinline__ f3 srgb_fast(f3 x) {
    f3 r;
    r.x = x.x * (x.x * (x.x * 0.305306011f + 0.682171111f) + 0.012522878f);
    r.y = x.y * (x.y * (x.y * 0.305306011f + 0.682171111f) + 0.012522878f);
    r.z = x.z * (x.z * (x.z * 0.305306011f + 0.682171111f) + 0.012522878f);
    return r;
}
// n / 255.0f, correctly rounded, for integers n in [0, 255]: the product with
// RN(1/255) is corrected by one fused multiply-add on its exact residual
// (checked against IEEE division for all 256 inputs).
__device__ __forceinline__ float div255(float n) {
    const float inv = 1.0f / 255.0f;
    float q = n * inv;
    return fmaf(fmaf(-255.0f, q, n), inv, q);
}
__device__ __forceinline__ f3 unpack_rgb(uint32_t c) {
    return {div255((float)((c >> 24) & 0xFFu)), div255((float)((c >> 16) & 0xFFu)), div255((float)((c >> 8) & 0xFFu))};
}
__device__ __forceinline__ float unpack_a(uint32_t c) { return div255((float)(c & 0xFFu)); }

template <class T>
__device__ __forceinline__ int find_last_le(const T *arr, int n, uint32_t key, uint32_t T::*field) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (arr[mid].*field <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// find_last_le where the active lanes' keys increase with the lane index
// (consecutive triangles or vertices): one wave-uniform search for the first
// active lane's key answers for every lane when the last active lane's key
// still lies before the next entry, which is the usual case; otherwise each
// lane searches.
template <class T>
__device__ __forceinline__ int find_last_le_consecutive(const T *arr, int n, uint32_t key, uint32_t T::*field) {
    uint64_t act = __ballot(1);
    uint32_t k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)key);
    uint32_t k1 = (uint32_t)__builtin_amdgcn_readlane((int)key, 63 - __clzll((long long)act));
    int d0 = find_last_le(arr, n, k0, field);
    if (d0 + 1 >= n || arr[d0 + 1].*field > k1) return d0;
    return find_last_le(arr, n, key, field);
}

// ---------------------------------------------------------------------------
// Clip-space vertex with attributes (Vertex, src/renderer.rs:30-37)
// ---------------------------------------------------------------------------
struct CV {
    float cx, cy, cz, cw;         // pos_clip
    float wx, wy, wz;             // pos_world
    float nx, ny, nz;             // normal
    float u, v;                   // uv
    float tx, ty, tz, tw;         // tangent
    float oow;                    // 1 / cw
};

// Transformed vertices are two arrays: clip-space positions (one float4 per
// vertex, all that setup and clipping classification read) and attributes
// (two float4: world position and normal x; normal y, z and uv).

// Plane dot products in glam's SSE Vec4::dot order (x*x' + z*z') + (y*y' + w*w')
// for the six clip planes of src/renderer.rs:581-588.
__device__ __forceinline__ float plane_dot(int p, float x, float y, float z, float w) {
    const float PX[6] = {0.0f, 0.0f, 1.0f, -1.0f, 0.0f, 0.0f};
    const float PY[6] = {0.0f, 0.0f, 0.0f, 0.0f, 1.0f, -1.0f};
    const float PZ[6] = {1.0f, -1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float xx = PX[p] * x, yy = PY[p] * y, zz = PZ[p] * z, ww = 1.0f * w;
    return (xx + zz) + (yy + ww);
}

__device__ __forceinline__ void lerp_cv(const CV &a, const CV &b, float t, CV &o) {
    o.cx = a.cx + (b.cx - a.cx) * t; o.cy = a.cy + (b.cy - a.cy) * t;
    o.cz = a.cz + (b.cz - a.cz) * t; o.cw = a.cw + (b.cw - a.cw) * t;
    o.wx = a.wx + (b.wx - a.wx) * t; o.wy = a.wy + (b.wy - a.wy) * t; o.wz = a.wz + (b.wz - a.wz) * t;
    o.nx = a.nx + (b.nx - a.nx) * t; o.ny = a.ny + (b.ny - a.ny) * t; o.nz = a.nz + (b.nz - a.nz) * t;
    o.tx = a.tx + (b.tx - a.tx) * t; o.ty = a.ty + (b.ty - a.ty) * t;
    o.tz = a.tz + (b.tz - a.tz) * t; o.tw = a.tw + (b.tw - a.tw) * t;
    o.u = a.u + (b.u - a.u) * t; o.v = a.v + (b.v - a.v) * t;
}

// Position-only pre-pass: planes whose test leaves the triangle untouched are
// skipped; returns -1 if the triangle passes all planes unchanged, -2 if it is
// rejected whole by a plane, or the first plane that cuts it.
__device__ __forceinline__ int clip_classify(const float c[3][4]) {
    for (int p = 0; p < 6; p++) {
        int in = 0;
        for (int k = 0; k < 3; k++) in += plane_dot(p, c[k][0], c[k][1], c[k][2], c[k][3]) >= 0.0f ? 1 : 0;
        if (in == 3) continue;
        if (in == 0) return -2;
        return p;
    }
    return -1;
}

// Clears the tiles' list bounds over the whole grid, in k_front (or in
// k_frame_begin when there are no vertices); the frame's status is cleared
// by the upload that precedes the first kernel.
// Stage stamps on the wall clock (StageClock): the kernel's first workgroup
// stamps its start (workgroups are dispatched in index order).
__device__ __forceinline__ void stamp_start(uint64_t *t) {
    if (t && blockIdx.x == 0 && threadIdx.x == 0) *t = (uint64_t)wall_clock64();
}
__device__ __forceinline__ void clock_start(StageClock *c, int which) {
    stamp_start(c ? &c->start[which] : nullptr);
}
// Every thread of the workgroup calls it after its work; the workgroup's
// end raises its end slot.
__device__ __forceinline__ void clock_end(StageClock *c, int which) {
    if (!c) return;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(&c->end[which][blockIdx.x & (CLOCK_SPREAD - 1)]),
                  (unsigned long long)wall_clock64());
}

struct FrameBegin {
    uint32_t *bin_start, *bin_end, *tr_count;
    uint32_t items, tr_slots;
    StageClock *clock;
};

__device__ __forceinline__ void frame_begin(const FrameBegin &F) {
    clock_start(F.clock, CLOCK_FRONT);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = i; k < F.items; k += gridDim.x * blockDim.x) {
        F.bin_start[k] = 0u;
        F.bin_end[k] = 0u;
    }
    if (F.clock)
        for (uint32_t k = i; k < CLOCK_ENDS * CLOCK_SPREAD; k += gridDim.x * blockDim.x) (&F.clock->end[0][0])[k] = 0ull;
    if (F.tr_count)
        for (uint32_t k = i; k < F.tr_slots; k += gridDim.x * blockDim.x) F.tr_count[k] = 0u;
}

__global__ void __launch_bounds__(256) k_frame_begin(FrameBegin F) { frame_begin(F); }

// ---------------------------------------------------------------------------
// Vertex transform (k_front): one thread per (draw, vertex)
// ---------------------------------------------------------------------------
// Clip position MVP p of source vertex s (render_primitive_inner,
// src/renderer.rs:496-521); the transform and the frustum classification
// both take it from here, so they agree bit for bit.
__device__ __forceinline__ void clip_position(const DevDraw &D, const DevScene &S, uint32_t s, float c[4]) {
    float x = S.px[s], y = S.py[s], z = S.pz[s];
    for (int r = 0; r < 4; r++) {
        float acc = D.mvp[r] * x;
        acc = acc + D.mvp[4 + r] * y;
        acc = acc + D.mvp[8 + r] * z;
        acc = acc + D.mvp[12 + r] * 1.0f;
        c[r] = acc;
    }
}

// Per-vertex part of bin_triangle (src/renderer.rs:699-709, 834-846): the
// snapped sub-pixel position round(((ndc + 1) W / 2) 16) (y flipped), 1/w and
// z/w as z (1/w).  Every operation depends on the vertex and the view only,
// so the transform computes it once per vertex for the triangles that need
// no clipping, and clipped fans compute it from their polygon's vertices.
struct SnapV {
    int32_t px, py;
    float oow, zow;
};
__device__ __forceinline__ SnapV snap_vertex(float Wf, float Hf, float cx, float cy, float cz, float cw) {
    SnapV v;
    float nxv = cx / cw, nyv = cy / cw;
    float sx = ((nxv + 1.0f) * Wf) / 2.0f;
    float sy = ((1.0f - nyv) * Hf) / 2.0f;
    v.px = sat_i32(roundf(sx * 16.0f));
    v.py = sat_i32(roundf(sy * 16.0f));
    v.oow = 1.0f / cw;
    v.zow = cz * v.oow;
    return v;
}

// Transform of vertex i of draw D (render_primitive_inner, src/renderer.rs:
// 496-521): clip = MVP p, world = M p, normal = M3 n, tangent = M3 t; and the
// vertex's snap for the view (xs: px, py as integer bits, 1/w, z/w).
__device__ __forceinline__ void transform_vertex(const DevDraw &D, const DevView &V, uint32_t i, float4 *xp,
                                                 float4 *xa, float4 *xt, float4 *xs) {
    const DevScene &S = *V.scene;
    uint32_t s = D.vtx_base + (i - D.xv_base);
    float x = S.px[s], y = S.py[s], z = S.pz[s];
    float nx = S.nx[s], ny = S.ny[s], nz = S.nz[s];
    float c[4], w[3], n[3];
    clip_position(D, S, s, c);
    for (int r = 0; r < 3; r++) {
        float acc = D.model[r] * x;
        acc = acc + D.model[4 + r] * y;
        acc = acc + D.model[8 + r] * z;
        acc = acc + D.model[12 + r] * 1.0f;
        w[r] = acc;
        float an = D.model[r] * nx;
        an = an + D.model[4 + r] * ny;
        an = an + D.model[8 + r] * nz;
        n[r] = an;
    }
    xp[i] = make_float4(c[0], c[1], c[2], c[3]);
    SnapV sv = snap_vertex(V.Wf, V.Hf, c[0], c[1], c[2], c[3]);
    xs[i] = make_float4(__int_as_float(sv.px), __int_as_float(sv.py), sv.oow, sv.zow);
    xa[2 * (size_t)i] = make_float4(w[0], w[1], w[2], n[0]);
    xa[2 * (size_t)i + 1] = make_float4(n[1], n[2], S.tu[s], S.tv[s]);
    if (xt) {
        if ((D.flags & SR_DRAW_TANGENTS) && S.tangent) {
            float4 t = reinterpret_cast<const float4 *>(S.tangent)[s];
            float tw[3];
            for (int r = 0; r < 3; r++) {
                float acc = D.model[r] * t.x;
                acc = acc + D.model[4 + r] * t.y;
                acc = acc + D.model[8 + r] * t.z;
                tw[r] = acc;
            }
            xt[i] = make_float4(tw[0], tw[1], tw[2], t.w);
        } else {
            xt[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

// Wave-uniform copy of a per-lane index when every active lane holds the same
// value (then loads through it are scalar), else -1.
__device__ __forceinline__ int wave_uniform(int d) {
    int d0 = __builtin_amdgcn_readfirstlane(d);
    return __ballot(d != d0) == 0ull ? d0 : -1;
}

// Vertices b*256+lane, stepping by nb*256, of the transform's nb workgroups.
__device__ __forceinline__ void transform_vertices(const DevView *views, const DevDraw *draws, int ndraws,
                                                   uint32_t total, float4 *xp, float4 *xa, float4 *xt, float4 *xs,
                                                   uint32_t b, uint32_t nb) {
    for (uint32_t i = b * blockDim.x + threadIdx.x; i < total; i += nb * blockDim.x) {
        int d = find_last_le_consecutive(draws, ndraws, i, &DevDraw::xv_base);
        int du = wave_uniform(d);
        if (du >= 0) {
            const DevDraw &D = draws[du];
            transform_vertex(D, views[D.view], i, xp, xa, xt, xs);
        } else {
            const DevDraw &D = draws[d];
            transform_vertex(D, views[D.view], i, xp, xa, xt, xs);
        }
    }
}

// Orders LDS writes of some lanes of a wave before reads by other lanes of it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Triangle classification (k_front): fan triangles per input triangle.
// Triangles inside the frustum make one, rejected ones none; those that
// cross a plane are queued for k_clip, which writes their count.
// ---------------------------------------------------------------------------
// Frustum classification of input triangle t of draw D (-1 = inside, or no
// clipping needed), from clip positions computed here as the transform
// computes them.
__device__ __forceinline__ int classify_input_triangle(const DevDraw &D, const DevView *views, uint32_t t) {
    if (!(D.flags & SR_DRAW_CLIP)) return -1;
    const DevScene &S = *views[D.view].scene;
    uint32_t lt = t - D.in_tri_base;
    float c[3][4];
    for (int k = 0; k < 3; k++) clip_position(D, S, D.vtx_base + S.indices[D.idx_base + 3 * lt + k], c[k]);
    return clip_classify(c);
}

// counts[t]: 1 inside the frustum, 0 outside; for a triangle that crosses
// it (written by k_clip) SR_COUNT_CLIPPED | clip-queue slot << 4 | fan size.
#define SR_COUNT_CLIPPED 0x80000000u

// Post-clip triangles that count stands for.
__device__ __forceinline__ uint32_t count_fans(uint32_t c) { return (c & SR_COUNT_CLIPPED) ? (c & 15u) : c; }

// chunk[c] gets the number of triangles of chunk c (input triangles 256c ..
// 256c + 255) that lie inside the frustum; k_clip adds the fans of the
// clipped ones, and k_scan_top turns the sums into each chunk's first id.
// Workgroup b of nb (256 threads) takes chunks b, b + nb, ...
__device__ __forceinline__ void classify_triangles(const DevView *views, const DevDraw *draws, int ndraws,
                                                   uint32_t total, uint32_t *counts, uint32_t *chunk,
                                                   uint32_t *queue, uint32_t qcap, DevStatus *status, uint32_t b,
                                                   uint32_t nb) {
    __shared__ uint32_t wsum[4];
    const uint32_t stride = nb * blockDim.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t base = b * blockDim.x; base < total; base += stride) {
        uint32_t t = base + threadIdx.x;
        bool q = false, in = false;
        if (t < total) {
            int d = find_last_le_consecutive(draws, ndraws, t, &DevDraw::in_tri_base);
            int du = wave_uniform(d);
            int cls = du >= 0 ? classify_input_triangle(draws[du], views, t)
                              : classify_input_triangle(draws[d], views, t);
            in = cls == -1;
            counts[t] = in ? 1u : 0u;
            q = cls >= 0;
        }
        uint32_t nin = (uint32_t)__popcll(__ballot(in));
        __syncthreads();                 // the previous chunk's sums are read
        if (lane == 0) wsum[wave] = nin;
        __syncthreads();
        if (threadIdx.x == 0) chunk[base >> 8] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
        uint64_t m = __ballot(q);
        if (m) {
            int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t slot0 = 0;
            if (lane == leader) slot0 = atomicAdd(&status->clip_queued, (uint32_t)__popcll(m));
            slot0 = (uint32_t)__shfl((int)slot0, leader, 64);
            if (q) {
                uint32_t slot = slot0 + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (slot < qcap) queue[slot] = t;
                else atomicOr(&status->overflow_clip, 1u);
            }
        }
    }
}

// The frame's first kernel: workgroups [0, tri_blocks) classify the input
// triangles, the others transform the vertices (no data passes between the
// two: the classification computes its clip positions itself), and every
// workgroup helps clear the tiles' list bounds.
__global__ void __launch_bounds__(256) k_front(const DevView *__restrict__ views, const DevDraw *__restrict__ draws,
                                               int ndraws, uint32_t total_in, uint32_t total_xv,
                                               float4 *__restrict__ xp, float4 *__restrict__ xa,
                                               float4 *__restrict__ xt, float4 *__restrict__ xs, uint32_t *counts,
                                               uint32_t *chunk, uint32_t *queue, uint32_t qcap, DevStatus *status,
                                               FrameBegin F, uint32_t tri_blocks) {
    frame_begin(F);
    if (blockIdx.x < tri_blocks)
        classify_triangles(views, draws, ndraws, total_in, counts, chunk, queue, qcap, status, blockIdx.x, tri_blocks);
    else
        transform_vertices(views, draws, ndraws, total_xv, xp, xa, xt, xs, blockIdx.x - tri_blocks,
                           gridDim.x - tri_blocks);
}

// ---------------------------------------------------------------------------
// k_clip: Sutherland-Hodgman against the six planes with all vertex
// attributes (clip_against_frustum, src/renderer.rs:590-655).  Sixteen lanes
// per queued triangle, one per vertex channel in CV order; the polygon lives
// in LDS and every lane makes the same inside/outside decisions from the
// shared positions, so each channel gets exactly the reference's
// a + (b - a) * t at every step.  The clipped polygon goes to `buf` and its
// fan size to counts.  In exact arithmetic a triangle clipped by six planes
// has at most 9 vertices; rounding can make the polygon slightly non-convex,
// so room is kept for CLIP_MAXV and a polygon that would outgrow it fails the
// frame (status clip_overflow) rather than being cut short.
// ---------------------------------------------------------------------------
#define CLIP_MAXV 16
static_assert(CLIP_MAXV - 2 < 16, "a fan size fits the count's low four bits");

__global__ void __launch_bounds__(256) k_clip(const DevView *views, const DevDraw *draws, int ndraws,
                                              const float4 *xp, const float4 *xa, const float4 *xt, uint32_t *counts,
                                              uint32_t *chunk, const uint32_t *queue, uint32_t qcap, float *buf,
                                              DevStatus *status) {
    __shared__ float sh[16][2][CLIP_MAXV][16];
    const int g = threadIdx.x >> 4, c = threadIdx.x & 15;
    uint32_t nq = min(status->clip_queued, qcap);
    for (uint32_t e = blockIdx.x * 16u + (uint32_t)g; e < nq; e += gridDim.x * 16u) {
        uint32_t t = queue[e];
        int d = find_last_le(draws, ndraws, t, &DevDraw::in_tri_base);
        const DevDraw &D = draws[d];
        const DevScene &S = *views[D.view].scene;
        uint32_t lt = t - D.in_tri_base;
        float(*P)[16] = sh[g][0];
        float(*N)[16] = sh[g][1];
        for (int k = 0; k < 3; k++) {
            uint32_t vi = D.xv_base + S.indices[D.idx_base + 3 * lt + k];
            float val;
            if (c < 4) val = reinterpret_cast<const float *>(xp)[4 * (size_t)vi + c];
            else if (c < 12) val = reinterpret_cast<const float *>(xa)[8 * (size_t)vi + c - 4];
            else val = ((D.flags & SR_DRAW_TANGENTS) && xt) ? reinterpret_cast<const float *>(xt)[4 * (size_t)vi + c - 12]
                                                            : 0.0f;
            P[k][c] = val;
        }
        wave_lds_sync();
        int len = 3;
        bool over = false;
        for (int p = 0; p < 6 && len > 0; p++) {
            // a plane that every vertex is inside leaves the polygon as it is
            // (the pass would copy it in order): the group's lanes test one
            // vertex each and skip the pass (a 16-vertex polygon still takes
            // it, which fails the frame as the pass would)
            float dv = 0.0f;
            if (c < len) dv = plane_dot(p, P[c][0], P[c][1], P[c][2], P[c][3]);
            uint64_t outm = __ballot(c < len && !(dv >= 0.0f));
            if (((outm >> (threadIdx.x & 48u)) & 0xFFFFull) == 0ull && len < CLIP_MAXV) continue;
            int nl = 0;
            for (int i = 0; i < len; i++) {
                int ip = i == 0 ? len - 1 : i - 1;
                float dc = plane_dot(p, P[i][0], P[i][1], P[i][2], P[i][3]);
                float dp = plane_dot(p, P[ip][0], P[ip][1], P[ip][2], P[ip][3]);
                bool ci = dc >= 0.0f, pi = dp >= 0.0f;
                if (nl + 2 > CLIP_MAXV) {
                    over = true;
                    break;
                }
                if (ci) {
                    if (!pi) {
                        float tt = dp / (dp - dc);
                        N[nl][c] = P[ip][c] + (P[i][c] - P[ip][c]) * tt;
                        nl++;
                    }
                    N[nl][c] = P[i][c];
                    nl++;
                } else if (pi) {
                    float tt = dp / (dp - dc);
                    N[nl][c] = P[ip][c] + (P[i][c] - P[ip][c]) * tt;
                    nl++;
                }
            }
            if (over) {
                nl = 0;
                if (c == 0) atomicOr(&status->clip_overflow, 1u);
            }
            float(*tmp)[16] = P;
            P = N;
            N = tmp;
            len = nl;
            wave_lds_sync();
        }
        if (len < 3) len = 0;
        for (int i = 0; i < len; i++) buf[((size_t)e * CLIP_MAXV + i) * 16 + c] = P[i][c];
        if (c == 0) {
            uint32_t fan = len ? (uint32_t)(len - 2) : 0u;
            counts[t] = SR_COUNT_CLIPPED | (e << 4) | fan;
            if (fan) atomicAdd(&chunk[t >> 8], fan);
        }
        wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (4096 elements per block)
// ---------------------------------------------------------------------------
#define SCAN_ITEMS 16
#define SCAN_BLOCK 256
#define SCAN_TILE (SCAN_ITEMS * SCAN_BLOCK)

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive block scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t &excl, uint32_t *sh) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < nw; w++) {
            uint32_t t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[16] = acc;
    }
    __syncthreads();
    excl = sh[wave] + inc - v;
    uint32_t total = sh[16];
    __syncthreads();
    return total;
}

// Each wave owns a contiguous segment of 64 x SCAN_ITEMS elements and reads
// it in SCAN_ITEMS coalesced rows (element seg + k * 64 + lane); v[k] gets
// the element and ex[k] its exclusive prefix within the segment.  Returns
// the segment total.
__device__ __forceinline__ uint32_t wave_scan_segment(const uint32_t *in, size_t seg, size_t n,
                                                      uint32_t ex[SCAN_ITEMS], uint32_t v[SCAN_ITEMS]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        size_t i = seg + (size_t)k * 64 + (size_t)lane;
        v[k] = i < n ? in[i] : 0u;
    }
    uint32_t carry = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        uint32_t inc = wave_incl_scan(v[k]);
        ex[k] = carry + inc - v[k];
        carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    }
    return carry;
}

// out gets the exclusive prefix; out2 (if not null) the same, or the
// inclusive one with SCAN_OUT2_INCLUSIVE.
#define SCAN_OUT2_INCLUSIVE 1u

__device__ __forceinline__ void wave_store_segment(uint32_t *out, uint32_t *out2, size_t seg, size_t n, uint32_t base,
                                                   const uint32_t ex[SCAN_ITEMS], const uint32_t v[SCAN_ITEMS],
                                                   uint32_t flags) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        size_t i = seg + (size_t)k * 64 + (size_t)lane;
        if (i < n) {
            out[i] = base + ex[k];
            if (out2) out2[i] = base + ex[k] + ((flags & SCAN_OUT2_INCLUSIVE) ? v[k] : 0u);
        }
    }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_partials(const uint32_t *in, uint32_t n, uint32_t *partials,
                                                               uint64_t *stamp) {
    __shared__ uint32_t sh[17];
    stamp_start(stamp);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    size_t seg = (size_t)blockIdx.x * SCAN_TILE + (size_t)wave * (64 * SCAN_ITEMS);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        size_t i = seg + (size_t)k * 64 + (size_t)lane;
        if (i < n) s += in[i];
    }
    uint32_t ex;
    uint32_t total = block_excl_scan(s, ex, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) k_scan_top(uint32_t *partials, uint32_t nparts, uint32_t *total_out,
                                                  uint32_t *total_out2) {
    __shared__ uint32_t sh[17];
    uint32_t carry = 0;
    // four consecutive elements per thread, all loaded before one workgroup
    // scan (4096 per round)
    for (uint32_t c = 0; c < nparts; c += 4096u) {
        const uint32_t i = c + 4u * threadIdx.x;
        uint32_t v[4], sum = 0u;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = i + k < nparts ? partials[i + k] : 0u;
            sum += v[k];
        }
        uint32_t ex;
        uint32_t tot = block_excl_scan(sum, ex, sh);
        uint32_t run = carry + ex;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (i + k < nparts) partials[i + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        if (total_out) *total_out = carry;
        if (total_out2) *total_out2 = carry;
    }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(const uint32_t *in, uint32_t n, const uint32_t *partials,
                                                           uint32_t *out, uint32_t *out2, uint32_t flags) {
    __shared__ uint32_t sh[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    size_t seg = (size_t)blockIdx.x * SCAN_TILE + (size_t)wave * (64 * SCAN_ITEMS);
    uint32_t ex[SCAN_ITEMS], v[SCAN_ITEMS];
    uint32_t tot = wave_scan_segment(in, seg, n, ex, v);
    if (lane == 0) sh[wave] = tot;
    __syncthreads();
    uint32_t base = partials[blockIdx.x];
    for (int w = 0; w < wave; w++) base += sh[w];
    wave_store_segment(out, out2, seg, n, base, ex, v, flags);
}

// One workgroup scans arrays of up to SCAN_SINGLE_MAX elements in a single
// launch (workgroup b takes job b, so independent small scans share one
// launch); larger arrays use the partials / top / apply sequence above.
#define SCAN_SINGLE_MAX 65536u

struct ScanJob {
    const uint32_t *in;
    uint32_t n;
    uint32_t *out, *out2, *total, *total2;
    uint32_t flags;
};

struct ScanJobs {
    ScanJob job[2];
    uint64_t *stamp;              // start stamp (or null)
};

__global__ void __launch_bounds__(1024) k_scan_single(ScanJobs J) {
    __shared__ uint32_t sh[16];
    stamp_start(J.stamp);
    const ScanJob &q = J.job[blockIdx.x];
    const uint32_t *in = q.in;
    const uint32_t n = q.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint32_t c = 0; c < n; c += 1024u * SCAN_ITEMS) {
        size_t seg = (size_t)c + (size_t)wave * (64 * SCAN_ITEMS);
        uint32_t ex[SCAN_ITEMS], v[SCAN_ITEMS];
        uint32_t tot = wave_scan_segment(in, seg, n, ex, v);
        if (lane == 0) sh[wave] = tot;
        __syncthreads();
        uint32_t base = carry, all = 0;
        for (int w = 0; w < 16; w++) {
            if (w < wave) base += sh[w];
            all += sh[w];
        }
        wave_store_segment(q.out, q.out2, seg, n, base, ex, v, q.flags);
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (q.total) *q.total = carry;
        if (q.total2) *q.total2 = carry;
    }
}

static ScanJob scan_job(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *out2, uint32_t *total,
                        uint32_t *total2, uint32_t flags) {
    ScanJob j;
    j.in = in;
    j.n = n;
    j.out = out;
    j.out2 = out2;
    j.total = total;
    j.total2 = total2;
    j.flags = flags;
    return j;
}

hipError_t launch_scan(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *out2, uint32_t *partials,
                       uint32_t *total, hipStream_t st, uint32_t flags, uint32_t *total2, uint64_t *start_stamp) {
    if (n <= SCAN_SINGLE_MAX) {
        ScanJobs J;
        J.job[0] = scan_job(in, n, out, out2, total, total2, flags);
        J.job[1] = J.job[0];
        J.stamp = start_stamp;
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, st, J);
        return hipGetLastError();
    }
    uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_partials, dim3(nb), dim3(SCAN_BLOCK), 0, st, in, n, partials, start_stamp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, partials, nb, total, total2);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_BLOCK), 0, st, in, n, partials, out, out2, flags);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Binning rectangle: GPU tiles touched by the quad-aligned pixel bbox
// ---------------------------------------------------------------------------
struct Rect { int gx0, gy0, gx1, gy1; };

__device__ __forceinline__ Rect bin_rect(const DevView &V, int mnx, int mny, int mxx, int mxy, int T) {
    // quads cover [min & ~1, round_up_even(max)) in each axis
    int qx0 = mnx & ~1, qy0 = mny & ~1;
    int qx1 = (mxx + 1) & ~1, qy1 = (mxy + 1) & ~1;
    int gw = V.tiles_x * T, gh = V.tiles_y * T;
    qx1 = qx1 < gw ? qx1 : gw;
    qy1 = qy1 < gh ? qy1 : gh;
    Rect r;
    r.gx0 = qx0 / T;
    r.gy0 = qy0 / T;
    r.gx1 = (qx1 + T - 1) / T;
    r.gy1 = (qy1 + T - 1) / T;
    return r;
}

// ---------------------------------------------------------------------------
// k_setup: clip with attributes, snap, cull, bbox, setup records, pair counts
// ---------------------------------------------------------------------------
struct SetupOut {
    SR_GLOBAL uint4 *geo;
    SR_GLOBAL float4 *raster;
    SR_GLOBAL uint32_t *chunk_pairs;      // (triangle, tile) pairs of each 256-triangle chunk's records
    uint32_t cap;
    SR_GLOBAL DevStatus *status;
    int T;
};

__device__ __forceinline__ int32_t top_left_bias(int32_t a, int32_t b) { return (a < 0 || (a == 0 && b > 0)) ? 0 : -1; }

// Edge functions of the snapped triangle (src/tilerasterizer.rs:139-164):
// edge 0 = v1->v2 (w0), 1 = v2->v0 (w1), 2 = v0->v1 (w2); i32 constants in
// wrapping arithmetic, f32 coefficients and 2-px steps.
__device__ __forceinline__ void edge_setup(int32_t p0x, int32_t p0y, int32_t p1x, int32_t p1y, int32_t p2x, int32_t p2y,
                                           float A[3], float B[3], float C[3], float SX[3], float SY[3]) {
    int32_t a01 = wsub(p1y, p0y), b01 = wsub(p0x, p1x);
    int32_t c01 = wadd(wsub(wmul(p1x, p0y), wmul(p0x, p1y)), top_left_bias(a01, b01));
    int32_t a12 = wsub(p2y, p1y), b12 = wsub(p1x, p2x);
    int32_t c12 = wadd(wsub(wmul(p2x, p1y), wmul(p1x, p2y)), top_left_bias(a12, b12));
    int32_t a20 = wsub(p0y, p2y), b20 = wsub(p2x, p0x);
    int32_t c20 = wadd(wsub(wmul(p0x, p2y), wmul(p2x, p0y)), top_left_bias(a20, b20));
    A[0] = (float)a12; B[0] = (float)b12; C[0] = (float)c12;
    A[1] = (float)a20; B[1] = (float)b20; C[1] = (float)c20;
    A[2] = (float)a01; B[2] = (float)b01; C[2] = (float)c01;
    SX[0] = (float)wmul(a12, 32); SY[0] = (float)wmul(b12, 32);
    SX[1] = (float)wmul(a20, 32); SY[1] = (float)wmul(b20, 32);
    SX[2] = (float)wmul(a01, 32); SY[2] = (float)wmul(b01, 32);
}

// A small opaque triangle that makes no fragment: its pixel box
// [pmnx, pmxx) x [pmny, pmxy) lies in one 64-px reference tile and spans at
// most two quads each way, so the reference walks it with the fine raster
// from the quad-aligned corner (src/tilerasterizer.rs:191-217, 293-383),
// and no sample of those quads passes all three edge tests.  The values are
// the raster's own: origin evaluation (A x + B y) + C, then +step_y for the
// second quad row, then +step_x for the second column.
__device__ __forceinline__ bool small_tri_uncovered(const int32_t px[3], const int32_t py[3], int32_t pmnx, int32_t pmny,
                                                    int32_t pmxx, int32_t pmxy) {
    int32_t ox = pmnx & ~1, oy = pmny & ~1;
    if (pmxx - ox > 4 || pmxy - oy > 4) return false;
    if ((pmnx >> 6) != ((pmxx - 1) >> 6) || (pmny >> 6) != ((pmxy - 1) >> 6)) return false;
    float A[3], B[3], C[3], SX[3], SY[3];
    edge_setup(px[0], py[0], px[1], py[1], px[2], py[2], A, B, C, SX, SY);
    bool col2 = ox + 2 < pmxx, row2 = oy + 2 < pmxy;
    bool covered = false;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        float X = (float)(ox * 16 + 8 + 16 * (s & 1)), Y = (float)(oy * 16 + 8 + 16 * (s >> 1));
        bool q00 = true, q10 = true, q01 = true, q11 = true;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            float v = (A[e] * X + B[e] * Y) + C[e];
            float vy = v + SY[e];
            q00 = q00 && v >= 0.0f;
            q10 = q10 && (v + SX[e]) >= 0.0f;
            q01 = q01 && vy >= 0.0f;
            q11 = q11 && (vy + SX[e]) >= 0.0f;
        }
        covered = covered || q00 || (col2 && q10) || (row2 && q01) || (col2 && row2 && q11);
    }
    return !covered;
}

// Snap, cull, bbox and setup of one (post-clip) triangle (bin_triangle,
// src/renderer.rs:668-831).  Writes the geometry and raster records and
// returns the bin rectangle (empty when culled, when the box is empty, or
// for a small opaque triangle that covers no sample: small_tri_uncovered;
// such a triangle keeps its id but gets no raster record and no bins).
// The shading interpolators of the reference's packet (:701-754) are not
// stored: tri_shade_attrs recomputes them, with the same operations, for the
// few triangles that win a pixel, from the vertices that `src` locates
// (SRC_* encoding below).
#define SRC_CLIPPED 0x80000000u       // src0 = SRC_CLIPPED | fan << 27 | clip slot
__device__ Rect emit_triangle(const DevView &V, int view, const SnapV sv[3], uint32_t id, int material,
                              uint32_t matflags, bool translucent, uint32_t src0, uint32_t src1, const SetupOut &O) {
    Rect none = {0, 0, 0, 0};
    if (id >= O.cap) {
        atomicOr(&O.status->overflow_records, 1u);
        return none;
    }
    int32_t px[3] = {sv[0].px, sv[1].px, sv[2].px}, py[3] = {sv[0].py, sv[1].py, sv[2].py};
    int32_t area = wsub(wmul(wsub(px[1], px[0]), wsub(py[2], py[0])), wmul(wsub(px[2], px[0]), wsub(py[1], py[0])));
    if (area > 0) {
        O.geo[id] = make_uint4(0u, 0u, SR_GEO_EMPTY, (uint32_t)view);
        return none;
    }
    int32_t mnx = max(min(min(px[0], px[1]), px[2]), 0);
    int32_t mny = max(min(min(py[0], py[1]), py[2]), 0);
    int32_t mxx = min(max(max(px[0], px[1]), px[2]), V.W_sub);
    int32_t mxy = min(max(max(py[0], py[1]), py[2]), V.H_sub);
    int32_t pmnx = mnx >> 4, pmny = mny >> 4;
    int32_t pmxx = wadd(mxx, 16) >> 4, pmxy = wadd(mxy, 16) >> 4;
    if (!(pmnx < pmxx && pmny < pmxy) || (!translucent && small_tri_uncovered(px, py, pmnx, pmny, pmxx, pmxy))) {
        O.geo[id] = make_uint4(0u, 0u, SR_GEO_EMPTY, (uint32_t)view);
        return none;
    }
    float ooa = 1.0f / (float)wabs(area);
    float oow[3] = {sv[0].oow, sv[1].oow, sv[2].oow}, zow[3] = {sv[0].zow, sv[1].zow, sv[2].zow};
    uint32_t flags = 0;
    if (matflags & SR_MAT_ALPHA_TESTED) flags |= SR_GEO_ALPHA;
    if (matflags & SR_MAT_NORMAL_MAP) flags |= SR_GEO_NORMAL_MAP;
    if (translucent) flags |= SR_GEO_TRANSLUCENT;
    O.geo[id] = make_uint4((uint32_t)pmnx | ((uint32_t)pmny << 16), (uint32_t)pmxx | ((uint32_t)pmxy << 16), flags,
                           (uint32_t)view);
    float4 *R = O.raster + 4 * (size_t)id;
    R[0] = make_float4(__int_as_float(px[0]), __int_as_float(py[0]), __int_as_float(px[1]), __int_as_float(py[1]));
    R[1] = make_float4(__int_as_float(px[2]), __int_as_float(py[2]), ooa, oow[0]);
    R[2] = make_float4(oow[1] - oow[0], oow[2] - oow[0], zow[0], zow[1] - zow[0]);
    R[3] = make_float4(zow[2] - zow[0], __int_as_float(material), __uint_as_float(src0), __uint_as_float(src1));
    if (translucent) return none;     // 64-px reference-tile lists instead (k_tr_bins)
    return bin_rect(V, pmnx, pmny, pmxx, pmxy, O.T);
}

// Sources of a set-up triangle's vertices.
struct TriSource {
    const SR_GLOBAL DevView *views;
    const SR_GLOBAL DevDraw *draws;
    const SR_GLOBAL float4 *xp;         // transformed clip-space positions
    const SR_GLOBAL float4 *xa;         // transformed attributes (2 per vertex)
    const SR_GLOBAL float4 *xt;         // transformed tangents (or null)
    const SR_GLOBAL float4 *xs;         // per-vertex snap (px, py, 1/w, z/w)
    const SR_GLOBAL float *clip_buf;    // clipped polygons
};

// Transformed vertices of input triangle src0 of draw D; the first float4 is
// the vertex's snap, whose z is 1/w (shading reads no clip position).
__device__ __forceinline__ void load_input_tri(const TriSource &T, const DevDraw &D, uint32_t src0, float4 q[3][4]) {
    const DevScene &S = *T.views[D.view].scene;
    uint32_t lt = src0 - D.in_tri_base;
    bool tan = (D.flags & SR_DRAW_TANGENTS) && T.xt;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        size_t i = (size_t)D.xv_base + S.indices[D.idx_base + 3 * lt + k];
        q[k][0] = T.xs[i]; q[k][1] = T.xa[2 * i]; q[k][2] = T.xa[2 * i + 1];
        q[k][3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (tan) q[k][3] = T.xt[i];
    }
}

// The three vertices of the triangle whose raster record is R, with their
// 1/w: the indexed transformed vertices of an unclipped input triangle
// (src0 = input triangle, src1 = draw; 1/w from the transform's snap), or
// vertices 0, f+1, f+2 of a clipped polygon (1/w computed here).  The clip
// position is not needed for shading and is left unset.
__device__ __forceinline__ void load_tri_cv(const TriSource &T, const float4 *R, CV v[3]) {
    uint32_t src0 = __float_as_uint(R[3].z), src1 = __float_as_uint(R[3].w);
    float4 q[3][4];
    if (src0 & SRC_CLIPPED) {
        uint32_t e = src0 & 0x07FFFFFFu, f = (src0 >> 27) & 15u;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            uint32_t vi = j == 0 ? 0u : f + (uint32_t)j;
            const float4 *p = reinterpret_cast<const float4 *>(T.clip_buf + ((size_t)e * 16u + vi) * 16u);
            q[j][0] = p[0]; q[j][1] = p[1]; q[j][2] = p[2]; q[j][3] = p[3];
            q[j][0].z = 1.0f / q[j][0].w;
        }
    } else {
        int du = wave_uniform((int)src1);
        if (du >= 0) load_input_tri(T, T.draws[du], src0, q);
        else load_input_tri(T, T.draws[src1], src0, q);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        v[j].oow = q[j][0].z;
        v[j].wx = q[j][1].x; v[j].wy = q[j][1].y; v[j].wz = q[j][1].z; v[j].nx = q[j][1].w;
        v[j].ny = q[j][2].x; v[j].nz = q[j][2].y; v[j].u = q[j][2].z; v[j].v = q[j][2].w;
        v[j].tx = q[j][3].x; v[j].ty = q[j][3].y; v[j].tz = q[j][3].z; v[j].tw = q[j][3].w;
    }
}

// The reference packet's shading interpolators (src/renderer.rs:701-754):
// normal, position/w, u/w, v/w as (a, b - a, c - a), UV derivatives, and the
// tangent frame's when normal mapped.  Sh[0..6] / Tn[0..2] as read by
// shade_pixel and alpha_pass.
__device__ __forceinline__ void tri_shade_attrs(const CV v[3], const float4 *R, bool tangents, float4 Sh[7],
                                                float4 Tn[3]) {
    float4 r0 = R[0], r1 = R[1];
    int32_t px[3] = {__float_as_int(r0.x), __float_as_int(r0.z), __float_as_int(r1.x)};
    int32_t py[3] = {__float_as_int(r0.y), __float_as_int(r0.w), __float_as_int(r1.y)};
    float ooa = r1.z;
    float oow[3], uw[3], vw[3];
    f3 pw[3];
#pragma unroll
    for (int k = 0; k < 3; k++) oow[k] = v[k].oow;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        uw[k] = v[k].u * oow[k];
        vw[k] = v[k].v * oow[k];
        pw[k] = mk3(v[k].wx * oow[k], v[k].wy * oow[k], v[k].wz * oow[k]);
    }
    float dx1 = (float)wsub(px[1], px[0]), dx2 = (float)wsub(px[2], px[0]);
    float dy1 = (float)wsub(py[1], py[0]), dy2 = (float)wsub(py[2], py[0]);
    float du1 = uw[1] - uw[0], du2 = uw[2] - uw[0];
    float dv1 = vw[1] - vw[0], dv2 = vw[2] - vw[0];
    float du_dx = du1 * dy2 - du2 * dy1;
    float du_dy = du2 * dx1 - du1 * dx2;
    float dv_dx = dv1 * dy2 - dv2 * dy1;
    float dv_dy = dv2 * dx1 - dv1 * dx2;
    float scale = ooa * 16.0f;
    f3 n0 = mk3(v[0].nx, v[0].ny, v[0].nz), n1 = mk3(v[1].nx, v[1].ny, v[1].nz), n2 = mk3(v[2].nx, v[2].ny, v[2].nz);
    f3 nda = sub3(n1, n0), ndb = sub3(n2, n0);
    f3 pda = sub3(pw[1], pw[0]), pdb = sub3(pw[2], pw[0]);
    Sh[0] = make_float4(n0.x, n0.y, n0.z, nda.x);
    Sh[1] = make_float4(nda.y, nda.z, ndb.x, ndb.y);
    Sh[2] = make_float4(ndb.z, pw[0].x, pw[0].y, pw[0].z);
    Sh[3] = make_float4(pda.x, pda.y, pda.z, pdb.x);
    Sh[4] = make_float4(pdb.y, pdb.z, uw[0], uw[1] - uw[0]);
    Sh[5] = make_float4(uw[2] - uw[0], vw[0], vw[1] - vw[0], vw[2] - vw[0]);
    Sh[6] = make_float4(du_dx * scale, du_dy * scale, dv_dx * scale, dv_dy * scale);
    if (tangents) {
        f3 t0 = mk3(v[0].tx, v[0].ty, v[0].tz), t1 = mk3(v[1].tx, v[1].ty, v[1].tz), t2 = mk3(v[2].tx, v[2].ty, v[2].tz);
        f3 tda = sub3(t1, t0), tdb = sub3(t2, t0);
        Tn[0] = make_float4(t0.x, t0.y, t0.z, tda.x);
        Tn[1] = make_float4(tda.y, tda.z, tdb.x, tdb.y);
        Tn[2] = make_float4(tdb.z, v[0].tw, v[1].tw - v[0].tw, v[2].tw - v[0].tw);
    }
}

__device__ __forceinline__ int wave_max(int v) {
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}


// The bins of a rectangle of GPU tiles (row-major over its view's tiles).
struct RectBins {
    int nb;            // bins in the rectangle (0 if empty)
    int rw;            // rectangle width in tiles
    uint32_t first;    // bin index of its top-left tile
    int stride;        // tiles per row of its view
};

__device__ __forceinline__ RectBins rect_bins(const DevView *views, int view, Rect r) {
    RectBins b;
    b.rw = r.gx1 - r.gx0;
    int rh = r.gy1 - r.gy0;
    b.nb = (b.rw > 0 && rh > 0) ? b.rw * rh : 0;
    b.first = 0u;
    b.stride = 0;
    if (b.nb > 0) {
        const DevView &V = views[view];
        b.stride = V.tiles_x;
        b.first = V.item_base + (uint32_t)(r.gy0 * V.tiles_x + r.gx0);
    }
    return b;
}

// Setup of input triangle t of draw D (index d), which the classification (k_front) found
// inside the frustum (the fans of those that cross it come from their
// clipped polygons, emit_clipped_fan).
__device__ __forceinline__ Rect setup_input_triangle(const DevDraw &D, int d, const DevView *views, uint32_t t,
                                                     const float4 *xs, uint32_t id, const SetupOut &O) {
    const DevView &V = views[D.view];
    const DevScene &S = *V.scene;
    uint32_t lt = t - D.in_tri_base;
    SnapV v[3];
    for (int k = 0; k < 3; k++) {
        uint32_t vi = S.indices[D.idx_base + 3 * lt + k];
        float4 q = xs[D.xv_base + vi];    // the transform's per-vertex snap
        v[k].px = __float_as_int(q.x); v[k].py = __float_as_int(q.y); v[k].oow = q.z; v[k].zow = q.w;
    }
    return emit_triangle(V, D.view, v, id, D.material, S.materials[D.material].flags, !(D.flags & SR_DRAW_OPAQUE),
                         t, (uint32_t)d, O);
}

// Fan triangle f, (v0, v_f+1, v_f+2), of the clipped polygon in clip-queue
// slot e (src/renderer.rs:657-666), set up as post-clip triangle id.
__device__ __forceinline__ Rect emit_clipped_fan(const DevView *views, const DevDraw &D, const float *buf, uint32_t e,
                                                  uint32_t f, uint32_t id, const SetupOut &O) {
    const DevView &V = views[D.view];
    const DevScene &S = *V.scene;
    SnapV v[3];
    const uint32_t src[3] = {0u, f + 1u, f + 2u};
    for (int j = 0; j < 3; j++) {
        float4 a = *reinterpret_cast<const float4 *>(buf + ((size_t)e * CLIP_MAXV + src[j]) * 16);
        v[j] = snap_vertex(V.Wf, V.Hf, a.x, a.y, a.z, a.w);
    }
    return emit_triangle(V, D.view, v, id, D.material, S.materials[D.material].flags, !(D.flags & SR_DRAW_OPAQUE),
                         SRC_CLIPPED | (f << 27) | e, 0u, O);
}

// The post-clip id of input triangle t is its chunk's first id plus the
// fan counts before it in the chunk (a workgroup scan); it is stored for
// the per-view offsets.  A triangle inside the frustum is set up from its
// clip positions; one that crosses it has its fan triangles set up from
// the polygon k_clip left in the clip buffer (ids tri_base[t] + f).
__global__ void __launch_bounds__(256) k_setup(const DevView *views, const DevDraw *draws, int ndraws, uint32_t total,
                                               const float4 *xs, const uint32_t *counts, const uint32_t *chunk,
                                               const float *clip_buf, uint32_t *tri_base, SetupOut O) {
    __shared__ uint32_t sh[17];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
        uint32_t t = base + threadIdx.x;
        uint32_t ex, c = t < total ? counts[t] : 0u;
        uint32_t nf = count_fans(c);
        block_excl_scan(nf, ex, sh);
        uint32_t id = chunk[base >> 8] + ex;
        Rect r = {0, 0, 0, 0};
        int view = 0, d = 0;
        if (t < total) {
            tri_base[t] = id;
            if (t == total - 1) tri_base[total] = id + nf;
            d = find_last_le_consecutive(draws, ndraws, t, &DevDraw::in_tri_base);
            int du = wave_uniform(d);
            if (du >= 0) {
                view = draws[du].view;
                if (c == 1u) r = setup_input_triangle(draws[du], du, views, t, xs, id, O);
            } else {
                view = draws[d].view;
                if (c == 1u) r = setup_input_triangle(draws[d], d, views, t, xs, id, O);
            }
        }
        uint32_t np = (uint32_t)rect_bins(views, view, r).nb;
        int nclip = (c & SR_COUNT_CLIPPED) ? (int)nf : 0;
        int maxf = wave_max(nclip);
        for (int f = 0; f < maxf; f++) {
            Rect rf = {0, 0, 0, 0};
            if (f < nclip)
                rf = emit_clipped_fan(views, draws[d], clip_buf, (c & ~SR_COUNT_CLIPPED) >> 4, (uint32_t)f,
                                      id + (uint32_t)f, O);
            np += (uint32_t)rect_bins(views, view, rf).nb;
        }
        uint32_t npx;
        uint32_t chunk_np = block_excl_scan(np, npx, sh);
        if (threadIdx.x == 0) O.chunk_pairs[base >> 8] = chunk_np;
    }
}

// ---------------------------------------------------------------------------
// Stable binning: every (triangle, tile) pair into its tile's list in
// submission order
// ---------------------------------------------------------------------------
// The reference's single-threaded front end pushes packets into each tile's
// queue in submission order (src/bumpqueue.rs:96-112).  Record ids are that
// order (node distance, opaque then translucent primitives, triangle,
// clip-fan index), and a record appears at most once in a bin, so a tile's
// list is its records in ascending id.  k_bin_expand writes every record's
// pairs (bin, record) in record order into one flat array, at offsets from
// the per-chunk pair counts that k_setup leaves; then a
// least-significant-digit radix sort on the global bin index, 8 bits per
// pass, orders them by bin.  Every pass is stable, so a bin keeps ascending
// record ids; the last pass writes the record ids (the lists) and their
// bins, and k_bin_bounds marks where each bin's run starts and ends.  No
// per-bin counts are taken beforehand (their atomics cost k_setup a third
// of its time).
//
// A pass is count, scan, scatter over tiles of 256 x IT elements (a
// contiguous quarter per wave): k_bin_hist counts each tile's digits into a
// digit-major matrix, k_bin_rowscan turns each digit's row into the tiles'
// offsets within the digit and leaves the digit's total, and k_bin_scatter
// scans the totals into the digits' first positions and writes each tile.
// Within a wave, elements with equal digits are ranked by lane with
// ballots; the waves of a tile are ordered by their per-digit counts.  No
// workgroup waits on another.
// Tiles hold 256 x IT elements (IT per thread): 4 up to BS_SMALL_PAIRS
// pairs of capacity, so a frame's few hundred thousand pairs still spread
// over several hundred workgroups, else 16.
#define BS_SMALL_PAIRS (4u << 20)

struct BinPass {
    const SR_GLOBAL DevView *views;
    const SR_GLOBAL uint4 *geo;                     // expansion: the records
    const SR_GLOBAL uint32_t *chunk_first;          // expansion: first record of each 256-triangle chunk
    const SR_GLOBAL uint32_t *chunk_pairs;          // expansion: first pair of each chunk
    uint32_t nchunks, parts;
    const SR_GLOBAL uint32_t *in_keys, *in_vals;    // radix passes: (bin, record) in the last pass's order
    SR_GLOBAL uint32_t *out_keys, *out_vals;        // the expansion and every pass (the last: keys only)
    SR_GLOBAL uint32_t *list;                       // the last pass: record ids per bin
    SR_GLOBAL uint32_t *hist;                       // 256 rows of hist_cols tile counts, then offsets
    uint32_t hist_cols;
    SR_GLOBAL DevStatus *status;
    const SR_GLOBAL BinSortState *bs;
    SR_GLOBAL uint32_t *dtotal;                     // this pass's count per digit (row scan -> scatter)
    uint32_t total_items, rec_cap, pair_cap;
    uint32_t tile_elems;                            // elements per radix tile (256 x IT)
    int T, pass, last, shift, bits;
};

// Pair kk of a rectangle (row-major over its tiles).
__device__ __forceinline__ uint32_t bs_pair_bin(uint32_t first, int rw, int stride, int kk) {
    return first + (uint32_t)((kk / rw) * stride + kk % rw);
}

// Expansion: the pairs of chunk c's records [chunk_first[c], next chunk's
// first) in record order.  A chunk's pairs are taken 256 at a time (a slab),
// one per thread; `parts` workgroups share a chunk, each taking a
// contiguous range of its slabs, so a chunk of large triangles (tens of
// pairs per record) spreads over several workgroups.  The host sets parts
// from the ratio of the pair and record capacities (1 for small triangles,
// up to BS_EXPAND_PARTS), so frames of small triangles launch no idle parts.  Each part
// loads the chunk's records 256 at a time and finds a pair's record by
// binary search over their pair counts' prefix in LDS.
#define BS_EXPAND_PARTS 8u

__global__ void __launch_bounds__(256) k_bin_expand(BinPass P) {
    __shared__ uint32_t sh[17];
    __shared__ uint32_t pex[257];
    __shared__ uint32_t rfirst[256];
    __shared__ int rrw[256], rstride[256];
    const uint32_t tid = threadIdx.x;
    const uint32_t total = min(P.status->total_records, P.rec_cap);
    const uint32_t parts = P.parts;
    const uint32_t items = P.nchunks * parts;
    // part-major order: the first parts of all chunks come first, so the
    // parts with work spread over the grid
    for (uint32_t w = blockIdx.x; w < items; w += gridDim.x) {
        const uint32_t c = w % P.nchunks, q = w / P.nchunks;
        const uint32_t cbase = P.chunk_pairs[c];
        const uint32_t cend = c + 1 < P.nchunks ? P.chunk_pairs[c + 1] : P.bs->chunk_total;
        const uint32_t nslabs = (cend - cbase + 255u) / 256u;
        const uint32_t per = (nslabs + parts - 1) / parts;
        const uint32_t jlo = q * per, jhi = min(jlo + per, nslabs);
        // past the capacity the frame runs again with larger buffers (the
        // host sees total_pairs > pair_cap)
        if (jlo >= jhi || cbase + jlo * 256u >= P.pair_cap) continue;
        const uint32_t first = P.chunk_first[c];
        const uint32_t last = min(c + 1 < P.nchunks ? P.chunk_first[c + 1] : P.status->total_records, total);
        uint32_t sbase = 0;     // the sub-batch's first pair, counted from the chunk's
        for (uint32_t r0 = first; r0 < last && sbase < jhi * 256u; r0 += 256u) {
            const uint32_t id = r0 + tid;
            Rect r = {0, 0, 0, 0};
            int view = 0;
            if (id < last) {
                uint4 g = P.geo[id];
                view = (int)g.w;
                if (!(g.z & (SR_GEO_EMPTY | SR_GEO_TRANSLUCENT))) {
                    int mnx = (int)(g.x & 0xFFFFu), mny = (int)(g.x >> 16), mxx = (int)(g.y & 0xFFFFu), mxy = (int)(g.y >> 16);
                    r = bin_rect(P.views[view], mnx, mny, mxx, mxy, P.T);
                }
            }
            const RectBins b = rect_bins((const DevView *)P.views, view, r);
            uint32_t ex;
            const uint32_t sub = block_excl_scan((uint32_t)b.nb, ex, sh);
            pex[tid] = ex;
            rfirst[tid] = b.first;
            rrw[tid] = b.rw;
            rstride[tid] = b.stride;
            if (tid == 0) pex[256] = sub;
            __syncthreads();
            const uint32_t ja = max(jlo, sbase / 256u);
            for (uint32_t j = ja; j < jhi && j * 256u < sbase + sub; j++) {
                const uint32_t p = j * 256u + tid;
                if (p >= sbase && p < sbase + sub) {
                    const uint32_t lp = p - sbase;
                    // the last record whose first pair is at or before lp
                    // (it has pairs: a record without any shares its
                    // prefix with the next)
                    uint32_t o = 0;
#pragma unroll
                    for (uint32_t step = 128; step > 0; step >>= 1)
                        if (pex[o + step] <= lp) o += step;
                    const int kk = (int)(lp - pex[o]);
                    const uint32_t pos = cbase + p;
                    if (pos < P.pair_cap) {
                        P.out_keys[pos] = bs_pair_bin(rfirst[o], rrw[o], rstride[o], kk);
                        P.out_vals[pos] = r0 + o;
                    }
                }
            }
            sbase += sub;
            __syncthreads();
        }
    }
}

// Valid lanes of the wave whose digit equals this lane's.
__device__ __forceinline__ uint64_t bs_match(bool valid, uint32_t d, int bits) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < bits; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t v = __ballot(bit);
        m &= bit ? v : ~v;
    }
    return m;
}

__device__ __forceinline__ uint32_t bs_elements(const BinPass &P) { return min(P.status->total_pairs, P.pair_cap); }

// This wave's 1024 keys of tile t (zero past the end), and their digits
// counted per wave into cnt[wave][digit] (cleared first; LDS adds, whose
// order does not matter here).
template <int IT>
__device__ __forceinline__ void bs_load_count(const BinPass &P, uint32_t t, uint32_t n, uint32_t key[IT],
                                              uint32_t (*cnt)[256]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t wbase = t * (256u * IT) + (uint32_t)wave * (64u * IT);
    const uint32_t mask = (1u << P.bits) - 1u;
#pragma unroll
    for (int s = 0; s < IT; s++) {
        uint32_t i = wbase + (uint32_t)(s * 64 + lane);
        key[s] = i < n ? P.in_keys[i] : 0u;
    }
    for (int k = lane; k < 256; k += 64) cnt[wave][k] = 0u;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int s = 0; s < IT; s++)
        if (wbase + (uint32_t)(s * 64 + lane) < n) atomicAdd(&cnt[wave][(key[s] >> P.shift) & mask], 1u);
}

// Count: digit counts of each tile into hist[digit][tile].
template <int IT>
__global__ void __launch_bounds__(256) k_bin_hist(BinPass P) {
    __shared__ uint32_t cnt[4][256];
    const uint32_t n = bs_elements(P);
    const uint32_t t = blockIdx.x;
    if (t * (256u * IT) >= n) return;
    uint32_t key[IT];
    bs_load_count<IT>(P, t, n, key, cnt);
    __syncthreads();
    const uint32_t d = threadIdx.x;
    P.hist[(size_t)d * P.hist_cols + t] = cnt[0][d] + cnt[1][d] + cnt[2][d] + cnt[3][d];
}

// Scan: workgroup d turns row d into the tiles' first output positions for
// digit d.
__global__ void __launch_bounds__(256) k_bin_rowscan(BinPass P) {
    __shared__ uint32_t sh[17];
    const uint32_t d = blockIdx.x;
    const uint32_t ntiles = (bs_elements(P) + P.tile_elems - 1) / P.tile_elems;
    // four consecutive tiles per thread, all loaded before one workgroup
    // scan (1024 tiles per round); the digit's total goes to dtotal
    SR_GLOBAL uint32_t *row = P.hist + (size_t)d * P.hist_cols;
    uint32_t base = 0u;
    for (uint32_t t0 = 0; t0 < ntiles; t0 += 1024u) {
        const uint32_t t = t0 + 4u * threadIdx.x;
        uint32_t v[4], sum = 0u;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = t + k < ntiles ? row[t + k] : 0u;
            sum += v[k];
        }
        uint32_t ex;
        uint32_t tot = block_excl_scan(sum, ex, sh);
        uint32_t run = base + ex;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (t + k < ntiles) row[t + k] = run;
            run += v[k];
        }
        base += tot;
    }
    if (threadIdx.x == 0) P.dtotal[d] = base;
}

// Scatter: each tile's elements to their positions, in input order within
// each digit.  The tile is first reordered in LDS by digit, so the global
// writes go out as one contiguous run per digit.  Workgroups are dispatched
// to the 8 XCDs round-robin; XCD x takes the consecutive tiles
// [x * per, (x + 1) * per), so the runs that consecutive tiles write into
// one digit's range meet in that XCD's L2.
template <int IT>
__global__ void __launch_bounds__(256) k_bin_scatter(BinPass P) {
    constexpr uint32_t TILE = 256u * IT;
    __shared__ uint32_t cnt[4][256];
    __shared__ uint32_t gofs[256];
    __shared__ uint32_t sh[17];
    __shared__ uint32_t skey[TILE], sval[TILE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t n = bs_elements(P);
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    const uint32_t per = (ntiles + 7u) / 8u, k = blockIdx.x >> 3;
    const uint32_t t = (blockIdx.x & 7u) * per + k;
    if (k >= per || t >= ntiles) return;
    uint32_t key[IT];
    bs_load_count<IT>(P, t, n, key, cnt);
    const uint32_t wbase = t * TILE + (uint32_t)wave * (64u * IT);
    uint32_t val[IT];
#pragma unroll
    for (int s = 0; s < IT; s++) {
        uint32_t i = wbase + (uint32_t)(s * 64 + lane);
        val[s] = i < n ? P.in_vals[i] : 0u;
    }
    __syncthreads();
    {
        // digit d's run in the tile starts at local index lstart; wave w's
        // elements of digit d follow those of the waves before it
        const uint32_t d = threadIdx.x;
        uint32_t c0 = cnt[0][d], c1 = cnt[1][d], c2 = cnt[2][d], c3 = cnt[3][d];
        uint32_t lstart, dbase;
        block_excl_scan(c0 + c1 + c2 + c3, lstart, sh);
        block_excl_scan(P.dtotal[d], dbase, sh);     // the digit's first position
        gofs[d] = dbase + P.hist[(size_t)d * P.hist_cols + t] - lstart;
        cnt[0][d] = lstart;
        cnt[1][d] = lstart + c0;
        cnt[2][d] = lstart + c0 + c1;
        cnt[3][d] = lstart + c0 + c1 + c2;
    }
    __syncthreads();
    // ranks: for each row s of the wave's elements, the group of lanes with
    // one digit takes its slots with one LDS fetch-add by its last lane;
    // the wave's fetch-adds run in issue order, so rows keep their order
    // and the adds need no wait between rows
    const uint32_t mask = (1u << P.bits) - 1u;
    uint32_t got[IT], rk[IT];
    int ld[IT];
#pragma unroll
    for (int s = 0; s < IT; s++) {
        bool valid = wbase + (uint32_t)(s * 64 + lane) < n;
        uint32_t d = (key[s] >> P.shift) & mask;
        uint64_t m = bs_match(valid, d, P.bits);
        rk[s] = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        ld[s] = valid ? 63 - __clzll(m) : lane;
        got[s] = 0u;
        if (valid && ld[s] == lane) got[s] = atomicAdd(&cnt[wave][d], (uint32_t)__popcll(m));
    }
#pragma unroll
    for (int s = 0; s < IT; s++) {
        uint32_t li = (uint32_t)__shfl((int)got[s], ld[s], 64) + rk[s];
        if (wbase + (uint32_t)(s * 64 + lane) < n) {
            skey[li] = key[s];
            sval[li] = val[s];
        }
    }
    __syncthreads();
    const uint32_t m_tile = min(n - t * TILE, TILE);
    for (uint32_t i = threadIdx.x; i < m_tile; i += 256u) {
        uint32_t kv = skey[i];
        uint32_t pos = gofs[(kv >> P.shift) & mask] + i;
        if (pos < P.pair_cap) {      // else the host sees total_pairs > pair_cap and runs the frame again
            if (P.last) P.list[pos] = sval[i];
            else P.out_vals[pos] = sval[i];
            P.out_keys[pos] = kv;
        }
    }
}

// Each tile's list bounds from the sorted bins: bin b's entries are
// [bin_start[b], bin_end[b]); both arrays are zeroed at the frame's start,
// so a bin without entries stays empty.  (The arguments come in a struct:
// a kernel parameter typed with an address space is mangled into the
// device-side name but not the host's, and the launch would find no
// kernel.)
struct BinBounds {
    const SR_GLOBAL uint32_t *key;      // the last pass's bins, in list order
    const SR_GLOBAL DevStatus *status;
    uint32_t pair_cap, items;
    SR_GLOBAL uint32_t *bin_start, *bin_end;
};

__global__ void __launch_bounds__(256) k_bin_bounds(BinBounds B) {
    const uint32_t n = min(B.status->total_pairs, B.pair_cap);
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t b = B.key[i];
        if (b >= B.items) continue;    // (only in a frame that overflowed and runs again)
        if (i == 0u || B.key[i - 1u] != b) B.bin_start[b] = i;
        if (i + 1u == n || B.key[i + 1u] != b) B.bin_end[b] = i + 1u;
    }
}

// ---------------------------------------------------------------------------
// Per-(triangle, tile) raster setup (src/tilerasterizer.rs:114-218)
// ---------------------------------------------------------------------------
struct TriSetup {
    float A[3], B[3], C[3], SX[3], SY[3];   // edges: 0 = v1->v2 (w0), 1 = v2->v0 (w1), 2 = v0->v1 (w2)
    float ooa, oa, oda, odb, za, zda, zdb;
    int cmnx, cmny, cmxx, cmxy, ox0, oy0;
    bool coarse;
};

__device__ __forceinline__ void tri_setup(const uint4 g, const float4 *R, int tx0, int ty0, TriSetup &S) {
    float4 r0 = R[0], r1 = R[1], r2 = R[2], r3 = R[3];
    int32_t p0x = __float_as_int(r0.x), p0y = __float_as_int(r0.y), p1x = __float_as_int(r0.z), p1y = __float_as_int(r0.w);
    int32_t p2x = __float_as_int(r1.x), p2y = __float_as_int(r1.y);
    S.ooa = r1.z; S.oa = r1.w; S.oda = r2.x; S.odb = r2.y; S.za = r2.z; S.zda = r2.w; S.zdb = r3.x;
    edge_setup(p0x, p0y, p1x, p1y, p2x, p2y, S.A, S.B, S.C, S.SX, S.SY);
    int mnx = (int)(g.x & 0xFFFFu), mny = (int)(g.x >> 16), mxx = (int)(g.y & 0xFFFFu), mxy = (int)(g.y >> 16);
    int rx = tx0 & ~(SR_REF_TILE - 1), ry = ty0 & ~(SR_REF_TILE - 1);
    S.cmnx = max(mnx, rx); S.cmny = max(mny, ry);
    S.cmxx = min(mxx, rx + SR_REF_TILE); S.cmxy = min(mxy, ry + SR_REF_TILE);
    S.coarse = (S.cmxx - S.cmnx > 16) || (S.cmxy - S.cmny > 16);
    S.ox0 = S.cmnx & ~1;
    S.oy0 = S.cmny & ~1;
}

// Origin of the fine-raster walk that covers quad (qpx, qpy) and whether the
// coarse 16px block containing it survives the corner test
// (src/tilerasterizer.rs:220-291).  Returns false if the quad is not walked.
__device__ __forceinline__ bool quad_origin(const TriSetup &S, int qpx, int qpy, int &bx, int &by) {
    if (!(qpx >= S.ox0 && qpx < S.cmxx && qpy >= S.oy0 && qpy < S.cmxy)) return false;
    bx = S.ox0;
    by = S.oy0;
    if (S.coarse) {
        bx = S.ox0 + ((qpx - S.ox0) & ~15);
        by = S.oy0 + ((qpy - S.oy0) & ~15);
    }
    return true;
}

__device__ __forceinline__ bool block_rejected(const TriSetup &S, int bx, int by) {
    float cx0 = (float)(bx * 16 + 8), cx1 = (float)(bx * 16 + 256 - 8);
    float cy0 = (float)(by * 16 + 8), cy1 = (float)(by * 16 + 256 - 8);
    bool out = false;
    for (int e = 0; e < 3; e++) {
        float w00 = (S.A[e] * cx0 + S.B[e] * cy0) + S.C[e];
        float w10 = (S.A[e] * cx1 + S.B[e] * cy0) + S.C[e];
        float w01 = (S.A[e] * cx0 + S.B[e] * cy1) + S.C[e];
        float w11 = (S.A[e] * cx1 + S.B[e] * cy1) + S.C[e];
        float m = maxps(maxps(w00, w10), maxps(w01, w11));
        out = out || (m < 0.0f);
    }
    return out;
}

// Edge value of edge e at lane (sx, sy) of the quad i steps right and j steps
// down from the walk origin (bx, by): origin evaluation, then j serial
// +step_y, then i serial +step_x (src/tilerasterizer.rs:307-381).
__device__ __forceinline__ float chain_serial(float v, float sy, float sx, int j, int i) {
    for (int k = 0; k < j; k++) v = v + sy;
    for (int k = 0; k < i; k++) v = v + sx;
    return v;
}

// Trailing zero bits of an integer-valued float (64 for zero).
__device__ __forceinline__ int ctz_intf(float x) {
    uint32_t b = __float_as_uint(x) & 0x7FFFFFFFu;
    if (b == 0u) return 64;
    return (int)(b >> 23) - 150 + __builtin_ctz((b & 0x7FFFFFu) | 0x800000u);
}

// The serial chain v + sy (j times) + sx (i times) in O(1).  v, sy and sx are
// integers.  If 2^t divides all three and every partial sum stays below
// 2^(24+t) in magnitude, every partial sum is representable, each addition is
// exact, and the chain equals v + j*sy + i*sx -- which two fused multiply-adds
// return exactly (one rounding of a representable value).  The partial sums
// lie on the segments v -> v+j*sy -> end, so bounding those three points
// bounds them all; rounding is monotone, so testing the rounded points is
// equivalent.  Clears ok when the shortcut does not apply.
__device__ __forceinline__ float chain_fast(float v, float sy, float sx, float fj, float fi, bool &ok) {
    float r1 = fmaf(fj, sy, v);
    float r2 = fmaf(fi, sx, r1);
    float m = fmaxf(fmaxf(fabsf(v), fabsf(r1)), fabsf(r2));
    if (!(m < 16777216.0f)) {
        int t = min(min(ctz_intf(v), min(ctz_intf(sy), ctz_intf(sx))), 100);
        ok = ok && (m < __uint_as_float((uint32_t)(151 + t) << 23));
    }
    return r2;
}

__device__ __forceinline__ float edge_chain_v(float A, float B, float C, float SX, float SY, int bx, int by, int sx,
                                             int sy, int i, int j) {
    float X = (float)(bx * 16 + 8 + 16 * sx);
    float Y = (float)(by * 16 + 8 + 16 * sy);
    float v = (A * X + B * Y) + C;
    bool ok = true;
    float r = chain_fast(v, SY, SX, (float)j, (float)i, ok);
    return ok ? r : chain_serial(v, SY, SX, j, i);
}

__device__ __forceinline__ float edge_chain(const TriSetup &S, int e, int bx, int by, int sx, int sy, int i, int j) {
    return edge_chain_v(S.A[e], S.B[e], S.C[e], S.SX[e], S.SY[e], bx, by, sx, sy, i, j);
}

// ---------------------------------------------------------------------------
// Texture sampling (src/texture.rs:578-863)
// ---------------------------------------------------------------------------
// x / d; for d a power of two (texture sizes usually are) the product with the
// exact reciprocal 2^-k, which rounds the same exact quotient.
__device__ __forceinline__ float div_dim(float x, float d) {
    uint32_t b = __float_as_uint(d);
    uint32_t e = (b >> 23) & 0xFFu;
    if ((b & 0x807FFFFFu) == 0u && e >= 1u && e <= 253u) return x * __uint_as_float((254u - e) << 23);
    return x / d;
}

__device__ __forceinline__ float apply_wrap(float texel, float dim, int mode) {
    float t2;
    if (mode == SR_WRAP_CLAMP_TO_EDGE) {
        t2 = texel;
    } else if (mode == SR_WRAP_REPEAT) {
        t2 = texel - floorf(div_dim(texel, dim)) * dim;
    } else {
        float two = dim * 2.0f;
        float t = texel - floorf(div_dim(texel, two)) * two;
        t2 = minps(t, two - t);
    }
    return minps(t2, dim - 1.0f);
}

__device__ __forceinline__ uint32_t compute_mip(const DevTexture &T, const float dd[4]) {
    float d0 = dd[0] * T.whv[0], d1 = dd[1] * T.whv[1], d2 = dd[2] * T.whv[2], d3 = dd[3] * T.whv[3];
    float dx2 = d0 * d0 + d2 * d2;
    float dy2 = d1 * d1 + d3 * d3;
    float fp = (dx2 + dy2) * 0.5f;
    fp = fmaxf(fp, 1.0f);   // f32::max: NaN -> 1.0
    uint32_t u = sat_u32(fp);
    uint32_t mip = (31u - (uint32_t)__builtin_clz(u)) >> 1;
    return mip < T.max_mip ? mip : T.max_mip;
}

__device__ __forceinline__ uint32_t texel_nearest(const DevScene &S, const DevTexRef &ref, float u, float v,
                                                  const float dd[4]) {
    const DevTexture &T = S.textures[ref.tex];
    uint32_t mip = compute_mip(T, dd);
    float wf = (float)T.mip_w[mip], hf = (float)T.mip_h[mip];
    uint32_t x = sat_u32(apply_wrap(floorf(u * wf), wf, ref.wrap_s));
    uint32_t y = sat_u32(apply_wrap(floorf(v * hf), hf, ref.wrap_t));
    // levels are stored in 4x4-texel blocks, row-major over the level
    // padded to multiples of 4
    uint32_t bw = (T.mip_w[mip] + 3u) >> 2;
    uint32_t idx = T.mip_off[mip] + ((((y >> 2) * bw) + (x >> 2)) << 4) + ((y & 3u) << 2) + (x & 3u);
    return S.texels[T.base + idx];
}

__device__ __forceinline__ void cube_uv(f3 n, float &uf, float &vf, uint32_t &slice) {
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    bool mx = (ax >= ay) && (ax >= az);
    bool my = (ay > ax) && (ay >= az);
    float sx = n.x >= 0.0f ? 1.0f : -1.0f;
    float sy = n.y >= 0.0f ? 1.0f : -1.0f;
    float sz = n.z >= 0.0f ? 1.0f : -1.0f;
    float u_x = -n.z * sx, v_x = -n.y, d_x = ax;
    float u_y = n.x, v_y = n.z * sy, d_y = ay;
    float u_z = n.x * sz, v_z = -n.y, d_z = az;
    float u = mx ? u_x : (my ? u_y : u_z);
    float v = mx ? v_x : (my ? v_y : v_z);
    float den = mx ? d_x : (my ? d_y : d_z);
    den = maxps(den, 1.0e-19f);
    uf = 0.5f * (u / den + 1.0f);
    vf = 0.5f * (v / den + 1.0f);
    float ix = n.x >= 0.0f ? 0.0f : 1.0f;
    float iy = n.y >= 0.0f ? 2.0f : 3.0f;
    float iz = n.z >= 0.0f ? 4.0f : 5.0f;
    slice = sat_u32(mx ? ix : (my ? iy : iz));
}

// sample_cubemap_rgb (nearest, full-resolution faces at every mip)
__device__ __forceinline__ f3 cube_nearest(const uint32_t *data, uint32_t size, uint32_t mip, f3 n) {
    float u, v;
    uint32_t slice;
    cube_uv(n, u, v, slice);
    float wf = (float)size;
    float x = roundf(u * (wf - 1.0f)), y = roundf(v * (wf - 1.0f));
    uint32_t xi = sat_u32(clampv(x, 0.0f, wf - 1.0f));
    uint32_t yi = sat_u32(clampv(y, 0.0f, wf - 1.0f));
    uint32_t off = mip * 6u * size * size + slice * (size * size);
    return unpack_rgb(data[off + yi * size + xi]);
}

// ---------------------------------------------------------------------------
// Voxel grid lookup (src/voxelgrid.rs:264-368)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tri_const(float k, float fx, float fy, float fz) {
    float a = k * (1.0f - fx) + k * fx;
    float b = a * (1.0f - fy) + a * fy;
    return b * (1.0f - fz) + b * fz;
}

__device__ __forceinline__ void voxel_lookup(const DevScene &S, f3 p, f3 rgb[4], float &w0, float &w1) {
    float vx = (p.x - S.vox_min[0]) / S.vox_size[0];
    float vy = (p.y - S.vox_min[1]) / S.vox_size[1];
    float vz = (p.z - S.vox_min[2]) / S.vox_size[2];
    float x0f = floorf(vx), y0f = floorf(vy), z0f = floorf(vz);
    uint32_t W = (uint32_t)S.vox_dims[0], H = (uint32_t)S.vox_dims[1], D = (uint32_t)S.vox_dims[2];
    uint32_t x0 = sat_index(x0f, W - 1), y0 = sat_index(y0f, H - 1), z0 = sat_index(z0f, D - 1);
    uint32_t x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, H - 1), z1 = min(z0 + 1, D - 1);
    float fx = vx - x0f, fy = vy - y0f, fz = vz - z0f;
    float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
    size_t sy = W, sz = (size_t)W * H;
    auto idx = [&](uint32_t x, uint32_t y, uint32_t z) { return (size_t)z * sz + (size_t)y * sy + x; };
    if (S.vox_sh4) {
        const float4 *g = reinterpret_cast<const float4 *>(S.vox_sh4);
        for (int c = 0; c < 4; c++) {
            float4 v000 = g[4 * idx(x0, y0, z0) + c], v001 = g[4 * idx(x0, y0, z1) + c];
            float4 v010 = g[4 * idx(x0, y1, z0) + c], v011 = g[4 * idx(x0, y1, z1) + c];
            float4 v100 = g[4 * idx(x1, y0, z0) + c], v101 = g[4 * idx(x1, y0, z1) + c];
            float4 v110 = g[4 * idx(x1, y1, z0) + c], v111 = g[4 * idx(x1, y1, z1) + c];
            float o[4];
            const float *a000 = &v000.x, *a001 = &v001.x, *a010 = &v010.x, *a011 = &v011.x;
            const float *a100 = &v100.x, *a101 = &v101.x, *a110 = &v110.x, *a111 = &v111.x;
            for (int ch = 0; ch < 4; ch++) {
                float v00 = a000[ch] * ox + a100[ch] * fx;
                float v01 = a001[ch] * ox + a101[ch] * fx;
                float v10 = a010[ch] * ox + a110[ch] * fx;
                float v11 = a011[ch] * ox + a111[ch] * fx;
                float v0 = v00 * oy + v10 * fy;
                float v1 = v01 * oy + v11 * fy;
                o[ch] = v0 * oz + v1 * fz;
            }
            rgb[c] = mk3(o[0], o[1], o[2]);
            if (c == 0) w0 = o[3];
            if (c == 1) w1 = o[3];
        }
        return;
    }
    for (int c = 0; c < 4; c++)
        rgb[c] = mk3(tri_const(S.sh_const[c][0], fx, fy, fz), tri_const(S.sh_const[c][1], fx, fy, fz),
                     tri_const(S.sh_const[c][2], fx, fy, fz));
    const float4 *L = reinterpret_cast<const float4 *>(S.vox_light) + 2 * idx(x0, y0, z0);
    float4 lo = L[0], hi = L[1];
    float v000 = lo.x, v001 = lo.y, v010 = lo.z, v011 = lo.w;
    float v100 = hi.x, v101 = hi.y, v110 = hi.z, v111 = hi.w;
    float v00 = v000 * ox + v100 * fx;
    float v01 = v001 * ox + v101 * fx;
    float v10 = v010 * ox + v110 * fx;
    float v11 = v011 * ox + v111 * fx;
    float v0 = v00 * oy + v10 * fy;
    float v1 = v01 * oy + v11 * fy;
    w0 = v0 * oz + v1 * fz;
    w1 = tri_const(S.sh_const[1][3], fx, fy, fz);
}

// ---------------------------------------------------------------------------
// PBR shading of one pixel (pbr_shader, src/shader.rs:110-309)
// b1[4], b2[4]: the quad's four barycentric pairs; lane: this pixel's lane.
// TR: the translucent form, which blends over cur (src/shader.rs:66-76).
// ---------------------------------------------------------------------------
// The quad's four w = 1 / interpolated 1/w of this pixel's triangle at each
// lane's barycentrics (src/shader.rs:123-130: du_dv takes all four lanes'
// w), from the quad's barycentric pairs b1[4], b2[4].
__device__ __forceinline__ void quad_w(float oa, float oda, float odb, const float b1[4], const float b2[4],
                                       float wv[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++) wv[j] = 1.0f / ((oa + b1[j] * oda) + b2[j] * odb);
}

// B1, B2: this lane's barycentrics; wv: the quad's four w (quad_w).
template <bool TR>
__device__ __forceinline__ f3 shade_pixel(const DevScene &S, const DevView &V, int material, const float4 *Sh,
                                          const float4 *Tn, bool nmap, float B1, float B2, const float wv[4], int lane,
                                          f3 cur) {
    const DevMaterial &M = S.materials[material];
    float4 s0 = Sh[0], s1 = Sh[1], s2 = Sh[2], s3 = Sh[3], s4 = Sh[4], s5 = Sh[5], s6 = Sh[6];
    float w = wv[lane];
    f3 na = mk3(s0.x, s0.y, s0.z), nda = mk3(s0.w, s1.x, s1.y), ndb = mk3(s1.z, s1.w, s2.x);
    f3 pa = mk3(s2.y, s2.z, s2.w), pda = mk3(s3.x, s3.y, s3.z), pdb = mk3(s3.w, s4.x, s4.y);
    f3 in_n = mk3((na.x + B1 * nda.x) + B2 * ndb.x, (na.y + B1 * nda.y) + B2 * ndb.y, (na.z + B1 * nda.z) + B2 * ndb.z);
    in_n = normalize3(in_n);
    f3 pos = mk3(((pa.x + B1 * pda.x) + B2 * pdb.x) * w, ((pa.y + B1 * pda.y) + B2 * pdb.y) * w,
                 ((pa.z + B1 * pda.z) + B2 * pdb.z) * w);
    float uvx = ((s4.z + B1 * s4.w) + B2 * s5.x) * w;
    float uvy = ((s5.y + B1 * s5.z) + B2 * s5.w) * w;
    float dd[4] = {s6.x * wv[0], s6.y * wv[1], s6.z * wv[2], s6.w * wv[3]};
    f3 nw = normalize3(in_n);
    if ((M.flags & SR_MAT_NORMAL_MAP) && nmap) {
        float4 t0 = Tn[0], t1 = Tn[1], t2 = Tn[2];
        f3 ta = mk3(t0.x, t0.y, t0.z), tda = mk3(t0.w, t1.x, t1.y), tdb = mk3(t1.z, t1.w, t2.x);
        f3 in_t = mk3((ta.x + B1 * tda.x) + B2 * tdb.x, (ta.y + B1 * tda.y) + B2 * tdb.y, (ta.z + B1 * tda.z) + B2 * tdb.z);
        in_t = normalize3(in_t);
        float tsign = (t2.y + B1 * t2.z) + B2 * t2.w;
        f3 tw = normalize3(sub3(in_t, muls(in_n, dot3(in_n, in_t))));
        float hand = tsign >= 0.0f ? 1.0f : -1.0f;
        f3 bw = muls(cross3(in_n, tw), hand);
        f3 tsn = unpack_rgb(texel_nearest(S, M.normal_tex, uvx, uvy, dd));
        tsn = mk3(tsn.x * 2.0f - 1.0f, tsn.y * 2.0f - 1.0f, tsn.z * 2.0f - 1.0f);
        nw = add3(add3(muls(tw, tsn.x), muls(bw, tsn.y)), muls(in_n, tsn.z));
        nw = normalize3(nw);
    }
    f3 L = mk3(S.light_dir[0], S.light_dir[1], S.light_dir[2]);
    f3 LC = mk3(S.light_color[0], S.light_color[1], S.light_color[2]);
    f3 vdir = sub3(mk3(V.cam_pos[0], V.cam_pos[1], V.cam_pos[2]), pos);
    f3 vn = normalize3(vdir);
    float ndl = maxps(dot3(nw, L), 0.0f);
    f3 hv = normalize3(add3(L, vn));
    float ndh = maxps(dot3(nw, hv), 0.0f);
    float ndv = maxps(dot3(nw, vn), 1.0e-4f);
    float vdh = maxps(dot3(vn, hv), 0.0f);
    f3 irr;
    float gw0, gw1;
    {
        // the irradiance right away: four SH coefficients need not stay live
        f3 gi[4];
        voxel_lookup(S, pos, gi, gw0, gw1);
        irr.x = ((gi[0].x + gi[1].x * nw.y) + gi[2].x * nw.z) + gi[3].x * nw.x;
        irr.y = ((gi[0].y + gi[1].y * nw.y) + gi[2].y * nw.z) + gi[3].y * nw.x;
        irr.z = ((gi[0].z + gi[1].z * nw.y) + gi[2].z * nw.z) + gi[3].z * nw.x;
        irr = max3s(irr, 0.0f);
    }
    float vli = clampv(gw0, 0.0f, 1.0f);
    float skyv = clampv(gw1, 0.0f, 1.0f);
    f3 base = mk3(M.base[0], M.base[1], M.base[2]);
    if (M.base_tex.tex >= 0) base = mul3(base, srgb_fast(unpack_rgb(texel_nearest(S, M.base_tex, uvx, uvy, dd))));
    float rough = M.roughness, metal = M.metallic;
    if (M.mr_tex.tex >= 0) {
        f3 mr = unpack_rgb(texel_nearest(S, M.mr_tex, uvx, uvy, dd));
        rough = rough * mr.y;
        metal = metal * mr.z;
    }
    rough = clampv(rough, 0.045f, 1.0f);
    metal = clampv(metal, 0.0f, 1.0f);
    float ao = 1.0f;
    if (M.occlusion_tex.tex >= 0) {
        ao = unpack_rgb(texel_nearest(S, M.occlusion_tex, uvx, uvy, dd)).x;
        ao = 1.0f + (ao - 1.0f) * M.occlusion_strength;
    }
    f3 f0 = mk3(0.04f + (base.x - 0.04f) * metal, 0.04f + (base.y - 0.04f) * metal, 0.04f + (base.z - 0.04f) * metal);
    float omvh = 1.0f - vdh;
    float omvh2 = omvh * omvh;
    float omvh4 = omvh2 * omvh2;
    float omvh5 = omvh4 * omvh;
    f3 bf = mk3(f0.x + (1.0f - f0.x) * omvh5, f0.y + (1.0f - f0.y) * omvh5, f0.z + (1.0f - f0.z) * omvh5);
    const float PI = 3.14159265358979323846f;
    const float EPS = 1e-6f;
    float alpha = rough * rough;
    float alpha2 = alpha * alpha;
    float ndh2 = ndh * ndh;
    float dden = ndh2 * (alpha2 - 1.0f) + 1.0f;
    float brdf_d = alpha2 / (PI * (dden * dden) + EPS);
    float k = rough + 1.0f;
    k = (k * k) * 0.125f;
    float gv = ndv / ((ndv * (1.0f - k) + k) + EPS);
    float gl = ndl / ((ndl * (1.0f - k) + k) + EPS);
    float brdf_g = gv * gl;
    float sdg = (brdf_d * brdf_g) / ((4.0f * ndl) * ndv + EPS);
    float om = 1.0f - metal;
    f3 kdd = mk3((1.0f - bf.x) * om, (1.0f - bf.y) * om, (1.0f - bf.z) * om);
    const float INV_PI = 1.0f / 3.14159265358979323846f;
    f3 lam = muls(base, INV_PI);
    f3 cdd = muls(muls(mul3(mul3(LC, kdd), lam), ndl), vli);
    f3 cds = muls(muls(muls(mul3(LC, bf), sdg), ndl), vli);
    f3 kdi = mk3((1.0f - f0.x) * om, (1.0f - f0.y) * om, (1.0f - f0.z) * om);
    f3 cid = muls(mul3(mul3(irr, base), kdi), INV_PI);
    f3 iv = muls(vn, -1.0f);
    f3 refl = sub3(iv, muls(muls(nw, dot3(iv, nw)), 2.0f));
    float smip = rough * S.spec_max_mip_f;
    // sample_cubemap_trilinear_rgb
    float mip = clampv(smip, 0.0f, S.spec_max_mip_f);
    float m0 = floorf(mip);
    float m1 = minps(m0 + 1.0f, S.spec_max_mip_f);
    float tt = mip - m0;
    uint32_t mi0 = min(sat_u32(m0), S.spec_mips - 1), mi1 = min(sat_u32(m1), S.spec_mips - 1);
    f3 c0 = cube_nearest(S.specular, S.spec_size, mi0, refl);
    f3 c1 = cube_nearest(S.specular, S.spec_size, mi1, refl);
    f3 env = mk3(c0.x + (c1.x - c0.x) * tt, c0.y + (c1.y - c0.y) * tt, c0.z + (c1.z - c0.z) * tt);
    // sample_bilinear_rgb on the BRDF LUT, mip 0, clamp-to-edge
    float lu = clampv(ndv, 0.0f, 1.0f), lv = clampv(rough, 0.0f, 1.0f);
    float lw = (float)S.lut_size;
    float xf = lu * lw - 0.5f, yf = lv * lw - 0.5f;
    float x0 = floorf(xf), y0 = floorf(yf);
    float x1 = x0 + 1.0f, y1 = y0 + 1.0f;
    float fx = xf - x0, fy = yf - y0;
    float omfx = 1.0f - fx, omfy = 1.0f - fy;
    uint32_t x0i = sat_u32(minps(x0, lw - 1.0f)), y0i = sat_u32(minps(y0, lw - 1.0f));
    uint32_t x1i = sat_u32(minps(x1, lw - 1.0f)), y1i = sat_u32(minps(y1, lw - 1.0f));
    f3 p00 = unpack_rgb(S.lut[y0i * S.lut_size + x0i]);
    f3 p10 = unpack_rgb(S.lut[y0i * S.lut_size + x1i]);
    f3 p01 = unpack_rgb(S.lut[y1i * S.lut_size + x0i]);
    f3 p11 = unpack_rgb(S.lut[y1i * S.lut_size + x1i]);
    float w00 = omfx * omfy, w10 = fx * omfy, w01 = omfx * fy, w11 = fx * fy;
    float lutx = ((p00.x * w00 + p10.x * w10) + p01.x * w01) + p11.x * w11;
    float luty = ((p00.y * w00 + p10.y * w10) + p01.y * w01) + p11.y * w11;
    f3 bsf = adds(muls(f0, lutx), luty);
    f3 cis = mul3(env, bsf);
    float ao_spec = 1.0f + (ao - 1.0f) * 0.5f;
    cis = muls(cis, ao_spec * skyv);
    f3 color;
    if (TR) {
        // transmission factor times the transmission texture's red channel
        float tr = M.transmission;
        if (M.transmission_tex.tex >= 0) tr = tr * unpack_rgb(texel_nearest(S, M.transmission_tex, uvx, uvy, dd)).x;
        tr = clampv(tr, 0.0f, 1.0f);
        float inv_tr = 1.0f - tr;
        color = add3(add3(muls(mul3(cur, base), tr), muls(add3(cdd, cid), inv_tr)), cds);
        color = add3(color, cis);
    } else {
        color = add3(add3(add3(cdd, cds), cid), cis);
    }
    f3 em = mk3(1.0f, 1.0f, 1.0f);
    if (M.emissive_tex.tex >= 0) em = srgb_fast(unpack_rgb(texel_nearest(S, M.emissive_tex, uvx, uvy, dd)));
    color = add3(color, mul3(em, mk3(M.emissive[0], M.emissive[1], M.emissive[2])));
    return color;
}

// compute_skybox (src/tilerasterizer.rs:478-508) for one lane of quad (qx, qy)
__device__ f3 skybox_pixel(const DevScene &S, const DevView &V, int qx, int qy, int sx, int sy) {
    float pxf = (float)qx + (sx ? 1.5f : 0.5f);
    float pyf = (float)qy + (sy ? 1.5f : 0.5f);
    float ndx = (pxf * V.one_over_w) * 2.0f - 1.0f;
    float ndy = (1.0f - pyf * V.one_over_h) * 2.0f - 1.0f;
    const float *m = V.sky;
    f3 d;
    d.x = ((ndx * m[0] + ndy * m[1]) + 1.0f * m[2]) + m[3];
    d.y = ((ndx * m[4] + ndy * m[5]) + 1.0f * m[6]) + m[7];
    d.z = ((ndx * m[8] + ndy * m[9]) + 1.0f * m[10]) + m[11];
    f3 n = normalize3(d);
    return srgb_fast(cube_nearest(S.cubemap, S.cube_size, 0, n));
}

// ---------------------------------------------------------------------------
// k_raster_shade: one workgroup per (view, tile)
// ---------------------------------------------------------------------------
struct RasterArgs {
    const SR_GLOBAL DevView *views;
    int32_t num_views;
    uint32_t total_items;
    const SR_GLOBAL uint32_t *bin_start;
    const SR_GLOBAL uint32_t *bin_end;
    const SR_GLOBAL uint32_t *list;
    uint32_t pair_cap, rec_cap;
    const SR_GLOBAL uint4 *geo;
    const SR_GLOBAL float4 *raster;
    TriSource src;
    const SR_GLOBAL uint32_t *tri_base;
    SR_GLOBAL uint32_t *out_ids;
    SR_GLOBAL float *out_depth;
    SR_GLOBAL float *out_color;
    uint64_t color_plane;
    SR_GLOBAL uint32_t *out_rgba;
    SR_GLOBAL float *out_lum;
    SR_GLOBAL uint32_t *center_ids;         // 4 winners per 64px tile: its centre quad, on screen or not
    SR_GLOBAL float4 *center_color;         // 4 (colour, depth) per 64px tile for the translucent pass (or null)
    SR_GLOBAL DevStatus *status;
    DevStatus *status_out;                  // pinned host memory for the frame's status
    StageClock *clock;                      // stage stamps (or null)
    int32_t persistent;
};

__device__ __forceinline__ uint32_t depth_key(float z) {
    if (z == 0.0f) z = 0.0f;  // -0 == +0
    uint32_t b = __float_as_uint(z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ bool alpha_pass(const DevScene &S, const TriSource &T, const float4 *R, float b1, float b2,
                                           float w) {
    int material = __float_as_int(R[3].y);
    const DevMaterial &M = S.materials[material];
    CV tv[3];
    float4 Sh[7], Tn[3];
    load_tri_cv(T, R, tv);
    tri_shade_attrs(tv, R, false, Sh, Tn);
    float4 s4 = Sh[4], s5 = Sh[5], s6 = Sh[6];
    float u = ((s4.z + b1 * s4.w) + b2 * s5.x) * w;
    float v = ((s5.y + b1 * s5.z) + b2 * s5.w) * w;
    float dd[4] = {s6.x, s6.y, s6.z, s6.w};
    float a = unpack_a(texel_nearest(S, M.base_tex, u, v, dd));
    return a >= M.alpha_cutoff;
}

// Per-wave staging of one batch of 64 bin entries (structure of arrays, one
// 64-wide row per field) from which the wave's lanes draw quad tasks.
enum {
    F_A = 0, F_B = 3, F_C = 6,
    F_OOA = 9, F_OA, F_ODA, F_ODB, F_ZA, F_ZDA, F_ZDB,
    F_ID, F_ORG, F_Q0, F_FLAGS, F_INCL, F_ROWS
};
// The walk steps are (float)(32 a) and (float)(32 b) in i32 arithmetic; they
// equal 32 A and 32 B exactly unless 32 a wraps (|a| >= 2^26), which the
// entry's STEP_WRAP flag marks and walk_quad then recomputes.
#define STEP_WRAP 4u
// After F_ROWS: inclusive quad counts per quad row (u16 pairs), then each
// row's first quad column (4-bit nibbles).
template <int T> struct Stage {
    static constexpr int NR = T / 2;
    static constexpr int CUM = F_ROWS;
    static constexpr int LO = F_ROWS + NR / 2;
    static constexpr int COUNT = LO + (NR + 7) / 8;
};

// f32 corner test of the 16px block at (bx, by) (src/tilerasterizer.rs:258-291)
__device__ __forceinline__ bool block_out(const TriSetup &ts, int bx, int by) {
    float cx0 = (float)(bx * 16 + 8), cx1 = (float)(bx * 16 + 256 - 8);
    float cy0 = (float)(by * 16 + 8), cy1 = (float)(by * 16 + 256 - 8);
    bool out = false;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        float w00 = (ts.A[e] * cx0 + ts.B[e] * cy0) + ts.C[e];
        float w10 = (ts.A[e] * cx1 + ts.B[e] * cy0) + ts.C[e];
        float w01 = (ts.A[e] * cx0 + ts.B[e] * cy1) + ts.C[e];
        float w11 = (ts.A[e] * cx1 + ts.B[e] * cy1) + ts.C[e];
        out = out || (maxps(maxps(w00, w10), maxps(w01, w11)) < 0.0f);
    }
    return out;
}

// Stages one bin entry and returns how many of its quads inside the tile
// can hold a covered sample.
//
// Walked quads are the reference's: quad-aligned from (ox0, oy0) up to the
// clipped bounding box, in 16px blocks whose corner test rejects them.  Rows
// are then narrowed to conservative column spans.  Let E = A X + B Y + C be
// the exact edge function of the f32 coefficients.  The reference's value
// at a sample is E after at most 20 roundings (products and sums of the
// origin evaluation, at most 16 chain additions), each off by at most 2^-24
// of a magnitude bounded by P = |A| Xm + |B| Ym + |C| + 512 (|A| + |B|), so
// it differs from E by less than M = 2^-19 P; a sample with E < -M is never
// covered.  The spans work in f32 relative to the tile's origin (Xt, Yt):
// E = A X' + B Y' + C' with C' = C + A Xt + B Yt, whose three roundings are
// off by at most 2^-22 P, so E < -M2 with M2 = 2^-18 P (P as computed, at
// most 5 roundings low) excludes the sample.  Per row, each edge with A != 0
// bounds X' on one side, -(M2 + C') / A - (B / A) Y', taken at the looser of
// the row's two sample lines; its operands and products are each off by a
// few 2^-24 (the reciprocal by one ulp), in all less than 2^-21 (|base| +
// |slope| Y'max), and the slack is twice that plus 16 (one pixel) for the
// rounding of the column bounds.  The argument needs SX = 32 A and SY = 32 B
// (no i32 wrap), else the full rows are kept.
template <int T, int SW>
__device__ __forceinline__ int stage_entry(const TriSetup &ts, uint32_t id, uint32_t gflags, int tx0, int ty0,
                                           uint32_t *st, int lane) {
    using SG = Stage<T>;
    constexpr int NR = SG::NR;
    int qx0 = max(ts.ox0, tx0), qy0 = max(ts.oy0, ty0);
    int qx1 = min(ts.cmxx, tx0 + T), qy1 = min(ts.cmxy, ty0 + T);
    int nqx = qx1 > qx0 ? (qx1 - qx0 + 1) >> 1 : 0;
    int nqy = qy1 > qy0 ? (qy1 - qy0 + 1) >> 1 : 0;
    bool wrap = false;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        st[(F_A + k) * SW + lane] = __float_as_uint(ts.A[k]);
        st[(F_B + k) * SW + lane] = __float_as_uint(ts.B[k]);
        st[(F_C + k) * SW + lane] = __float_as_uint(ts.C[k]);
        wrap = wrap || ts.SX[k] != ts.A[k] * 32.0f || ts.SY[k] != ts.B[k] * 32.0f;
    }
    st[F_OOA * SW + lane] = __float_as_uint(ts.ooa);
    st[F_OA * SW + lane] = __float_as_uint(ts.oa);
    st[F_ODA * SW + lane] = __float_as_uint(ts.oda);
    st[F_ODB * SW + lane] = __float_as_uint(ts.odb);
    st[F_ZA * SW + lane] = __float_as_uint(ts.za);
    st[F_ZDA * SW + lane] = __float_as_uint(ts.zda);
    st[F_ZDB * SW + lane] = __float_as_uint(ts.zdb);
    st[F_ID * SW + lane] = id;
    st[F_ORG * SW + lane] = (uint32_t)ts.ox0 | ((uint32_t)ts.oy0 << 16);
    st[F_Q0 * SW + lane] = (uint32_t)qx0 | ((uint32_t)qy0 << 16);

    // rejected 16px blocks among the (at most 3x3) that meet the tile
    uint32_t flags = (ts.coarse ? 1u : 0u) | ((gflags & SR_GEO_ALPHA) ? 2u : 0u) | (wrap ? STEP_WRAP : 0u);
    if (ts.coarse && nqx > 0 && nqy > 0) {
        int kx0 = (qx0 - ts.ox0) >> 4, kx1 = (qx0 + 2 * (nqx - 1) - ts.ox0) >> 4;
        int ky0 = (qy0 - ts.oy0) >> 4, ky1 = (qy0 + 2 * (nqy - 1) - ts.oy0) >> 4;
        uint32_t mask = 0u;
        for (int ky = ky0; ky <= ky1; ky++)
            for (int kx = kx0; kx <= kx1; kx++)
                if (block_out(ts, ts.ox0 + 16 * kx, ts.oy0 + 16 * ky)) mask |= 1u << ((ky - ky0) * 3 + (kx - kx0));
        flags |= (mask << 8) | ((uint32_t)kx0 << 20) | ((uint32_t)ky0 << 24);
    }
    st[F_FLAGS * SW + lane] = flags;

    bool spans = nqx * nqy > 4;
#pragma unroll
    for (int e = 0; e < 3; e++) spans = spans && fabsf(ts.A[e]) < 67108864.0f && fabsf(ts.B[e]) < 67108864.0f;
    float base[3], slope[3], slack[3];
    int kind[3];
    const int qxr = qx0 - tx0, qyr = qy0 - ty0;     // first quad relative to the tile (pixels)
    if (spans) {
        const float Xt = (float)(tx0 * 16), Yt = (float)(ty0 * 16);
        const float Xm = (float)((tx0 + T) * 16 + 24), Ym = (float)((ty0 + T) * 16 + 24);
        const float Ymr = (float)(T * 16 + 24);
#pragma unroll
        for (int e = 0; e < 3; e++) {
            float A = ts.A[e], B = ts.B[e], C = ts.C[e];
            float P = ((fabsf(A) * Xm + fabsf(B) * Ym) + fabsf(C)) + 512.0f * (fabsf(A) + fabsf(B));
            float M2 = 0x1p-18f * P;
            float Cr = (C + A * Xt) + B * Yt;
            kind[e] = A > 0.0f ? 1 : (A < 0.0f ? -1 : 0);
            if (kind[e] != 0) {
                // E >= -M2  <=>  A X' >= -(M2 + C' + B Y')
                float inv = __builtin_amdgcn_rcpf(A);
                base[e] = -(M2 + Cr) * inv;
                slope[e] = -B * inv;
            } else {
                base[e] = M2 + Cr;    // feasible where B Y' + C' + M2 >= 0
                slope[e] = B;
            }
            slack[e] = 16.0f + 0x1p-20f * (fabsf(base[e]) + fabsf(slope[e]) * Ymr);
        }
    }
    int cum = 0;
    uint32_t cw = 0u, lw = 0u;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        int l = 0, n = 0;
        if (r < nqy) {
            if (!spans) {
                n = nqx;
            } else {
                float Y0 = (float)((qyr + 2 * r) * 16 + 8), Y1 = Y0 + 16.0f;
                float xlo = -INFINITY, xhi = INFINITY;
                bool feas = true;
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    float b0 = fmaf(slope[e], Y0, base[e]), b1 = fmaf(slope[e], Y1, base[e]);
                    if (kind[e] > 0) xlo = fmaxf(xlo, fminf(b0, b1) - slack[e]);
                    else if (kind[e] < 0) xhi = fminf(xhi, fmaxf(b0, b1) + slack[e]);
                    else feas = feas && fmaxf(b0, b1) >= -slack[e];
                }
                // quad column c has samples at X' = (qxr + 2c) * 16 + 8 and + 24
                float clo = fmaxf(ceilf((xlo - 24.0f - 16.0f * (float)qxr) * (1.0f / 32.0f)), 0.0f);
                float chi = fminf(floorf((xhi - 8.0f - 16.0f * (float)qxr) * (1.0f / 32.0f)), (float)(nqx - 1));
                if (feas && clo <= chi) {
                    l = (int)clo;
                    n = (int)chi - l + 1;
                }
            }
        }
        cum += n;
        cw |= (uint32_t)cum << (16 * (r & 1));
        lw |= (uint32_t)l << (4 * (r & 7));
        if (r & 1) {
            st[(SG::CUM + r / 2) * SW + lane] = cw;
            cw = 0u;
        }
        if ((r & 7) == 7 || r == NR - 1) {
            st[(SG::LO + r / 8) * SW + lane] = lw;
            lw = 0u;
        }
    }
    return cum;
}

// Coverage and depth of the local-th candidate quad of staged entry o: the
// reference's block choice and rejection, then the four samples' edge
// chains (src/tilerasterizer.rs:220-381); covered samples with a non-NaN
// depth enter the 64-bit minimum.
template <int T, int SW>
__device__ __forceinline__ void walk_quad(const DevScene &S, const RasterArgs &P, const uint32_t *st, int o, int local,
                                          int tx0, int ty0, unsigned long long *keys, uint32_t zbound) {
    using SG = Stage<T>;
    constexpr int NR = SG::NR;
#define SF(f) __uint_as_float(st[(f) * SW + o])
    // row = number of rows whose inclusive count is <= local
    int row = 0;
#pragma unroll
    for (int r = 0; r < NR; r += 2) {
        uint32_t cw = st[(SG::CUM + r / 2) * SW + o];
        row += ((int)(cw & 0xFFFFu) <= local) + ((int)(cw >> 16) <= local);
    }
    int before = 0;
    if (row > 0) before = (int)((st[(SG::CUM + (row - 1) / 2) * SW + o] >> (16 * ((row - 1) & 1))) & 0xFFFFu);
    int col = (int)((st[(SG::LO + row / 8) * SW + o] >> (4 * (row & 7))) & 15u) + (local - before);
    uint32_t q0 = st[F_Q0 * SW + o];
    int qpx = (int)(q0 & 0xFFFFu) + 2 * col, qpy = (int)(q0 >> 16) + 2 * row;
    if (zbound) {
        unsigned long long *kz = keys + (qpy - ty0) * T + (qpx - tx0);
        auto depth_of = [](unsigned long long *k) {
            return (uint32_t)(__hip_atomic_load(k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> 32);
        };
        uint32_t qmax = max(max(depth_of(kz), depth_of(kz + 1)), max(depth_of(kz + T), depth_of(kz + T + 1)));
        if (zbound > qmax) return;
    }

    uint32_t org = st[F_ORG * SW + o], flags = st[F_FLAGS * SW + o];
    int ox0 = (int)(org & 0xFFFFu), oy0 = (int)(org >> 16);
    int bx = ox0, by = oy0;
    if (flags & 1u) {
        int kx = (qpx - ox0) >> 4, ky = (qpy - oy0) >> 4;
        int bit = (ky - (int)((flags >> 24) & 15u)) * 3 + (kx - (int)((flags >> 20) & 15u));
        if ((flags >> (8 + bit)) & 1u) return;
        bx = ox0 + 16 * kx;
        by = oy0 + 16 * ky;
    }
    int i = (qpx - bx) >> 1, j = (qpy - by) >> 1;
    float fi = (float)i, fj = (float)j;
    float X0 = (float)(bx * 16 + 8), X1 = (float)(bx * 16 + 24);
    float Y0 = (float)(by * 16 + 8), Y1 = (float)(by * 16 + 24);
    float v[3][4], w[3][4], SXe[3], SYe[3];
#pragma unroll
    for (int e = 0; e < 3; e++) {
        SXe[e] = SF(F_A + e) * 32.0f;
        SYe[e] = SF(F_B + e) * 32.0f;
    }
    if (flags & STEP_WRAP) {
        // rare: recompute the i32 steps from the snapped vertices
        uint32_t id = st[F_ID * SW + o];
        float4 r0 = P.raster[4 * (size_t)id], r1 = P.raster[4 * (size_t)id + 1];
        int32_t p0x = __float_as_int(r0.x), p0y = __float_as_int(r0.y), p1x = __float_as_int(r0.z);
        int32_t p1y = __float_as_int(r0.w), p2x = __float_as_int(r1.x), p2y = __float_as_int(r1.y);
        SXe[0] = (float)wmul(wsub(p2y, p1y), 32); SYe[0] = (float)wmul(wsub(p1x, p2x), 32);
        SXe[1] = (float)wmul(wsub(p0y, p2y), 32); SYe[1] = (float)wmul(wsub(p2x, p0x), 32);
        SXe[2] = (float)wmul(wsub(p1y, p0y), 32); SYe[2] = (float)wmul(wsub(p0x, p1x), 32);
    }
    bool ok = true;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        float A = SF(F_A + e), B = SF(F_B + e), Cc = SF(F_C + e);
        float sx = SXe[e], sy = SYe[e];
        float ax0 = A * X0, ax1 = A * X1, by0 = B * Y0, by1 = B * Y1;
        v[e][0] = (ax0 + by0) + Cc;
        v[e][1] = (ax1 + by0) + Cc;
        v[e][2] = (ax0 + by1) + Cc;
        v[e][3] = (ax1 + by1) + Cc;
#pragma unroll
        for (int s = 0; s < 4; s++) w[e][s] = chain_fast(v[e][s], sy, sx, fj, fi, ok);
    }
    if (!ok) {
#pragma unroll
        for (int e = 0; e < 3; e++) {
#pragma unroll
            for (int s = 0; s < 4; s++) w[e][s] = chain_serial(v[e][s], SYe[e], SXe[e], j, i);
        }
    }
    uint32_t id = st[F_ID * SW + o];
    bool alpha = (flags & 2u) != 0;
    float ooa = SF(F_OOA), oa = SF(F_OA), oda = SF(F_ODA), odb = SF(F_ODB);
    float za = SF(F_ZA), zda = SF(F_ZDA), zdb = SF(F_ZDB);
    unsigned long long *kq = keys + (qpy - ty0) * T + (qpx - tx0);
#pragma unroll
    for (int s = 0; s < 4; s++) {
        if (!(w[0][s] >= 0.0f && w[1][s] >= 0.0f && w[2][s] >= 0.0f)) continue;
        float b1 = w[1][s] * ooa, b2 = w[2][s] * ooa;
        float wv = 1.0f / ((oa + b1 * oda) + b2 * odb);
        float z = ((za + b1 * zda) + b2 * zdb) * wv;
        if (z != z) continue;
        if (alpha && !alpha_pass(S, P.src, P.raster + 4 * (size_t)id, b1, b2, wv)) continue;
        unsigned long long key = ((unsigned long long)depth_key(z) << 32) | (unsigned long long)(~id);
        atomicMin(&kq[(s >> 1) * T + (s & 1)], key);
    }
#undef SF
}

// A depth key no fragment of the triangle goes below, or 0 (no bound).  The
// depth ((za + b1 zda) + b2 zdb) / ((oa + b1 oda) + b2 odb) is a ratio of
// affine functions of the barycentrics, so over the triangle it lies between
// its values at the vertices when the denominator keeps its sign.  Covered
// samples have b1, b2 >= 0 and b1 + b2 <= 1 to within 4e-7 (the edge values
// and 1/area rounded once each) and are evaluated to a few ulps; with the
// vertices' denominators within a factor 16 of each other the depth's
// derivative in b is below 48 times the largest vertex depth, so a margin of
// 1e-4 of that depth covers both with room to spare.
__device__ __forceinline__ uint32_t tri_zmin_key(float za, float zda, float zdb, float oa, float oda, float odb) {
    float d0 = oa, d1 = oa + oda, d2 = oa + odb;
    float dmin = fminf(d0, fminf(d1, d2)), dmax = fmaxf(d0, fmaxf(d1, d2));
    if (!(dmin > 0.0f) || !(dmax <= 16.0f * dmin)) return 0u;
    float z0 = za / d0, z1 = (za + zda) / d1, z2 = (za + zdb) / d2;
    float zmin = fminf(z0, fminf(z1, z2));
    float zabs = fmaxf(fabsf(z0), fmaxf(fabsf(z1), fabsf(z2)));
    float lo = zmin - (zabs * 1e-4f + 1e-30f);
    if (!(lo == lo) || !(zabs < 1e30f)) return 0u;
    return depth_key(lo);
}

__device__ __forceinline__ int find_view(const RasterArgs &P, uint32_t item) {
    int lo = 0, hi = P.num_views - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (P.views[mid].item_base <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The 2x2 quad at the origin of a tile whose origin is (64k, 64m + 32) is the
// centre quad of its 64px reference tile (src/tilerasterizer.rs:103-106).
__device__ __forceinline__ bool centre_tile(int tx0, int ty0) { return ((tx0 & 63) == 0) && ((ty0 & 63) == 32); }

// ---------------------------------------------------------------------------
// k_raster: one workgroup per (view, tile): coverage and depth of every bin
// entry into an LDS depth buffer, then the visibility buffer (triangle id per
// pixel, or none).
// ---------------------------------------------------------------------------
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b % 8).  The per-tile kernels take their items in runs of XCD_RUN
// consecutive ones per XCD (run r on XCD r % 8): neighbouring tiles, which
// read the same triangles, texels and voxels, then share one XCD's L2,
// while the runs still interleave over the whole chip.  The grid is a
// multiple of 8 x XCD_RUN; workgroups past the last item do nothing.
#define XCD_RUN 8u
__device__ __forceinline__ uint32_t xcd_item(uint32_t b) {
    const uint32_t x = b & 7u, k = b >> 3;
    return ((k / XCD_RUN) * 8u + x) * XCD_RUN + (k % XCD_RUN);
}

static inline uint32_t xcd_grid(uint32_t items) { return (items + 8u * XCD_RUN - 1) / (8u * XCD_RUN) * (8u * XCD_RUN); }

template <int T, int NT, bool HIZ>
__device__ void raster_item(const RasterArgs &P, uint32_t item, unsigned long long *keys, uint32_t *stage) {
    constexpr int NW = NT / 64;               // waves
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const DevView &V = P.views[find_view(P, item)];
    const DevScene &S = *V.scene;
    uint32_t tile = item - V.item_base;
    int tx0 = (int)(tile % (uint32_t)V.tiles_x) * T, ty0 = (int)(tile / (uint32_t)V.tiles_x) * T;

    for (int i = tid; i < T * T; i += NT) keys[i] = SR_EMPTY_KEY;
    __syncthreads();

    // ---- phase 1: coverage and depth ----
    // The workgroup takes NT bin entries at a time: every thread loads and
    // sets up one triangle, stages it in LDS, and counts the quads of its
    // bounding box inside this tile that can hold a covered sample.  A
    // workgroup prefix sum turns the counts into a dense list of (entry,
    // quad) tasks that all NT threads then walk, so one-pixel and full-tile
    // triangles, and tiles with few or many entries, keep every lane busy.
    __shared__ int wtot[NW];
    __shared__ uint32_t zmin_s[HIZ ? NT : 1];
    // A frame whose records or list entries outgrew their buffers is run
    // again with larger ones; its lists are incomplete (the binning keeps
    // only entries that fit), so it rasterizes nothing.  Every id read is
    // also checked against the records.
    const uint32_t nrec = P.status->total_records, npairs = P.status->total_pairs;
    uint32_t start = P.bin_start[item];
    uint32_t end = P.bin_end[item];
    if (nrec > P.rec_cap || npairs > P.pair_cap) end = start;
    if (end > P.pair_cap) end = P.pair_cap;
    for (uint32_t c0 = start; c0 < end; c0 += (uint32_t)NT) {
        uint32_t e = c0 + (uint32_t)tid;
        int nq = 0;
        if (e < end && P.list[e] < nrec) {
            uint32_t id = P.list[e];
            uint4 g = P.geo[id];
            TriSetup ts;
            tri_setup(g, P.raster + 4 * (size_t)id, tx0, ty0, ts);
            nq = stage_entry<T, NT>(ts, id, g.z, tx0, ty0, stage, tid);
        }
        int incl = nq;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63) wtot[wave] = incl;
        __syncthreads();
        int total = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            if (w < wave) incl += wtot[w];
            total += wtot[w];
        }
        stage[F_INCL * NT + tid] = (uint32_t)incl;
        // HIZ: a quad task is dropped when its entry's depth bound lies
        // behind all four of the quad's stored depths (the entry could not
        // win any of them; a stale read only keeps more work).
        if constexpr (HIZ) {
#define SFE(f) __uint_as_float(stage[(f) * NT + tid])
            zmin_s[tid] = nq > 0 ? tri_zmin_key(SFE(F_ZA), SFE(F_ZDA), SFE(F_ZDB), SFE(F_OA), SFE(F_ODA), SFE(F_ODB)) : 0u;
#undef SFE
        }
        __syncthreads();
        for (int k0 = 0; k0 < total; k0 += NT) {
            int k = k0 + tid;
            if (k < total) {
                // owner = number of entries whose inclusive count is <= k
                int o = 0;
#pragma unroll
                for (int step = NT / 2; step > 0; step >>= 1)
                    if ((int)stage[F_INCL * NT + o + step - 1] <= k) o += step;
                int before = o ? (int)stage[F_INCL * NT + o - 1] : 0;
                uint32_t zb = 0u;
                if constexpr (HIZ) zb = zmin_s[o];
                walk_quad<T, NT>(S, P, stage, o, k - before, tx0, ty0, keys, zb);
            }
        }
        __syncthreads();
    }

    // ---- resolve: canonical triangle id per pixel ----
    uint32_t rec0 = P.tri_base[V.in_tri_base];
    bool centre = centre_tile(tx0, ty0);
    for (int p = tid; p < T * T; p += NT) {
        int px = tx0 + (p % T), py = ty0 + (p / T);
        unsigned long long key = keys[p];
        uint32_t out = key != SR_EMPTY_KEY ? ~(uint32_t)key - rec0 : 0xFFFFFFFFu;
        if (px < V.W && py < V.H) P.out_ids[V.fb_offset + (size_t)py * (size_t)V.W + (size_t)px] = out;
        if (centre && px < tx0 + 2 && py < ty0 + 2) {
            int ri = (ty0 / 64) * V.ref_tiles_x + (tx0 / 64);
            P.center_ids[4 * (V.lum_offset + (uint32_t)ri) + (uint32_t)((py - ty0) * 2 + (px - tx0))] = out;
        }
    }
}

template <int T, int NT, bool HIZ = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(T == 32 ? 4 : 5))) k_raster(RasterArgs P) {
    __shared__ unsigned long long keys[T * T];
    __shared__ uint32_t stage[Stage<T>::COUNT * NT];
    __shared__ uint32_t next;
    clock_start(P.clock, CLOCK_RASTER);
    if (!P.persistent) {
        const uint32_t item = xcd_item(blockIdx.x);
        if (item < P.total_items) raster_item<T, NT, HIZ>(P, item, keys, stage);
    } else {
        for (;;) {
            if (threadIdx.x == 0) next = atomicAdd(&P.status->work_counter, 1u);
            __syncthreads();
            uint32_t item = next;
            __syncthreads();
            if (item >= P.total_items) break;
            raster_item<T, NT, HIZ>(P, item, keys, stage);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// k_shade: one workgroup per (view, tile), one thread per pixel in 2x2-quad
// order: barycentrics and depth of the visible triangle, PBR shading or
// skybox, colour planes or the fused tonemapped RGBA8, and the centre-quad
// luminance of each 64px tile (src/tilerasterizer.rs:384-476).
// ---------------------------------------------------------------------------
// Value of lane K of this lane's quad (lanes 4q..4q+3 of the wave), by a DPP
// quad permutation (one VALU move, no LDS round trip).  All four lanes of
// the quad must be active.
template <int K> __device__ __forceinline__ float quad_lane(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xF, 0xF, false));
}

// The tile's triangle table: when a tile shows at most CAP distinct
// triangles, each is set up once, by one thread, into an LDS row (TT_WORDS
// words: the two edges and the depth interpolators the barycentric chains
// need, then the shading interpolators and tangent frame of
// tri_shade_attrs), and every pixel of the triangle reads its row.  Ids are
// deduplicated through an open-addressing hash of 2 x NT slots and ranked
// densely.  Denser tiles (a triangle or two per pixel, where a row would
// serve almost no other pixel) set up each pixel's triangle in the pixel's
// thread.
#define TT_WORDS 60
template <int NT> struct ShadeLds {
    static constexpr int HS = 2 * NT;                         // hash slots
    // table rows: 256-thread workgroups stay within 32 KB of LDS, so five
    // fit a CU's 160 KB (five waves per SIMD, as the VGPR budget allows)
    static constexpr int CAP = NT == 64 ? 32 : (NT == 256 ? 112 : 256);
    uint32_t key[HS];
    uint16_t rank[HS];
    uint32_t tri[CAP];
    float4 tab[CAP][TT_WORDS / 4];
    uint32_t wsum[NT / 64];
};

__device__ __forceinline__ uint32_t tt_hash(uint32_t id, int hs) { return (id * 2654435761u) & (uint32_t)(hs - 1); }

// Row layout: [0] A1 B1 C1 SX1, [1] SY1 A2 B2 C2, [2] SX2 SY2 ooa oa,
// [3] oda odb za zda, [4] zdb origin(ox0 | oy0 << 16) flags(coarse 1,
// normal map 2) material, [5..11] shading interpolators Sh[0..6],
// [12..14] tangent frame Tn[0..2] (normal-mapped triangles only).
__device__ __forceinline__ void tt_fill(const RasterArgs &P, uint32_t id, int tx0, int ty0, float4 *row) {
    const float4 *R = P.raster + 4 * (size_t)id;
    uint4 g = P.geo[id];
    bool nmap = (g.z & SR_GEO_NORMAL_MAP) != 0;
    {
        TriSetup ts;
        tri_setup(g, R, tx0, ty0, ts);
        row[0] = make_float4(ts.A[1], ts.B[1], ts.C[1], ts.SX[1]);
        row[1] = make_float4(ts.SY[1], ts.A[2], ts.B[2], ts.C[2]);
        row[2] = make_float4(ts.SX[2], ts.SY[2], ts.ooa, ts.oa);
        row[3] = make_float4(ts.oda, ts.odb, ts.za, ts.zda);
        row[4] = make_float4(ts.zdb, __uint_as_float((uint32_t)ts.ox0 | ((uint32_t)ts.oy0 << 16)),
                             __uint_as_float((ts.coarse ? 1u : 0u) | (nmap ? 2u : 0u)), R[3].y);
    }
    CV tv[3];
    float4 Sh[7], Tn[3];
    load_tri_cv(P.src, R, tv);
    tri_shade_attrs(tv, R, nmap, Sh, Tn);
#pragma unroll
    for (int k = 0; k < 7; k++) row[5 + k] = Sh[k];
    if (nmap) {
#pragma unroll
        for (int k = 0; k < 3; k++) row[12 + k] = Tn[k];
    }
}

template <int T, int NT>
__device__ void shade_item(const RasterArgs &P, uint32_t item) {
    constexpr int QW = T / 2;                 // quads per row
    // one thread per pixel, in whole waves: the DPP quad exchanges below
    // read all four lanes of each quad, so every lane must be active
    static_assert(T * T == NT && NT % 64 == 0, "one thread per pixel in whole waves");
    using L = ShadeLds<NT>;
    __shared__ L sm;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const DevView &V = P.views[find_view(P, item)];
    const DevScene &S = *V.scene;
    uint32_t tile = item - V.item_base;
    int tx0 = (int)(tile % (uint32_t)V.tiles_x) * T, ty0 = (int)(tile / (uint32_t)V.tiles_x) * T;
    uint32_t rec0 = P.tri_base[V.in_tri_base];
    bool fused = (V.flags & 1u) != 0;
    const int p = tid;
    int q = p >> 2, s = p & 3;
    int qx = q % QW, qy = q / QW;
    int sx = s & 1, sy = s >> 1;
    int px = tx0 + 2 * qx + sx, py = ty0 + 2 * qy + sy;
    int qpx = tx0 + 2 * qx, qpy = ty0 + 2 * qy;
    bool on_screen = px < V.W && py < V.H;
    bool center = centre_tile(tx0, ty0) && q == 0;
    size_t o = V.fb_offset + (size_t)py * (size_t)V.W + (size_t)px;
    uint32_t vid = 0xFFFFFFFFu;
    if (center) vid = P.center_ids[4 * (V.lum_offset + (uint32_t)((ty0 / 64) * V.ref_tiles_x + (tx0 / 64))) + s];
    else if (on_screen) vid = P.out_ids[o];
    bool has = vid != 0xFFFFFFFFu;
    uint32_t id = has ? vid + rec0 : 0u;

    // ---- the tile's distinct triangles: hash, dense ranks, rows ----
    for (int k = tid; k < L::HS; k += NT) sm.key[k] = 0xFFFFFFFFu;
    __syncthreads();
    uint32_t h = 0;
    bool owner = false;
    if (has) {
        h = tt_hash(id, L::HS);
        for (;;) {      // at most NT ids in 2 NT slots: a free slot is always found
            uint32_t old = atomicCAS(&sm.key[h], 0xFFFFFFFFu, id);
            if (old == 0xFFFFFFFFu || old == id) {
                owner = old == 0xFFFFFFFFu;
                break;
            }
            h = (h + 1u) & (uint32_t)(L::HS - 1);
        }
    }
    uint64_t om = __ballot(owner);
    if (lane == 0) sm.wsum[wave] = (uint32_t)__popcll(om);
    __syncthreads();
    uint32_t before = 0, ntri = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
        uint32_t c = sm.wsum[w];
        if (w < wave) before += c;
        ntri += c;
    }
    const bool table = ntri <= (uint32_t)L::CAP;     // the same in every thread
    if (table && owner) {
        uint32_t r = before + (uint32_t)__popcll(om & ((1ull << lane) - 1ull));
        sm.rank[h] = (uint16_t)r;
        sm.tri[r] = id;
    }
    __syncthreads();
    if (table && (uint32_t)tid < ntri) tt_fill(P, sm.tri[tid], tx0, ty0, sm.tab[tid]);
    __syncthreads();
    const float4 *rw = sm.tab[(table && has) ? sm.rank[h] : 0];

    // ---- barycentrics and depth of the visible triangle ----
    float b1 = 0.0f, b2 = 0.0f, z = INFINITY, w = 0.0f;
    float oa = 0.0f, oda = 0.0f, odb = 0.0f;
    if (has) {
        float A1, B1, C1, SX1, SY1, A2, B2, C2, SX2, SY2, ooa, za, zda, zdb;
        int ox0, oy0;
        bool coarse;
        if (table) {
            float4 r0 = rw[0], r1 = rw[1], r2 = rw[2], r3 = rw[3], r4 = rw[4];
            A1 = r0.x; B1 = r0.y; C1 = r0.z; SX1 = r0.w;
            SY1 = r1.x; A2 = r1.y; B2 = r1.z; C2 = r1.w;
            SX2 = r2.x; SY2 = r2.y; ooa = r2.z; oa = r2.w;
            oda = r3.x; odb = r3.y; za = r3.z; zda = r3.w;
            zdb = r4.x;
            uint32_t org = __float_as_uint(r4.y);
            ox0 = (int)(org & 0xFFFFu);
            oy0 = (int)(org >> 16);
            coarse = (__float_as_uint(r4.z) & 1u) != 0u;
        } else {
            TriSetup ts;
            tri_setup(P.geo[id], P.raster + 4 * (size_t)id, tx0, ty0, ts);
            A1 = ts.A[1]; B1 = ts.B[1]; C1 = ts.C[1]; SX1 = ts.SX[1]; SY1 = ts.SY[1];
            A2 = ts.A[2]; B2 = ts.B[2]; C2 = ts.C[2]; SX2 = ts.SX[2]; SY2 = ts.SY[2];
            ooa = ts.ooa; oa = ts.oa; oda = ts.oda; odb = ts.odb; za = ts.za; zda = ts.zda; zdb = ts.zdb;
            ox0 = ts.ox0; oy0 = ts.oy0; coarse = ts.coarse;
        }
        // the walk origin of the pixel's quad (quad_origin; the quad is walked)
        int bx = ox0, by = oy0;
        if (coarse) {
            bx = ox0 + ((qpx - ox0) & ~15);
            by = oy0 + ((qpy - oy0) & ~15);
        }
        int i = (qpx - bx) >> 1, j = (qpy - by) >> 1;
        float w1 = edge_chain_v(A1, B1, C1, SX1, SY1, bx, by, sx, sy, i, j);
        float w2 = edge_chain_v(A2, B2, C2, SX2, SY2, bx, by, sx, sy, i, j);
        b1 = w1 * ooa;
        b2 = w2 * ooa;
        w = 1.0f / ((oa + b1 * oda) + b2 * odb);
        z = ((za + b1 * zda) + b2 * zdb) * w;
    }
    // every lane of the wave is active here (one thread per pixel).  When
    // the quad shows one triangle, its lanes' own w are the shading's four
    // w bit for bit (the same expression on the same barycentrics);
    // otherwise each lane evaluates them for its triangle.
    float wv[4] = {quad_lane<0>(w), quad_lane<1>(w), quad_lane<2>(w), quad_lane<3>(w)};
    {
        float qb1[4] = {quad_lane<0>(b1), quad_lane<1>(b1), quad_lane<2>(b1), quad_lane<3>(b1)};
        float qb2[4] = {quad_lane<0>(b2), quad_lane<1>(b2), quad_lane<2>(b2), quad_lane<3>(b2)};
        uint32_t q0 = __float_as_uint(quad_lane<0>(__uint_as_float(vid)));
        bool one_tri = __float_as_uint(quad_lane<1>(__uint_as_float(vid))) == q0 &&
                       __float_as_uint(quad_lane<2>(__uint_as_float(vid))) == q0 &&
                       __float_as_uint(quad_lane<3>(__uint_as_float(vid))) == q0;
        if (has && !one_tri) quad_w(oa, oda, odb, qb1, qb2, wv);
    }
    f3 color = mk3(0.0f, 0.0f, 0.0f);
    if (on_screen || center) {
        bool covered = has && z != INFINITY;
        if (covered) {
            float4 Sh[7], Tn[3];
            bool nmap;
            int material;
            if (table) {
#pragma unroll
                for (int k = 0; k < 7; k++) Sh[k] = rw[5 + k];
                float4 r4 = rw[4];
                nmap = (__float_as_uint(r4.z) & 2u) != 0;
                material = __float_as_int(r4.w);
                if (nmap) {
#pragma unroll
                    for (int k = 0; k < 3; k++) Tn[k] = rw[12 + k];
                }
            } else {
                const float4 *R = P.raster + 4 * (size_t)id;
                CV tv[3];
                nmap = (P.geo[id].z & SR_GEO_NORMAL_MAP) != 0;
                material = __float_as_int(R[3].y);
                load_tri_cv(P.src, R, tv);
                tri_shade_attrs(tv, R, nmap, Sh, Tn);
            }
            color = shade_pixel<false>(S, V, material, Sh, Tn, nmap, b1, b2, wv, s, color);
        } else {
            color = skybox_pixel(S, V, qpx, qpy, sx, sy);
        }
    }
    if (on_screen) {
        P.out_depth[o] = z;
        if (fused) {
            float k = 0.2f;
            float cr = color.x * V.exposure, cg = color.y * V.exposure, cb = color.z * V.exposure;
            cr = (cr / (cr + k)) * (1.0f + k);
            cg = (cg / (cg + k)) * (1.0f + k);
            cb = (cb / (cb + k)) * (1.0f + k);
            P.out_rgba[o] = (sat_u8(cr * 255.0f) << 24) | (sat_u8(cg * 255.0f) << 16) | (sat_u8(cb * 255.0f) << 8) |
                            0xFFu;
        } else {
            P.out_color[o] = color.x;
            P.out_color[P.color_plane + o] = color.y;
            P.out_color[2 * P.color_plane + o] = color.z;
        }
    }
    if (center && P.center_color) {
        int ri = (ty0 / 64) * V.ref_tiles_x + (tx0 / 64);
        P.center_color[4 * (V.lum_offset + (uint32_t)ri) + s] = make_float4(color.x, color.y, color.z, z);
    }
    // centre-quad luminance of the 64px tile (src/tilerasterizer.rs:103-106)
    float lum = (color.x * 0.2126f + color.y * 0.7152f) + color.z * 0.0722f;
    float l0 = quad_lane<0>(lum), l1 = quad_lane<1>(lum), l2 = quad_lane<2>(lum), l3 = quad_lane<3>(lum);
    if (center && s == 0) {
        int ri = (ty0 / 64) * V.ref_tiles_x + (tx0 / 64);
        P.out_lum[V.lum_offset + (uint32_t)ri] = (((l0 + l1) + l2) + l3) * 0.25f;
    }
}

// One pixel per thread at five waves per SIMD (up to 96 VGPRs, no spill)
// hides more of the shading chain's dependent memory latency than four
// (measured 0.90x the time at 1080p, 0.85x at 2160p), and beats six, whose
// 80 VGPRs spill 56 bytes per thread to scratch (0.98x the time at 1080p).
// 32-px tiles take 1024 threads (2.5% faster batches than 256 threads
// looping over four pixels each).
template <int T, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(T * T > NT ? 4 : 5))) k_shade(RasterArgs P) {
    // the frame's status is final once shading starts (no later kernel
    // writes it): the first workgroup stores it in the host's pinned copy
    // with system-scope stores (after its pixels, where no register is
    // live), so no read-back copy follows the frame
    const uint32_t item = xcd_item(blockIdx.x);
    if (item < P.total_items) shade_item<T, NT>(P, item);
    constexpr uint32_t STATUS_WORDS = sizeof(DevStatus) / 4;
    static_assert(STATUS_WORDS <= NT, "one thread per status word");
    if (blockIdx.x == 0 && threadIdx.x < STATUS_WORDS && P.status_out)
        __hip_atomic_store(reinterpret_cast<uint32_t *>(P.status_out) + threadIdx.x,
                           reinterpret_cast<const SR_GLOBAL uint32_t *>(P.status)[threadIdx.x], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    clock_end(P.clock, CLOCK_SHADE_END);
}

// ---------------------------------------------------------------------------
// Translucent triangles: 64-px reference-tile lists, a back-to-front sort per
// tile and forward shading over the opaque frame (src/renderer.rs:765-829,
// src/tilerasterizer.rs:92-101, src/shader.rs:66-76).
// ---------------------------------------------------------------------------
// The reference's bin loop for one triangle (src/renderer.rs:756-793): the
// 64-px tile range of its pixel box, bin index y * tiles_x + x in wrapping
// i32 with its bounds check, and the clip rectangle's emptiness test.  A box
// that reaches the right edge of a screen whose width is a multiple of 64
// runs one column past the last tile; that index names the first tile of the
// next row, so the triangle is listed there twice and blends twice, as in the
// reference.
template <class F>
__device__ __forceinline__ void ref_tile_bins(const DevView &V, uint4 g, F f) {
    int mnx = (int)(g.x & 0xFFFFu), mny = (int)(g.x >> 16), mxx = (int)(g.y & 0xFFFFu), mxy = (int)(g.y >> 16);
    int bx0 = mnx / SR_REF_TILE, by0 = mny / SR_REF_TILE;
    int bx1 = wadd(mxx, SR_REF_TILE - 1) / SR_REF_TILE, by1 = wadd(mxy, SR_REF_TILE - 1) / SR_REF_TILE;
    int ntiles = V.ref_tiles_x * V.ref_tiles_y;
    for (int y = by0; y < by1; y++)
        for (int x = bx0; x < bx1; x++) {
            int bi = wadd(wmul(y, V.ref_tiles_x), x);
            if (bi < 0 || bi >= ntiles) continue;
            int tx = (bi % V.ref_tiles_x) * SR_REF_TILE, ty = (bi / V.ref_tiles_x) * SR_REF_TILE;
            if (min(mxx, tx + SR_REF_TILE) - max(mnx, tx) < 1 || min(mxy, ty + SR_REF_TILE) - max(mny, ty) < 1) continue;
            f((uint32_t)bi);
        }
}

// Count (list null) or scatter pass over the set-up records: translucent
// ones into the lists of their 64-px tiles (slot = view's first tile + bin).
__global__ void __launch_bounds__(256) k_tr_bins(const DevView *views, const uint4 *geo, uint32_t rec_cap,
                                                 uint32_t *counts, uint32_t *list, uint32_t cap, DevStatus *status) {
    uint32_t total = min(status->total_records, rec_cap);
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < total; id += gridDim.x * blockDim.x) {
        uint4 g = geo[id];
        if ((g.z & (SR_GEO_TRANSLUCENT | SR_GEO_EMPTY)) != SR_GEO_TRANSLUCENT) continue;
        const DevView &V = views[g.w];
        uint32_t *c = counts + V.lum_offset;
        ref_tile_bins(V, g, [&](uint32_t bi) {
            uint32_t pos = atomicAdd(&c[bi], 1u);
            if (!list) return;
            if (pos < cap) list[pos] = id;
            else atomicOr(&status->overflow_tr, 1u);
        });
    }
}

// avg_z of the packet: mean clip-space z of the three vertices
// (src/renderer.rs:765-775).
__device__ __forceinline__ float tri_avg_z(const TriSource &T, const float4 *R) {
    uint32_t src0 = __float_as_uint(R[3].z), src1 = __float_as_uint(R[3].w);
    float z[3];
    if (src0 & SRC_CLIPPED) {
        uint32_t e = src0 & 0x07FFFFFFu, f = (src0 >> 27) & 15u;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            uint32_t vi = j == 0 ? 0u : f + (uint32_t)j;
            z[j] = T.clip_buf[((size_t)e * 16u + vi) * 16u + 2u];
        }
    } else {
        const DevDraw &D = T.draws[src1];
        const DevScene &S = *T.views[D.view].scene;
        uint32_t lt = src0 - D.in_tri_base;
#pragma unroll
        for (int k = 0; k < 3; k++) z[k] = T.xp[(size_t)D.xv_base + S.indices[D.idx_base + 3 * lt + k]].z;
    }
    return ((z[0] + z[1]) + z[2]) / 3.0f;
}

// Key that orders z descending under OrderedFloat (src/tilerasterizer.rs:
// 93-94: b.avg_z.cmp(&a.avg_z)): every NaN first (greatest), -0 == +0.
__device__ __forceinline__ uint32_t z_desc_key(float z) {
    if (z != z) return 0u;
    if (z == 0.0f) z = 0.0f;
    uint32_t b = __float_as_uint(z);
    uint32_t asc = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ~asc;
}

// Ascending bitonic sort of v[0..P), P a power of two, by the whole workgroup.
__device__ void bitonic_sort(unsigned long long *v, int P) {
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    unsigned long long a = v[i], b = v[ixj];
                    if ((a > b) == ((i & k) == 0)) {
                        v[i] = b;
                        v[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
}

// Ascending heap sort of v[0..n) by one thread.
__device__ void heap_sort(unsigned long long *v, int n) {
    auto sift = [&](int root, int end) {
        for (;;) {
            int c = 2 * root + 1;
            if (c >= end) return;
            if (c + 1 < end && v[c] < v[c + 1]) c++;
            if (!(v[root] < v[c])) return;
            unsigned long long t = v[root];
            v[root] = v[c];
            v[c] = t;
            root = c;
        }
    };
    for (int r = n / 2 - 1; r >= 0; r--) sift(r, n);
    for (int end = n - 1; end > 0; end--) {
        unsigned long long t = v[0];
        v[0] = v[end];
        v[end] = t;
        sift(0, end);
    }
}

// BumpQueue::sort_by (src/bumpqueue.rs:155-200) by one thread: Lomuto
// quicksort with the rightmost element as pivot, an element moving left when
// it compares Less, i.e. its z key is below the pivot's (its avg_z above).
// The two parts of a partition are disjoint, so their order of processing
// does not change the result: the smaller is sorted first and the larger
// stacked, which bounds the stack by log2(n) ranges.  Keys in the low 32 bits.
__device__ void lomuto_sort(unsigned long long *v, int n, int *stack) {
    int sp = 0, lo = 0, hi = n - 1;
    for (;;) {
        while (lo < hi) {
            uint32_t pivot = (uint32_t)v[hi];
            int i = lo;
            for (int j = lo; j < hi; j++) {
                unsigned long long x = v[j];
                if ((uint32_t)x < pivot) {
                    if (i != j) {
                        v[j] = v[i];
                        v[i] = x;
                    }
                    i++;
                }
            }
            if (i != hi) {
                unsigned long long t = v[i];
                v[i] = v[hi];
                v[hi] = t;
            }
            if (i - lo < hi - i) {
                stack[sp++] = i + 1;
                stack[sp++] = hi;
                hi = i - 1;
            } else {
                stack[sp++] = lo;
                stack[sp++] = i - 1;
                lo = i + 1;
            }
        }
        if (sp == 0) break;
        hi = stack[--sp];
        lo = stack[--sp];
    }
}

// k_tr_sort: one workgroup per 64-px tile orders its list back to front as
// the reference does.  With no two different triangles on the same key every
// correct sort gives the reference's order, and a bitonic sort of
// (z key, id) in LDS is final.  Otherwise Lomuto's result depends on the
// starting order, which is the canonical submission order (ascending record
// id, the copies of a wrapped entry adjacent and identical): the list is put
// in that order and quicksorted as the reference does.  Lists longer than
// the LDS holds are sorted in global scratch by one thread.
#define TR_SORT_MAX 4096
__global__ void __launch_bounds__(256) k_tr_sort(const uint32_t *tr_start, const uint32_t *tr_end, uint32_t *list,
                                                 uint32_t cap, unsigned long long *scratch, const float4 *raster,
                                                 TriSource src) {
    __shared__ unsigned long long sk[TR_SORT_MAX];
    __shared__ int stack[2 * 40];
    __shared__ int serial;
    uint32_t slot = blockIdx.x;
    uint32_t s = tr_start[slot], e = min(tr_end[slot], cap);
    if (e <= s + 1) return;
    int n = (int)(e - s);
    if (n <= TR_SORT_MAX) {
        int P = 1;
        while (P < n) P <<= 1;
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
            unsigned long long key = ~0ull;
            if (i < n) {
                uint32_t id = list[s + i];
                key = ((unsigned long long)z_desc_key(tri_avg_z(src, raster + 4 * (size_t)id)) << 32) | id;
            }
            sk[i] = key;
        }
        if (threadIdx.x == 0) serial = 0;
        __syncthreads();
        bitonic_sort(sk, P);
        for (int i = threadIdx.x; i + 1 < n; i += blockDim.x)
            if ((sk[i] >> 32) == (sk[i + 1] >> 32) && (uint32_t)sk[i] != (uint32_t)sk[i + 1]) serial = 1;
        __syncthreads();
        if (!serial) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) list[s + i] = (uint32_t)sk[i];
            return;
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            unsigned long long x = sk[i];
            sk[i] = (x << 32) | (x >> 32);          // (id, z key)
        }
        __syncthreads();
        bitonic_sort(sk, P);
        if (threadIdx.x == 0) lomuto_sort(sk, n, stack);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) list[s + i] = (uint32_t)(sk[i] >> 32);
        return;
    }
    unsigned long long *v = scratch + s;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t id = list[s + i];
        v[i] = ((unsigned long long)id << 32) | z_desc_key(tri_avg_z(src, raster + 4 * (size_t)id));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        heap_sort(v, n);
        lomuto_sort(v, n, stack);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) list[s + i] = (uint32_t)(v[i] >> 32);
}

struct TrArgs {
    const SR_GLOBAL uint32_t *start;
    const SR_GLOBAL uint32_t *end;
    const SR_GLOBAL uint32_t *list;
    uint32_t cap;
};

// Staged entry fields for k_translucent (one 256-wide row each).
enum {
    G_A = 0, G_B = 3, G_C = 6, G_SX = 9, G_SY = 12,
    G_OOA = 15, G_OA, G_ODA, G_ODB, G_ZA, G_ZDA, G_ZDB, G_ID, G_ORG, G_LIM, G_COARSE, G_COUNT
};

__device__ __forceinline__ int find_view_tile(const RasterArgs &P, uint32_t slot) {
    int lo = 0, hi = P.num_views - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (P.views[mid].lum_offset <= slot) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// k_translucent: the forward pass of each 64-px tile's sorted list
// (src/tilerasterizer.rs:96-101 with TranslucentForwardShader,
// src/shader.rs:66-76).  Sixteen workgroups per tile, each a 16x16 pixel
// block with one thread per pixel in 2x2-quad order, walk the list in order:
// the reference's quad walk and coverage of the entry (the same walk as the
// opaque pass), the depth test against the opaque depth without writing it,
// and on passing samples the translucent shader over the current colour.
// A pixel's colour depends only on its own samples and its quad's
// barycentrics, so threads never wait on each other.  The block holding the
// tile's centre quad also updates it (on screen or not) and the tile
// luminance.
__device__ __forceinline__ void translucent_block(const RasterArgs &P, const TrArgs &Q, uint32_t *st) {
    uint32_t slot = blockIdx.x >> 4, sub = blockIdx.x & 15u;
    uint32_t s = Q.start[slot], e = min(Q.end[slot], Q.cap);
    if (e <= s) return;
    const DevView &V = P.views[find_view_tile(P, slot)];
    const DevScene &S = *V.scene;
    uint32_t ri = slot - V.lum_offset;
    int rx0 = (int)(ri % (uint32_t)V.ref_tiles_x) * SR_REF_TILE, ry0 = (int)(ri / (uint32_t)V.ref_tiles_x) * SR_REF_TILE;
    int bx0 = rx0 + 16 * (int)(sub & 3u), by0 = ry0 + 16 * (int)(sub >> 2);
    bool centre_block = sub == 8u;            // holds the quad at (0, 32) of the tile
    if (!(bx0 < V.W && by0 < V.H) && !centre_block) return;
    const int tid = threadIdx.x, lane = tid & 63, qbase = lane & ~3;
    const int q = tid >> 2, sl = tid & 3, sx = sl & 1, sy = sl >> 1;
    const int qpx = bx0 + 2 * (q & 7), qpy = by0 + 2 * (q >> 3);
    const int px = qpx + sx, py = qpy + sy;
    const bool on_screen = px < V.W && py < V.H;
    const bool centre = centre_block && q == 0;
    const bool valid = on_screen || centre;
    size_t o = V.fb_offset + (size_t)py * (size_t)V.W + (size_t)px;
    float4 *cc = P.center_color + 4 * (size_t)slot + sl;
    f3 cur = mk3(0.0f, 0.0f, 0.0f);
    float depth = INFINITY;
    if (centre) {
        float4 c = *cc;
        cur = mk3(c.x, c.y, c.z);
        depth = c.w;
    } else if (on_screen) {
        cur = mk3(P.out_color[o], P.out_color[P.color_plane + o], P.out_color[2 * P.color_plane + o]);
        depth = P.out_depth[o];
    }
    for (uint32_t b0 = s; b0 < e; b0 += 256u) {
        uint32_t nb = min(e - b0, 256u);
        __syncthreads();
        if ((uint32_t)tid < nb) {
            uint32_t id = Q.list[b0 + (uint32_t)tid];
            TriSetup ts;
            tri_setup(P.geo[id], P.raster + 4 * (size_t)id, rx0, ry0, ts);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                st[(G_A + k) * 256 + tid] = __float_as_uint(ts.A[k]);
                st[(G_B + k) * 256 + tid] = __float_as_uint(ts.B[k]);
                st[(G_C + k) * 256 + tid] = __float_as_uint(ts.C[k]);
                st[(G_SX + k) * 256 + tid] = __float_as_uint(ts.SX[k]);
                st[(G_SY + k) * 256 + tid] = __float_as_uint(ts.SY[k]);
            }
            st[G_OOA * 256 + tid] = __float_as_uint(ts.ooa);
            st[G_OA * 256 + tid] = __float_as_uint(ts.oa);
            st[G_ODA * 256 + tid] = __float_as_uint(ts.oda);
            st[G_ODB * 256 + tid] = __float_as_uint(ts.odb);
            st[G_ZA * 256 + tid] = __float_as_uint(ts.za);
            st[G_ZDA * 256 + tid] = __float_as_uint(ts.zda);
            st[G_ZDB * 256 + tid] = __float_as_uint(ts.zdb);
            st[G_ID * 256 + tid] = id;
            st[G_ORG * 256 + tid] = (uint32_t)ts.ox0 | ((uint32_t)ts.oy0 << 16);
            st[G_LIM * 256 + tid] = (uint32_t)ts.cmxx | ((uint32_t)ts.cmxy << 16);
            st[G_COARSE * 256 + tid] = ts.coarse ? 1u : 0u;
        }
        __syncthreads();
        for (uint32_t k = 0; k < nb; k++) {
#define GF(f) __uint_as_float(st[(f) * 256 + k])
            uint32_t org = st[G_ORG * 256 + k], lim = st[G_LIM * 256 + k];
            int ox0 = (int)(org & 0xFFFFu), oy0 = (int)(org >> 16), cmxx = (int)(lim & 0xFFFFu), cmxy = (int)(lim >> 16);
            if (!(max(ox0, bx0) < min(cmxx, bx0 + 16) && max(oy0, by0) < min(cmxy, by0 + 16))) continue;
            bool fm = false;
            float b1 = 0.0f, b2 = 0.0f;
            if (qpx >= ox0 && qpx < cmxx && qpy >= oy0 && qpy < cmxy) {
                TriSetup ts;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    ts.A[m] = GF(G_A + m);
                    ts.B[m] = GF(G_B + m);
                    ts.C[m] = GF(G_C + m);
                    ts.SX[m] = GF(G_SX + m);
                    ts.SY[m] = GF(G_SY + m);
                }
                bool coarse = st[G_COARSE * 256 + k] != 0u;
                int bx = ox0, by = oy0;
                if (coarse) {
                    bx = ox0 + ((qpx - ox0) & ~15);
                    by = oy0 + ((qpy - oy0) & ~15);
                }
                if (!(coarse && block_out(ts, bx, by))) {
                    int i = (qpx - bx) >> 1, j = (qpy - by) >> 1;
                    float w0 = edge_chain(ts, 0, bx, by, sx, sy, i, j);
                    float w1 = edge_chain(ts, 1, bx, by, sx, sy, i, j);
                    float w2 = edge_chain(ts, 2, bx, by, sx, sy, i, j);
                    float ooa = GF(G_OOA);
                    b1 = w1 * ooa;
                    b2 = w2 * ooa;
                    float wv = 1.0f / ((GF(G_OA) + b1 * GF(G_ODA)) + b2 * GF(G_ODB));
                    float z = ((GF(G_ZA) + b1 * GF(G_ZDA)) + b2 * GF(G_ZDB)) * wv;
                    fm = valid && w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f && z <= depth;
                }
            }
#undef GF
            if (__ballot(fm)) {
                float qb1[4], qb2[4];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    qb1[m] = __shfl(b1, qbase + m, 64);
                    qb2[m] = __shfl(b2, qbase + m, 64);
                }
                if (fm) {
                    uint32_t id = st[G_ID * 256 + k];
                    const float4 *R = P.raster + 4 * (size_t)id;
                    CV tv[3];
                    float4 Sh[7], Tn[3];
                    load_tri_cv(P.src, R, tv);
                    bool nmap = (P.geo[id].z & SR_GEO_NORMAL_MAP) != 0;
                    tri_shade_attrs(tv, R, nmap, Sh, Tn);
                    float wv[4];
                    quad_w(__uint_as_float(st[G_OA * 256 + k]), __uint_as_float(st[G_ODA * 256 + k]),
                           __uint_as_float(st[G_ODB * 256 + k]), qb1, qb2, wv);
                    cur = shade_pixel<true>(S, V, __float_as_int(R[3].y), Sh, Tn, nmap, qb1[sl], qb2[sl], wv, sl,
                                            cur);
                }
            }
        }
    }
    if (on_screen) {
        P.out_color[o] = cur.x;
        P.out_color[P.color_plane + o] = cur.y;
        P.out_color[2 * P.color_plane + o] = cur.z;
    }
    if (centre) *cc = make_float4(cur.x, cur.y, cur.z, depth);
    // centre-quad luminance after the translucent pass (src/tilerasterizer.rs:103-106)
    float lum = (cur.x * 0.2126f + cur.y * 0.7152f) + cur.z * 0.0722f;
    float l0 = __shfl(lum, qbase, 64), l1 = __shfl(lum, qbase + 1, 64);
    float l2 = __shfl(lum, qbase + 2, 64), l3 = __shfl(lum, qbase + 3, 64);
    if (centre && sl == 0) P.out_lum[slot] = (((l0 + l1) + l2) + l3) * 0.25f;
}

__global__ void __launch_bounds__(256) k_translucent(RasterArgs P, TrArgs Q) {
    __shared__ uint32_t st[G_COUNT * 256];
    translucent_block(P, Q, st);
    clock_end(P.clock, CLOCK_TR_END);
}

// ---------------------------------------------------------------------------
// k_blit: exposure, tonemap, quantize, pack (src/renderer.rs:293-355).  With
// frame_bytes the word is stored byte-swapped, so memory holds r, g, b, a per
// pixel: the presented frame (rgba8_unpack_u8 into the pixel bytes,
// src/main.rs:574-590) without a host pass.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_blit(const float *color, uint64_t plane, int W, int H, float exposure,
                                              const float *exposure_dev, uint32_t *out, int frame_bytes) {
    if (exposure_dev) exposure = exposure_dev[2];
    uint64_t n = (uint64_t)W * (uint64_t)(H & ~1);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float k = 0.2f;
        float c[3];
        for (int ch = 0; ch < 3; ch++) {
            float x = color[ch * plane + i] * exposure;
            c[ch] = (x / (x + k)) * (1.0f + k);
        }
        uint32_t v = (sat_u8(c[0] * 255.0f) << 24) | (sat_u8(c[1] * 255.0f) << 16) | (sat_u8(c[2] * 255.0f) << 8) | 0xFFu;
        out[i] = frame_bytes ? __builtin_bswap32(v) : v;
    }
}

// Renderer::update_auto_exposure (src/renderer.rs:258-290) on the device: one
// workgroup sorts the tiles' log luminance in LDS (bitonic, on the order of
// f32::total_cmp), one lane sums the trimmed middle in sorted order and
// steps the exposure.  state = {ev, target, exposure}; with use_host the
// step starts from (host_ev, host_target) instead of the stored state.
// The device's log2f / expf / powf may differ from the host libm's by an
// ulp, so the exposure agrees with the host step to a few ulps.
__global__ void __launch_bounds__(1024) k_exposure(const float *lum, uint32_t n, float dt, int use_host, float host_ev,
                                                   float host_target, float *state) {
    __shared__ int32_t key[SR_EXPOSURE_MAX];
    float ev = use_host ? host_ev : state[0];
    float target = use_host ? host_target : state[1];
    float dtc = maxps(dt, 0.0f);
    float tau = maxps(1.0f, 1e-4f);
    float alpha = 1.0f - expf(-(dtc / tau));
    if (alpha == 0.0f) {    // a zero step: the exposure is 2^ev whatever the target
        if (threadIdx.x == 0) {
            state[0] = ev;
            state[1] = target;
            state[2] = powf(2.0f, ev);
        }
        return;
    }
    uint32_t m = 1;
    while (m < n) m <<= 1;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        int32_t k = 0x7FFFFFFF;
        if (i < n) {
            float sv = log2f(maxps(lum[i], 1e-4f));
            int32_t b = __float_as_int(sv);
            k = b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
        }
        key[i] = k;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
                uint32_t l = i ^ j;
                if (l > i) {
                    int32_t a = key[i], b = key[l];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        key[i] = b;
                        key[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    if (threadIdx.x == 0) {
        uint32_t trim = sat_u32(floorf((float)n * 0.10f));
        trim = min(trim, (n - 1) / 2);
        float sum = 0.0f;
        for (uint32_t i = trim; i < n - trim; i++) {
            int32_t k = key[i];
            sum += __int_as_float(k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1));
        }
        float mean = sum / (float)(n - 2 * trim);
        float y = minps(maxps(0.45f, 0.0f), 1.0f);
        float kk = 0.2f;
        float denom = maxps(1.0f + kk - y, 1e-6f);
        float meter = maxps((y * kk) / denom, 1e-4f);
        float tev = log2f(meter) - mean;
        float t = minps(maxps(powf(2.0f, tev), 0.05f), 32.0f);
        float nev = ev + (log2f(t) - ev) * alpha;
        state[0] = nev;
        state[1] = t;
        state[2] = powf(2.0f, nev);
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

int bin_sort_bits(uint32_t bins) {
    int nbits = 1;
    while (nbits < 32 && (1ull << nbits) < (unsigned long long)bins) nbits++;
    return nbits;
}

int bin_sort_passes(uint32_t bins) {
    int n = (bin_sort_bits(bins) + 7) / 8;
    return n <= 3 ? n : 0;
}

// Workgroups per chunk in the expansion: about one per 1024 pairs a chunk
// of 256 triangles holds, judged from the capacities (the pair capacity
// follows the frames' pair counts).
static uint32_t bin_expand_parts(uint32_t rec_cap, uint32_t pair_cap) {
    uint64_t per_chunk = (uint64_t)pair_cap * 256u / std::max(rec_cap, 1u);
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(per_chunk / 1024u, 1u), BS_EXPAND_PARTS);
}

static uint32_t bin_sort_items(uint32_t pair_cap) { return pair_cap <= BS_SMALL_PAIRS ? 4u : 16u; }

uint32_t bin_sort_tiles(uint32_t pair_cap) {
    uint32_t te = 256u * bin_sort_items(pair_cap);
    return std::max((uint32_t)(((uint64_t)pair_cap + te - 1) / te), 1u);
}

// A launch that stamps the stage-timing events from its own dispatch (start
// when the kernel begins, stop when it ends) where the frame asks for them:
// a separate event record between two kernels holds the queue for several
// microseconds.
#define SR_LAUNCH_TIMED(e0, e1, kern, grid, block, st, ...)                                      \
    do {                                                                                        \
        if ((e0) || (e1)) hipExtLaunchKernelGGL(kern, grid, block, 0, st, e0, e1, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(kern, grid, block, 0, st, __VA_ARGS__);                          \
    } while (0)

hipError_t launch_frame(const FrameArgs &a, hipStream_t st) {
    const uint32_t GRID_CAP = 256 * 16;
    FrameBegin fb;
    fb.bin_start = a.bin_start;
    fb.bin_end = a.bin_end;
    fb.items = a.total_items;
    fb.tr_count = a.has_translucent ? a.tr_count : nullptr;
    fb.tr_slots = a.tr_slots;
    fb.clock = a.clock;
    uint32_t tri_blocks = a.total_in > 0 ? grid_for(a.total_in, 256, GRID_CAP) : 0u;
    uint32_t vtx_blocks = a.total_xv > 0 ? grid_for(a.total_xv, 256, GRID_CAP) : 0u;
    hipEvent_t ev_none = nullptr;
    if (tri_blocks + vtx_blocks > 0)
        SR_LAUNCH_TIMED(a.events[0], ev_none, k_front, dim3(tri_blocks + vtx_blocks), dim3(256), st, a.views, a.draws,
                        a.num_draws, a.total_in, a.total_xv, a.xv, a.xv + a.total_xv, a.any_tangents ? a.xt : nullptr,
                        a.xv + 3 * (size_t)a.total_xv, a.tri_counts, a.tri_chunks, a.clip_queue, a.clip_cap,
                        a.status, fb, tri_blocks);
    else
        SR_LAUNCH_TIMED(a.events[0], ev_none, k_frame_begin, dim3(grid_for(a.total_items, 256, GRID_CAP)), dim3(256),
                        st, fb);
    if (a.total_in > 0) {
        hipLaunchKernelGGL(k_clip, dim3(512), dim3(256), 0, st, a.views, a.draws, a.num_draws, a.xv, a.xv + a.total_xv,
                           a.any_tangents ? a.xt : nullptr, a.tri_counts, a.tri_chunks, a.clip_queue, a.clip_cap,
                           a.clip_buf, a.status);
    }
    // first post-clip id of every 256-triangle chunk; the total is also
    // tri_base[total_in], the end of the last view's ids (k_setup writes it
    // too).  One workgroup up to SCAN_SINGLE_MAX chunks, else the
    // three-pass scan in place.
    uint32_t nchunks = (a.total_in + 255u) / 256u;
    hipError_t e;
    if (nchunks <= SCAN_SINGLE_MAX) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, a.tri_chunks, nchunks, &a.status->total_records,
                           a.tri_base + a.total_in);
        e = hipGetLastError();
    } else {
        e = launch_scan(a.tri_chunks, nchunks, a.tri_chunks, nullptr, a.scan_partials, &a.status->total_records, st);
    }
    if (e != hipSuccess) return e;
    SetupOut so;
    so.geo = (SR_GLOBAL uint4 *)a.rec_geo; so.raster = (SR_GLOBAL float4 *)a.rec_raster;
    so.cap = a.rec_cap; so.status = (SR_GLOBAL DevStatus *)a.status;
    so.T = a.tile_size;
    so.chunk_pairs = (SR_GLOBAL uint32_t *)a.chunk_pairs;
    if (a.total_in > 0) {
        SR_LAUNCH_TIMED(ev_none, a.events[1], k_setup, dim3(grid_for(a.total_in, 256, GRID_CAP)), dim3(256), st,
                        a.views, a.draws, a.num_draws, a.total_in, a.xv + 3 * (size_t)a.total_xv, a.tri_counts,
                        a.tri_chunks, a.clip_buf, a.tri_base, so);
    } else if (a.events[1]) {
        hipEventRecord(a.events[1], st);
    }
    // the chunks' first pairs (their total is the frame's pair count), then
    // the stable binning and the lists' bounds; the scan's start is the end
    // of setup on the stage clock
    uint64_t *setup_end = a.clock ? &a.clock->start[CLOCK_SETUP_END] : nullptr;
    const int npass = bin_sort_passes(a.total_items);
    if (npass < 1) return hipErrorInvalidValue;
    const int nbits = bin_sort_bits(a.total_items);
    if (a.total_in > 0) {
        if (nchunks <= SCAN_SINGLE_MAX) {
            ScanJobs J;
            J.job[0] = scan_job(a.chunk_pairs, nchunks, a.chunk_pairs, nullptr, &a.sort->chunk_total,
                                &a.status->total_pairs, 0u);
            J.job[1] = J.job[0];
            J.stamp = setup_end;
            hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, st, J);
            e = hipGetLastError();
        } else {
            e = launch_scan(a.chunk_pairs, nchunks, a.chunk_pairs, nullptr, a.scan_partials, &a.sort->chunk_total, st,
                            0u, &a.status->total_pairs, setup_end);
        }
        if (e != hipSuccess) return e;
        BinPass bp = {};
        bp.views = (const SR_GLOBAL DevView *)a.views;
        bp.geo = (const SR_GLOBAL uint4 *)a.rec_geo;
        bp.chunk_first = (const SR_GLOBAL uint32_t *)a.tri_chunks;
        bp.chunk_pairs = (const SR_GLOBAL uint32_t *)a.chunk_pairs;
        bp.nchunks = nchunks;
        bp.list = (SR_GLOBAL uint32_t *)a.list;
        bp.hist = (SR_GLOBAL uint32_t *)a.sort_hist;
        bp.hist_cols = bin_sort_tiles(a.pair_cap);
        bp.tile_elems = 256u * bin_sort_items(a.pair_cap);
        bp.status = (SR_GLOBAL DevStatus *)a.status;
        bp.bs = (const SR_GLOBAL BinSortState *)a.sort;
        bp.total_items = a.total_items;
        bp.rec_cap = a.rec_cap;
        bp.pair_cap = a.pair_cap;
        bp.T = a.tile_size;
        // expansion into E (pair_keys/vals); the radix passes go E -> A ->
        // E ..., the last one's record ids into the lists and its bins into
        // the key array it did not read
        uint32_t *keys[2] = {a.pair_keys, a.pair_keys2}, *vals[2] = {a.pair_vals, a.pair_vals2};
        bp.out_keys = (SR_GLOBAL uint32_t *)keys[0];
        bp.out_vals = (SR_GLOBAL uint32_t *)vals[0];
        bp.parts = bin_expand_parts(a.rec_cap, a.pair_cap);
        hipLaunchKernelGGL(k_bin_expand, dim3(std::min(nchunks * bp.parts, 65536u)), dim3(256), 0, st, bp);
        for (int p = 0; p < npass; p++) {
            bp.pass = p;
            bp.last = p == npass - 1;
            bp.shift = 8 * p;
            bp.bits = std::min(8, nbits - 8 * p);
            bp.dtotal = (SR_GLOBAL uint32_t *)&a.sort->digit_total[p][0];
            bp.in_keys = (const SR_GLOBAL uint32_t *)keys[p & 1];
            bp.in_vals = (const SR_GLOBAL uint32_t *)vals[p & 1];
            bp.out_keys = (SR_GLOBAL uint32_t *)keys[(p + 1) & 1];
            bp.out_vals = (SR_GLOBAL uint32_t *)(bp.last ? nullptr : vals[(p + 1) & 1]);
            if (bp.tile_elems == 1024u) hipLaunchKernelGGL(k_bin_hist<4>, dim3(bp.hist_cols), dim3(256), 0, st, bp);
            else hipLaunchKernelGGL(k_bin_hist<16>, dim3(bp.hist_cols), dim3(256), 0, st, bp);
            hipLaunchKernelGGL(k_bin_rowscan, dim3(256), dim3(256), 0, st, bp);
            if (bp.tile_elems == 1024u)
                hipLaunchKernelGGL(k_bin_scatter<4>, dim3(bp.hist_cols + 8u), dim3(256), 0, st, bp);
            else hipLaunchKernelGGL(k_bin_scatter<16>, dim3(bp.hist_cols + 8u), dim3(256), 0, st, bp);
        }
        BinBounds bb;
        bb.key = (const SR_GLOBAL uint32_t *)keys[npass & 1];
        bb.status = (const SR_GLOBAL DevStatus *)a.status;
        bb.pair_cap = a.pair_cap;
        bb.items = a.total_items;
        bb.bin_start = (SR_GLOBAL uint32_t *)a.bin_start;
        bb.bin_end = (SR_GLOBAL uint32_t *)a.bin_end;
        hipLaunchKernelGGL(k_bin_bounds, dim3(grid_for(a.pair_cap, 256, GRID_CAP)), dim3(256), 0, st, bb);
    }
    TriSource src;
    src.views = (const SR_GLOBAL DevView *)a.views;
    src.draws = (const SR_GLOBAL DevDraw *)a.draws;
    src.xp = (const SR_GLOBAL float4 *)a.xv;
    src.xa = (const SR_GLOBAL float4 *)(a.xv + a.total_xv);
    src.xt = (const SR_GLOBAL float4 *)(a.any_tangents ? a.xt : nullptr);
    src.xs = (const SR_GLOBAL float4 *)(a.xv + 3 * (size_t)a.total_xv);
    src.clip_buf = (const SR_GLOBAL float *)a.clip_buf;
    if (a.has_translucent) {
        uint32_t g = grid_for(a.rec_cap, 256, GRID_CAP);
        hipLaunchKernelGGL(k_tr_bins, dim3(g), dim3(256), 0, st, a.views, a.rec_geo, a.rec_cap, a.tr_count, nullptr,
                           0u, a.status);
        e = launch_scan(a.tr_count, a.tr_slots, a.tr_start, a.tr_end, a.scan_partials, &a.status->total_tr_pairs, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_tr_bins, dim3(g), dim3(256), 0, st, a.views, a.rec_geo, a.rec_cap, a.tr_end, a.tr_list,
                           a.tr_cap, a.status);
        hipLaunchKernelGGL(k_tr_sort, dim3(a.tr_slots), dim3(256), 0, st, a.tr_start, a.tr_end, a.tr_list, a.tr_cap,
                           a.tr_scratch, a.rec_raster, src);
    }
    RasterArgs P;
    P.views = (const SR_GLOBAL DevView *)a.views; P.num_views = a.num_views; P.total_items = a.total_items;
    P.bin_start = (const SR_GLOBAL uint32_t *)a.bin_start; P.bin_end = (const SR_GLOBAL uint32_t *)a.bin_end;
    P.list = (const SR_GLOBAL uint32_t *)a.list; P.pair_cap = a.pair_cap; P.rec_cap = a.rec_cap;
    P.geo = (const SR_GLOBAL uint4 *)a.rec_geo; P.raster = (const SR_GLOBAL float4 *)a.rec_raster;
    P.src = src;
    P.tri_base = (const SR_GLOBAL uint32_t *)a.tri_base; P.out_ids = (SR_GLOBAL uint32_t *)a.out_ids;
    P.out_depth = (SR_GLOBAL float *)a.out_depth; P.out_color = (SR_GLOBAL float *)a.out_color;
    P.color_plane = a.color_plane; P.out_rgba = (SR_GLOBAL uint32_t *)a.out_rgba;
    P.out_lum = (SR_GLOBAL float *)a.out_lum; P.status = (SR_GLOBAL DevStatus *)a.status;
    P.status_out = a.status_out;
    P.clock = a.clock;
    P.center_ids = (SR_GLOBAL uint32_t *)a.center_ids;
    P.center_color = (SR_GLOBAL float4 *)(a.has_translucent ? a.center_color : nullptr);
    P.persistent = a.persistent_blocks > 0;
    uint32_t grid = P.persistent ? (uint32_t)a.persistent_blocks : xcd_grid(a.total_items);
    // stage events: bins done = the raster's start, the raster's end, and
    // the frame's end = the shading's end (or the translucent pass's)
    hipEvent_t e2 = a.events[2], e4 = a.events[4], e3 = a.has_translucent ? ev_none : a.events[3];
    switch (a.tile_size) {
        case 8: SR_LAUNCH_TIMED(e2, e4, (k_raster<8, 64>), dim3(grid), dim3(64), st, P); break;
        case 32: SR_LAUNCH_TIMED(e2, e4, (k_raster<32, 512>), dim3(grid), dim3(512), st, P); break;
        default:
            if (a.hiz) SR_LAUNCH_TIMED(e2, e4, (k_raster<16, 256, true>), dim3(grid), dim3(256), st, P);
            else SR_LAUNCH_TIMED(e2, e4, (k_raster<16, 256>), dim3(grid), dim3(256), st, P);
            break;
    }
    switch (a.tile_size) {
        case 8: SR_LAUNCH_TIMED(ev_none, e3, (k_shade<8, 64>), dim3(xcd_grid(a.total_items)), dim3(64), st, P); break;
        case 32:
            SR_LAUNCH_TIMED(ev_none, e3, (k_shade<32, 1024>), dim3(xcd_grid(a.total_items)), dim3(1024), st, P);
            break;
        default:
            SR_LAUNCH_TIMED(ev_none, e3, (k_shade<16, 256>), dim3(xcd_grid(a.total_items)), dim3(256), st, P);
            break;
    }
    if (a.has_translucent) {
        TrArgs Q;
        Q.start = (const SR_GLOBAL uint32_t *)a.tr_start; Q.end = (const SR_GLOBAL uint32_t *)a.tr_end;
        Q.list = (const SR_GLOBAL uint32_t *)a.tr_list; Q.cap = a.tr_cap;
        SR_LAUNCH_TIMED(ev_none, a.events[3], k_translucent, dim3(a.tr_slots * 16u), dim3(256), st, P, Q);
    }
    return hipGetLastError();
}

hipError_t launch_blit(const float *color, uint64_t plane, int32_t W, int32_t H, float exposure, uint32_t *out,
                       hipStream_t st, const float *exposure_dev, bool frame_bytes) {
    uint64_t n = (uint64_t)W * (uint64_t)(H & ~1);
    hipLaunchKernelGGL(k_blit, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, st, color, plane, W, H, exposure,
                       exposure_dev, out, frame_bytes ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_exposure(const float *lum, uint32_t n, float dt, bool use_host, float ev, float target, float *state,
                           hipStream_t st) {
    if (n == 0 || n > SR_EXPOSURE_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_exposure, dim3(1), dim3(1024), 0, st, lum, n, dt, use_host ? 1 : 0, ev, target, state);
    return hipGetLastError();
}

}  // namespace sr
