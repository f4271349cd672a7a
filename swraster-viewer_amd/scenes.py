"""Synthetic, seeded scenes for the benchmark configurations.

Each generator returns ``(SceneData, camera_spec)``.  camera_spec is a dict
``{"position", "yaw", "pitch", "fov", "width", "height", "far"}`` in the
reference's camera convention (src/rendercamera.rs:126-135: the view rotation
is roll * pitch * yaw, so a positive pitch field looks *down*).

  C1  16 icospheres (subdivision 3: 642 vertices / 1,280 triangles each),
      centres U[-4,4]^3, radii U[0.5,1.0], spherical UVs, one material per
      sphere (baseColor U[0,1]^3, metallic/roughness U[0,1]), one shared
      256^2 RGBA8 uniform-noise base-colour texture; camera (0,0,12) looking
      at the origin, 60 deg, 1280x720, far = 2 x scene diagonal.
  C2  724 x 724 heightfield over [-50,50]^2 (1,045,458 triangles), heights
      5 x 4-octave value noise, one 2048^2 value-noise texture, metallic 0,
      roughness 0.5; camera at (0,20,60) pitched 30 deg down, 60 deg,
      1920x1080, far = 2 x diagonal.
  C3  100,000 independent front-facing triangles at 3840x2160: screen-space
      centroids uniform over the viewport, two edges with log-uniform lengths
      in [1,400] px and an opening angle U[0.1pi,0.9pi], random orientation,
      view depth U[1,100] with +-5% per-vertex jitter, unprojected through a
      camera at the origin (60 deg, far 150).  Untextured; the triangles are
      split over 1,024 primitives with one random material each (materials
      are per primitive in the reference, so colour varies per primitive).
  C4  256 views over a pool of heightfield scenes built as C2's (same terrain
      seed, one 2048^2 texture each) with triangle counts log-uniform in
      10k-2M; each view picks a scene, a resolution uniformly from
      {320x240, 640x480, 1280x720, 1920x1080} and an orbit camera aimed at
      the terrain centre (azimuth U[0, 2pi), distance U[45, 75], height
      U[12, 30]).  The scenes share the terrain function, so the voxel grid
      built for C2 lights all of them.
  C5  eight resident scenes mixing the three configurations, 12.3M
      triangles and 240 distinct 2048^2 RGBA8 uniform-noise textures
      (4.03 GB of base levels): icosphere clusters of 64 and 256 spheres at
      subdivision 4 (one textured material per sphere, 48 textures each),
      heightfields of 724^2, 1448^2 and 1535^2 vertices cut into 8 x 6
      patches (one textured material per patch), a 512^2 untextured
      heightfield and two 100,000-triangle overdraw scenes (C3's generator,
      seeds 3 and 31).
"""
import json
import os
import struct

import numpy as np

from . import (SR_ALPHA_OPAQUE, SR_TEX_SRGB, SR_WRAP_REPEAT, SceneData, pack_rgba)


def camera_spec(position, yaw, pitch, fov, width, height, far):
    return {"position": [float(v) for v in position], "yaw": float(yaw), "pitch": float(pitch),
            "fov": float(fov), "width": int(width), "height": int(height), "far": float(far)}


def make_camera(spec, width=None, height=None):
    from . import RenderCamera
    w = width or spec["width"]
    h = height or spec["height"]
    cam = RenderCamera.from_pose(spec["position"], spec["yaw"], spec["pitch"], spec["fov"], w, h, spec["far"])
    return cam.update_matrices()


def _translate_scale(c, r):
    m = np.eye(4, dtype=np.float32)
    m[0, 0] = m[1, 1] = m[2, 2] = r
    m[:3, 3] = c
    return m


# ---------------------------------------------------------------------------
# C1: icospheres
# ---------------------------------------------------------------------------
def icosphere(subdiv=3):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts, np.float32)
    f = np.array(faces, np.uint32)
    # outward counter-clockwise winding
    p0, p1, p2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(p1 - p0, p2 - p0), p0) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def make_c1(seed=1, width=1280, height=720, spheres=16, subdiv=3, tex_size=256):
    rng = np.random.Generator(np.random.PCG64(seed))
    v, f = icosphere(subdiv)
    uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi), np.arccos(np.clip(v[:, 1], -1, 1)) / np.pi],
                  1).astype(np.float32)
    centres = rng.uniform(-4, 4, (spheres, 3)).astype(np.float32)
    radii = rng.uniform(0.5, 1.0, spheres).astype(np.float32)
    base = rng.uniform(0, 1, (spheres, 3))
    metal = rng.uniform(0, 1, spheres)
    rough = rng.uniform(0, 1, spheres)
    noise = rng.integers(0, 256, (tex_size, tex_size, 3), dtype=np.uint8)
    rgba = np.concatenate([noise, np.full((tex_size, tex_size, 1), 255, np.uint8)], -1)
    tex = (tex_size, tex_size, SR_TEX_SRGB, pack_rgba(rgba).ravel())
    prims, meshes, nodes, mats = [], [], [], []
    for i in range(spheres):
        prims.append({"positions": v, "normals": v, "texcoords": uv, "tangents": None,
                      "indices": f.ravel(), "material": i})
        meshes.append((i, 1))
        nodes.append((_translate_scale(centres[i], radii[i]), i))
        mats.append({"base_color": (base[i, 0], base[i, 1], base[i, 2], 1.0), "metallic": metal[i],
                     "roughness": rough[i], "base_color_tex": (0, SR_WRAP_REPEAT, SR_WRAP_REPEAT),
                     "alpha_mode": SR_ALPHA_OPAQUE})
    scene = SceneData(prims, meshes, nodes, mats, [tex], name=f"c1_s{seed}")
    _, _, _, diag = scene.bounds()
    cam = camera_spec((0, 0, 12), 0.0, 0.0, np.pi / 3, width, height, 2.0 * diag)
    return scene, cam


# ---------------------------------------------------------------------------
# C2: heightfield
# ---------------------------------------------------------------------------
def value_noise(shape, rng, octaves=4, base_cells=8):
    """Sum of `octaves` bilinear-smoothstep value-noise layers in [-1, 1]."""
    h, w = shape
    out = np.zeros(shape, np.float64)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        cells = base_cells * (2 ** o)
        lat = rng.uniform(-1, 1, (cells + 1, cells + 1))
        ys = np.linspace(0, cells, h, endpoint=False) if h > 1 else np.zeros(1)
        xs = np.linspace(0, cells, w, endpoint=False) if w > 1 else np.zeros(1)
        yi, xi = np.floor(ys).astype(int), np.floor(xs).astype(int)
        fy, fx = ys - yi, xs - xi
        sy, sx = fy * fy * (3 - 2 * fy), fx * fx * (3 - 2 * fx)
        a = lat[yi][:, xi]
        b = lat[yi][:, xi + 1]
        c = lat[yi + 1][:, xi]
        d = lat[yi + 1][:, xi + 1]
        top = a + (b - a) * sx[None, :]
        bot = c + (d - c) * sx[None, :]
        out += amp * (top + (bot - top) * sy[:, None])
        total += amp
        amp *= 0.5
    return out / total


def heightfield(n, extent, amplitude, rng):
    hgt = amplitude * value_noise((n, n), rng)
    xs = np.linspace(-extent, extent, n)
    X, Z = np.meshgrid(xs, xs, indexing="ij")  # [i, j] -> x_i, z_j
    pos = np.stack([X, hgt, Z], -1).reshape(-1, 3).astype(np.float32)
    dx = xs[1] - xs[0]
    gx = np.gradient(hgt, dx, axis=0)
    gz = np.gradient(hgt, dx, axis=1)
    nrm = np.stack([-gx, np.ones_like(gx), -gz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    uv = np.stack([(X + extent) / (2 * extent), (Z + extent) / (2 * extent)], -1).reshape(-1, 2).astype(np.float32)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (i * n + j).ravel()
    v01 = (i * n + j + 1).ravel()
    v10 = ((i + 1) * n + j).ravel()
    v11 = ((i + 1) * n + j + 1).ravel()
    idx = np.stack([np.stack([v00, v01, v10], 1), np.stack([v10, v01, v11], 1)], 1).reshape(-1).astype(np.uint32)
    return pos, nrm.reshape(-1, 3).astype(np.float32), uv, idx


def make_c2(seed=2, width=1920, height=1080, grid=724, tex_size=2048):
    rng = np.random.Generator(np.random.PCG64(seed))
    pos, nrm, uv, idx = heightfield(grid, 50.0, 5.0, rng)
    tex = value_noise((tex_size, tex_size), rng, octaves=4, base_cells=16)
    g = ((tex * 0.5 + 0.5) * 255).clip(0, 255).astype(np.uint8)
    tint = np.array([0.55, 0.75, 0.45])
    rgb = (g[..., None] * tint[None, None, :]).clip(0, 255).astype(np.uint8)
    rgba = np.concatenate([rgb, np.full(rgb.shape[:2] + (1,), 255, np.uint8)], -1)
    tex = (tex_size, tex_size, SR_TEX_SRGB, pack_rgba(rgba).ravel())
    prims = [{"positions": pos, "normals": nrm, "texcoords": uv, "tangents": None, "indices": idx, "material": 0}]
    mats = [{"base_color": (1.0, 1.0, 1.0, 1.0), "metallic": 0.0, "roughness": 0.5,
             "base_color_tex": (0, SR_WRAP_REPEAT, SR_WRAP_REPEAT)}]
    scene = SceneData(prims, [(0, 1)], [(np.eye(4, dtype=np.float32), 0)], mats, [tex],
                      name=f"c2_s{seed}_g{grid}")
    _, _, _, diag = scene.bounds()
    cam = camera_spec((0, 20, 60), 0.0, np.pi / 6, np.pi / 3, width, height, 2.0 * diag)
    return scene, cam


# ---------------------------------------------------------------------------
# C3: independent triangles, overdraw
# ---------------------------------------------------------------------------
def make_c3(seed=3, width=3840, height=2160, triangles=100_000, primitives=1024, fov=np.pi / 3, far=150.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = triangles
    cx = rng.uniform(0, width, n)
    cy = rng.uniform(0, height, n)
    e1 = np.exp(rng.uniform(np.log(1.0), np.log(400.0), n))
    e2 = np.exp(rng.uniform(np.log(1.0), np.log(400.0), n))
    phi = rng.uniform(0, 2 * np.pi, n)
    theta = rng.uniform(0.1 * np.pi, 0.9 * np.pi, n)
    a = np.zeros((n, 2))
    b = np.stack([e1 * np.cos(phi), e1 * np.sin(phi)], 1)
    c = np.stack([e2 * np.cos(phi + theta), e2 * np.sin(phi + theta)], 1)
    pts = np.stack([a, b, c], 1)
    pts -= pts.mean(1, keepdims=True)
    pts[:, :, 0] += cx[:, None]
    pts[:, :, 1] += cy[:, None]
    # screen space has y down: front-facing == negative signed area
    d1 = pts[:, 1] - pts[:, 0]
    d2 = pts[:, 2] - pts[:, 0]
    area = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    swap = area > 0
    pts[swap] = pts[swap][:, [0, 2, 1]]
    depth = rng.uniform(1, 100, n)[:, None] * rng.uniform(0.95, 1.05, (n, 3))
    aspect = width / height
    ty = np.tan(fov / 2)
    ndx = 2 * pts[:, :, 0] / width - 1
    ndy = 1 - 2 * pts[:, :, 1] / height
    X = ndx * ty * aspect * depth
    Y = ndy * ty * depth
    Z = -depth
    P = np.stack([X, Y, Z], -1).astype(np.float32)  # (n, 3, 3)
    fn = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    fn /= np.linalg.norm(fn, axis=-1, keepdims=True) + 1e-30
    order = rng.permutation(n)
    groups = np.array_split(order, primitives)
    prims, mats = [], []
    for gi, g in enumerate(groups):
        pos = P[g].reshape(-1, 3)
        nrm = np.repeat(fn[g], 3, axis=0).astype(np.float32)
        prims.append({"positions": pos, "normals": nrm, "texcoords": None, "tangents": None,
                      "indices": np.arange(len(pos), dtype=np.uint32), "material": gi})
        col = rng.uniform(0, 1, 3)
        mats.append({"base_color": (col[0], col[1], col[2], 1.0), "metallic": rng.uniform(0, 1),
                     "roughness": rng.uniform(0.05, 1)})
    scene = SceneData(prims, [(0, len(prims))], [(np.eye(4, dtype=np.float32), 0)], mats, [],
                      name=f"c3_s{seed}_n{n}")
    cam = camera_spec((0, 0, 0), 0.0, 0.0, fov, width, height, far)
    return scene, cam


CONFIGS = {"c1": make_c1, "c2": make_c2, "c3": make_c3}


# ---------------------------------------------------------------------------
# C5: resident scenes of the served configuration
# ---------------------------------------------------------------------------
def noise_texture(rng, size=2048):
    """size^2 RGBA8 uniform noise (opaque), packed, sRGB base colour."""
    t = rng.integers(0, 1 << 32, size * size, dtype=np.uint64).astype(np.uint32) | np.uint32(0xFF)
    return (size, size, SR_TEX_SRGB, t)


def textured_material(rng, tex):
    col = rng.uniform(0.2, 1.0, 3)
    return {"base_color": (col[0], col[1], col[2], 1.0), "metallic": rng.uniform(0, 1),
            "roughness": rng.uniform(0.05, 1), "base_color_tex": (tex, SR_WRAP_REPEAT, SR_WRAP_REPEAT),
            "alpha_mode": SR_ALPHA_OPAQUE}


def make_icosphere_cluster(seed, spheres, subdiv, textures, tex_size=2048, name=None):
    """C1's construction with more spheres: centres U[-4,4]^3, radii
    U[0.5,1], spherical UVs, one material per sphere; materials take the
    `textures` distinct noise textures in turn."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v, f = icosphere(subdiv)
    uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi), np.arccos(np.clip(v[:, 1], -1, 1)) / np.pi],
                  1).astype(np.float32)
    texs = [noise_texture(rng, tex_size) for _ in range(textures)]
    prims, meshes, nodes, mats = [], [], [], []
    for i in range(spheres):
        c = rng.uniform(-4, 4, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.5, 1.0))
        prims.append({"positions": v, "normals": v, "texcoords": uv, "tangents": None,
                      "indices": f.ravel(), "material": i})
        meshes.append((i, 1))
        nodes.append((_translate_scale(c, r), i))
        mats.append(textured_material(rng, i % textures))
    return SceneData(prims, meshes, nodes, mats, texs, name=name or f"c5_ico_s{seed}_n{spheres}")


def make_patched_heightfield(seed, grid, patches=(8, 6), textures=True, tex_size=2048, name=None):
    """C2's terrain (seed 2 heights) at `grid`^2 vertices cut into patches
    (px x pz grid cells each, neighbours sharing their border vertices), one
    material per patch with its own noise texture spanning the patch."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos, nrm, _, _ = heightfield(grid, 50.0, 5.0, np.random.Generator(np.random.PCG64(2)))
    P = pos.reshape(grid, grid, 3)
    N = nrm.reshape(grid, grid, 3)
    cuts_i = np.linspace(0, grid - 1, patches[0] + 1).round().astype(int)
    cuts_j = np.linspace(0, grid - 1, patches[1] + 1).round().astype(int)
    prims, mats, texs = [], [], []
    for a in range(patches[0]):
        for b in range(patches[1]):
            i0, i1, j0, j1 = cuts_i[a], cuts_i[a + 1], cuts_j[b], cuts_j[b + 1]
            ni, nj = i1 - i0 + 1, j1 - j0 + 1
            p = np.ascontiguousarray(P[i0:i1 + 1, j0:j1 + 1].reshape(-1, 3))
            n = np.ascontiguousarray(N[i0:i1 + 1, j0:j1 + 1].reshape(-1, 3))
            ii, jj = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
            uv = np.stack([ii / (ni - 1), jj / (nj - 1)], -1).reshape(-1, 2).astype(np.float32)
            qi, qj = np.meshgrid(np.arange(ni - 1), np.arange(nj - 1), indexing="ij")
            v00 = (qi * nj + qj).ravel()
            v01 = (qi * nj + qj + 1).ravel()
            v10 = ((qi + 1) * nj + qj).ravel()
            v11 = ((qi + 1) * nj + qj + 1).ravel()
            idx = np.stack([np.stack([v00, v01, v10], 1), np.stack([v10, v01, v11], 1)], 1).reshape(-1)
            m = len(mats)
            prims.append({"positions": p, "normals": n, "texcoords": uv, "tangents": None,
                          "indices": idx.astype(np.uint32), "material": m})
            if textures:
                texs.append(noise_texture(rng, tex_size))
                mats.append(textured_material(rng, len(texs) - 1))
            else:
                col = rng.uniform(0.2, 1.0, 3)
                mats.append({"base_color": (col[0], col[1], col[2], 1.0), "metallic": 0.0, "roughness": 0.5})
    return SceneData(prims, [(0, len(prims))], [(np.eye(4, dtype=np.float32), 0)], mats, texs,
                     name=name or f"c5_hf_s{seed}_g{grid}")


C5_SCENES = (
    ("ico64", lambda: make_icosphere_cluster(51, 64, 4, 48)),
    ("ico256", lambda: make_icosphere_cluster(52, 256, 4, 48)),
    ("hf724", lambda: make_patched_heightfield(53, 724)),
    ("hf1448", lambda: make_patched_heightfield(54, 1448)),
    ("hf1535", lambda: make_patched_heightfield(55, 1535)),
    ("hf512", lambda: make_patched_heightfield(56, 512, textures=False)),
    ("c3a", lambda: make_c3(seed=3)[0]),
    ("c3b", lambda: make_c3(seed=31)[0]),
)


def make_c5_scenes(names=None):
    """The served configuration's resident scenes (all, or those named)."""
    return [fn() for name, fn in C5_SCENES if names is None or name in names]


C5_RESOLUTIONS = ((1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))


def make_c5_requests(datas, n, rate, seed):
    """n requests [(arrival s, scene index, camera, width, height)] with
    Poisson arrivals at `rate`/s: a random scene, an orbit camera around its
    bounds (azimuth U[0, 2pi), elevation U[0.15, 0.6] rad, distance
    U[0.45, 0.8] x the scene diagonal, aimed at the centre, 60 deg, far =
    2 x diagonal) and a resolution from {720p, 1080p, 1440p, 2160p}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        k = int(rng.integers(len(datas)))
        w, h = C5_RESOLUTIONS[int(rng.integers(len(C5_RESOLUTIONS)))]
        _, _, centre, diag = datas[k].bounds()
        az, el = rng.uniform(0, 2 * np.pi), rng.uniform(0.15, 0.6)
        dist = diag * rng.uniform(0.45, 0.8)
        pos = (centre[0] + dist * np.cos(el) * np.sin(az), centre[1] + dist * np.sin(el),
               centre[2] + dist * np.cos(el) * np.cos(az))
        yaw, pitch = look_pose(pos, centre)
        spec = camera_spec(pos, yaw, pitch, np.pi / 3, w, h, 2.0 * diag)
        reqs.append((float(arrivals[i]), k, make_camera(spec), w, h))
    return reqs


# ---------------------------------------------------------------------------
# C4: batched views over a pool of heightfields
# ---------------------------------------------------------------------------
C4_RESOLUTIONS = ((320, 240), (640, 480), (1280, 720), (1920, 1080))


def look_pose(position, target):
    """(yaw, pitch) in the reference convention whose view direction points
    from position to target: forward = (sin y cos p, -sin p, -cos y cos p)."""
    d = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    d /= np.linalg.norm(d)
    return float(np.arctan2(d[0], -d[2])), float(np.arcsin(-d[1]))


def make_c4_pool(seed=4, scenes=16, tris_min=10_000, tris_max=2_000_000, tex_size=2048):
    """Heightfield scenes with log-uniform triangle counts; each samples the
    C2 terrain (seed 2) at its own grid resolution."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.sort(np.exp(rng.uniform(np.log(tris_min), np.log(tris_max), scenes)))
    tex_rng = np.random.Generator(np.random.PCG64(seed + 1000))
    t = value_noise((tex_size, tex_size), tex_rng, octaves=4, base_cells=16)
    g = ((t * 0.5 + 0.5) * 255).clip(0, 255).astype(np.uint8)
    rgb = (g[..., None] * np.array([0.55, 0.75, 0.45])[None, None, :]).clip(0, 255).astype(np.uint8)
    rgba = np.concatenate([rgb, np.full(rgb.shape[:2] + (1,), 255, np.uint8)], -1)
    tex = (tex_size, tex_size, SR_TEX_SRGB, pack_rgba(rgba).ravel())
    pool = []
    for c in counts:
        n = max(3, int(round(np.sqrt(c / 2.0))) + 1)
        pos, nrm, uv, idx = heightfield(n, 50.0, 5.0, np.random.Generator(np.random.PCG64(2)))
        prims = [{"positions": pos, "normals": nrm, "texcoords": uv, "tangents": None, "indices": idx,
                  "material": 0}]
        mats = [{"base_color": (1.0, 1.0, 1.0, 1.0), "metallic": 0.0, "roughness": 0.5,
                 "base_color_tex": (0, SR_WRAP_REPEAT, SR_WRAP_REPEAT)}]
        pool.append(SceneData(prims, [(0, 1)], [(np.eye(4, dtype=np.float32), 0)], mats, [tex],
                              name=f"c4_g{n}"))
    return pool


def prepare_c4_pool(pool, environment, device=None):
    """Environment and voxel grid for every pool scene: the grid is baked
    once, from the largest scene, and shared (the scenes sample one terrain)."""
    src = pool[-1]
    src.prepare(environment, device=device)
    for p in pool[:-1]:
        p.attach(src.environment, src.voxels)
    return pool


def make_c4_views(pool, seed=4, views=256, fov=np.pi / 3):
    """[(scene index, camera_spec)] for the batched configuration."""
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    out = []
    for _ in range(views):
        k = int(rng.integers(len(pool)))
        w, h = C4_RESOLUTIONS[int(rng.integers(len(C4_RESOLUTIONS)))]
        az, dist, hgt = rng.uniform(0, 2 * np.pi), rng.uniform(45, 75), rng.uniform(12, 30)
        pos = (dist * np.sin(az), hgt, dist * np.cos(az))
        yaw, pitch = look_pose(pos, (0.0, 0.0, 0.0))
        _, _, _, diag = pool[k].bounds()
        out.append((k, camera_spec(pos, yaw, pitch, fov, w, h, 2.0 * diag)))
    return out


# ---------------------------------------------------------------------------
# glTF export (the form the reference loader reads: .gltf + .bin + .png by URI)
# ---------------------------------------------------------------------------
def write_gltf(scene, path):
    """Write `scene` as glTF 2.0 (embedded-in-file buffers are not used: the
    reference ignores bufferView images, src/scene.rs:686)."""
    from PIL import Image
    base = os.path.splitext(path)[0]
    binp = base + ".bin"
    blob = bytearray()
    views, accessors = [], []

    def add(arr, target, comp, typ, minmax=False):
        arr = np.ascontiguousarray(arr)
        while len(blob) % 4:
            blob.append(0)
        off = len(blob)
        blob.extend(arr.tobytes())
        views.append({"buffer": 0, "byteOffset": off, "byteLength": arr.nbytes, "target": target})
        acc = {"bufferView": len(views) - 1, "componentType": comp, "count": int(arr.shape[0]), "type": typ}
        if minmax:
            acc["min"] = [float(x) for x in arr.min(0)]
            acc["max"] = [float(x) for x in arr.max(0)]
        accessors.append(acc)
        return len(accessors) - 1

    meshes = []
    prim_acc = []
    for p in scene.primitives:
        a = {"POSITION": add(p["positions"], 34962, 5126, "VEC3", True),
             "NORMAL": add(p["normals"], 34962, 5126, "VEC3")}
        if p.get("texcoords") is not None:
            a["TEXCOORD_0"] = add(p["texcoords"], 34962, 5126, "VEC2")
        ind = add(p["indices"].astype(np.uint32), 34963, 5125, "SCALAR")
        prim_acc.append({"attributes": a, "indices": ind, "material": int(p["material"])})
    for first, count in scene.meshes:
        meshes.append({"primitives": prim_acc[first:first + count]})
    nodes = []
    for t, m in scene.nodes:
        col = np.asarray(t, np.float32).reshape(4, 4).T.ravel()
        node = {"matrix": [float(x) for x in col]}
        if m >= 0:
            node["mesh"] = int(m)
        nodes.append(node)
    images, textures = [], []
    for i, (w, h, _ty, data) in enumerate(scene.textures):
        png = f"{os.path.basename(base)}_tex{i}.png"
        rgba = np.stack([(data >> 24) & 255, (data >> 16) & 255, (data >> 8) & 255, data & 255], -1)
        Image.fromarray(rgba.reshape(h, w, 4).astype(np.uint8), "RGBA").save(os.path.join(os.path.dirname(path), png))
        images.append({"uri": png})
        textures.append({"source": i})
    materials = []
    for m in scene.materials:
        pbr = {"baseColorFactor": [float(x) for x in m["base_color"]], "metallicFactor": float(m.get("metallic", 1)),
               "roughnessFactor": float(m.get("roughness", 1))}
        if m.get("base_color_tex") is not None:
            pbr["baseColorTexture"] = {"index": int(m["base_color_tex"][0])}
        materials.append({"pbrMetallicRoughness": pbr})
    doc = {"asset": {"version": "2.0", "generator": "swraster-viewer_amd scenes.py"}, "scene": 0,
           "scenes": [{"nodes": list(range(len(nodes)))}], "nodes": nodes, "meshes": meshes,
           "materials": materials, "accessors": accessors, "bufferViews": views,
           "buffers": [{"uri": os.path.basename(binp), "byteLength": len(blob)}]}
    if images:
        doc["images"], doc["textures"] = images, textures
    with open(binp, "wb") as fh:
        fh.write(bytes(blob))
    with open(path, "w") as fh:
        json.dump(doc, fh)
    return path


def array_digest(scene):
    """SHA-256 over every generated array, pinning the generator output."""
    import hashlib
    h = hashlib.sha256()
    for p in scene.primitives:
        for k in ("positions", "normals", "texcoords", "indices"):
            if p.get(k) is not None:
                h.update(np.ascontiguousarray(p[k]).tobytes())
    for _, _, _, d in scene.textures:
        h.update(np.ascontiguousarray(d).tobytes())
    for t, m in scene.nodes:
        h.update(np.asarray(t, np.float32).tobytes() + struct.pack("<i", m))
    return h.hexdigest()
