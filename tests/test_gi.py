"""CPU tests of the voxel GI bake on the host (sr_prepare_gi with device -1;
bake_voxel_gi, src/gi.rs:316-412), pinned by what the geometry fixes:
no surfaces leave the scene SH untouched; an open uniform sky gives
irradiance pi * L; a closed emissive room gives pi * E and no sky; light
from one wall lands in that wall's SH axis; a room that returns nothing is
floored at 0.005 of the scene SH; a sunlit floor bounces albedo * sun *
cos / pi.  The GPU bake is compared with this one in test_gpu_gi.py."""
import ctypes as C

import numpy as np
import pytest

from util import dark_environment, dummy_environment

PI = np.pi


def _scene(sr, tris, mats, bounds=None):
    """One primitive per material group; tris: list of ((3,3) vertices,
    material).  bounds: (lo, hi) widened by a degenerate triangle (the ray
    tracer skips degenerate faces, the scene bounds keep them)."""
    prims, meshes, nodes = [], [], []
    groups = {}
    for v, m in tris:
        groups.setdefault(m, []).append(np.asarray(v, np.float32))
    if bounds is not None:
        lo, hi = np.asarray(bounds[0], np.float32), np.asarray(bounds[1], np.float32)
        groups.setdefault(0, []).append(np.stack([lo, hi, (lo + hi) * np.float32(0.5)]))
    for i, (m, vs) in enumerate(sorted(groups.items())):
        pos = np.concatenate(vs).astype(np.float32)
        prims.append({"positions": pos, "normals": np.tile(np.float32([0, 1, 0]), (len(pos), 1)),
                      "texcoords": None, "tangents": None, "indices": np.arange(len(pos), dtype=np.uint32),
                      "material": m})
        meshes.append((i, 1))
        nodes.append((np.eye(4, dtype=np.float32), i))
    return sr.SceneData(prims, meshes, nodes, mats, [], name="gi_test")


def _quad(a, b, c, d, m):
    return [(np.array([a, b, c]), m), (np.array([a, c, d]), m)]


def _box(lo, hi, mats):
    """Six faces of an axis-aligned box; mats: material per face in the
    order -x, +x, -y, +y, -z, +z."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        ((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)),
        ((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0)),
        ((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)),
        ((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)),
        ((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)),
        ((x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1)),
    ]
    out = []
    for f, m in zip(faces, mats):
        out += _quad(*f, m)
    return out


def _mat(base=(0.0, 0.0, 0.0), emissive=(0.0, 0.0, 0.0)):
    return {"base_color": (*base, 1.0), "metallic": 0.0, "roughness": 1.0, "emissive": emissive}


def _gi(sr, scene, env, grid, device=-1, threads=8):
    scene.environment = env
    scene.desc.environment = C.pointer(env.desc)
    g = sr.sr_voxel_grid()
    sr.check(sr.lib().sr_prepare_gi(C.byref(scene.desc), grid, device, threads, C.byref(g)))
    n = g.dims[0] * g.dims[1] * g.dims[2]
    light = np.ctypeslib.as_array(g.light_intensity, shape=(n,)).copy().reshape(grid, grid, grid)
    sh4 = np.ctypeslib.as_array(g.sh4, shape=(n * 16,)).copy().reshape(grid, grid, grid, 4, 4)
    assert g.sh_scale == 1.0
    sr.lib().sr_voxel_grid_free(C.byref(g))
    return light, sh4          # indexed [z, y, x]


def _srgb_decode(c):
    s = np.float64(c) / 255.0
    return s / 12.92 if s <= 0.04045 else ((s + 0.055) / 1.055) ** 2.4


ROOM = ((8.5, 8.5, 8.5), (12.5, 12.5, 12.5))     # interior voxels 9..11 of a 24^3 grid over [0, 24]^3
BOUNDS = ((0.0, 0.0, 0.0), (24.0, 24.0, 24.0))


def test_gi_needs_environment(sr):
    scene = _scene(sr, _box(*ROOM, [0] * 6), [_mat()], BOUNDS)
    g = sr.sr_voxel_grid()
    assert sr.lib().sr_prepare_gi(C.byref(scene.desc), 8, -1, 1, C.byref(g)) == -1
    assert sr.lib().sr_prepare_gi(C.byref(scene.desc), 0, -1, 1, C.byref(g)) == -1


def test_gi_without_surfaces_keeps_scene_sh(sr):
    """No (non-degenerate) triangle: nothing is active and every voxel keeps
    the scene SH with w0 = sun visibility and w1 = 1, not blurred
    (src/gi.rs:336-341)."""
    env = dummy_environment(sr)
    scene = _scene(sr, [], [_mat()], BOUNDS)
    light, sh4 = _gi(sr, scene, env, 6)
    sh = np.asarray(env.sh, np.float32)
    for c in range(4):
        assert np.array_equal(sh4[..., c, :3], np.broadcast_to(sh[c], sh4[..., c, :3].shape))
    assert np.array_equal(sh4[..., 0, 3], light)
    assert np.all(light == 1.0)
    assert np.all(sh4[..., 1, 3] == 1.0) and np.all(sh4[..., 2:, 3] == 0.0)


def test_gi_open_sky_irradiance(sr):
    """A voxel whose rays all escape sees the uniform skybox of linear
    radiance L: SH0 irradiance = pi * 0.282095^2 * 4 pi * L = pi * L, the
    linear terms only sampling noise, sky visibility 1.  One tiny triangle
    makes the voxels around it active."""
    sky = (128, 160, 200)
    env = dummy_environment(sr, color=sky)
    c = np.float32(12.5)
    e = np.float32(1e-3)
    tri = np.array([[c - e, c, c - e], [c + e, c, c - e], [c, c, c + e]], np.float32)
    scene = _scene(sr, [(tri, 0)], [_mat((0.5, 0.5, 0.5))], BOUNDS)
    light, sh4 = _gi(sr, scene, env, 24)
    v = sh4[12, 12, 12]
    assert v[1, 3] == 1.0                      # every ray of all 27 voxels escaped
    assert v[0, 3] == light[12, 12, 12] == 1.0
    for k in range(3):
        L = _srgb_decode(sky[k])
        assert abs(v[0, k] / (PI * L) - 1.0) < 2e-4
        for i in (1, 2, 3):
            assert 0.0 <= v[i, k] < 0.1 * v[0, k]


def test_gi_closed_emissive_room(sr):
    """Inside a closed room whose walls emit E (albedo 0, black sky, no
    SH), every ray returns E: SH0 = pi * E, sky visibility 0; the blurred
    sun visibility is the 3x3x3 mean of the light grid."""
    E = (0.5, 0.25, 0.125)
    env = dark_environment(sr, sky=(0, 0, 0))
    scene = _scene(sr, _box(*ROOM, [0] * 6), [_mat(emissive=E)], BOUNDS)
    light, sh4 = _gi(sr, scene, env, 24)
    v = sh4[10, 10, 10]
    assert v[1, 3] == 0.0
    for k in range(3):
        assert abs(v[0, k] / (PI * E[k]) - 1.0) < 1e-4
    acc = np.float32(0.0)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                acc = np.float32(acc + light[10 + dz, 10 + dy, 10 + dx])
    assert v[0, 3] == np.float32(acc / np.float32(27.0))


@pytest.mark.parametrize("face,axis", [(1, 3), (3, 1), (5, 2)])
def test_gi_light_lands_on_its_axis(sr, face, axis):
    """Only the +x, +y or +z wall emits: the irradiance gradient is in SH
    coefficient 3, 1 or 2 (basis 0.282095, 0.488603 y, 0.488603 z,
    0.488603 x); the opposite wall alone gives a negative gradient, which
    the bake clamps to zero."""
    env = dark_environment(sr, sky=(0, 0, 0))
    mats = [_mat(), _mat(emissive=(1.0, 1.0, 1.0))]
    walls = [0] * 6
    walls[face] = 1
    _, sh4 = _gi(sr, _scene(sr, _box(*ROOM, walls), mats, BOUNDS), env, 24)
    v = sh4[10, 10, 10, :, 0]
    others = [i for i in (1, 2, 3) if i != axis]
    assert v[axis] > 0.2 and v[axis] > 4 * max(v[others[0]], v[others[1]])
    walls = [0] * 6
    walls[face - 1] = 1
    _, sh4 = _gi(sr, _scene(sr, _box(*ROOM, walls), mats, BOUNDS), env, 24)
    assert sh4[10, 10, 10, axis, 0] == 0.0 and sh4[10, 10, 10, 0, 0] > 0.2


def test_gi_min_brightness_floor(sr):
    """A closed room with black walls (albedo 0, no emission) returns zero
    radiance: every active voxel becomes 0.005 * scene SH
    (apply_min_brightness_floor, src/gi.rs:630-665)."""
    env = dummy_environment(sr, color=(0, 0, 0))
    scene = _scene(sr, _box(*ROOM, [0] * 6), [_mat()], BOUNDS)
    _, sh4 = _gi(sr, scene, env, 24)
    sh = np.asarray(env.sh, np.float32) * np.float32(0.005)
    np.testing.assert_allclose(sh4[10, 10, 10, :, :3], sh, rtol=1e-6, atol=0)


def _floor_scene(sr, albedo):
    y = 8.5
    q = _quad((0.0, y, 0.0), (0.0, y, 24.0), (24.0, y, 24.0), (24.0, y, 0.0), 0)
    return _scene(sr, q, [_mat((albedo, albedo, albedo))], BOUNDS)


def test_gi_sunlit_floor_bounce(sr):
    """Sunlit floor, black sky, no SH: a hit returns albedo * sun * (n.L) / pi;
    the voxel layer one above the floor sees it over the lower hemisphere
    (irradiance pi * R / 2, less the grazing rays that pass the floor's
    edge), the layer through the floor sees nothing.  Halving the albedo
    halves every coefficient exactly."""
    env = dark_environment(sr, sky=(0, 0, 0))
    scene = _floor_scene(sr, 0.5)
    light_dir, light_color = scene.light_dir, scene.light_color
    _, a = _gi(sr, scene, env, 24)
    _, b = _gi(sr, _floor_scene(sr, 0.25), env, 24)
    assert np.array_equal(a[..., :3], b[..., :3] * np.float32(2.0))
    assert np.array_equal(a[..., 3], b[..., 3])
    # voxel (12, 9, 12): layers y = 8 (on the floor: nothing), 9 and 10
    for k in range(3):
        R = 0.5 * light_color[k] * light_dir[1] / PI
        expect = (2.0 / 3.0) * PI * R / 2.0
        assert 0.75 < a[12, 9, 12, 0, k] / expect < 1.02
    assert a[12, 9, 12, 1, 0] == 0.0       # all light from below: clamped y gradient
