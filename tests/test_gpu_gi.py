"""GPU tests of the voxel GI bake (sr_prepare_gi on a HIP device;
bake_voxel_gi, src/gi.rs:316-412) against the host bake (test_gi.py pins
that one): identical sh4 and light grids.  Then a frame lit by the baked
grid renders on the GPU as the oracle renders it."""
import time

import numpy as np
import pytest

from test_gi import BOUNDS, ROOM, _box, _gi, _mat, _scene
from test_gpu_parity import compare, render_gpu, render_oracle
from util import dark_environment

pytestmark = pytest.mark.gpu


def _both(sr, name, scene, env, grid):
    t0 = time.perf_counter()
    ld, sd = _gi(sr, scene, env, grid, device=0)
    t1 = time.perf_counter()
    lh, sh = _gi(sr, scene, env, grid, device=-1, threads=0)
    t2 = time.perf_counter()
    print(f"[gi] {name} grid {grid}: device {(t1 - t0) * 1e3:.1f} ms, host {(t2 - t1) * 1e3:.1f} ms "
          f"({(t2 - t1) / (t1 - t0):.1f}x)", flush=True)
    assert np.array_equal(ld.view(np.uint32), lh.view(np.uint32))
    assert np.array_equal(sd.view(np.uint32), sh.view(np.uint32)), \
        f"{name}: {int((sd.view(np.uint32) != sh.view(np.uint32)).sum())} sh4 floats differ"
    return sd


def test_device_gi_rooms(sr):
    env = dark_environment(sr, sky=(0, 0, 0))
    mats = [_mat(), _mat(emissive=(1.0, 0.5, 0.25)), _mat((0.6, 0.6, 0.6))]
    _both(sr, "emissive room", _scene(sr, _box(*ROOM, [1, 2, 2, 2, 2, 2]), mats, BOUNDS), env, 24)
    q = [((np.float32([0, 8.5, 0]), np.float32([0, 8.5, 24]), np.float32([24, 8.5, 24])), 2),
         ((np.float32([0, 8.5, 0]), np.float32([24, 8.5, 24]), np.float32([24, 8.5, 0])), 2)]
    _both(sr, "sunlit floor", _scene(sr, q, mats, BOUNDS), env, 24)


@pytest.mark.parametrize("config,grid", [("c1", 32), ("c3", 24), ("c2", 16)])
def test_device_gi_matches_host(sr, scenes, env, config, grid):
    scene, _ = {"c1": scenes.make_c1, "c2": scenes.make_c2, "c3": scenes.make_c3}[config]()
    sd = _both(sr, config, scene, env, grid)
    assert np.all(np.isfinite(sd))


def test_device_gi_emissive_texture_and_translucency(sr, scenes, env):
    """C1 with emissive textures under mirrored and clamped wrap and a third
    of the spheres translucent (transmittance rays multiply through them)."""
    scene, _ = scenes.make_c1()
    for i, m in enumerate(scene.materials):
        if i % 2 == 0:
            m["emissive"] = (0.6, 0.4, 0.2)
            m["emissive_tex"] = (0, sr.SR_WRAP_MIRRORED_REPEAT, sr.SR_WRAP_CLAMP_TO_EDGE)
        if i % 3 == 1:
            m["is_translucent"] = 1
            m["transmission"] = 0.3 + 0.05 * i
    scene.rebuild()
    _both(sr, "c1 emissive+translucent", scene, env, 32)


def test_gi_lit_frame_parity(sr, orc, scenes, env):
    """C1 lit by a GI grid baked on the device: the GPU frame matches the
    oracle's (visibility identical, RGBA8 within 1 LSB) and differs from the
    frame without GI."""
    W, H = 640, 360
    scene, spec = scenes.make_c1(width=W, height=H)
    cam = scenes.make_camera(spec)
    scene.prepare(env, grid=64, device=0)
    plain = render_gpu(sr, sr.Scene(scene), cam, W, H)
    scene.prepare(env, grid=64, device=0, gi=True)
    assert scene.voxels.sh4 is not None
    g = render_gpu(sr, sr.Scene(scene), cam, W, H)
    o = render_oracle(orc, scene, cam, W, H)
    compare(sr, g, o, "c1 gi")
    changed = int((plain["px"] != g["px"]).sum())
    print(f"[gi] {changed} pixels change with GI", flush=True)
    assert changed > 10000


def test_device_gi_full_grid_properties(sr, scenes, env):
    """The reference's 128^3 grid on C1 (device only; the host bake takes
    seconds): its sun visibility is the sun bake's grid exactly, every
    coefficient is finite, SH0 is non-negative (baked terms are clamped, and
    the scene's SH0 is positive; the linear terms keep the scene's signs), the
    sky visibility is a mean of fractions, and SH0's luminance stays at or
    above 3/4 of the floor (0.005 of the scene SH): the floor's blend gives
    cur + (floor - cur)^2 / floor >= 3/4 floor, and the blur averages."""
    scene, _ = scenes.make_c1()
    light, sh4 = _gi(sr, scene, env, 128, device=0)
    d = sr.sr_voxel_grid()
    sr.check(sr.lib().sr_prepare_voxel_grid_device(sr.C.byref(scene.desc), 128, 0, sr.C.byref(d)))
    sun = np.ctypeslib.as_array(d.light_intensity, shape=(128 ** 3,)).copy().reshape(128, 128, 128)
    sr.lib().sr_voxel_grid_free(sr.C.byref(d))
    assert np.array_equal(light.view(np.uint32), sun.view(np.uint32))
    assert np.all(np.isfinite(sh4))
    assert np.all(sh4[..., 0, :3] >= 0.0)
    sky = sh4[..., 1, 3]
    assert np.all((sky >= 0.0) & (sky <= 1.0))
    floor = np.asarray(env.sh, np.float32)[0] * np.float32(0.005)
    lum = lambda c: 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]
    assert lum(sh4[..., 0, :3]).min() >= 0.75 * lum(floor) * (1 - 1e-5)
