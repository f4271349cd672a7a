"""GPU tests of the voxel sun-visibility bake on the device
(sr_prepare_voxel_grid_device; src/gi.rs:151-314, src/raytracer.rs:172-260,
src/voxelgrid.rs:371-419) against the host bake: identical grids."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bake(sr, scene, grid, device):
    d = sr.sr_voxel_grid()
    t0 = time.perf_counter()
    if device is None:
        sr.check(sr.lib().sr_prepare_voxel_grid(sr.C.byref(scene.desc), grid, 0, sr.C.byref(d)))
    else:
        sr.check(sr.lib().sr_prepare_voxel_grid_device(sr.C.byref(scene.desc), grid, device, sr.C.byref(d)))
    dt = time.perf_counter() - t0
    n = d.dims[0] * d.dims[1] * d.dims[2]
    light = np.ctypeslib.as_array(d.light_intensity, shape=(n,)).copy()
    wmin, wmax = list(d.world_min), list(d.world_max)
    sr.lib().sr_voxel_grid_free(sr.C.byref(d))
    return light, wmin, wmax, dt


@pytest.mark.parametrize("config,grid", [("c1", 128), ("c1", 37), ("c3", 128), ("c2", 128)])
def test_device_bake_matches_host(sr, scenes, config, grid):
    scene, _ = {"c1": scenes.make_c1, "c2": scenes.make_c2, "c3": scenes.make_c3}[config]()
    _bake(sr, scene, 8, 0)                                  # first-use set-up outside the timing
    dev, dmin, dmax, t_dev = _bake(sr, scene, grid, 0)
    host, hmin, hmax, t_host = _bake(sr, scene, grid, None)
    lit = int((host > 0).sum())
    print(f"[bake] {config} grid {grid}: device {t_dev * 1e3:.1f} ms, host {t_host * 1e3:.1f} ms "
          f"({t_host / t_dev:.1f}x); {lit} lit of {host.size} voxels")
    assert dmin == hmin and dmax == hmax
    assert np.array_equal(dev.view(np.uint32), host.view(np.uint32))
    assert lit > 0
    if config != "c2":           # the heightfield casts no shadow at this sun angle
        assert np.any(host < 1.0)


def test_device_bake_translucent_shadows(sr, scenes):
    """Translucent casters: transmissions multiply along the sun ray."""
    scene, _ = scenes.make_c1()
    for i, m in enumerate(scene.materials):
        if i % 2:
            m["is_translucent"] = 1
            m["transmission"] = 0.3 + 0.04 * i
    scene.rebuild()
    dev, _, _, _ = _bake(sr, scene, 96, 0)
    host, _, _, _ = _bake(sr, scene, 96, None)
    assert np.array_equal(dev.view(np.uint32), host.view(np.uint32))
    # partial shadow values exist (neither 0 nor 1 before the blur)
    assert np.any((host > 0.01) & (host < 0.99))


def _environment(sr, faces, device):
    env = sr.sr_environment()
    t0 = time.perf_counter()
    if device is None:
        sr.check(sr.lib().sr_prepare_environment(sr._uptr(faces), faces.shape[-1], 0, sr.C.byref(env)))
    else:
        sr.check(sr.lib().sr_prepare_environment_device(sr._uptr(faces), faces.shape[-1], device, sr.C.byref(env)))
    dt = time.perf_counter() - t0
    size = env.cubemap_size
    spec = np.ctypeslib.as_array(env.specular, shape=(env.specular_mips * 6 * size * size,)).copy()
    lut = np.ctypeslib.as_array(env.brdf_lut, shape=(env.brdf_lut_size ** 2,)).copy()
    sh = np.array([[env.irradiance_sh[i][k] for k in range(3)] for i in range(4)], np.float32)
    sr.lib().sr_environment_free(sr.C.byref(env))
    return spec, lut, sh, dt


def test_device_environment_matches_host(sr):
    """GGX-prefiltered specular cubemap, BRDF LUT and irradiance SH of the
    skybox (src/texture.rs:199-420) on the device against the host: the LUT
    bit-identical, the specular cubemap within 1 LSB per channel, the SH
    within float rounding."""
    faces = sr.cubemap_faces()
    _environment(sr, faces[:, :16, :16].copy(), 0)               # first-use set-up outside the timing
    spec_d, lut_d, sh_d, t_dev = _environment(sr, faces, 0)
    spec_h, lut_h, sh_h, t_host = _environment(sr, faces, None)
    assert np.array_equal(lut_d, lut_h)
    a, b = sr.unpack_rgba(spec_d).astype(int), sr.unpack_rgba(spec_h).astype(int)
    diff = np.abs(a - b).max(-1)
    rel = np.abs(sh_d.astype(np.float64) - sh_h) / np.maximum(np.abs(sh_h), 1e-12)
    print(f"[environment] device {t_dev * 1e3:.0f} ms, host {t_host * 1e3:.0f} ms ({t_host / t_dev:.0f}x); "
          f"specular texels differing {(diff > 0).sum()} of {diff.size} (max {diff.max()} LSB); "
          f"SH max rel diff {rel.max():.2e}")
    assert diff.max() <= 1
    assert (diff > 0).mean() < 1e-3
    assert rel.max() < 1e-5
