"""Pins of the oracle's shading against closed forms (CPU only).

Each test evaluates one part of the oracle's restatement of the reference's
shading on hand-set inputs and compares it with the published model written
here independently in float64, chosen so that a plausible mistake (a wrong
exponent, a dropped or swapped term, a transposed matrix, a swapped SH axis)
fails:
  - pbr_shader's direct term (src/shader.rs:159-237): Schlick Fresnel with
    exponent 5, the GGX normal distribution, Smith-Schlick geometry with
    k = (r + 1)^2 / 8 and the Lambert diffuse weighted by (1 - F)(1 - metal),
    at random normal / light / view configurations, and the N = L = V case
    where they reduce to F = f0, D = a^2 / (pi a^4 + eps), G = 1 / (1 + eps)^2;
  - metallic 1 leaves no diffuse term; roughness clamps to [0.045, 1];
  - the SH irradiance (src/shader.rs:240-250): c0 + c1 n.y + c2 n.z + c3 n.x,
    clamped at zero, on axis normals;
  - the split-sum specular (src/shader.rs:251-262): env x (f0 LUT.x + LUT.y)
    and the prefiltered mip = roughness x max mip;
  - sample_cubemap_trilinear_rgb (src/texture.rs:665-678) between two
    constant mips, and sample_bilinear_rgb (src/texture.rs:730-790) on a
    linear ramp, clamped at the edges;
  - compute_skybox's view direction (src/tilerasterizer.rs:478-505) against
    the camera basis for yaw, pitch and both;
  - update_auto_exposure (src/renderer.rs:258-290): trimmed mean of log2
    luminance and the exponential step, on hand-set tile luminances.
"""
import math

import numpy as np
import pytest

from util import dummy_voxels

F32 = np.float32
PI_F = float(F32(3.14159265358979323846))
EPS_F = float(F32(1e-6))
INV_PI_F = float(F32(1.0) / F32(3.14159265358979323846))


def solid(sr, rgb):
    return int(sr.pack_rgba(np.array([[rgb[0], rgb[1], rgb[2], 255]], np.uint8))[0])


def make_env(sr, sky=(128, 160, 200), spec=((0, 0, 0),), lut=None, sh=None, face=4):
    """Environment of face^2 cubemap faces: sky colour, one constant colour per
    specular mip, a LUT (default constant) and SH coefficients (default 0)."""
    cube = np.full((6, face, face), solid(sr, sky), np.uint32)
    specular = np.concatenate([np.full(6 * face * face, solid(sr, c), np.uint32) for c in spec])
    if lut is None:
        lut = np.full(64, solid(sr, (200, 40, 0)), np.uint32)
    if sh is None:
        sh = np.zeros((4, 3), np.float32)
    return sr.Environment(cube, specular, len(spec), lut, sh)


def probe_scene(sr, orc, materials, env, light_dir=(0.0, 1.0, 0.0), light_color=(1.0, 1.0, 1.0)):
    """A one-triangle scene carrying the materials, environment, light and a
    voxel grid of sun visibility 1 (sky visibility 1, SH scale 0.25)."""
    pos = np.array([[-1, -1, -5], [1, -1, -5], [0, 1, -5]], np.float32)
    prim = {"positions": pos, "normals": np.tile(np.array([0, 0, 1], np.float32), (3, 1)), "texcoords": None,
            "tangents": None, "indices": np.arange(3, dtype=np.uint32), "material": 0}
    scene = sr.SceneData([prim], [(0, 1)], [(np.eye(4, dtype=np.float32), 0)], materials, [], name="probe")
    scene.light_dir = np.asarray(light_dir, np.float32)
    scene.light_color = np.asarray(light_color, np.float32)
    scene.rebuild()
    scene.attach(env, dummy_voxels(sr, scene))
    return scene, orc.OracleScene(scene)


def camera_at(sr, pos):
    return sr.RenderCamera.from_pose(list(pos), 0.0, 0.0, math.pi / 3, 640, 360, 100.0).update_matrices()


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def cook_torrance(n, l, v, base, metal, rough, light_color):
    """Direct sun term of the reference's model in float64 (constants as the
    f32 values the shader uses)."""
    ndl = max(float(n @ l), 0.0)
    ndv = max(float(n @ v), 1e-4)
    h = unit(l + v)
    ndh = max(float(n @ h), 0.0)
    vdh = max(float(v @ h), 0.0)
    r = min(max(rough, 0.045), 1.0)
    m = min(max(metal, 0.0), 1.0)
    base = np.asarray(base, np.float64)
    f0 = 0.04 + (base - 0.04) * m
    fres = f0 + (1.0 - f0) * (1.0 - vdh) ** 5
    a2 = (r * r) ** 2
    d = a2 / (PI_F * (ndh * ndh * (a2 - 1.0) + 1.0) ** 2 + EPS_F)
    k = (r + 1.0) ** 2 / 8.0
    g = ndv / (ndv * (1.0 - k) + k + EPS_F) * (ndl / (ndl * (1.0 - k) + k + EPS_F))
    spec = d * g / (4.0 * ndl * ndv + EPS_F)
    lc = np.asarray(light_color, np.float64)
    diffuse = lc * (1.0 - fres) * (1.0 - m) * base * INV_PI_F * ndl
    return diffuse + lc * fres * spec * ndl


def material(base, metal, rough):
    return {"base_color": (base[0], base[1], base[2], 1.0), "metallic": float(metal), "roughness": float(rough)}


def test_direct_term_matches_cook_torrance(sr, orc):
    """Random normal, light and view directions, base colours, metallic and
    roughness (a dark environment: black specular cubemap, zero SH)."""
    rng = np.random.Generator(np.random.PCG64(21))
    checked = 0
    for trial in range(40):
        n = unit(rng.normal(size=3))
        l = unit(n + rng.normal(size=3) * 0.6)
        v = unit(n + rng.normal(size=3) * 0.6)
        if n @ l < 0.05 or n @ v < 0.05:
            continue
        base, metal, rough = rng.uniform(0.05, 1, 3), rng.uniform(0, 1), rng.uniform(0.1, 1)
        lc = rng.uniform(0.5, 5, 3)
        scene, osc = probe_scene(sr, orc, [material(base, metal, rough)], make_env(sr), light_dir=l,
                                 light_color=lc)
        p = np.array([0.3, -0.2, -5.0])
        cam = camera_at(sr, p + 7.0 * v)
        # the view direction the shader sees is normalize(cam - p) in f32
        vv = unit(np.asarray(cam.c.position[:], np.float64) - p)
        want = cook_torrance(n, np.asarray(scene.light_dir, np.float64), vv, base, metal, rough, lc)
        c32, c64 = orc.shade_probe(osc, cam, 0, n, p)
        np.testing.assert_allclose(c64, want, rtol=2e-5, atol=1e-7, err_msg=f"trial {trial} (f64)")
        np.testing.assert_allclose(c32, want, rtol=1e-4, atol=1e-6, err_msg=f"trial {trial} (f32)")
        checked += 1
    assert checked >= 25


@pytest.mark.parametrize("rough", [0.045, 0.3, 0.8, 1.0])
def test_normal_light_view_aligned(sr, orc, rough):
    """N = L = V: F = f0, D = a^2 / (pi a^4 + eps), G = 1 / (1 + eps)^2."""
    n = unit([0.2, 0.9, -0.3])
    base, metal = np.array([0.7, 0.4, 0.2]), 0.35
    scene, osc = probe_scene(sr, orc, [material(base, metal, rough)], make_env(sr), light_dir=n.astype(np.float32))
    p = np.array([0.0, 0.0, -5.0])
    cam = camera_at(sr, p + 5.0 * n)
    _, c64 = orc.shade_probe(osc, cam, 0, n, p)
    f0 = 0.04 + (base - 0.04) * metal
    a = rough * rough
    d = a * a / (PI_F * (a * a) ** 2 + EPS_F)
    g = (1.0 / (1.0 + EPS_F)) ** 2
    want = (1.0 - f0) * (1.0 - metal) * base * INV_PI_F + f0 * d * g / (4.0 + EPS_F)
    np.testing.assert_allclose(c64, want, rtol=2e-6)


def test_metallic_one_has_no_diffuse(sr, orc):
    """metallic = 1: the sun's Lambert term and the SH diffuse vanish; what
    is left is the specular lobe with f0 = base colour."""
    n, l, v = unit([0, 1, 0]), unit([0.3, 1, 0.1]), unit([-0.2, 1, 0.4])
    base = np.array([0.9, 0.6, 0.3])
    env = make_env(sr, sh=np.array([[0.5, 0.5, 0.5], [0.2, 0.2, 0.2], [0.1, 0.1, 0.1], [0.0, 0.0, 0.0]]))
    scene, osc = probe_scene(sr, orc, [material(base, 1.0, 0.5)], env, light_dir=l)
    p = np.array([0.0, 0.0, -5.0])
    cam = camera_at(sr, p + 6.0 * v)
    vv = unit(np.asarray(cam.c.position[:], np.float64) - p)
    _, c64 = orc.shade_probe(osc, cam, 0, n, p)
    ll = np.asarray(scene.light_dir, np.float64)
    h = unit(ll + vv)
    fres = base + (1.0 - base) * (1.0 - max(vv @ h, 0.0)) ** 5
    spec_only = cook_torrance(n, ll, vv, base, 1.0, 0.5, (1, 1, 1))
    np.testing.assert_allclose(c64, spec_only, rtol=2e-5)
    assert np.all(fres > 0)


def test_roughness_clamps(sr, orc):
    """Roughness below 0.045 shades as 0.045 and above 1 as 1, exactly."""
    n, l, v = unit([0, 1, 0]), unit([0.2, 1, 0.3]), unit([-0.1, 1, -0.2])
    mats = [material((0.5, 0.5, 0.5), 0.2, r) for r in (0.0, 0.045, 0.01, 1.0, 3.0)]
    scene, osc = probe_scene(sr, orc, mats, make_env(sr), light_dir=l)
    p = np.array([0.0, 0.0, -5.0])
    cam = camera_at(sr, p + 6.0 * v)
    out = [orc.shade_probe(osc, cam, i, n, p)[0] for i in range(len(mats))]
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[2], out[1])
    assert np.array_equal(out[4], out[3])
    assert not np.array_equal(out[1], out[3])


def test_sh_irradiance_axes(sr, orc):
    """No sun (black light colour), black specular: the colour is the SH
    diffuse max(c0 + c1 n.y + c2 n.z + c3 n.x, 0) x base x (1 - f0)(1 - m) / pi,
    with the grid's coefficients c = 0.25 x the environment's SH."""
    sh = np.array([[0.8, 0.6, 0.4], [0.3, 0.2, 0.1], [-0.25, 0.15, 0.05], [0.12, -0.3, 0.2]], np.float32)
    env = make_env(sr, sh=sh)
    base = np.array([0.9, 0.7, 0.5])
    scene, osc = probe_scene(sr, orc, [material(base, 0.0, 0.5)], env, light_color=(0.0, 0.0, 0.0))
    c = 0.25 * sh.astype(np.float64)
    kd = (1.0 - (0.04 + (base - 0.04) * 0.0)) * 1.0
    p = np.array([0.0, 0.0, -5.0])
    for axis, sign, coeff in ((1, 1, 1), (1, -1, 1), (2, 1, 2), (2, -1, 2), (0, 1, 3), (0, -1, 3)):
        n = np.zeros(3)
        n[axis] = sign
        cam = camera_at(sr, p + 5.0 * unit(n + np.array([0.05, 0.05, 0.05])))
        _, c64 = orc.shade_probe(osc, cam, 0, n, p)
        irr = np.maximum(c[0] + sign * c[coeff], 0.0)
        np.testing.assert_allclose(c64, irr * base * kd * INV_PI_F, rtol=1e-6, atol=1e-9,
                                   err_msg=f"normal {n}")


def test_split_sum_specular(sr, orc):
    """No sun and zero SH: the colour is the prefiltered environment times
    (f0 LUT.x + LUT.y); a constant specular cubemap E and a constant LUT
    (a, b) give E/255 (f0 a/255 + b/255)."""
    E, a, b = (180, 90, 30), 200, 40
    env = make_env(sr, spec=(E, E, E), lut=np.full(64, solid(sr, (a, b, 0)), np.uint32))
    base, metal = np.array([0.8, 0.5, 0.2]), 0.6
    scene, osc = probe_scene(sr, orc, [material(base, metal, 0.4)], env, light_color=(0.0, 0.0, 0.0))
    p = np.array([0.0, 0.0, -5.0])
    n = unit([0.1, 1.0, 0.2])
    cam = camera_at(sr, p + 5.0 * unit([0.3, 1.0, -0.1]))
    _, c64 = orc.shade_probe(osc, cam, 0, n, p)
    f0 = 0.04 + (base - 0.04) * metal
    want = np.array(E, np.float64) / 255.0 * (f0 * (a / 255.0) + b / 255.0)
    np.testing.assert_allclose(c64, want, rtol=1e-6)


def test_specular_mip_follows_roughness(sr, orc):
    """Specular mips constant A, B, B: roughness r reads mip r x 2, so the
    environment term is A + (B - A) min(2r, 1) (LUT (255, 0): bsf = f0)."""
    A, B = (200, 100, 50), (20, 60, 240)
    env = make_env(sr, spec=(A, B, B), lut=np.full(64, solid(sr, (255, 0, 0)), np.uint32))
    roughs = (0.1, 0.25, 0.4, 0.5, 0.9)
    mats = [material((1.0, 1.0, 1.0), 1.0, r) for r in roughs]
    scene, osc = probe_scene(sr, orc, mats, env, light_color=(0.0, 0.0, 0.0))
    p = np.array([0.0, 0.0, -5.0])
    n = unit([0.0, 1.0, 0.0])
    cam = camera_at(sr, p + 5.0 * unit([0.2, 1.0, 0.1]))
    a, bb = np.array(A) / 255.0, np.array(B) / 255.0
    for i, r in enumerate(roughs):
        _, c64 = orc.shade_probe(osc, cam, i, n, p)
        t = min(2 * r, 1.0)
        np.testing.assert_allclose(c64, a + (bb - a) * t, rtol=1e-6, err_msg=f"roughness {r}")


def test_cubemap_trilinear_between_constant_mips(sr, orc):
    """sample_cubemap_trilinear_rgb: level t in [0, 1] between constant mips
    A and B gives A + (B - A) t; levels clamp to [0, max]."""
    A, B = (250, 10, 120), (30, 200, 60)
    env = make_env(sr, spec=(A, B))
    _, osc = probe_scene(sr, orc, [material((1, 1, 1), 0, 0.5)], env)
    a, b = np.array(A) / 255.0, np.array(B) / 255.0
    for d in (unit([1, 0.2, 0.1]), unit([-0.3, -1, 0.2]), unit([0.1, 0.3, -1])):
        for t in (0.0, 0.125, 0.5, 0.7, 1.0):
            got = orc.env_sample(osc, "specular", d, t)
            np.testing.assert_allclose(got, a + (b - a) * t, rtol=1e-6, atol=1e-7)
        np.testing.assert_array_equal(orc.env_sample(osc, "specular", d, -3.0), orc.env_sample(osc, "specular", d, 0.0))
        np.testing.assert_array_equal(orc.env_sample(osc, "specular", d, 9.0), orc.env_sample(osc, "specular", d, 1.0))


def test_lut_bilinear_reproduces_a_ramp(sr, orc):
    """sample_bilinear_rgb on a 16^2 LUT whose red byte is 16 x and green
    byte 16 y: inside the texel centres it returns the ramp 16 (s N - 1/2) / 255
    exactly up to rounding; outside, clamp-to-edge holds the edge texel."""
    N = 16
    y, x = np.mgrid[0:N, 0:N]
    rgba = np.stack([16 * x, 16 * y, np.zeros_like(x), np.full_like(x, 255)], -1).astype(np.uint8)
    lut = sr.pack_rgba(rgba.reshape(-1, 4)).astype(np.uint32)
    env = make_env(sr, lut=lut)
    _, osc = probe_scene(sr, orc, [material((1, 1, 1), 0, 0.5)], env)
    rng = np.random.Generator(np.random.PCG64(3))

    def ramp(s):
        return 16.0 * min(max(s * N - 0.5, 0.0), N - 1.0) / 255.0

    for u, v in list(rng.uniform(0, 1, (40, 2))) + [(0.0, 0.0), (1.0, 1.0), (0.01, 0.99), (0.5, 0.03125)]:
        got = orc.env_sample(osc, "lut", (u, v))
        np.testing.assert_allclose(got[:2], [ramp(u), ramp(v)], atol=2e-6, err_msg=f"uv ({u}, {v})")
        assert got[2] == 0.0


@pytest.mark.parametrize("yaw,pitch", [(0.0, 0.0), (0.7, 0.0), (math.pi / 2, 0.0), (0.0, 0.35), (-1.1, -0.4)])
def test_skybox_direction(sr, orc, yaw, pitch):
    """compute_skybox's direction for pixel (x, y) equals the camera basis
    combination x_c right + y_c up + forward (normalized) with
    x_c = ndc_x tan(fov/2) aspect, y_c = ndc_y tan(fov/2) and
    forward = (sin yaw cos pitch, -sin pitch, -cos yaw cos pitch),
    right = (cos yaw, 0, sin yaw), up = right x forward (the reference's
    positive pitch looks down)."""
    W, H, fov = 320, 200, 1.0
    cam = sr.RenderCamera.from_pose([1.0, 2.0, 3.0], yaw, pitch, fov, W, H, 100.0).update_matrices()
    fwd = np.array([math.sin(yaw) * math.cos(pitch), -math.sin(pitch), -math.cos(yaw) * math.cos(pitch)])
    right = np.array([math.cos(yaw), 0.0, math.sin(yaw)])
    up = np.cross(right, fwd)
    ty = math.tan(fov / 2)
    for (x, y) in ((0, 0), (158, 98), (318, 198), (40, 150)):
        got = orc.skybox_dir(cam, x, y)
        for lane in range(4):
            px, py = x + (lane & 1) + 0.5, y + (lane >> 1) + 0.5
            ndx, ndy = 2 * px / W - 1, 1 - 2 * py / H
            want = unit(ndx * ty * (W / H) * right + ndy * ty * up + fwd)
            np.testing.assert_allclose(got[lane], want, atol=3e-6, err_msg=f"pixel ({x}, {y}) lane {lane}")


def exposure_reference(lum, dt, ev):
    s = np.sort(np.log2(np.maximum(np.asarray(lum, np.float64), 1e-4)))
    n = len(s)
    trim = min(int(math.floor(n * 0.1)), (n - 1) // 2)
    mean = s[trim:n - trim].mean()
    meter = (0.45 * 0.2) / (1.0 + 0.2 - 0.45)
    target = min(max(2.0 ** (math.log2(meter) - mean), 0.05), 32.0)
    alpha = 1.0 - math.exp(-max(dt, 0.0))
    ev2 = ev + (math.log2(target) - ev) * alpha
    return ev2, target, 2.0 ** ev2


@pytest.mark.parametrize("case", ["twenty", "ten", "three", "one", "dark", "bright", "zeros", "settle", "still"])
def test_exposure_step(orc, case):
    rng = np.random.Generator(np.random.PCG64(9))
    lum, dt, ev = rng.uniform(0.01, 2.0, 20), 0.25, 1.0
    if case == "ten":
        lum = rng.uniform(0.1, 0.5, 10)
    elif case == "three":
        lum = np.array([0.2, 5.0, 0.01])
    elif case == "one":
        lum = np.array([0.3])
    elif case == "dark":
        lum = np.full(30, 1e-6)             # target clamps at 32
    elif case == "bright":
        lum = np.full(12, 500.0)            # target clamps at 0.05
    elif case == "zeros":
        lum = np.concatenate([np.zeros(5), rng.uniform(0.1, 1, 15)])
    elif case == "settle":
        dt = 1e6                            # alpha = 1: ev jumps to the target
    elif case == "still":
        dt = 0.0                            # alpha = 0: ev stays
    lum = np.asarray(lum, np.float32)
    got = orc.exposure_step(lum, dt, ev, 2.0)
    want = exposure_reference(lum, dt, ev)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-6)
    if case == "dark":
        assert got[1] == 32.0
    if case == "bright":
        assert got[1] == np.float32(0.05)
    if case == "still":
        assert got[0] == 1.0 and got[2] == 2.0


def test_exposure_trim_removes_outliers(orc):
    """Two extreme tiles among twenty fall in the trimmed 10% at each end and
    do not move the exposure."""
    base = np.full(20, 0.25, np.float32)
    out = base.copy()
    out[3], out[11] = 1e4, 1e-8
    a = orc.exposure_step(base, 0.5, 0.0, 2.0)
    b = orc.exposure_step(out, 0.5, 0.0, 2.0)
    assert a == b
