"""CPU tests of the product library's host side: C ABI exports, argument
validation, scene bounds, and the load-time input producers (environment
and voxel grid) against closed forms.  No GPU call is made here."""
import ctypes as C
import os
import re

import numpy as np

from util import dummy_voxels

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    hdr = open(os.path.join(REPO, "include", "swraster.h")).read()
    return re.findall(r"SR_API\s+[\w\s\*]+?\b(sr_\w+)\s*\(", hdr)


def test_library_exports_every_declared_symbol(sr):
    names = declared_functions()
    assert len(names) >= 40
    L = C.CDLL(sr.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    assert set(names) == set(sr._SIGS), "ctypes signatures out of sync with the header"


def test_kernel_names_match_between_host_and_device(sr):
    """Every kernel the host registers exists under the same name in the
    gfx950 code object.  A kernel parameter typed with an address space
    (address_space(1) pointers, which the device pass uses) is mangled into
    the device-side name only (`U3AS1`), so the host's launch would find no
    kernel and the runtime aborts: kernels take such pointers inside
    structs."""
    data = open(sr.LIB_PATH, "rb").read()
    kernels = set(re.findall(rb"_ZN2sr[0-9]+k_[A-Za-z0-9_]+", data))
    assert len(kernels) >= 20
    with_as = sorted(k.decode() for k in kernels if b"U3AS" in k)
    assert not with_as, f"kernels with address-space parameters: {with_as}"


def test_version_and_device_count(sr):
    a, b = C.c_int32(), C.c_int32()
    sr.check(sr.lib().sr_version(C.byref(a), C.byref(b)))
    assert (a.value, b.value) >= (0, 1)
    assert sr.device_count() >= 0


def test_argument_validation(sr):
    L = sr.lib()
    h = C.c_void_p()
    assert L.sr_renderer_create(0, 10, None, C.byref(h)) == -1
    assert L.sr_renderer_create(64, 64, None, None) == -1
    assert L.sr_scene_create(None, 0, C.byref(h)) == -1
    assert "null" in L.sr_last_error().decode()
    assert L.sr_camera_update_matrices(None) == -1
    assert L.sr_batch_render(None, None, 0) == -1
    # a descriptor whose primitive indexes past its vertices is rejected
    pos = np.zeros((3, 3), np.float32)
    prim = {"positions": pos, "normals": pos, "texcoords": None, "tangents": None,
            "indices": np.array([0, 1, 7], np.uint32), "material": 0}
    scene = sr.SceneData([prim], [(0, 1)], [(np.eye(4, dtype=np.float32), 0)],
                         [{"base_color": (1, 1, 1, 1)}], [])
    from util import dummy_environment
    scene.attach(dummy_environment(sr), sr.VoxelGrid((2, 2, 2), [0, 0, 0], [1, 1, 1], np.ones(8, np.float32)))
    assert L.sr_scene_create(C.byref(scene.desc), 0, C.byref(h)) == -1
    assert "out of range" in L.sr_last_error().decode()
    prim["indices"] = np.array([0, 1, 2], np.uint32)
    prim["material"] = 3
    scene = sr.SceneData([prim], [(0, 1)], [(np.eye(4, dtype=np.float32), 0)], [{"base_color": (1, 1, 1, 1)}], [])
    scene.attach(dummy_environment(sr), sr.VoxelGrid((2, 2, 2), [0, 0, 0], [1, 1, 1], np.ones(8, np.float32)))
    assert L.sr_scene_create(C.byref(scene.desc), 0, C.byref(h)) == -1
    assert "material" in L.sr_last_error().decode()


def test_scene_bounds_match_numpy(sr, scenes):
    scene, _ = scenes.make_c1()
    mn, mx, ce, dg = scene.bounds()
    pts = []
    for (t, m) in scene.nodes:
        p = scene.primitives[m]["positions"].astype(np.float64)
        pts.append((np.c_[p, np.ones(len(p))] @ np.asarray(t, np.float64).T)[:, :3])
    pts = np.concatenate(pts)
    np.testing.assert_allclose(mn, pts.min(0), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(mx, pts.max(0), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(ce, (pts.min(0) + pts.max(0)) / 2, rtol=1e-6, atol=1e-6)
    assert abs(dg - np.linalg.norm(pts.max(0) - pts.min(0))) < 1e-4


def test_environment_uniform_cubemap_closed_forms(sr):
    """For a uniform sky of sRGB value c: irradiance SH band 0 equals
    pi * L (L = linear radiance, after the Lambert convolution and basis
    fold of src/texture.rs:316-327), bands 1-3 vanish, and every specular
    texel is L; the BRDF LUT tends to (1, 0) at normal incidence on a smooth
    surface and stays within [0, 1]."""
    size = 32
    c = 150
    px = sr.pack_rgba(np.array([[c, c, c, 255]], np.uint8))[0]
    faces = np.full((6, size, size), px, np.uint32)
    env = sr.sr_environment()
    sr.check(sr.lib().sr_prepare_environment(faces.ctypes.data_as(C.POINTER(C.c_uint32)), size, 0, C.byref(env)))
    try:
        s = c / 255.0
        lin = ((s + 0.055) / 1.055) ** 2.4
        sh = np.array([[env.irradiance_sh[i][k] for k in range(3)] for i in range(4)])
        np.testing.assert_allclose(sh[0], np.pi * lin, rtol=2e-3)
        assert np.all(np.abs(sh[1:]) < 2e-3 * lin)
        n = env.specular_mips * 6 * size * size
        spec = sr.unpack_rgba(np.ctypeslib.as_array(env.specular, shape=(n,)))[:, :3].astype(int)
        assert env.specular_mips == 1 + int(np.log2(size))
        assert np.all(np.abs(spec - int(lin * 255)) <= 1)
        lut = sr.unpack_rgba(np.ctypeslib.as_array(env.brdf_lut, shape=(128 * 128,))).reshape(128, 128, 4)
        # x = N.V, y = roughness; texel (x = 127, y = 0): smooth, normal incidence
        assert lut[0, 127, 0] >= 250 and lut[0, 127, 1] <= 3
        assert np.all(lut[..., 2] == 0) and np.all(lut[..., 3] == 255)
    finally:
        sr.lib().sr_environment_free(C.byref(env))


def test_voxel_sun_visibility_shadow(sr):
    """An opaque horizontal plate over a floor: voxels above the plate see
    the sun (1), the floor voxels under it are shadowed (< 1), and values
    stay in [0, 1] (src/gi.rs:267-314, blur + square src/voxelgrid.rs:371-419)."""
    def quad(y, half, mat):
        p = np.array([[-half, y, -half], [-half, y, half], [half, y, -half], [half, y, half]], np.float32)
        return {"positions": p, "normals": np.tile(np.float32([0, 1, 0]), (4, 1)), "texcoords": None,
                "tangents": None, "indices": np.array([0, 1, 2, 2, 1, 3], np.uint32), "material": mat}
    prims = [quad(0.0, 4.0, 0), quad(2.0, 1.0, 0)]
    scene = sr.SceneData(prims, [(0, 2)], [(np.eye(4, dtype=np.float32), 0)], [{"base_color": (1, 1, 1, 1)}], [])
    g = sr.sr_voxel_grid()
    sr.check(sr.lib().sr_prepare_voxel_grid(C.byref(scene.desc), 32, 0, C.byref(g)))
    try:
        light = np.ctypeslib.as_array(g.light_intensity, shape=(32 ** 3,)).reshape(32, 32, 32)  # [z, y, x]
        wmin, wmax = np.array(g.world_min[:]), np.array(g.world_max[:])
        assert np.all(light >= 0) and np.all(light <= 1)
        size = (wmax - wmin) / 32

        def at(x, y, z):
            i = ((np.array([x, y, z]) - wmin) / size).astype(int).clip(0, 31)
            return light[i[2], i[1], i[0]]
        assert at(3.5, 1.0, -3.5) > 0.99          # floor far from the plate
        d = np.array([-0.2, 1.0, 0.5]) / np.linalg.norm([-0.2, 1.0, 0.5])
        # only voxels near surfaces are traced: probe the floor point whose sun
        # ray passes through the plate centre
        under = np.array([0.0, 2.0, 0.0]) - d * (2.0 / d[1])
        assert abs(under[1]) < 1e-9
        assert at(under[0], 0.01, under[2]) < 0.5
    finally:
        sr.lib().sr_voxel_grid_free(C.byref(g))


def test_dummy_voxels_helper(sr, scenes):
    scene, _ = scenes.make_c1()
    v = dummy_voxels(sr, scene)
    assert v.dims == (2, 2, 2)


def _fma32(a, b, c):
    """Correctly rounded float32 fused multiply-add, from exact rationals."""
    from fractions import Fraction
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    cand = np.float32(float(exact))
    best = None
    for f in (np.nextafter(cand, np.float32(-np.inf)), cand, np.nextafter(cand, np.float32(np.inf))):
        d = abs(Fraction(float(f)) - exact)
        if best is None or d < best[0] or (d == best[0] and (f.view(np.uint32) & 1) == 0):
            best = (d, f)
    return best[1]


def test_div255_identity():
    """The shading kernels turn a texel byte n into n / 255 with
    q = n * RN(1/255), then q + (n - 255 q) * RN(1/255) in two fused
    multiply-adds; that equals IEEE division for every n in [0, 255]."""
    inv = np.float32(1.0) / np.float32(255.0)
    plain_wrong = 0
    for n in range(256):
        x = np.float32(n)
        ref = x / np.float32(255.0)
        q = x * inv
        plain_wrong += int(q != ref)
        q2 = _fma32(_fma32(np.float32(-255.0), q, x), inv, q)
        assert q2 == ref, n
    assert plain_wrong > 0          # the correction step is needed


def test_power_of_two_reciprocal_division():
    """x / 2^k equals x * 2^-k (both round the same exact value), which the
    texture wrap uses for power-of-two sizes; spot-check across magnitudes."""
    rng = np.random.Generator(np.random.PCG64(4))
    x = (rng.standard_normal(100000) * np.exp2(rng.integers(-140, 120, 100000))).astype(np.float32)
    for k in (0, 1, 5, 11, 12, 20):
        d = np.float32(2.0 ** k)
        assert np.array_equal((x / d).view(np.uint32), (x * np.float32(2.0 ** -k)).view(np.uint32))


def test_srgb_decode_power_is_correctly_rounded():
    """The device IBL precompute evaluates the sRGB decode's x^2.4 in double
    and rounds once.  Over every float input the decode can take
    (x in ((0.04045 + 0.055) / 1.055, 1]) that equals x^2.4 computed in
    80-bit extended precision and rounded to float, and no exact value lies
    within 2.5e-16 (relative) of a float rounding boundary, so any double
    power accurate to about one ulp gives the same float."""
    lo = np.float32((np.float32(0.04045) + np.float32(0.055)) / np.float32(1.055))
    b0, b1 = int(lo.view(np.uint32)) - 16, int(np.float32(1.0).view(np.uint32))
    e24 = np.longdouble(np.float32(2.4))
    assert np.finfo(np.longdouble).nmant >= 63, "needs x87 extended precision"
    mismatches, closest = 0, 1.0
    for s in range(b0, b1 + 1, 1 << 22):
        x = np.arange(s, min(s + (1 << 22), b1 + 1), dtype=np.uint32).view(np.float32)
        d = np.power(x.astype(np.float64), np.float64(np.float32(2.4))).astype(np.float32)
        e = np.power(x.astype(np.longdouble), e24)
        f = e.astype(np.float32)
        mismatches += int((f != d).sum())
        toward = np.where(e > f.astype(np.longdouble), np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
        mid = (f.astype(np.longdouble) + np.nextafter(f, toward).astype(np.longdouble)) / 2
        closest = min(closest, float((np.abs(e - mid) / e).min()))
    assert mismatches == 0
    assert closest > 2.5e-16
