"""CPU closed-form tests of the oracle's samplers, voxel lookup, cubemap
addressing, tonemap/quantize, camera and mip generation (and of the
product's host-side camera and mip code against the oracle's)."""
import ctypes as C

import numpy as np
import pytest


def tex_desc(sr, w, h, ty, texels):
    d = sr.sr_texture_desc()
    d.width, d.height, d.type = w, h, ty
    d.texels = texels.ctypes.data_as(C.POINTER(C.c_uint32))
    return d


def sample4(sr, orc, tex, w, h, ty, wrap_s, wrap_t, u, v, dd):
    d = tex_desc(sr, w, h, ty, tex)
    out = np.zeros(12, np.float32)
    f4 = lambda a: (C.c_float * 4)(*[float(x) for x in a])  # noqa: E731
    L = orc.lib()
    L.orc_sample4.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    mip = L.orc_sample4(C.addressof(d), wrap_s, wrap_t, f4(u), f4(v), f4(dd), out.ctypes.data)
    return mip, out.reshape(4, 3)


def wrap_index(t, n, mode, sr):
    if mode == sr.SR_WRAP_REPEAT:
        return t % n
    if mode == sr.SR_WRAP_MIRRORED_REPEAT:
        m = t % (2 * n)
        return min(min(m, 2 * n - m), n - 1)
    return max(0, min(t, n - 1))


@pytest.mark.parametrize("mode", ["repeat", "mirror", "clamp"])
def test_nearest_sampler_wrap_modes(sr, orc, mode):
    """sample4_rgb at mip 0: texel = wrap(floor(u * w)), closed form."""
    rng = np.random.Generator(np.random.PCG64(3))
    m = {"repeat": sr.SR_WRAP_REPEAT, "mirror": sr.SR_WRAP_MIRRORED_REPEAT, "clamp": sr.SR_WRAP_CLAMP_TO_EDGE}[mode]
    w, h = 13, 7
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    tex = sr.pack_rgba(rgba).ravel()
    for _ in range(200):
        # texel-centred coordinates, away from boundaries, including negatives
        tx = rng.integers(-3 * w, 3 * w, 4) + 0.5
        ty = rng.integers(-3 * h, 3 * h, 4) + 0.5
        mip, rgb = sample4(sr, orc, tex, w, h, sr.SR_TEX_LINEAR, m, m, tx / w, ty / h, [0, 0, 0, 0])
        assert mip == 0
        for lane in range(4):
            x = wrap_index(int(np.floor(tx[lane])), w, m, sr)
            y = wrap_index(int(np.floor(ty[lane])), h, m, sr)
            np.testing.assert_array_equal(rgb[lane], rgba[y, x, :3].astype(np.float32) / np.float32(255))


def test_mip_level_selection(sr, orc):
    """compute_mip_level = ilog2(max((dx^2 + dy^2) / 2, 1)) >> 1, clamped to
    the last level, with dx^2 = (du_dx w)^2 + (dv_dx h)^2 and
    dy^2 = (du_dy w)^2 + (dv_dy h)^2 (src/texture.rs:851-863)."""
    rng = np.random.Generator(np.random.PCG64(4))
    w, h = 64, 32  # 7 levels
    tex = np.arange(w * h * 2, dtype=np.uint32)[: w * h * 2][: w * h]
    for _ in range(300):
        dd = rng.uniform(-3, 3, 4) * 10.0 ** rng.uniform(-4, 0)
        dx2 = (dd[0] * w) ** 2 + (dd[2] * h) ** 2
        dy2 = (dd[1] * w) ** 2 + (dd[3] * h) ** 2
        fp = max((dx2 + dy2) / 2, 1.0)
        if abs(np.log2(fp) - np.round(np.log2(fp))) < 1e-4:
            continue
        want = min(int(np.floor(np.log2(fp))) >> 1, 6)
        mip, _ = sample4(sr, orc, tex, w, h, sr.SR_TEX_LINEAR, 0, 0, [0.1] * 4, [0.1] * 4, dd)
        assert mip == want


def test_mip_chain_closed_forms_and_product_agreement(sr, orc):
    """Level sizes halve with max(1) for 1 + ilog2(max(w, h)) levels; a
    constant linear texture stays constant; the metallic-roughness rule keeps
    red/alpha of the top-left texel, roughness = sqrt(mean of squares),
    metallic = mean.  The product's mip builder (used for the GPU upload)
    agrees bit for bit with the oracle's."""
    rng = np.random.Generator(np.random.PCG64(9))
    L = sr.lib()
    for (w, h) in [(5, 3), (16, 16), (1, 7), (37, 64), (256, 1)]:
        for ty in (sr.SR_TEX_SRGB, sr.SR_TEX_NORMAL, sr.SR_TEX_METALLIC_ROUGHNESS, sr.SR_TEX_LINEAR):
            data = sr.pack_rgba(rng.integers(0, 256, (h, w, 4), dtype=np.uint8)).ravel()
            d = tex_desc(sr, w, h, ty, data)
            ref = orc.mip_chain(d)
            tot, lv = C.c_uint32(), C.c_uint32()
            off, ws, hs = (C.c_uint32 * 32)(), (C.c_uint32 * 32)(), (C.c_uint32 * 32)()
            sr.check(L.sr_texture_mip_chain(C.byref(d), None, C.byref(tot), C.byref(lv), off, ws, hs))
            got = np.empty(tot.value, np.uint32)
            sr.check(L.sr_texture_mip_chain(C.byref(d), got.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(tot),
                                            C.byref(lv), off, ws, hs))
            np.testing.assert_array_equal(got, ref)
            n = 1 + int(np.floor(np.log2(max(w, h))))
            assert lv.value == n
            for k in range(n):
                assert ws[k] == max(w >> k, 1) and hs[k] == max(h >> k, 1)
    const = np.full(8 * 8, sr.pack_rgba(np.array([[10, 200, 30, 255]], np.uint8))[0], np.uint32)
    chain = orc.mip_chain(tex_desc(sr, 8, 8, sr.SR_TEX_LINEAR, const))
    # x/255*255 rounds back to x for these values, so levels stay constant
    assert np.all(chain == const[0])
    vals = np.array([[[0, 40, 100, 7], [0, 120, 20, 9]], [[0, 200, 60, 11], [0, 10, 250, 13]]], np.uint8)
    chain = orc.mip_chain(tex_desc(sr, 2, 2, sr.SR_TEX_METALLIC_ROUGHNESS, sr.pack_rgba(vals).ravel()))
    top = sr.unpack_rgba(chain[4:5])[0].astype(float)
    g = vals[..., 1].astype(np.float64) / 255
    b = vals[..., 2].astype(np.float64) / 255
    assert top[0] == 0 and top[3] == 7
    assert abs(top[1] - np.sqrt((g ** 2).mean()) * 255) <= 1.0
    assert abs(top[2] - b.mean() * 255) <= 1.0


def test_voxel_lookup_linear_and_clamp(sr, orc):
    """Trilinear lookup reproduces a linear function of voxel coordinates,
    clamps beyond the far faces, and below the near faces mixes voxels 0 and
    1 with the fractional part of the unclamped coordinate (`as usize`
    saturates negatives to 0; src/voxelgrid.rs:271-304)."""
    n = 8
    lo, hi = np.array([-2.0, 0.0, 1.0]), np.array([6.0, 4.0, 9.0])
    size = (hi - lo) / n
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f = 0.5 * i - 0.25 * j + 0.125 * k + 1.0  # indexed [x, y, z]
    grid = np.zeros((n, n, n, 4, 4), np.float32)  # [z, y, x, coeff, ch]
    grid[..., 0, 0] = np.transpose(f, (2, 1, 0))
    grid[..., 2, 3] = 3.0
    L = orc.lib()
    L.orc_voxel_lookup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
    lo32, hi32 = lo.astype(np.float32), hi.astype(np.float32)
    rng = np.random.Generator(np.random.PCG64(2))
    for _ in range(50):
        vc = rng.uniform([0, 0, 0], [n - 1, n - 1, n - 1], (4, 3))
        pos = (lo + vc * size).astype(np.float32)
        rgb = np.zeros(48, np.float32)
        w = np.zeros(16, np.float32)
        L.orc_voxel_lookup(n, n, n, lo32.ctypes.data, hi32.ctypes.data, grid.ctypes.data, pos.ctypes.data,
                           rgb.ctypes.data, w.ctypes.data)
        rgb = rgb.reshape(4, 3, 4)
        want = 0.5 * vc[:, 0] - 0.25 * vc[:, 1] + 0.125 * vc[:, 2] + 1.0
        np.testing.assert_allclose(rgb[0, 0], want, rtol=2e-6, atol=2e-6)
        np.testing.assert_allclose(w.reshape(4, 4)[2], 3.0, rtol=1e-6)
    # far side: clamped to the last voxel; near side: the saturating cast quirk
    pos = np.array([[100, 100, 100], [-2 - 0.3 * size[0], 0, 1], [-2, 0, 1], [5.9, 3.9, 8.9]], np.float32)
    rgb = np.zeros(48, np.float32)
    w = np.zeros(16, np.float32)
    L.orc_voxel_lookup(n, n, n, lo32.ctypes.data, hi32.ctypes.data, grid.ctypes.data, pos.ctypes.data,
                       rgb.ctypes.data, w.ctypes.data)
    rgb = rgb.reshape(4, 3, 4)
    assert abs(rgb[0, 0, 0] - f[n - 1, n - 1, n - 1]) < 1e-5
    # x coordinate -0.3 voxels: x0 = 0, x1 = 1, fx = frac(-0.3) = 0.7
    assert abs(rgb[0, 0, 1] - (f[0, 0, 0] * 0.3 + f[1, 0, 0] * 0.7)) < 1e-5
    assert abs(rgb[0, 0, 2] - f[0, 0, 0]) < 1e-6


def test_cubemap_face_selection(sr, orc):
    """cubemap_uv_from_normal: major axis selects the face (+X,-X,+Y,-Y,+Z,-Z),
    uv = 0.5 * (sc / |ma| + 1) with the reference's sign table."""
    L = orc.lib()
    L.orc_cubemap_uv.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.Generator(np.random.PCG64(6))
    for _ in range(100):
        n = rng.normal(size=(4, 3)).astype(np.float32)
        a = np.abs(n)
        srt = np.sort(a, 1)
        if np.any(srt[:, 2] - srt[:, 1] < 1e-3):
            continue
        uv = np.zeros(8, np.float32)
        face = np.zeros(4, np.uint32)
        L.orc_cubemap_uv(n.ctypes.data, uv.ctypes.data, face.ctypes.data)
        for lane in range(4):
            x, y, z = n[lane].astype(np.float64)
            ax, ay, az = abs(x), abs(y), abs(z)
            if ax >= ay and ax >= az:
                f, u, v, m = (0 if x >= 0 else 1), -z * np.sign(x), -y, ax
            elif ay >= az:
                f, u, v, m = (2 if y >= 0 else 3), x, z * np.sign(y), ay
            else:
                f, u, v, m = (4 if z >= 0 else 5), x * np.sign(z), -y, az
            assert face[lane] == f
            np.testing.assert_allclose(uv[2 * lane:2 * lane + 2], [0.5 * (u / m + 1), 0.5 * (v / m + 1)], atol=1e-6)


def test_tonemap_quantize_closed_form(orc):
    """blit: x = c * exposure; y = x / (x + 0.2) * 1.2; byte = trunc(255 y)
    saturating, NaN -> 0 (src/util.rs:37-41, src/renderer.rs:327-344)."""
    rng = np.random.Generator(np.random.PCG64(7))
    for c in np.concatenate([rng.uniform(0, 4, 500), [0.0, 1e-8, 1e6, np.inf]]):
        for e in (0.5, 2.0, 7.3):
            x = np.float64(np.float32(c)) * e
            y = x / (x + 0.2) * 1.2 if np.isfinite(x) else np.nan
            want = 0 if not np.isfinite(y) else min(255, int(np.floor(255 * y)))
            got = orc.quantize(float(c), e)
            frac = 255 * y - np.floor(255 * y) if np.isfinite(y) else 0.5
            if 1e-4 < frac < 1 - 1e-4:
                assert got == want, (c, e, got, want)
    assert orc.quantize(float("nan"), 2.0) == 0
    # negative input: x = -2, x / (x + 0.2) * 1.2 = 1.33 -> saturates to 255
    assert orc.quantize(-1.0, 2.0) == 255
    assert orc.quantize(-0.05, 2.0) == 0
    assert orc.quantize(1e9, 2.0) == 255


def test_camera_closed_forms_and_product_agreement(sr, orc):
    """RenderCamera (src/rendercamera.rs): yaw = pitch = 0 gives a pure
    translation view; yaw of +90 deg looks along +x; the projection follows
    perspective_rh; the product's camera equals the oracle's bit for bit."""
    pos = [1.5, -2.0, 3.0]
    cam = sr.RenderCamera.from_pose(pos, 0.0, 0.0, np.pi / 3, 640, 480, 100.0).update_matrices()
    V = cam.matrix("view_matrix")
    np.testing.assert_allclose(V, [[1, 0, 0, -1.5], [0, 1, 0, 2.0], [0, 0, 1, -3.0], [0, 0, 0, 1]], atol=1e-7)
    P = cam.matrix("projection_matrix")
    f = 1 / np.tan(np.pi / 6)
    near, far = 1.0, 100.0
    np.testing.assert_allclose(P, [[f / (640 / 480), 0, 0, 0], [0, f, 0, 0],
                                   [0, 0, far / (near - far), near * far / (near - far)], [0, 0, -1, 0]],
                               rtol=1e-6, atol=1e-6)
    c90 = sr.RenderCamera.new([0, 0, 0], [5, 0, 0], np.pi / 2, 100, 100, 50).update_matrices()
    p = c90.matrix("view_project_matrix") @ np.array([10.0, 0.0, 0.0, 1.0])
    assert abs(p[0] / p[3]) < 1e-6 and abs(p[1] / p[3]) < 1e-6 and p[3] > 0
    # pitch quirk (Q1): pitch = +30 deg looks down
    cp = sr.RenderCamera.from_pose([0, 0, 0], 0.0, np.pi / 6, np.pi / 2, 100, 100, 50).update_matrices()
    q = cp.matrix("view_project_matrix") @ np.array([0.0, -np.tan(np.pi / 6) * 10, -10.0, 1.0])
    assert abs(q[1] / q[3]) < 1e-5
    for look in ([0, 0, -1], [3, 1, -2], [-4, -2, 1]):
        a = sr.RenderCamera.new([0.5, 1.0, 2.0], look, 1.1, 800, 600, 40.0).update_matrices().c
        b = orc.camera_update(orc.camera_new(sr.sr_camera, [0.5, 1.0, 2.0], look, 1.1, 800, 600, 40.0))
        assert bytes(a) == bytes(b)


def test_light_default(sr):
    d, c = sr.default_light()
    v = np.array([-0.2, 1.0, 0.5])
    np.testing.assert_allclose(d, v / np.linalg.norm(v), rtol=1e-6)
    np.testing.assert_allclose(c, [5.0, 5.0, 4.75], rtol=1e-7)
