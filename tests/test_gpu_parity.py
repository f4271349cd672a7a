"""GPU parity: the HIP pipeline against the CPU oracle on the same scenes.

Bar: visibility buffer (canonical triangle id + depth bits) bit-identical;
RGBA8 within +-1 LSB per channel; linear colour and 64px-tile centre
luminance reported (and expected identical, since both sides evaluate the
same f32 expressions with IEEE division and square root and no FMA).
"""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def render_gpu(sr, gscene, cam, w, h, tile=0, dt=0.0):
    r = sr.Renderer(w, h, tile_size=tile)
    r.render_scene(gscene, cam)
    ids, depth = r.read_vbuffer()
    rgb = r.read_color()
    lum = r.read_tile_luminance()
    r.update_auto_exposure(dt)
    px = r.blit_to_buffer()
    st = r.stats()
    exposure = r.exposure
    r.close()
    return {"ids": ids, "depth": depth, "rgb": rgb, "lum": lum, "px": px, "stats": st, "exposure": exposure}


def render_oracle(orc, scene, cam, w, h, dt=0.0, threads=8):
    o = orc.OracleRenderer(w, h)
    o.render_scene(orc.OracleScene(scene), cam, threads=threads)
    f = o.read_frame()
    f["lum"] = o.tile_luminance()
    o.update_auto_exposure(dt)
    f["px"] = o.blit()
    f["exposure"] = o.exposure
    f["setup"] = o.setup_count()
    return f


def compare(sr, g, o, name):
    ids_eq = np.array_equal(g["ids"], o["ids"])
    d_eq = np.array_equal(bits(g["depth"]), bits(o["depth"]))
    nbad = int((g["ids"] != o["ids"]).sum() + (bits(g["depth"]) != bits(o["depth"])).sum())
    if not (ids_eq and d_eq):
        bad = np.argwhere((g["ids"] != o["ids"]) | (bits(g["depth"]) != bits(o["depth"])))
        y, x = bad[0]
        print(f"[{name}] first mismatch at ({x},{y}): gpu id {g['ids'][y, x]} depth {g['depth'][y, x]!r}, "
              f"oracle id {o['ids'][y, x]} depth {o['depth'][y, x]!r}; {nbad} mismatching values")
    assert ids_eq and d_eq, f"{name}: visibility buffer differs in {nbad} values"
    rgba_g, rgba_o = sr.unpack_rgba(g["px"]).astype(int), sr.unpack_rgba(o["px"]).astype(int)
    lsb = int(np.abs(rgba_g - rgba_o).max())
    rgb_err = float(np.nanmax(np.abs(g["rgb"].astype(np.float64) - o["rgb"])))
    print(f"[{name}] vbuffer identical; covered {(g['ids'] != sr.NO_ID).sum()} px; max RGBA8 diff {lsb} LSB; "
          f"max linear diff {rgb_err:.3g}; colour bit-identical {np.array_equal(bits(g['rgb']), bits(o['rgb']))}")
    assert lsb <= 1, f"{name}: RGBA8 differs by {lsb} LSB"
    assert np.array_equal(bits(g["lum"]), bits(o["lum"])) or np.allclose(g["lum"], o["lum"], rtol=1e-6)
    assert g["exposure"] == o["exposure"]


@pytest.fixture(scope="module")
def c1(sr, scenes, env):
    scene, spec = scenes.make_c1()
    scene.prepare(env, cache_key=scene.name)
    return scene, spec, sr.Scene(scene)


def test_c1_parity(sr, orc, scenes, c1):
    scene, spec, gscene = c1
    cam = scenes.make_camera(spec)
    w, h = spec["width"], spec["height"]
    g = render_gpu(sr, gscene, cam, w, h)
    o = render_oracle(orc, scene, cam, w, h)
    assert g["stats"]["setup_triangles"] == o["setup"]
    compare(sr, g, o, "c1")


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_c1_tile_sizes_identical(sr, orc, scenes, c1, tile):
    scene, spec, gscene = c1
    cam = scenes.make_camera(spec, 640, 360)
    g = render_gpu(sr, gscene, cam, 640, 360, tile=tile)
    o = render_oracle(orc, scene, cam, 640, 360)
    compare(sr, g, o, f"c1 tile {tile}")


@pytest.mark.parametrize("size", [(97, 61), (333, 202), (64, 64), (1, 1), (130, 2)])
def test_c1_odd_viewports(sr, orc, scenes, c1, size):
    scene, spec, gscene = c1
    w, h = size
    cam = scenes.make_camera(spec, w, h)
    compare(sr, render_gpu(sr, gscene, cam, w, h), render_oracle(orc, scene, cam, w, h), f"c1 {w}x{h}")


@pytest.mark.parametrize("seed", range(8))
def test_c1_random_cameras(sr, orc, scenes, c1, seed):
    """Random poses around and inside the sphere cluster (positions in a
    box twice the cluster's size, looking at a random point near its centre
    give or take 0.3 rad of yaw and pitch, field of view 20-110 degrees, far
    5-60, so near and far planes cut spheres) at random viewport sizes up to
    400x300, 8-, 16- or 32-px tiles: visibility buffer, setup count and
    pixels equal the oracle's in every case."""
    scene, spec, gscene = c1
    rng = np.random.Generator(np.random.PCG64(1000 + seed))
    w, h = int(rng.integers(16, 401)), int(rng.integers(16, 301))
    pos = rng.uniform(-9.0, 9.0, 3)
    yaw, pitch = scenes.look_pose(pos, rng.uniform(-3.0, 3.0, 3))
    s = scenes.camera_spec(pos, yaw + rng.uniform(-0.3, 0.3), float(np.clip(pitch + rng.uniform(-0.3, 0.3), -1.4, 1.4)),
                           rng.uniform(np.radians(20), np.radians(110)), w, h, rng.uniform(5.0, 60.0))
    cam = scenes.make_camera(s)
    tile = int(rng.choice([8, 16, 32]))
    g = render_gpu(sr, gscene, cam, w, h, tile=tile)
    o = render_oracle(orc, scene, cam, w, h)
    assert g["stats"]["setup_triangles"] == o["setup"]
    compare(sr, g, o, f"c1 camera {seed} ({w}x{h}, tile {tile})")


@pytest.mark.parametrize("name,over", [("fov 1 deg", {"fov": np.radians(1.0)}),
                                       ("fov 179 deg", {"fov": np.radians(179.0)}),
                                       ("far 1e6", {"far": 1e6}), ("far 1e30", {"far": 1e30}),
                                       ("far 0.05", {"far": 0.05}), ("inside a sphere", {"position": [2.0, 1.0, 0.5]})])
def test_c1_extreme_cameras(sr, orc, scenes, c1, name, over):
    """Extreme fields of view (at 1 degree a sphere fills the frame, at 179
    every sphere is below a pixel) and far planes (near = 0.01 far, from
    5e-4 to 1e28), and a camera inside a sphere: the visibility buffer and
    pixels still equal the oracle's."""
    scene, spec, gscene = c1
    s = dict(spec)
    s.update(over)
    cam = scenes.make_camera(s, 320, 180)
    compare(sr, render_gpu(sr, gscene, cam, 320, 180), render_oracle(orc, scene, cam, 320, 180), f"c1 {name}")


def test_c1_inside_cluster_clipping(sr, orc, scenes, c1):
    """Camera inside the sphere cluster: near/side-plane clipping and
    negative depths (view-space cull near plane sits behind the eye)."""
    scene, spec, gscene = c1
    s = dict(spec)
    s["position"] = [0.3, 0.2, 1.0]
    s["yaw"] = 0.4
    cam = scenes.make_camera(s, 800, 600)
    g = render_gpu(sr, gscene, cam, 800, 600)
    o = render_oracle(orc, scene, cam, 800, 600)
    assert o["setup"] > 0
    compare(sr, g, o, "c1 inside")


@pytest.mark.parametrize("seed", [31, 32])
def test_clip_stress(sr, orc, scenes, seed):
    """Thousands of random triangles around the camera, many crossing the
    near and side planes (2-7 fan triangles each), plus near-collinear
    slivers: setup count, visibility buffer and pixels match the oracle."""
    from util import attach_dummy
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 3000
    p = rng.uniform([-6, -5, -12], [6, 5, 1], (n, 3, 3))
    sl = rng.random(n) < 0.2            # slivers: third vertex near the first edge
    p[sl, 2] = p[sl, 0] + (p[sl, 1] - p[sl, 0]) * rng.uniform(0, 1, (sl.sum(), 1)) + rng.normal(0, 1e-4, (sl.sum(), 3))
    pos = p.reshape(-1, 3).astype(np.float32)
    e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
    fn = np.cross(e1, e2)
    fn /= np.linalg.norm(fn, axis=-1, keepdims=True) + 1e-30
    prim = {"positions": pos, "normals": np.repeat(fn, 3, axis=0).astype(np.float32), "texcoords": None,
            "tangents": None, "indices": np.arange(len(pos), dtype=np.uint32), "material": 0}
    scene = sr.SceneData([prim], [(0, 1)], [(np.eye(4, dtype=np.float32), 0)],
                         [{"base_color": (0.8, 0.5, 0.3, 1), "metallic": 0.1, "roughness": 0.6}], [])
    attach_dummy(sr, scene)
    cam = scenes.make_camera({"position": [0.0, 0.0, 0.0], "yaw": 0.0, "pitch": 0.0, "fov": np.pi / 3,
                              "width": 320, "height": 240, "far": 50.0})
    g = render_gpu(sr, sr.Scene(scene), cam, 320, 240)
    o = render_oracle(orc, scene, cam, 320, 240)
    assert o["setup"] > n // 2
    assert g["stats"]["setup_triangles"] == o["setup"]
    compare(sr, g, o, f"clip stress {seed}")


def test_c1_exposure_adaptation(sr, orc, scenes, c1):
    scene, spec, gscene = c1
    cam = scenes.make_camera(spec, 640, 360)
    compare(sr, render_gpu(sr, gscene, cam, 640, 360, dt=0.25), render_oracle(orc, scene, cam, 640, 360, dt=0.25),
            "c1 dt=0.25")


def test_c2_parity(sr, orc, scenes, env):
    scene, spec = scenes.make_c2()
    scene.prepare(env, cache_key=scene.name)
    gscene = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    w, h = spec["width"], spec["height"]
    g = render_gpu(sr, gscene, cam, w, h)
    o = render_oracle(orc, scene, cam, w, h, threads=16)
    compare(sr, g, o, "c2")


@pytest.fixture(scope="module")
def c2(sr, scenes, env):
    scene, spec = scenes.make_c2()
    scene.prepare(env, cache_key=scene.name)
    return scene, spec, sr.Scene(scene)


@pytest.mark.parametrize("seed", range(4))
def test_c2_random_cameras(sr, orc, scenes, c2, seed):
    """Random poses over the heightfield (1,045,458 triangles, 2048^2
    texture): 3-30 units above it, looking at a random point of it give or
    take 0.2 rad, field of view 30-90 degrees, so grazing views clip the
    terrain at the near plane and distant ones shade from small mips; at
    random viewport sizes up to 480x270 and 8-, 16- or 32-px tiles the
    visibility buffer, setup count and pixels equal the oracle's."""
    scene, spec, gscene = c2
    rng = np.random.Generator(np.random.PCG64(2000 + seed))
    w, h = int(rng.integers(32, 481)), int(rng.integers(32, 271))
    pos = np.array([rng.uniform(-45.0, 45.0), rng.uniform(3.0, 30.0), rng.uniform(-45.0, 45.0)])
    target = np.array([rng.uniform(-40.0, 40.0), 0.0, rng.uniform(-40.0, 40.0)])
    yaw, pitch = scenes.look_pose(pos, target)
    s = scenes.camera_spec(pos, yaw + rng.uniform(-0.2, 0.2), float(np.clip(pitch + rng.uniform(-0.2, 0.2), -1.5, 1.5)),
                           rng.uniform(np.radians(30), np.radians(90)), w, h, spec["far"])
    cam = scenes.make_camera(s)
    tile = int(rng.choice([8, 16, 32]))
    g = render_gpu(sr, gscene, cam, w, h, tile=tile)
    o = render_oracle(orc, scene, cam, w, h, threads=16)
    assert g["stats"]["setup_triangles"] == o["setup"]
    compare(sr, g, o, f"c2 camera {seed} ({w}x{h}, tile {tile})")


def test_c3_parity(sr, orc, scenes, env):
    scene, spec = scenes.make_c3()
    scene.prepare(env, cache_key=scene.name)
    gscene = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    w, h = spec["width"], spec["height"]
    g = render_gpu(sr, gscene, cam, w, h)
    o = render_oracle(orc, scene, cam, w, h, threads=16)
    compare(sr, g, o, "c3")


def _two_frames(sr, gscene, cam, w, h):
    """The same frame twice on one renderer: the second takes the raster
    variant chosen from the first frame's statistics."""
    r = sr.Renderer(w, h)
    out = []
    for _ in range(2):
        r.render_scene(gscene, cam)
        ids, depth = r.read_vbuffer()
        out.append((ids, depth, r.stats()["depth_rejection"]))
    r.close()
    return out


def test_c3_depth_rejection_parity(sr, orc, scenes, env):
    """C3's overdraw selects the raster that rejects entries behind the
    tile's stored depth; its visibility buffer is the oracle's."""
    scene, spec = scenes.make_c3()
    scene.prepare(env, cache_key=scene.name)
    gscene = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    w, h = spec["width"], spec["height"]
    (i1, d1, h1), (i2, d2, h2) = _two_frames(sr, gscene, cam, w, h)
    # (the first frame may already take it: a frame that outgrows its
    # buffers is rendered again with the statistics it gathered)
    assert h2 == 1
    o = render_oracle(orc, scene, cam, w, h, threads=16)
    for ids, depth in ((i1, d1), (i2, d2)):
        assert np.array_equal(ids, o["ids"]) and np.array_equal(bits(depth), bits(o["depth"]))


@pytest.mark.parametrize("seed", [5, 6])
def test_depth_rejection_tilted_overdraw(sr, orc, scenes, seed):
    """Large overlapping triangles whose vertices lie at different depths
    (1 to 100, so the rejection bound's margin matters; some with a depth
    ratio above the bound's limit of 16): the depth-rejecting raster matches
    the oracle bit for bit."""
    from util import screen_triangle_scene
    W, H = 640, 360
    rng = np.random.default_rng(seed)
    n = 3000
    c = rng.uniform([0, 0], [W * 16, H * 16], (n, 1, 2))
    ang = rng.uniform(0, 2 * np.pi, (n, 3))
    rad = rng.uniform(60, 250, (n, 3)) * 16
    tris = np.round(c + np.stack([np.cos(ang), np.sin(ang)], -1) * rad[..., None])
    tris = np.clip(tris, -2000, (W + 100) * 16)
    a = (tris[:, 1, 0] - tris[:, 0, 0]) * (tris[:, 2, 1] - tris[:, 0, 1]) - \
        (tris[:, 2, 0] - tris[:, 0, 0]) * (tris[:, 1, 1] - tris[:, 0, 1])
    tris[a > 0] = tris[a > 0][:, ::-1]
    depths = rng.uniform(1.0, 100.0, (n, 3))
    scene, cam = screen_triangle_scene(sr, tris, depths, W, H)
    from util import attach_dummy
    attach_dummy(sr, scene)
    gscene = sr.Scene(scene)
    camera = scenes.make_camera(cam)
    (i1, d1, h1), (i2, d2, h2) = _two_frames(sr, gscene, camera, W, H)
    assert h2 == 1, "the overdraw did not select the depth-rejecting raster"
    o = render_oracle(orc, scene, camera, W, H)
    for ids, depth in ((i1, d1), (i2, d2)):
        assert np.array_equal(ids, o["ids"]) and np.array_equal(bits(depth), bits(o["depth"]))


@pytest.mark.parametrize("W,H,seed", [(256, 144, 7), (130, 66, 8)])
def test_tiny_triangles_parity(sr, orc, scenes, W, H, seed):
    """Dense sub-pixel to 4-px triangles (most of them cover no pixel centre,
    and setup drops those before binning), over the whole screen including
    the partial right and bottom tiles and the 64-px tile seams: the
    visibility buffer is the oracle's bit for bit."""
    from util import attach_dummy, screen_triangle_scene
    rng = np.random.default_rng(seed)
    n = 6000
    c = rng.uniform([-16, -16], [(W + 1) * 16, (H + 1) * 16], (n, 1, 2))
    ang = rng.uniform(0, 2 * np.pi, (n, 3))
    rad = rng.uniform(2, 40, (n, 3))           # subpixels: 1/8 px to 2.5 px from the centre
    tris = np.round(c + np.stack([np.cos(ang), np.sin(ang)], -1) * rad[..., None])
    a = (tris[:, 1, 0] - tris[:, 0, 0]) * (tris[:, 2, 1] - tris[:, 0, 1]) - \
        (tris[:, 2, 0] - tris[:, 0, 0]) * (tris[:, 1, 1] - tris[:, 0, 1])
    tris[a > 0] = tris[a > 0][:, ::-1]
    depths = rng.uniform(1.0, 100.0, (n, 3))
    scene, cam = screen_triangle_scene(sr, tris, depths, W, H)
    attach_dummy(sr, scene)
    gscene = sr.Scene(scene)
    camera = scenes.make_camera(cam)
    r = sr.Renderer(W, H)
    r.render_scene(gscene, camera)
    ids, depth = r.read_vbuffer()
    st = r.stats()
    r.close()
    o = render_oracle(orc, scene, camera, W, H)
    assert np.array_equal(ids, o["ids"]) and np.array_equal(bits(depth), bits(o["depth"]))
    covered = int((ids != 0xFFFFFFFF).sum())
    print(f"[tiny] {W}x{H}: {st['setup_triangles']} set up, {st['binned_pairs']} bin entries, "
          f"{covered} covered pixels", flush=True)
    assert st["binned_pairs"] < st["setup_triangles"]     # most of them were dropped before binning


def test_determinism(sr, scenes, c1):
    scene, spec, gscene = c1
    cam = scenes.make_camera(spec)
    a = render_gpu(sr, gscene, cam, spec["width"], spec["height"])
    b = render_gpu(sr, gscene, cam, spec["width"], spec["height"])
    assert np.array_equal(a["ids"], b["ids"]) and np.array_equal(bits(a["depth"]), bits(b["depth"]))
    assert np.array_equal(a["px"], b["px"])


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_batch_matches_plain(sr, scenes, c1, tile):
    """Views of several sizes in one batch, at each screen-tile size, with
    the persistent raster and with one raster workgroup per (view, tile):
    every view is the plain renderer's frame bit for bit."""
    scene, spec, gscene = c1
    views, plain = [], []
    for i, (w, h) in enumerate([(320, 240), (640, 480), (1280, 720), (200, 150)]):
        s = dict(spec)
        s["yaw"] = 0.1 * i
        cam = scenes.make_camera(s, w, h)
        views.append((gscene, cam, w, h, 2.0))
        plain.append(render_gpu(sr, gscene, cam, w, h))
    b = sr.Batch(tile_size=tile)
    for persistent in (0, -1):          # the default persistent raster, then one workgroup per item
        b.configure(persistent_blocks=persistent)
        b.render(views)
        for i, p in enumerate(plain):
            ids, depth = b.read_vbuffer(i)
            assert np.array_equal(ids, p["ids"]) and np.array_equal(bits(depth), bits(p["depth"]))
            assert np.array_equal(b.read_pixels(i), p["px"])
    b.close()


def test_batch_with_translucent_views_matches_plain(sr, scenes, env, c1):
    """A batch that mixes translucent and opaque scenes at several sizes
    (colour planes, back-to-front blending per view, then one blit per view)
    gives each view the plain renderer's visibility buffer and pixels."""
    _, spec1, gc1 = c1
    tscene, tspec = _translucent_c1(sr, scenes, 640, 360)
    tscene.prepare(env, cache_key=None, grid=32)
    gt = sr.Scene(tscene)
    views, plain = [], []
    for i, (g, spec, w, h) in enumerate([(gt, tspec, 640, 360), (gc1, spec1, 320, 240), (gt, tspec, 256, 200),
                                         (gt, tspec, 640, 336)]):
        s = dict(spec)
        s["yaw"] = spec.get("yaw", 0.0) + 0.05 * i
        cam = scenes.make_camera(s, w, h)
        views.append((g, cam, w, h, 2.0))
        plain.append(render_gpu(sr, g, cam, w, h))
    for use_graph in (False, True):
        b = sr.Batch()
        b.configure(use_graph=use_graph)
        for _ in range(2 if use_graph else 1):
            b.render(views)
            for i, p in enumerate(plain):
                ids, depth = b.read_vbuffer(i)
                assert np.array_equal(ids, p["ids"]) and np.array_equal(bits(depth), bits(p["depth"])), f"view {i}"
                assert np.array_equal(b.read_pixels(i), p["px"]), f"view {i} pixels"
        b.close()


def test_batch_graph_replays_match_eager(sr, scenes, c1):
    """The batch's launch sequence captured as a hipGraph, replayed three
    times with eager frames and synchronisations in between: every replay
    gives the eager batch's visibility buffers and pixels bit for bit."""
    scene, spec, gscene = c1
    sizes = [(320, 240), (97, 61), (640, 360), (128, 128), (333, 202), (64, 40)]
    views = []
    for i, (w, h) in enumerate(sizes):
        s = dict(spec)
        s["yaw"] = 0.15 * i - 0.3
        views.append((gscene, scenes.make_camera(s, w, h), w, h, 0.0))
    eager = sr.Batch()
    eager.render(views)
    want = [(eager.read_vbuffer(i), eager.read_pixels(i)) for i in range(len(views))]
    eager.close()
    g = sr.Batch()
    g.configure(use_graph=True)
    other = sr.Renderer(200, 150)
    for rep in range(3):
        g.render(views)
        for i in range(len(views)):
            (ids, depth), px = g.read_vbuffer(i), g.read_pixels(i)
            (wids, wdepth), wpx = want[i]
            assert np.array_equal(ids, wids), f"replay {rep} view {i}: ids differ"
            assert np.array_equal(bits(depth), bits(wdepth)), f"replay {rep} view {i}: depth differs"
            assert np.array_equal(px, wpx), f"replay {rep} view {i}: pixels differ"
        other.render_scene(gscene, scenes.make_camera(spec, 200, 150))     # eager work between replays
        other.blit_to_buffer()
    other.close()
    g.close()


def test_autotuner_picks_a_measured_config(sr, scenes, c1, tmp_path):
    import importlib
    at = importlib.import_module(sr.__name__ + ".autotune")
    scene, spec, gscene = c1
    views = [(gscene, scenes.make_camera(spec, 320, 240), 320, 240, 0.0)] * 4
    r = at.tune(sr, views, tiles=(8, 16), workgroups=(256, 1024), reps=1, cache_path=str(tmp_path / "t.json"))
    assert r["tile_size"] in (8, 16) and r["persistent_blocks"] in (256, 1024)
    assert len(r["table"]) == 4 and r["ms"] == min(t["ms"] for t in r["table"])
    again = at.tune(sr, views, tiles=(8, 16), workgroups=(256, 1024), reps=1, cache_path=str(tmp_path / "t.json"))
    assert again["cached"] and again["tile_size"] == r["tile_size"]


def test_server_same_alone_and_in_the_mix(sr, scenes, c1, env):
    """Requests served one at a time and the same requests submitted together
    (batched continuously over two overlapping slots, with a small admission
    budget that splits them into several batches) produce identical
    visibility buffers and pixels, and equal the plain renderer's."""
    scene, spec, gscene = c1
    c3s, c3spec = scenes.make_c3(width=640, height=360, triangles=5000, primitives=16)
    c3s.prepare(env, cache_key=None, grid=16)
    g3 = sr.Scene(c3s)
    reqs = []
    sizes = [(320, 240), (640, 360), (97, 61), (256, 144), (640, 480), (128, 72), (333, 202), (160, 90)]
    for i, (w, h) in enumerate(sizes):
        sp = dict(spec if i % 3 else c3spec)
        sp["yaw"] = sp["yaw"] + 0.07 * i
        reqs.append((0 if i % 3 else 1, scenes.make_camera(sp, w, h), w, h))
    srv = sr.Server(admission_bytes=64 << 20, max_batch_views=3, streams=2)
    idx = [srv.add_scene(gscene), srv.add_scene(g3)]
    alone = []
    for k, cam, w, h in reqs:
        rid = srv.submit(idx[k], cam, w, h)
        srv.drain()
        alone.append((srv.read_vbuffer(rid), srv.read_pixels(rid)))
        srv.release(rid)
    rids = [srv.submit(idx[k], cam, w, h, tag=i) for i, (k, cam, w, h) in enumerate(reqs)]
    srv.drain()
    done = srv.poll(64)
    done += srv.poll(64)
    assert sorted(d["request_id"] for d in done if d["request_id"] in rids) == sorted(rids)
    assert all(d["status"] == 0 for d in done)
    for i, rid in enumerate(rids):
        (ids, depth), px = srv.read_vbuffer(rid), srv.read_pixels(rid)
        (aids, adepth), apx = alone[i]
        assert np.array_equal(ids, aids) and np.array_equal(bits(depth), bits(adepth)), f"request {i}: vbuffer"
        assert np.array_equal(px, apx), f"request {i}: pixels"
        k, cam, w, h = reqs[i]
        r = sr.Renderer(w, h)
        r.render_scene([gscene, g3][k], cam)
        pids, pdepth = r.read_vbuffer()
        r.update_auto_exposure(0.0)
        ppx = r.blit_to_buffer()
        r.close()
        assert np.array_equal(ids, pids) and np.array_equal(bits(depth), bits(pdepth)), f"request {i}: vs plain"
        assert np.array_equal(px[: h & ~1], ppx[: h & ~1]), f"request {i}: pixels vs plain"
        srv.release(rid)
    srv.close()


def test_server_translucent_requests(sr, scenes, c1, env):
    """Requests on a translucent scene mixed with opaque ones in the server's
    batches (colour planes, per-view blending and blits) equal the plain
    renderer's frames."""
    _, spec, gscene = c1
    tscene, tspec = _translucent_c1(sr, scenes, 640, 360)
    tscene.prepare(env, cache_key=None, grid=32)
    gt = sr.Scene(tscene)
    srv = sr.Server(admission_bytes=256 << 20, max_batch_views=4, streams=2)
    idx = [srv.add_scene(gscene), srv.add_scene(gt)]
    reqs = []
    for i, (w, h) in enumerate([(640, 360), (320, 240), (256, 144), (640, 336), (160, 90), (333, 202)]):
        sp = dict(tspec if i % 2 == 0 else spec)
        sp["yaw"] = sp["yaw"] + 0.03 * i
        reqs.append((1 if i % 2 == 0 else 0, scenes.make_camera(sp, w, h), w, h))
    rids = [srv.submit(idx[k], cam, w, h, tag=i) for i, (k, cam, w, h) in enumerate(reqs)]
    srv.drain()
    done = srv.poll(64)
    assert all(d["status"] == 0 for d in done), [d["status"] for d in done]
    for i, rid in enumerate(rids):
        (ids, depth), px = srv.read_vbuffer(rid), srv.read_pixels(rid)
        k, cam, w, h = reqs[i]
        r = sr.Renderer(w, h)
        r.render_scene([gscene, gt][k], cam)
        pids, pdepth = r.read_vbuffer()
        r.update_auto_exposure(0.0)
        ppx = r.blit_to_buffer()
        r.close()
        assert np.array_equal(ids, pids) and np.array_equal(bits(depth), bits(pdepth)), f"request {i}: vbuffer"
        assert np.array_equal(px[: h & ~1], ppx[: h & ~1]), f"request {i}: pixels"
        srv.release(rid)
    srv.close()


def _alpha_c1(sr, scenes, W, H):
    """C1 with half the spheres alpha-masked (different cutoffs) and the
    shared texture's alpha made of 8x8-texel blocks of noise."""
    scene, spec = scenes.make_c1(width=W, height=H)
    rng = np.random.Generator(np.random.PCG64(77))
    w, h, fmt, px = scene.textures[0]
    rgba = sr.unpack_rgba(np.asarray(px, np.uint32)).reshape(h, w, 4).copy()
    blocks = rng.integers(0, 256, (h // 8, w // 8), dtype=np.uint8)
    rgba[..., 3] = np.kron(blocks, np.ones((8, 8), np.uint8))
    scene.textures[0] = (w, h, fmt, sr.pack_rgba(rgba.reshape(-1, 4)).ravel())
    for i, m in enumerate(scene.materials):
        if i % 2 == 0:
            m["alpha_mode"] = sr.SR_ALPHA_MASK
            m["alpha_cutoff"] = 0.25 + 0.05 * i
    scene.rebuild()
    return scene, spec


def test_alpha_tested_materials(sr, orc, scenes, env):
    """MASK materials (src/shader.rs:40-61, :311-329): the base texture's
    nearest-sampled alpha against the cutoff decides coverage; alpha from
    block noise so whole regions pass or fail and edges fall inside
    triangles.  Half the spheres are masked, with different cutoffs."""
    scene, spec = _alpha_c1(sr, scenes, 640, 360)
    scene.prepare(env, cache_key=None, grid=32)
    cam = scenes.make_camera(spec)
    g = render_gpu(sr, sr.Scene(scene), cam, 640, 360)
    o = render_oracle(orc, scene, cam, 640, 360)
    compare(sr, g, o, "alpha-tested c1")
    plain, _ = scenes.make_c1(width=640, height=360)
    plain.prepare(env, cache_key=plain.name)
    opaque = render_oracle(orc, plain, cam, 640, 360)
    assert (g["ids"] != sr.NO_ID).sum() < (opaque["ids"] != sr.NO_ID).sum(), "the mask removed nothing"


def _world_quad(x0, x1, y0, y1, z):
    """Axis-aligned rectangle at world z facing +z (toward the C1 camera)."""
    pos = np.array([[x0, y1, z], [x1, y1, z], [x1, y0, z], [x0, y0, z]], np.float32)
    return {"positions": pos, "normals": np.tile(np.array([0, 0, 1], np.float32), (4, 1)),
            "texcoords": np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), "tangents": None,
            "indices": np.array([0, 2, 1, 0, 3, 2], np.uint32)}


def _translucent_c1(sr, scenes, W, H):
    """C1 with a third of the spheres translucent (one through a
    transmission texture), two overlapping camera-facing glass quads at one
    depth and a glass band wider than the frustum."""
    scene, spec = scenes.make_c1(width=W, height=H)
    for i, m in enumerate(scene.materials):
        if i % 3 == 1:
            m["is_translucent"] = 1
            m["transmission"] = 0.2 + 0.05 * i
            if i == 4:
                m["transmission_texture_tex"] = (0, sr.SR_WRAP_REPEAT, sr.SR_WRAP_REPEAT)

    def add(prim, mat):
        prim["material"] = len(scene.materials)
        scene.materials.append(mat)
        scene.primitives.append(prim)
        scene.meshes.append((len(scene.primitives) - 1, 1))
        scene.nodes.append((np.eye(4, dtype=np.float32), len(scene.meshes) - 1))

    def glass(base, t):
        return {"base_color": base + (1.0,), "metallic": 0.1, "roughness": 0.3, "is_translucent": 1,
                "transmission": t, "base_color_tex": (0, sr.SR_WRAP_REPEAT, sr.SR_WRAP_REPEAT)}

    add(_world_quad(-2.0, 0.5, -1.5, 1.0, 8.0), glass((0.9, 0.3, 0.2), 0.6))
    add(_world_quad(-0.5, 2.0, -1.0, 1.5, 8.0), glass((0.2, 0.4, 0.9), 0.5))
    add(_world_quad(-40.0, 40.0, -0.75, 0.25, 6.0), glass((0.7, 0.7, 0.3), 0.35))
    scene.rebuild()
    return scene, spec


@pytest.mark.parametrize("size", [(640, 360), (640, 336)])
def test_translucent_materials(sr, orc, scenes, env, size):
    """Translucent primitives (src/tilerasterizer.rs:92-101, src/shader.rs:
    66-76): per 64-px tile lists sorted back to front (src/bumpqueue.rs:
    155-200) and blended over the opaque frame.  C1 with a third of the
    spheres translucent (one through a transmission texture), two
    camera-facing quads at one depth that overlap (tied sort keys, so the
    quicksort's own order decides the blend), and a band wider than the
    frustum (clipped; a width of 640 = 10 x 64 lists it a second time in the
    first tile of the next row).  336 rows leave the last tile row's centre
    quads off screen."""
    W, H = size
    scene, spec = _translucent_c1(sr, scenes, W, H)
    scene.prepare(env, cache_key=None, grid=32)
    cam = scenes.make_camera(spec)
    g = render_gpu(sr, sr.Scene(scene), cam, W, H)
    o = render_oracle(orc, scene, cam, W, H)
    compare(sr, g, o, f"translucent c1 {W}x{H}")
    assert np.array_equal(bits(g["rgb"]), bits(o["rgb"])), "linear colour differs"
    assert np.array_equal(bits(g["lum"]), bits(o["lum"])), "tile luminance differs"
    # the layers are there: dropping the translucent quads changes many pixels
    environment, voxels = scene.environment, scene.voxels
    for k in range(3):
        scene.nodes.pop()
    scene.rebuild()
    scene.attach(environment, voxels)        # the same lighting inputs
    g0 = render_gpu(sr, sr.Scene(scene), cam, W, H)
    changed = int(np.any(g0["rgb"] != g["rgb"], -1).sum())
    print(f"[translucent c1 {W}x{H}] the three quads change {changed} px")
    assert changed > 20000


def test_gltf_scene_renders_identically(sr, scenes, env, tmp_path):
    """C1 written as glTF and read back by the C++ loader (sr_gltf_load)
    renders bit-identically to the same scene given as arrays."""
    scene, spec = scenes.make_c1(width=640, height=360)
    loaded = sr.SceneData.from_gltf(scenes.write_gltf(scene, str(tmp_path / "c1.gltf")))
    scene.prepare(env, cache_key=scene.name)
    loaded.prepare(env, cache_key=scene.name)
    cam = scenes.make_camera(spec)
    a = render_gpu(sr, sr.Scene(scene), cam, 640, 360)
    b = render_gpu(sr, sr.Scene(loaded), cam, 640, 360)
    assert np.array_equal(a["ids"], b["ids"])
    assert np.array_equal(bits(a["depth"]), bits(b["depth"]))
    assert np.array_equal(bits(a["rgb"]), bits(b["rgb"]))
    assert np.array_equal(a["px"], b["px"])


def test_gpu_matches_committed_c1_frame(sr, scenes, env):
    """The GPU reproduces the committed oracle frame (tests/golden/) without
    running the oracle."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_160x90.npz"))
    scene, spec = scenes.make_c1(width=160, height=90)
    scene.prepare(env, cache_key=scene.name)
    out = render_gpu(sr, sr.Scene(scene), scenes.make_camera(spec), 160, 90)
    assert np.array_equal(out["ids"], g["ids"])
    assert np.array_equal(bits(out["depth"]), bits(g["depth"]))
    assert np.array_equal(bits(out["lum"]), bits(g["tile_luminance"]))
    assert np.array_equal(out["px"], g["pixels"])


def test_fp32_shading_error_vs_fp64(sr, orc, scenes, c1):
    """FP32 shading (GPU) against the FP64 restatement of the shader on the
    same visibility buffer; the error is reported (nearest-texel and mip
    choices can flip at boundaries, so the bound is loose)."""
    scene, spec, gscene = c1
    cam = scenes.make_camera(spec, 640, 360)
    g = render_gpu(sr, gscene, cam, 640, 360)
    o = orc.OracleRenderer(640, 360)
    o.render_scene(orc.OracleScene(scene), cam, threads=8, fp64=True)
    f = o.read_frame(fp64=True)
    assert np.array_equal(g["ids"], f["ids"])
    err = np.abs(g["rgb"].astype(np.float64) - f["rgb64"])
    covered = g["ids"] != sr.NO_ID
    print(f"[fp64] max abs error {err.max():.3e}, covered-pixel p99.9 {np.percentile(err[covered], 99.9):.3e}, "
          f"mean {err[covered].mean():.3e}")
    assert np.percentile(err[covered], 99.0) < 1e-3
