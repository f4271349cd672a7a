"""GPU tests at the batched (C4) and served (C5) configurations' real shapes,
and the equal-depth tie rule on the device.

C4: views of the real heightfield pool (10k-2M triangles, every resolution
from 320x240 to 1920x1080, 1080p on the two largest scenes) in one batch
equal the plain renderer's frames bit for bit; two views equal the oracle.
C5: the eight resident scenes (12.3M triangles, 240 distinct 2048^2
textures) served at 720p-2160p give each request the same visibility buffer
and pixels alone, in the mix and in the plain renderer; two requests equal
the oracle.  Frames whose tiles get no list entry at all; zero-area
triangles.  Ties: duplicated
coplanar triangles spanning many 16-px tiles,
and a layer at exactly zero depth, resolve to the later triangle as the
reference's `z <= current` test leaves them (src/tilerasterizer.rs:511-523).
"""
import numpy as np
import pytest

from conftest import bits
from test_gpu_parity import compare, render_gpu, render_oracle
from util import attach_dummy, screen_triangle_scene

pytestmark = pytest.mark.gpu


def plain_frame(sr, gscene, cam, w, h, tile=0):
    r = sr.Renderer(w, h, tile_size=tile)
    r.render_scene(gscene, cam)
    ids, depth = r.read_vbuffer()
    r.update_auto_exposure(0.0)
    px = r.blit_to_buffer()
    r.close()
    return ids, depth, px


# ---------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4(sr, scenes, env):
    pool = scenes.prepare_c4_pool(scenes.make_c4_pool(), env, device=0)
    gpool = [sr.Scene(p) for p in pool]
    views = scenes.make_c4_views(pool)
    # four views of each resolution, then 1080p on the two largest scenes
    chosen = []
    for res in scenes.C4_RESOLUTIONS:
        chosen += [(k, s) for k, s in views if (s["width"], s["height"]) == res][:4]
    hd = [s for _, s in views if (s["width"], s["height"]) == (1920, 1080)]
    for k, s in ((len(pool) - 1, hd[4]), (len(pool) - 2, hd[5])):
        s = dict(s)
        s["far"] = 2.0 * float(pool[k].bounds()[3])
        chosen.append((k, s))
    return pool, gpool, chosen


@pytest.mark.timeout(600)
@pytest.mark.parametrize("persistent", [-1, 0])
def test_c4_real_views_batch_matches_plain(sr, scenes, c4, persistent):
    """A batch of 18 C4 views (16-px tiles; one raster workgroup per
    (view, tile), then the persistent raster) equals 18 plain frames."""
    pool, gpool, chosen = c4
    views, want = [], []
    for k, s in chosen:
        cam = scenes.make_camera(s)
        views.append((gpool[k], cam, s["width"], s["height"], 2.0))
        want.append(plain_frame(sr, gpool[k], cam, s["width"], s["height"]))
    b = sr.Batch(tile_size=16)
    b.configure(persistent_blocks=persistent, use_graph=True)
    for rep in range(2):        # capture, then a replay
        b.render(views)
        for i, (ids, depth, px) in enumerate(want):
            gids, gdepth = b.read_vbuffer(i)
            k, s = chosen[i]
            name = f"rep {rep} view {i} (scene {k}, {pool[k].triangles} tris, {s['width']}x{s['height']})"
            assert np.array_equal(gids, ids), f"{name}: ids differ"
            assert np.array_equal(bits(gdepth), bits(depth)), f"{name}: depth differs"
            assert np.array_equal(b.read_pixels(i), px), f"{name}: pixels differ"
    st = b.stats()
    print(f"[c4 batch persistent={persistent}] {st['views']} views, {st['input_triangles']} triangles, "
          f"{st['pixels']} px, {st['ms_total']:.2f} ms")
    b.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("which", ["smallest_view", "largest_scene_1080p"])
def test_c4_views_match_oracle(sr, orc, scenes, c4, which):
    pool, gpool, chosen = c4
    if which == "smallest_view":
        k, s = chosen[0]
    else:
        k, s = chosen[-2]
    cam = scenes.make_camera(s)
    w, h = s["width"], s["height"]
    g = render_gpu(sr, gpool[k], cam, w, h)
    o = render_oracle(orc, pool[k], cam, w, h, threads=16)
    compare(sr, g, o, f"c4 scene {k} ({pool[k].triangles} tris) {w}x{h}")


# ---------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c5(sr, scenes, env):
    datas = scenes.make_c5_scenes()
    for d in datas:
        d.prepare(env, device=0)
    gs = [sr.Scene(d) for d in datas]
    # one request per scene and resolution pair from the served stream's
    # generator, topped up so every scene and every resolution appears
    reqs = scenes.make_c5_requests(datas, 64, 1000.0, 5)
    picked, seen_k, seen_r = [], set(), set()
    for r in reqs:
        _, k, cam, w, h = r
        if k not in seen_k or (w, h) not in seen_r:
            picked.append(r)
            seen_k.add(k)
            seen_r.add((w, h))
    assert len(seen_k) == len(datas) and len(seen_r) == 4
    return datas, gs, picked


@pytest.mark.timeout(900)
def test_c5_served_same_alone_in_the_mix_and_plain(sr, scenes, c5):
    """Every request of a mixed C5 stream (all eight scenes, 720p-2160p):
    served alone, served together with the others (continuous batching over
    two slots, up to eight views per batch) and rendered by the plain
    renderer give the same visibility buffer and pixels."""
    datas, gs, reqs = c5
    srv = sr.Server(admission_bytes=64 << 30, max_batch_views=8, streams=2)
    idx = [srv.add_scene(g) for g in gs]
    alone = []
    for _, k, cam, w, h in reqs:
        rid = srv.submit(idx[k], cam, w, h)
        srv.drain()
        alone.append((srv.read_vbuffer(rid), srv.read_pixels(rid)))
        srv.release(rid)
    rids = [srv.submit(idx[k], cam, w, h, tag=i) for i, (_, k, cam, w, h) in enumerate(reqs)]
    srv.drain()
    done = []
    while len(done) < len(rids):
        got = srv.poll(64)
        assert got, "server drained without completing every request"
        done += got
    assert all(d["status"] == 0 for d in done), [d["status"] for d in done]
    for i, rid in enumerate(rids):
        _, k, cam, w, h = reqs[i]
        name = f"request {i} ({datas[k].name}, {w}x{h})"
        (ids, depth), px = srv.read_vbuffer(rid), srv.read_pixels(rid)
        (aids, adepth), apx = alone[i]
        assert np.array_equal(ids, aids) and np.array_equal(bits(depth), bits(adepth)), f"{name}: mix vs alone"
        assert np.array_equal(px, apx), f"{name}: pixels mix vs alone"
        pids, pdepth, ppx = plain_frame(sr, gs[k], cam, w, h)
        assert np.array_equal(ids, pids) and np.array_equal(bits(depth), bits(pdepth)), f"{name}: vs plain"
        assert np.array_equal(px[: h & ~1], ppx[: h & ~1]), f"{name}: pixels vs plain"
        print(f"[c5] {name}: identical alone, in the mix and plain; {(ids != sr.NO_ID).sum()} px covered")
        srv.release(rid)
    srv.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("scene_name", ["c5_ico_s51_n64", "c3_s31_n100000"])
def test_c5_requests_match_oracle(sr, orc, scenes, c5, scene_name):
    datas, gs, reqs = c5
    k = [d.name for d in datas].index(scene_name)
    _, _, cam, w, h = next(r for r in reqs if r[1] == k)
    g = render_gpu(sr, gs[k], cam, w, h)
    o = render_oracle(orc, datas[k], cam, w, h, threads=16)
    compare(sr, g, o, f"c5 {scene_name} {w}x{h}")


# ---------------------------------------------------------------------------
# equal-depth ties on the device
# ---------------------------------------------------------------------------
def big_front_triangles(rng, n, w, h, lo, hi):
    """n front-facing (negative signed area on screen, y down) triangles with
    vertices lo-hi px from their centre, kept inside the viewport so none is
    clipped (post-clip ids are then submission indices); subpixel units."""
    out = []
    while len(out) < n:
        c = rng.uniform([0.1 * w, 0.1 * h], [0.9 * w, 0.9 * h])
        a = rng.uniform(0, 2 * np.pi)
        r = rng.uniform(lo, hi, 3)
        ang = a + np.array([0.0, 2.1, 4.2]) + rng.uniform(-0.3, 0.3, 3)
        p = np.stack([c[0] + r * np.cos(ang), c[1] + r * np.sin(ang)], 1)
        p = np.clip(p, [1.0, 1.0], [w - 1.0, h - 1.0])
        tri = np.round(p * 16)
        area = (tri[1, 0] - tri[0, 0]) * (tri[2, 1] - tri[0, 1]) - (tri[2, 0] - tri[0, 0]) * (tri[1, 1] - tri[0, 1])
        if area > 0:
            tri = tri[[0, 2, 1]]
        if abs(area) > 256.0 * 400:       # at least ~400 px^2
            out.append(tri)
    return np.array(out)


def test_gpu_equal_depth_later_wins(sr, orc, scenes):
    """Large triangles (50-250 px, each over many 16-px tiles), every one
    submitted twice with identical vertices at one depth: each covered
    pixel's winner is a second copy (z <= current lets the later of two
    equal depths win), on the GPU as in the oracle, bit for bit."""
    rng = np.random.Generator(np.random.PCG64(11))
    w, h = 640, 360
    base = big_front_triangles(rng, 12, w, h, 50, 250)
    tris = np.concatenate([base, base])
    scene, cam_spec = screen_triangle_scene(sr, tris, np.full(len(tris), 7.0), w, h)
    attach_dummy(sr, scene)
    cam = scenes.make_camera(cam_spec)
    g = render_gpu(sr, sr.Scene(scene), cam, w, h)
    o = render_oracle(orc, scene, cam, w, h)
    hit = g["ids"] != sr.NO_ID
    assert hit.sum() > 20000
    assert np.all(g["ids"][hit] >= len(base)), "an earlier duplicate won an equal-depth tie"
    compare(sr, g, o, "equal-depth duplicates")


def degenerate_triangles(rng, n, w, h):
    """n zero-area triangles in subpixel units: three collinear vertices
    (integer multiples of one integer step, some spanning many tiles), two
    coincident vertices, or all three coincident."""
    lo, hi = np.array([16.0, 16.0]), np.array([16.0 * (w - 1), 16.0 * (h - 1)])
    out = []
    while len(out) < n:
        a = np.round(rng.uniform(lo, hi))
        kind = len(out) % 3
        if kind == 0:
            d = rng.integers(-40, 41, 2).astype(float)
            if not d.any():
                continue
            m = rng.choice(np.arange(-60, 61), 3, replace=False).astype(float)
            tri = a[None, :] + m[:, None] * d[None, :]
        elif kind == 1:
            b = np.round(rng.uniform(lo, hi))
            tri = np.stack([a, a, b])[rng.permutation(3)]
        else:
            tri = np.stack([a, a, a])
        if np.all(tri >= lo) and np.all(tri <= hi):      # inside the viewport: no clipping
            out.append(tri)
    return np.array(out)


def test_gpu_zero_area_triangles(sr, orc, scenes):
    """Zero-area triangles are kept by setup, as the reference keeps them
    (it culls signed area > 0 only, src/renderer.rs:677-682; 1/|A| is then
    +inf), and win no pixel: a sample the rounded edge functions let through
    gets a NaN depth.  Submitted in front of large triangles, with their
    vertices at different depths, they leave the visibility buffer and the
    pixels equal to the oracle's."""
    rng = np.random.Generator(np.random.PCG64(21))
    w, h = 640, 360
    back = big_front_triangles(rng, 30, w, h, 40, 220)
    degen = degenerate_triangles(rng, 600, w, h)
    for t in degen:
        area = (t[1, 0] - t[0, 0]) * (t[2, 1] - t[0, 1]) - (t[2, 0] - t[0, 0]) * (t[1, 1] - t[0, 1])
        assert area == 0.0
    tris = np.concatenate([back, degen])
    depths = np.concatenate([np.repeat(rng.uniform(20, 40, (len(back), 1)), 3, 1),
                             rng.uniform(2, 10, (len(degen), 3))])
    scene, cam_spec = screen_triangle_scene(sr, tris, depths, w, h)
    attach_dummy(sr, scene)
    cam = scenes.make_camera(cam_spec)
    g = render_gpu(sr, sr.Scene(scene), cam, w, h)
    o = render_oracle(orc, scene, cam, w, h)
    hit = g["ids"] != sr.NO_ID
    assert hit.sum() > 20000
    assert np.all(g["ids"][hit] < len(back)), "a zero-area triangle won a pixel"
    assert g["stats"]["setup_triangles"] == o["setup"]
    compare(sr, g, o, "zero-area triangles")


@pytest.mark.parametrize("case", ["backfacing", "looking_away"])
def test_gpu_frames_without_list_entries(sr, orc, scenes, case):
    """A frame that sets up triangles but bins none (every triangle faces
    away and is culled at setup: the radix binning sorts zero entries and
    every tile's list stays empty) and a frame whose camera sees no
    triangle at all: sky everywhere, equal to the oracle, and empty lists."""
    rng = np.random.Generator(np.random.PCG64(5))
    w, h = 640, 360
    tris = big_front_triangles(rng, 12, w, h, 50, 250)
    if case == "backfacing":
        tris = tris[:, [0, 2, 1]]          # clockwise on screen
    scene, cam_spec = screen_triangle_scene(sr, tris, np.full(len(tris), 7.0), w, h)
    attach_dummy(sr, scene)
    if case == "looking_away":
        cam_spec = dict(cam_spec, yaw=float(np.pi))
    cam = scenes.make_camera(cam_spec)
    gscene = sr.Scene(scene)
    g = render_gpu(sr, gscene, cam, w, h)
    o = render_oracle(orc, scene, cam, w, h)
    assert np.all(g["ids"] == sr.NO_ID)
    compare(sr, g, o, f"no list entries ({case})")
    r = sr.Renderer(w, h)
    r.render_scene(gscene, cam)
    start, end, lst = r.read_bins()
    st = r.stats()
    r.close()
    assert lst.size == 0 and st["binned_pairs"] == 0
    assert np.all(start == 0) and np.all(end == 0)
    assert st["setup_triangles"] == (len(tris) if case == "backfacing" else 0)


def test_gpu_zero_depth_ties(sr, orc, scenes):
    """A layer exactly at view depth = near, where clip z = r (-near) +
    r near cancels to +0, so every fragment's depth is +0 (its vertices'
    z/w are +0 and interpolation keeps +0); it is submitted three times and
    overlaps a layer behind it.  The last copy wins every pixel it covers
    and the depth bits are +0.  (-0 cannot arise on this path: clip z is a
    sum ending in the nonzero r near term, so it is never -0, and a sum or
    interpolation of non-negative-zero terms never produces -0.)"""
    rng = np.random.Generator(np.random.PCG64(12))
    w, h, far = 640, 360, 150.0
    near = float(np.float32(far) * np.float32(0.01))
    base = big_front_triangles(rng, 8, w, h, 60, 200)
    back = big_front_triangles(rng, 8, w, h, 60, 200)
    tris = np.concatenate([back, base, base, base])
    depths = np.concatenate([np.full(8, 3.0), np.full(24, near)])
    scene, cam_spec = screen_triangle_scene(sr, tris, depths, w, h, far=far)
    attach_dummy(sr, scene)
    cam = scenes.make_camera(cam_spec)
    g = render_gpu(sr, sr.Scene(scene), cam, w, h)
    o = render_oracle(orc, scene, cam, w, h)
    front = (g["ids"] >= 8) & (g["ids"] != sr.NO_ID)
    assert front.sum() > 20000
    assert np.all(g["ids"][front] >= 24), "an earlier copy won a tie at depth +0"
    assert np.all(bits(g["depth"])[front] == 0), "the zero-depth layer's depth is not +0"
    compare(sr, g, o, "zero-depth ties")


# ---------------------------------------------------------------------------
# bins in submission order
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("config", ["c1", "c2", "c3_small", "c1_tiny", "c1_8k"])
def test_bins_hold_submission_order(sr, scenes, env, config):
    """Every tile's list holds ascending set-up triangle ids (the order the
    reference's single-threaded front end queues packets), each once, and
    two renders of a frame give the same lists bit for bit.  The frames
    cover one radix pass over the bin index (c1_tiny: 96 bins), two (c1,
    c2, c3_small) and three (c1_8k: 7680x4320, 129,600 bins)."""
    if config == "c3_small":
        scene, spec = scenes.make_c3(width=1280, height=720, triangles=20000, primitives=64)
    elif config == "c1_tiny":
        scene, spec = scenes.make_c1(width=160, height=90)
    elif config == "c1_8k":
        scene, spec = scenes.make_c1(width=7680, height=4320)
    else:
        scene, spec = scenes.CONFIGS[config]()
    scene.prepare(env, device=0)
    g = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    out = []
    for _ in range(2):
        r = sr.Renderer(spec["width"], spec["height"])
        r.render_scene(g, cam)
        start, end, lst = r.read_bins()
        st = r.stats()
        r.close()
        out.append((start, end, lst))
        assert lst.size == st["binned_pairs"] > 0
        assert np.array_equal(start[1:], end[:-1]) and end[-1] == lst.size
        runs = np.diff(lst.astype(np.int64))
        inside = np.ones(lst.size - 1, bool)
        inside[end[:-1][(end[:-1] > 0) & (end[:-1] < lst.size)] - 1] = False   # boundaries between tiles
        assert np.all(runs[inside] > 0), "a tile's ids are not strictly ascending"
        assert lst.max() < st["setup_triangles"] or st["setup_triangles"] == 0
    assert all(np.array_equal(a, b) for a, b in zip(out[0], out[1])), "lists differ between two renders"
