"""CPU tests pinning the oracle's rasterization.

On viewports up to 128x128 every edge-function term and every step of the
reference's f32 evaluation chain is an integer below 2^24, so the f32
walk (block origins, coarse-block rejection, serial steps) must agree
exactly with plain integer arithmetic.  These tests compare the oracle's
visibility buffer with a brute-force exact-integer rasterizer, check
watertightness on a tessellated screen, the depth-test tie rule, and the
clipper's fan counts against an independent float64 Sutherland-Hodgman."""
import numpy as np
import pytest

from util import attach_dummy, exact_coverage, screen_triangle_scene, signed_area

NO_ID = 0xFFFFFFFF


def random_front_triangles(rng, n, width, height, max_edge_px):
    tris = []
    while len(tris) < n:
        c = rng.uniform([2, 2], [width - 2, height - 2]) * 16
        e1 = rng.uniform(1, max_edge_px) * 16
        e2 = rng.uniform(1, max_edge_px) * 16
        phi = rng.uniform(0, 2 * np.pi)
        th = rng.uniform(0.15 * np.pi, 0.85 * np.pi)
        p = np.array([[0, 0], [e1 * np.cos(phi), e1 * np.sin(phi)],
                      [e2 * np.cos(phi + th), e2 * np.sin(phi + th)]])
        p = np.round(p - p.mean(0) + c)
        # strictly inside the screen so no vertex meets a clip plane
        p[:, 0] = np.clip(p[:, 0], 8, width * 16 - 8)
        p[:, 1] = np.clip(p[:, 1], 8, height * 16 - 8)
        a = signed_area(p)
        if a == 0:
            continue
        if a > 0:
            p = p[[0, 2, 1]]
        tris.append(p)
    return np.array(tris)


def render(sr, orc, scenes, scene, cam):
    c = scenes.make_camera(cam)
    o = orc.OracleRenderer(cam["width"], cam["height"])
    o.render_scene(orc.OracleScene(scene), c)
    return o.read_frame(), o


def clip_depth(cam, d):
    """Stored depth: perspective-correct clip-space z = interp(z/w) * w, i.e.
    r * (near - d) for glam perspective_rh with r = far / (near - far)."""
    far = cam["far"]
    near = 0.01 * far
    r = far / (near - far)
    return r * (near - d)


@pytest.mark.parametrize("seed,size,max_edge", [(1, (128, 128), 90), (2, (100, 70), 40), (3, (64, 64), 70),
                                                (4, (128, 96), 12), (5, (37, 29), 25)])
def test_exact_integer_rasterizer(sr, orc, scenes, seed, size, max_edge):
    rng = np.random.Generator(np.random.PCG64(seed))
    w, h = size
    tris = random_front_triangles(rng, 40, w, h, max_edge)
    depths = rng.permutation(np.linspace(2.0, 60.0, len(tris)))
    scene, cam = screen_triangle_scene(sr, tris, depths, w, h)
    attach_dummy(sr, scene)
    f, _ = render(sr, orc, scenes, scene, cam)
    best = np.full((h, w), np.inf)
    want = np.full((h, w), NO_ID, np.uint64)
    for i, t in enumerate(tris):
        cov = exact_coverage(t, w, h)
        win = cov & (depths[i] < best)
        best[win] = depths[i]
        want[win] = i
    assert np.array_equal(f["ids"].astype(np.uint64), want), \
        f"{int((f['ids'] != want).sum())} pixels differ from the exact-integer rasterizer"
    covered = want != NO_ID
    zexp = clip_depth(cam, best[covered])
    assert np.allclose(f["depth"][covered], zexp, rtol=2e-5, atol=2e-6)
    assert np.all(np.isinf(f["depth"][~covered]))


def test_watertight_tessellation(sr, orc, scenes):
    """A perturbed grid of triangles covering the screen: every pixel centre
    is covered by exactly one triangle (exact top-left rule), and the oracle
    assigns every pixel to that triangle."""
    rng = np.random.Generator(np.random.PCG64(11))
    w, h = 96, 80
    nx, ny = 9, 7
    xs = np.linspace(0, w * 16, nx + 1)
    ys = np.linspace(0, h * 16, ny + 1)
    V = np.stack(np.meshgrid(xs, ys, indexing="xy"), -1)
    V[1:-1, 1:-1] += rng.uniform(-40, 40, V[1:-1, 1:-1].shape)
    V = np.round(V)
    tris = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = V[j, i], V[j, i + 1], V[j + 1, i], V[j + 1, i + 1]
            for t in ((a, c, b), (b, c, d)):
                t = np.array(t)
                if signed_area(t) > 0:
                    t = t[[0, 2, 1]]
                tris.append(t)
    tris = np.array(tris)
    scene, cam = screen_triangle_scene(sr, tris, np.full(len(tris), 10.0), w, h)
    attach_dummy(sr, scene)
    count = np.zeros((h, w), int)
    owner = np.full((h, w), -1)
    for i, t in enumerate(tris):
        cov = exact_coverage(t, w, h)
        count += cov
        owner[cov] = i
    assert np.all(count == 1), "tessellation is not watertight under the exact top-left rule"
    f, _ = render(sr, orc, scenes, scene, cam)
    assert np.array_equal(f["ids"].astype(np.int64), owner)


def test_equal_depth_later_wins(sr, orc, scenes):
    """Duplicated coplanar triangles: the later one in submission order wins
    every pixel (z <= current, src/tilerasterizer.rs:516)."""
    rng = np.random.Generator(np.random.PCG64(5))
    w, h = 120, 90
    base = random_front_triangles(rng, 6, w, h, 60)
    tris = np.concatenate([base, base])
    scene, cam = screen_triangle_scene(sr, tris, np.full(len(tris), 7.0), w, h)
    attach_dummy(sr, scene)
    f, _ = render(sr, orc, scenes, scene, cam)
    ids = f["ids"]
    hit = ids != NO_ID
    assert hit.any()
    assert np.all(ids[hit] >= len(base)), "an earlier duplicate won an equal-depth tie"


def test_backface_and_degenerate(sr, orc, scenes):
    """Counter-clockwise (screen, y down) triangles are culled; zero-area
    triangles are kept but cover nothing (src/renderer.rs:677-682)."""
    w, h = 64, 64
    front = np.array([[[100, 100], [100, 800], [800, 100]]], float)  # signed area < 0
    assert signed_area(front[0]) < 0
    back = front[:, [0, 2, 1]]
    degen = np.array([[[200, 200], [400, 400], [600, 600]]], float)
    scene, cam = screen_triangle_scene(sr, np.concatenate([back, degen, front]), [5.0, 6.0, 9.0], w, h)
    attach_dummy(sr, scene)
    f, o = render(sr, orc, scenes, scene, cam)
    ids = f["ids"]
    assert set(np.unique(ids)) <= {2, NO_ID}
    assert (ids == 2).sum() == exact_coverage(front[0], w, h).sum()
    assert o.setup_count() == 3


def sh_clip_count(tri):
    """Independent float64 Sutherland-Hodgman against z+w, -z+w, +-x+w, +-y+w."""
    planes = [(0, 0, 1, 1), (0, 0, -1, 1), (1, 0, 0, 1), (-1, 0, 0, 1), (0, 1, 0, 1), (0, -1, 0, 1)]
    poly = [np.asarray(v, np.float64) for v in tri]
    for p in planes:
        p = np.asarray(p, np.float64)
        out = []
        for i in range(len(poly)):
            cur, prev = poly[i], poly[i - 1]
            dc, dp = p @ cur, p @ prev
            if dc >= 0:
                if dp < 0:
                    out.append(prev + (cur - prev) * (dp / (dp - dc)))
                out.append(cur)
            elif dp >= 0:
                out.append(prev + (cur - prev) * (dp / (dp - dc)))
        poly = out
        if not poly:
            return 0
    return max(len(poly) - 2, 0)


def test_clip_fan_counts(sr, orc, scenes):
    """Triangles straddling the view frustum: the oracle's number of emitted
    fan triangles equals an independent float64 clipper's (cases near
    degeneracy are excluded), and never exceeds 7."""
    rng = np.random.Generator(np.random.PCG64(21))
    cam = {"position": [0.0, 0.0, 0.0], "yaw": 0.0, "pitch": 0.0, "fov": np.pi / 3, "width": 160,
           "height": 120, "far": 50.0}
    c = scenes.make_camera(cam)
    vp = np.array(c.c.view_project_matrix, np.float64).reshape(4, 4).T
    tris = []
    expect = 0
    while len(tris) < 60:
        p = rng.uniform([-6, -5, -12], [6, 5, 1], (3, 3))
        clip = (vp @ np.c_[p, np.ones(3)].T).T
        # skip near-degenerate plane crossings
        planes = np.array([(0, 0, 1, 1), (0, 0, -1, 1), (1, 0, 0, 1), (-1, 0, 0, 1), (0, 1, 0, 1), (0, -1, 0, 1)])
        if np.min(np.abs(clip @ planes.T)) < 1e-3:
            continue
        tris.append(p)
        expect += sh_clip_count(clip)
    pos = np.array(tris, np.float32).reshape(-1, 3)
    prim = {"positions": pos, "normals": np.tile(np.float32([0, 0, 1]), (len(pos), 1)), "texcoords": None,
            "tangents": None, "indices": np.arange(len(pos), dtype=np.uint32), "material": 0}
    scene = sr.SceneData([prim], [(0, 1)], [(np.eye(4, dtype=np.float32), 0)],
                         [{"base_color": (1, 1, 1, 1), "metallic": 0.0, "roughness": 0.5}], [])
    attach_dummy(sr, scene)
    o = orc.OracleRenderer(cam["width"], cam["height"])
    o.render_scene(orc.OracleScene(scene), c)
    assert o.setup_count() == expect


def test_inside_triangles_unchanged_by_clipping(sr, orc, scenes):
    """A triangle fully inside the frustum is rasterized identically whether
    its primitive is classified Inside (no clipper) or Intersecting (passes
    through the clipper unchanged)."""
    rng = np.random.Generator(np.random.PCG64(8))
    w, h = 128, 96
    tris = random_front_triangles(rng, 12, w, h, 40)
    depths = rng.uniform(3, 20, len(tris))
    scene_a, cam = screen_triangle_scene(sr, tris, depths, w, h)
    attach_dummy(sr, scene_a)
    fa, _ = render(sr, orc, scenes, scene_a, cam)
    # the same triangles plus one far outside the frustum (forces Intersecting)
    out = unproject_outside = np.array([[[-9000, -9000], [-9000, -8000], [-8000, -9000]]], float)
    scene_b, _ = screen_triangle_scene(sr, np.concatenate([tris, out]), np.concatenate([depths, [5.0]]), w, h)
    attach_dummy(sr, scene_b)
    fb, ob = render(sr, orc, scenes, scene_b, cam)
    assert np.array_equal(fa["ids"], fb["ids"])
    assert np.array_equal(fa["depth"].view(np.uint32), fb["depth"].view(np.uint32))
    del unproject_outside


def test_canonical_determinism_and_threaded_agreement(sr, orc, scenes, env):
    scene, spec = scenes.make_c1(width=320, height=180)
    scene.prepare(env, cache_key=scene.name)
    cam = scenes.make_camera(spec)
    osc = orc.OracleScene(scene)
    frames = []
    for threads, canonical in ((1, True), (8, True), (8, False)):
        o = orc.OracleRenderer(320, 180)
        o.render_scene(osc, cam, threads=threads, canonical=canonical)
        o.update_auto_exposure(0.0)
        frames.append((o.read_frame(), o.blit()))
    (a, pa), (b, pb), (c, pc) = frames
    assert np.array_equal(a["ids"], b["ids"]) and np.array_equal(pa, pb)
    # threaded front end: ids are not canonical, images agree except on
    # equal-depth ties, which this scene does not have
    assert np.array_equal(pa, pc)
