"""Parity sweeps over random poses.  Against the oracle: many small frames of the sphere cluster
(also with alpha-masked and with translucent materials), the heightfield
and a 20,000-triangle overdraw scene, each at a random
viewport (8-640 x 8-400), tile size (8, 16, 32) and exposure step (0 or
0.1 s), compared with the oracle: visibility buffer bit for bit, setup
count, RGBA8 within 1 LSB.  SR_SWEEP_SCALE multiplies the number of frames
(the default takes about 20 s; 5 gives about 2,600 frames).  Against the
plain renderer: random views over five scenes in one batch and as served
requests."""
import os

import numpy as np
import pytest

from conftest import bits
from test_gpu_parity import _alpha_c1, _translucent_c1, render_gpu, render_oracle

pytestmark = pytest.mark.gpu

SCALE = float(os.environ.get("SR_SWEEP_SCALE", "1"))


def random_pose(scenes, cfg, spec, rng, w, h):
    if cfg.startswith("c1"):      # around and inside the cluster, aimed near its centre
        pos = rng.uniform(-9.0, 9.0, 3)
        target = rng.uniform(-3.0, 3.0, 3)
        fov, far = rng.uniform(np.radians(15), np.radians(120)), rng.uniform(3.0, 80.0)
    elif cfg == "c2":     # 1-40 units over the terrain, aimed at a point of it
        pos = np.array([rng.uniform(-45.0, 45.0), rng.uniform(1.0, 40.0), rng.uniform(-45.0, 45.0)])
        target = np.array([rng.uniform(-40.0, 40.0), 0.0, rng.uniform(-40.0, 40.0)])
        fov, far = rng.uniform(np.radians(20), np.radians(100)), spec["far"]
    else:                 # near the overdraw scene's camera, looking into it
        pos = np.array(spec["position"]) + rng.normal(0.0, 2.0, 3)
        target = pos + np.array([rng.normal(0.0, 0.4), rng.normal(0.0, 0.4), -1.0])
        fov, far = spec["fov"], spec["far"]
    yaw, pitch = scenes.look_pose(pos, target)
    pitch = float(np.clip(pitch + rng.uniform(-0.3, 0.3), -1.5, 1.5))
    return scenes.make_camera(scenes.camera_spec(pos, yaw + rng.uniform(-0.3, 0.3), pitch, fov, w, h, far))


@pytest.mark.timeout(900)
@pytest.mark.parametrize("cfg,frames", [("c1", 300), ("c2", 40), ("c3_small", 60), ("c1_alpha", 60),
                                        ("c1_translucent", 60)])
def test_random_frames_match_oracle(sr, orc, scenes, env, cfg, frames):
    if cfg == "c1":
        scene, spec = scenes.make_c1()
    elif cfg == "c2":
        scene, spec = scenes.make_c2()
    elif cfg == "c3_small":
        scene, spec = scenes.make_c3(width=640, height=360, triangles=20000, primitives=64)
    elif cfg == "c1_alpha":
        scene, spec = _alpha_c1(sr, scenes, 640, 360)
    else:
        scene, spec = _translucent_c1(sr, scenes, 640, 360)
    if cfg in ("c1_alpha", "c1_translucent"):
        scene.prepare(env, cache_key=None, grid=32)
    else:
        scene.prepare(env, cache_key=scene.name)
    gscene = sr.Scene(scene)
    bad, covered = [], 0
    for seed in range(int(frames * SCALE)):
        rng = np.random.Generator(np.random.PCG64(50000 + seed))
        w, h = int(rng.integers(8, 641)), int(rng.integers(8, 401))
        cam = random_pose(scenes, cfg, spec, rng, w, h)
        tile = int(rng.choice([8, 16, 32]))
        dt = float(rng.choice([0.0, 0.0, 0.1]))
        g = render_gpu(sr, gscene, cam, w, h, tile=tile, dt=dt)
        o = render_oracle(orc, scene, cam, w, h, dt=dt, threads=16)
        same = np.array_equal(g["ids"], o["ids"]) and np.array_equal(bits(g["depth"]), bits(o["depth"]))
        lsb = int(np.abs(sr.unpack_rgba(g["px"]).astype(int) - sr.unpack_rgba(o["px"]).astype(int)).max())
        if not (same and lsb <= 1 and g["stats"]["setup_triangles"] == o["setup"] and g["exposure"] == o["exposure"]):
            bad.append(f"seed {seed} {w}x{h} tile {tile} dt {dt}: vbuffer {same}, {lsb} LSB")
        covered += int((g["ids"] != sr.NO_ID).any())
    print(f"[{cfg}] {int(frames * SCALE)} frames, {covered} with coverage, {len(bad)} differing")
    assert covered > frames * SCALE * 0.8
    assert not bad, "; ".join(bad[:5])


def five_scenes(sr, scenes, env):
    """The cluster (plain, alpha-masked, translucent), the heightfield and a
    20,000-triangle overdraw scene, uploaded, with their camera specs."""
    built = {}
    for cfg in ("c1", "c2", "c3_small", "c1_alpha", "c1_translucent"):
        if cfg == "c1":
            scene, spec = scenes.make_c1()
        elif cfg == "c2":
            scene, spec = scenes.make_c2()
        elif cfg == "c3_small":
            scene, spec = scenes.make_c3(width=640, height=360, triangles=20000, primitives=64)
        elif cfg == "c1_alpha":
            scene, spec = _alpha_c1(sr, scenes, 640, 360)
        else:
            scene, spec = _translucent_c1(sr, scenes, 640, 360)
        if cfg in ("c1_alpha", "c1_translucent"):
            scene.prepare(env, cache_key=None, grid=32)
        else:
            scene.prepare(env, cache_key=scene.name)
        built[cfg] = (sr.Scene(scene), spec)
    return built


def random_views(sr, scenes, built, seed, n):
    rng = np.random.Generator(np.random.PCG64(seed))
    names = list(built)
    views, plain = [], []
    for _ in range(n):
        cfg = names[int(rng.integers(len(names)))]
        g, spec = built[cfg]
        w, h = int(rng.integers(8, 641)), int(rng.integers(8, 401))
        cam = random_pose(scenes, cfg, spec, rng, w, h)
        views.append((cfg, g, cam, w, h))
        plain.append(render_gpu(sr, g, cam, w, h))
    assert sum(int((p["ids"] != sr.NO_ID).any()) for p in plain) > 0.75 * n
    return views, plain


@pytest.mark.timeout(900)
def test_random_batch_matches_plain(sr, scenes, env):
    """48 random views over five scenes (the cluster plain, alpha-masked and
    translucent, the heightfield, the overdraw scene) at random sizes in one
    batch: launched eagerly, as a hipGraph replayed twice, and with the
    persistent raster, every view's visibility buffer and pixels equal the
    plain renderer's."""
    built = five_scenes(sr, scenes, env)
    vs, plain = random_views(sr, scenes, built, 6000, 48)
    views = [(g, cam, w, h, 2.0) for _, g, cam, w, h in vs]
    for persistent, use_graph in ((-1, False), (-1, True), (0, False)):
        b = sr.Batch(tile_size=16)
        b.configure(persistent_blocks=persistent, use_graph=use_graph)
        for rep in range(2 if use_graph else 1):
            b.render(views)
            for i, p in enumerate(plain):
                ids, depth = b.read_vbuffer(i)
                where = f"view {i} (persistent {persistent}, graph {use_graph}, replay {rep})"
                assert np.array_equal(ids, p["ids"]) and np.array_equal(bits(depth), bits(p["depth"])), where
                assert np.array_equal(b.read_pixels(i), p["px"]), where + " pixels"
        b.close()


@pytest.mark.timeout(900)
def test_random_requests_served_match_plain(sr, scenes, env):
    """48 random requests over the same five scenes submitted at once to the
    render server (batches of at most 8 over two overlapping slots, with an
    admission budget that splits them further): every request's visibility
    buffer and pixels equal the plain renderer's."""
    built = five_scenes(sr, scenes, env)
    vs, plain = random_views(sr, scenes, built, 7000, 48)
    srv = sr.Server(admission_bytes=256 << 20, max_batch_views=8, streams=2)
    idx = {cfg: srv.add_scene(g) for cfg, (g, _) in built.items()}
    rids = [srv.submit(idx[cfg], cam, w, h, tag=i) for i, (cfg, _, cam, w, h) in enumerate(vs)]
    srv.drain()
    done = srv.poll(64)
    done += srv.poll(64)
    assert sorted(d["request_id"] for d in done if d["request_id"] in rids) == sorted(rids)
    assert all(d["status"] == 0 for d in done)
    for i, rid in enumerate(rids):
        (ids, depth), px = srv.read_vbuffer(rid), srv.read_pixels(rid)
        p, h = plain[i], vs[i][4]
        assert np.array_equal(ids, p["ids"]) and np.array_equal(bits(depth), bits(p["depth"])), f"request {i}"
        assert np.array_equal(px[: h & ~1], p["px"][: h & ~1]), f"request {i} pixels"
        srv.release(rid)
    srv.close()
