"""Parity sweep over random poses: many small frames of the sphere cluster
(also with alpha-masked and with translucent materials), the heightfield
and a 20,000-triangle overdraw scene, each at a random
viewport (8-640 x 8-400), tile size (8, 16, 32) and exposure step (0 or
0.1 s), compared with the oracle: visibility buffer bit for bit, setup
count, RGBA8 within 1 LSB.  SR_SWEEP_SCALE multiplies the number of frames
(the default takes about 20 s; 5 gives about 2,600 frames)."""
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
