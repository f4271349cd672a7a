"""Committed fixtures: the scene generator's arrays and a small oracle
frame are frozen; the oracle and the generator must keep reproducing them."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_scene_generator_digests(sr, scenes):
    import hashlib
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    want = json.load(open(os.path.join(HERE, "scene_digests.json")))
    got = mg.digests(sr, scenes)
    for k in ("c1", "c2", "c3"):
        assert got[k]["sha256"] == want[k]["sha256"], f"{k} generator output changed"
        assert got[k]["triangles"] == want[k]["triangles"]
    assert got["cubemap_faces_sha256"] == want["cubemap_faces_sha256"]
    assert hashlib.sha256(sr.cubemap_faces().tobytes()).hexdigest() == want["cubemap_faces_sha256"]
    # full-size triangle counts of the benchmark configurations
    assert want["c1"]["triangles"] == 16 * 1280
    assert 2 * 723 * 723 == 1_045_458


def test_oracle_matches_committed_c1_frame(sr, orc, scenes, env):
    g = np.load(os.path.join(HERE, "c1_160x90.npz"))
    scene, spec = scenes.make_c1(width=160, height=90)
    scene.prepare(env, cache_key=scene.name)
    cam = scenes.make_camera(spec)
    o = orc.OracleRenderer(160, 90)
    o.render_scene(orc.OracleScene(scene), cam)
    f = o.read_frame()
    assert np.array_equal(f["ids"], g["ids"])
    assert np.array_equal(f["depth"].view(np.uint32), g["depth"].view(np.uint32))
    assert np.array_equal(o.tile_luminance().view(np.uint32), g["tile_luminance"].view(np.uint32))
    o.update_auto_exposure(0.0)
    assert np.array_equal(o.blit(), g["pixels"])
