"""Stage timing of the per-frame renderer (the reference's clip/bin and
raster timers, src/renderer.rs:206-219): stage times from the kernels'
wall-clock stamps (the default), from HIP events carried by the stage
launches, or none.  The mode changes no pixel, both timings split the frame
into setup, binning and raster-with-shading times that add up to its total
(events also split the raster from the shading), and they agree with each
other to within timing noise.
"""
import numpy as np
import pytest

from conftest import bits
from test_gpu_parity import _translucent_c1

pytestmark = pytest.mark.gpu

STAGES = ("ms_setup", "ms_bin", "ms_raster", "ms_shade", "ms_raster_shade")


@pytest.mark.parametrize("translucent", [False, True])
def test_stage_timing_modes(sr, scenes, env, translucent):
    if translucent:     # the frame ends with the translucent pass
        scene, spec = _translucent_c1(sr, scenes, 640, 360)
    else:
        scene, spec = scenes.make_c3(width=1280, height=720, triangles=20000, primitives=64)
    scene.prepare(env, device=0)
    g = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    r = sr.Renderer(spec["width"], spec["height"])
    frames, times = [], {}
    for mode in ("clock", "events", "off", "clock"):
        r.set_stage_timing(mode)
        for _ in range(3):
            r.render_scene(g, cam)
        st = r.stats()
        ids, depth = r.read_vbuffer()
        frames.append((ids, bits(depth)))
        times.setdefault(mode, st)
        print(f"[timing {mode}] " + " ".join(f"{k} {st[k]:.4f}" for k in STAGES + ("ms_total",)))
    r.close()
    assert all(np.array_equal(frames[0][0], f[0]) and np.array_equal(frames[0][1], f[1]) for f in frames[1:])
    for mode in ("clock", "events"):
        t = times[mode]
        parts = ("ms_setup", "ms_bin", "ms_raster_shade")
        assert all(t[k] > 0.0 for k in parts), f"{mode}: a stage has no time"
        total = sum(t[k] for k in parts)
        assert abs(total - t["ms_total"]) <= 1e-4 + 1e-4 * t["ms_total"], f"{mode}: stages do not add up"
    # the raster / shading split comes with events only
    t = times["events"]
    assert t["ms_raster"] > 0.0 and t["ms_shade"] > 0.0
    assert abs(t["ms_raster"] + t["ms_shade"] - t["ms_raster_shade"]) <= 1e-4 + 1e-4 * t["ms_total"]
    assert times["clock"]["ms_raster"] == 0.0 and times["clock"]["ms_shade"] == 0.0
    assert all(times["off"][k] == 0.0 for k in STAGES + ("ms_total", "ms_raster_shade"))
    ratio = times["clock"]["ms_total"] / times["events"]["ms_total"]
    assert 0.5 < ratio < 2.0, f"wall-clock and event totals differ by {ratio:.2f}x"
