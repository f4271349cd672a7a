"""GPU test of the device auto-exposure step (sr_renderer_update_auto_
exposure_device; Renderer::update_auto_exposure, src/renderer.rs:258-290)
against the host step, which the oracle tests pin bit for bit: over a run
of frames with adaptation the exposures agree to 1e-5 relative (the
device's log2f / expf / powf against the host libm's) and the blits to
1 LSB; a zero step gives exactly 2^ev on both."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(sr, gscene, cams, W, H, device_step, dts):
    r = sr.Renderer(W, H)
    out = []
    for cam, dt in zip(cams, dts):
        r.render_scene(gscene, cam)
        if device_step:
            r.update_auto_exposure_device(dt)
        else:
            r.update_auto_exposure(dt)
        px = r.blit_to_buffer()
        out.append((r.exposure, px))
    r.close()
    return out


def test_device_exposure_matches_host(sr, scenes, env):
    W, H = 640, 360
    scene, spec = scenes.make_c1(width=W, height=H)
    scene.prepare(env, cache_key=scene.name)
    gscene = sr.Scene(scene)
    cams = []
    for k in range(8):
        s = dict(spec)
        s["yaw"] = spec.get("yaw", 0.0) + 0.15 * k
        cams.append(scenes.make_camera(s))
    dts = [1 / 60, 0.1, 0.25, 0.0, 1.0, 1 / 30, 2.0, 0.5]
    host = _frames(sr, gscene, cams, W, H, False, dts)
    dev = _frames(sr, gscene, cams, W, H, True, dts)
    exact = 0
    for k, ((eh, ph), (ed, pd)) in enumerate(zip(host, dev)):
        assert abs(ed - eh) <= 1e-5 * eh, f"frame {k}: device {ed!r} host {eh!r}"
        exact += ed == eh
        d = np.abs(sr.unpack_rgba(ph).astype(int) - sr.unpack_rgba(pd).astype(int)).max()
        assert d <= 1, f"frame {k}: blit differs by {d} LSB"
    print(f"[exposure] {exact} of {len(dts)} device exposures bit-identical to the host's", flush=True)


def test_device_exposure_zero_step_and_handover(sr, scenes, env):
    """dt = 0: exposure stays 2^ev (exactly the host's); set_exposure after a
    device step wins; a host step after a device step starts from the
    device's state."""
    W, H = 320, 180
    scene, spec = scenes.make_c1(width=W, height=H)
    scene.prepare(env, cache_key=scene.name)
    gscene = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    r = sr.Renderer(W, H)
    r.render_scene(gscene, cam)
    r.update_auto_exposure_device(0.0)
    assert r.exposure == 2.0
    r.update_auto_exposure_device(0.5)
    e1 = r.exposure
    assert e1 != 2.0
    r.set_exposure(3.0)
    r.update_auto_exposure_device(0.0)
    assert abs(r.exposure - 3.0) <= 3e-7 * 3.0
    r.update_auto_exposure_device(0.5)
    r.update_auto_exposure(0.5)       # host step continues from the device state
    h = sr.Renderer(W, H)
    h.render_scene(gscene, cam)
    h.set_exposure(3.0)
    h.update_auto_exposure(0.5)
    h.update_auto_exposure(0.5)
    assert abs(r.exposure - h.exposure) <= 1e-5 * h.exposure
    r.close()
    h.close()
