"""GPU test of the presented frame (sr_renderer_present / _to_device): the
blit with App::render_frame's copy into the pixels frame fused in
(src/main.rs:574-590: each packed pixel split by rgba8_unpack_u8,
src/util.rs:102-109, into the bytes r, g, b, a).  The frame equals the
blit's words split that way (the blit itself is pinned against the oracle
by the parity tests), on the host and in device memory, after a host or a
device exposure step; an odd last row is not written, as in the blit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _split(px):
    """rgba8_unpack_u8 of every word: (r, g, b, a) = bits 31-24, 23-16, 15-8, 7-0."""
    return np.stack([(px >> 24) & 255, (px >> 16) & 255, (px >> 8) & 255, px & 255], axis=-1).astype(np.uint8)


@pytest.mark.parametrize("W,H,device_step", [(640, 361, False), (333, 202, True)])
def test_present_matches_blit(sr, scenes, env, W, H, device_step):
    scene, spec = scenes.make_c1(width=W, height=H)
    scene.prepare(env, cache_key=scene.name)
    gscene = sr.Scene(scene)
    r = sr.Renderer(W, H)
    r.render_scene(gscene, scenes.make_camera(spec))
    if device_step:
        r.update_auto_exposure_device(0.25)
    else:
        r.update_auto_exposure(0.25)
    px = r.blit_to_buffer()
    frame = np.full((H, W, 4), 0xAB, np.uint8)
    r.present(frame)
    rows = H & ~1
    assert np.array_equal(frame[:rows], _split(px[:rows]))
    assert np.all(frame[rows:] == 0xAB)
    assert np.all(frame[:rows, :, 3] == 255)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), W * H * 4) == 0
    try:
        assert hip.hipMemset(dev, 0xCD, W * H * 4) == 0
        r.present_to_device(dev.value)
        d = np.empty((H, W, 4), np.uint8)
        assert hip.hipMemcpy(d.ctypes.data, dev, W * H * 4, 2) == 0      # hipMemcpyDeviceToHost
    finally:
        hip.hipFree(dev)
    assert np.array_equal(d[:rows], frame[:rows])
    assert np.all(d[rows:] == 0xCD)
    r.close()


def test_present_rejects_bad_frames(sr, scenes, env):
    W, H = 160, 90
    r = sr.Renderer(W, H)
    with pytest.raises(sr.SwrasterError):
        r.present()                                   # no frame rendered yet
    scene, spec = scenes.make_c1(width=W, height=H)
    scene.prepare(env, cache_key=scene.name)
    r.render_scene(sr.Scene(scene), scenes.make_camera(spec))
    with pytest.raises(ValueError):
        r.present(np.zeros((H, W, 3), np.uint8))
    rc = sr.lib().sr_renderer_present(r.handle, np.zeros(W * H * 4, np.uint8).ctypes.data_as(
        sr.C.POINTER(sr.C.c_uint8)), W + 1, H)
    assert rc == -1                               # SR_ERR_INVALID_ARGUMENT
    r.close()


def test_queued_frame_rerun_replays_exposure_and_blit(sr, scenes, env):
    """render_scene queues the frame; the first read checks it.  C3's first
    frame on a fresh renderer outgrows its bin-pair buffer, so the check
    reruns it: a device exposure step and a blit queued behind it are
    replayed, giving the bits and exposure of a frame checked before them
    (synchronize) and of a warm renderer's frame."""
    scene, spec = scenes.make_c3(width=1280, height=720, triangles=40_000)
    scene.prepare(env, cache_key=None)
    gscene = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    W, H = spec["width"], spec["height"]

    def frame(r, sync_first):
        r.render_scene(gscene, cam)
        if sync_first:
            r.synchronize()
        r.update_auto_exposure_device(0.5)
        px = r.blit_to_buffer()
        return px, r.exposure, r.stats()

    r1 = sr.Renderer(W, H)
    px1, e1, st = frame(r1, False)
    n_in = st["input_triangles"]
    first_pair_cap = 2 * (n_in + n_in // 8 + 1024)
    assert st["binned_pairs"] > first_pair_cap, "the first frame must outgrow its buffers"
    r2 = sr.Renderer(W, H)
    px2, e2, _ = frame(r2, True)
    r3 = sr.Renderer(W, H)
    r3.render_scene(gscene, cam)
    r3.synchronize()
    px3, e3, _ = frame(r3, False)
    assert e1 == e2 == e3
    assert np.array_equal(px1, px2) and np.array_equal(px1, px3)
    # the scene is in use while a frame is queued
    r1.render_scene(gscene, cam)
    with pytest.raises(sr.SwrasterError):
        sr.check(sr.lib().sr_scene_destroy(gscene.handle))
    r1.synchronize()
    for r in (r1, r2, r3):
        r.close()
