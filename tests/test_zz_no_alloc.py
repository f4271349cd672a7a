"""sr_render_frame_noalloc, the entry point that allocates no device memory,
against the allocating path (render_scene + update_auto_exposure(0) + blit
and read_vbuffer of the plain renderer): triangle ids, depth bits and RGBA8
pixels identical, on two calls in a row with the second smaller, then the
1080p heightfield; a frame too large for the scratch array and a translucent
scene are rejected with clear errors before any launch."""
import ctypes as C

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

SR_ERR_UNSUPPORTED, SR_ERR_CAPACITY = -4, -5        # include/swraster.h


class DeviceFrame:
    """Caller-owned device outputs (ids, depth, rgba) of one frame."""

    def __init__(self, w, h):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.w, self.h = w, h
        self.ptrs = []
        for _ in range(3):
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), w * h * 4) == 0
            self.ptrs.append(p)

    def read(self):
        out = [np.empty((self.h, self.w), t) for t in (np.uint32, np.float32, np.uint32)]
        for a, p in zip(out, self.ptrs):
            assert self.hip.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0     # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def noalloc(sr, gscene, cam, w, h, exposure=0.0):
    f = DeviceFrame(w, h)
    try:
        rc = sr.lib().sr_render_frame_noalloc(gscene.handle, C.byref(cam.c), w, h, exposure, f.ptrs[0], f.ptrs[1],
                                             f.ptrs[2], None)
        msg = sr.lib().sr_last_error().decode(errors="replace")
        return rc, msg, (f.read() if rc == 0 else None)
    finally:
        f.free()


def plain(sr, gscene, cam, w, h):
    r = sr.Renderer(w, h)
    r.render_scene(gscene, cam)
    ids, depth = r.read_vbuffer()
    r.update_auto_exposure(0.0)
    px = r.blit_to_buffer()
    r.close()
    return ids, depth, px


def check_same(sr, got, want, name):
    ids, depth, px = got
    wids, wdepth, wpx = want
    h = ids.shape[0]
    assert np.array_equal(ids, wids), f"{name}: triangle ids differ"
    assert np.array_equal(bits(depth), bits(wdepth)), f"{name}: depth differs"
    assert np.array_equal(px[: h & ~1], wpx[: h & ~1]), f"{name}: pixels differ"
    print(f"[noalloc] {name}: identical to the allocating path, {(ids != sr.NO_ID).sum()} px covered")


def test_noalloc_matches_allocating_path_twice(sr, scenes, env):
    scene, spec = scenes.make_c1()
    scene.prepare(env, device=0)
    g = sr.Scene(scene)
    s2 = dict(spec)
    s2["yaw"] = 0.3
    s2["position"] = [1.0, 0.5, 10.0]
    for name, sp, (w, h) in (("c1 640x360", spec, (640, 360)), ("c1 320x240, moved", s2, (320, 240))):
        cam = scenes.make_camera(sp, w, h)
        rc, msg, got = noalloc(sr, g, cam, w, h)
        assert rc == 0, msg
        check_same(sr, got, plain(sr, g, cam, w, h), name)


def test_noalloc_heightfield_1080p(sr, scenes, env):
    scene, spec = scenes.make_c2()
    scene.prepare(env, device=0)
    g = sr.Scene(scene)
    cam = scenes.make_camera(spec)
    rc, msg, got = noalloc(sr, g, cam, 1920, 1080)
    assert rc == 0, msg
    check_same(sr, got, plain(sr, g, cam, 1920, 1080), "c2 1920x1080")


def test_noalloc_rejects_what_does_not_fit(sr, scenes, env):
    """A 2M-triangle heightfield needs more than the scratch array holds: the
    call fails with SR_ERR_CAPACITY before launching anything; a translucent
    scene fails with SR_ERR_UNSUPPORTED."""
    pool = scenes.make_c4_pool()
    big = pool[-1]
    big.prepare(env, device=0)
    gb = sr.Scene(big)
    _, spec = scenes.make_c2()
    cam = scenes.make_camera(spec, 640, 360)
    rc, msg, _ = noalloc(sr, gb, cam, 640, 360)
    assert rc == SR_ERR_CAPACITY and "scratch" in msg, (rc, msg)
    scene, spec = scenes.make_c1(width=320, height=240)
    scene.materials[0]["is_translucent"] = 1
    scene.materials[0]["transmission"] = 0.5
    scene.rebuild()
    scene.prepare(env, grid=16)
    gt = sr.Scene(scene)
    rc, msg, _ = noalloc(sr, gt, scenes.make_camera(spec), 320, 240)
    assert rc == SR_ERR_UNSUPPORTED, (rc, msg)
