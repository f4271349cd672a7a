"""ctypes wrapper of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  The product path
never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [P]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_renderer_create.restype = P
        L.orc_renderer_create.argtypes = [C.c_int, C.c_int]
        L.orc_renderer_destroy.argtypes = [P]
        L.orc_render.argtypes = [P, P, P, C.c_int, C.c_int, C.c_int]
        L.orc_stage_ms.restype = C.c_double
        L.orc_stage_ms.argtypes = [P, C.c_int]
        L.orc_setup_count.restype = C.c_uint64
        L.orc_setup_count.argtypes = [P]
        L.orc_read_frame.argtypes = [P, P, P, P, P, P]
        L.orc_update_auto_exposure.argtypes = [P, C.c_float]
        L.orc_exposure.restype = C.c_float
        L.orc_exposure.argtypes = [P]
        L.orc_blit.argtypes = [P, P, C.c_size_t, C.c_size_t]
        L.orc_tile_luminance.argtypes = [P, P]
        L.orc_camera_new.argtypes = [P, P, C.c_float, C.c_float, C.c_float, C.c_float, P]
        L.orc_camera_update.argtypes = [P]
        L.orc_quantize.restype = C.c_uint32
        L.orc_quantize.argtypes = [C.c_float, C.c_float]
        L.orc_mip_chain.argtypes = [P, P, P]
        L.orc_sort_translucent.argtypes = [P, P, C.c_size_t]
        L.orc_shade_probe.argtypes = [P, P, C.c_int, P, P, P, P, P, P]
        L.orc_skybox_dir.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_env_sample.argtypes = [P, C.c_int, P, C.c_float, P]
        L.orc_exposure_step.argtypes = [P, C.c_size_t, C.c_float, C.c_float, C.c_float, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleScene:
    def __init__(self, scene_data):
        self.data = scene_data
        self.handle = lib().orc_scene_create(C.addressof(scene_data.desc))

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_scene_destroy(self.handle)
            self.handle = None


class OracleRenderer:
    """CPU restatement of Renderer; canonical=True gives the deterministic
    submission order the GPU vbuffer is compared against."""

    def __init__(self, width, height):
        self.width, self.height = width, height
        self.handle = lib().orc_renderer_create(width, height)

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_renderer_destroy(self.handle)
            self.handle = None

    def render_scene(self, scene, camera, threads=1, fp64=False, canonical=True):
        lib().orc_render(self.handle, scene.handle, C.addressof(camera.c), threads, int(fp64), int(canonical))

    def stage_ms(self):
        return lib().orc_stage_ms(self.handle, 0), lib().orc_stage_ms(self.handle, 1)

    def setup_count(self):
        return lib().orc_setup_count(self.handle)

    def read_frame(self, fp64=False, bary=False):
        n = self.width * self.height
        ids = np.empty(n, np.uint32)
        depth = np.empty(n, np.float32)
        rgb = np.empty(n * 3, np.float32)
        rgb64 = np.empty(n * 3, np.float64) if fp64 else None
        b = np.empty(n * 2, np.float32) if bary else None
        lib().orc_read_frame(self.handle, _p(ids), _p(depth), _p(rgb), _p(rgb64), _p(b))
        out = {"ids": ids.reshape(self.height, self.width), "depth": depth.reshape(self.height, self.width),
               "rgb": rgb.reshape(self.height, self.width, 3)}
        if fp64:
            out["rgb64"] = rgb64.reshape(self.height, self.width, 3)
        if bary:
            out["bary"] = b.reshape(self.height, self.width, 2)
        return out

    def update_auto_exposure(self, dt):
        lib().orc_update_auto_exposure(self.handle, float(dt))

    @property
    def exposure(self):
        return lib().orc_exposure(self.handle)

    def blit(self):
        px = np.zeros((self.height, self.width), np.uint32)
        lib().orc_blit(self.handle, _p(px), self.width, self.height)
        return px

    def tile_luminance(self):
        tx = (self.width + 63) // 64
        ty = (self.height + 63) // 64
        out = np.empty(tx * ty, np.float32)
        lib().orc_tile_luminance(self.handle, _p(out))
        return out


def camera_new(sr_camera_cls, position, look_at, fov, width, height, far_plane):
    c = sr_camera_cls()
    lib().orc_camera_new((C.c_float * 3)(*position), (C.c_float * 3)(*look_at), fov, width, height, far_plane,
                         C.byref(c))
    return c


def camera_update(cam_struct):
    lib().orc_camera_update(C.byref(cam_struct))
    return cam_struct


def quantize(c, exposure):
    return lib().orc_quantize(c, exposure)


def mip_chain(tex_desc):
    n = C.c_uint32()
    lib().orc_mip_chain(C.byref(tex_desc), None, C.byref(n))
    out = np.empty(n.value, np.uint32)
    lib().orc_mip_chain(C.byref(tex_desc), _p(out), C.byref(n))
    return out


def sort_translucent(avg_z):
    """The oracle's per-tile translucent packet order for packets with these
    avg_z values in queue order (returns the permutation)."""
    z = np.ascontiguousarray(avg_z, np.float32)
    order = np.empty(len(z), np.uint32)
    lib().orc_sort_translucent(_p(z), _p(order), len(z))
    return order


def _f32(a, n):
    a = np.ascontiguousarray(a, np.float32).ravel()
    assert a.size == n
    return a


def shade_probe(scene, camera, material, normal, pos, uv=(0.0, 0.0), du_dv=(0.0, 0.0, 0.0, 0.0)):
    """pbr_shader at one point of a packet with constant interpolators
    (normal, world position with 1/w = 1, uv): (f32 rgb, f64 rgb)."""
    n, p, t, d = _f32(normal, 3), _f32(pos, 3), _f32(uv, 2), _f32(du_dv, 4)
    o32 = np.empty(3, np.float32)
    o64 = np.empty(3, np.float64)
    rc = lib().orc_shade_probe(scene.handle, C.addressof(camera.c), int(material), _p(n), _p(p), _p(t), _p(d),
                               _p(o32), _p(o64))
    assert rc == 0
    return o32, o64


def skybox_dir(camera, x, y):
    """compute_skybox's normalized view directions of the quad at (x, y): (4, 3)."""
    out = np.empty(12, np.float32)
    lib().orc_skybox_dir(C.addressof(camera.c), int(x), int(y), _p(out))
    return out.reshape(4, 3)


def env_sample(scene, which, a, level=0.0):
    """which: 'specular' (trilinear at direction a, mip level), 'lut'
    (bilinear, clamp-to-edge, at (a[0], a[1])), 'sky' (nearest, mip 0)."""
    k = {"specular": 0, "lut": 1, "sky": 2}[which]
    v = np.zeros(3, np.float32)
    v[: len(a)] = a
    out = np.empty(3, np.float32)
    assert lib().orc_env_sample(scene.handle, k, _p(v), float(level), _p(out)) == 0
    return out


def exposure_step(lum, dt, ev, target):
    """The auto-exposure step on tile luminances: (ev, target, exposure)."""
    l = np.ascontiguousarray(lum, np.float32)
    st = np.empty(3, np.float32)
    assert lib().orc_exposure_step(_p(l), l.size, float(dt), float(ev), float(target), _p(st)) == 0
    return float(st[0]), float(st[1]), float(st[2])
