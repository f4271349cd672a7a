"""swraster-viewer_amd -- Python mirror of the MI355X rasterization pipeline.

The product is ``lib/libswraster.so`` (C ABI in ``include/swraster.h``): a C++
host runtime plus hand-written gfx950 HIP kernels.  This module binds it with
ctypes and mirrors the reference's per-frame interface:

  RenderCamera.new / update_matrices   src/rendercamera.rs:28, :78
  Renderer(width, height)              src/renderer.rs:165
  Renderer.render_scene(scene, camera) src/renderer.rs:201
  Renderer.update_auto_exposure(dt)    src/renderer.rs:258
  Renderer.blit_to_buffer(buffer)      src/renderer.rs:293

Errors raise ``SwrasterError`` carrying the library's status and text.  There
is no CPU fallback: without the built library every entry point raises.
"""
import ctypes as C
import os
import sys

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libswraster.so")

SR_OK = 0
SR_TEX_SRGB, SR_TEX_NORMAL, SR_TEX_METALLIC_ROUGHNESS, SR_TEX_LINEAR = 0, 1, 2, 3
SR_WRAP_REPEAT, SR_WRAP_MIRRORED_REPEAT, SR_WRAP_CLAMP_TO_EDGE = 0, 1, 2
SR_ALPHA_OPAQUE, SR_ALPHA_MASK, SR_ALPHA_BLEND = 0, 1, 2
NO_ID = 0xFFFFFFFF
STAGE_TIMING = {"clock": 0, "events": 1, "off": 2}   # SR_STAGE_TIMING_*


class SwrasterError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"swraster error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------------------
# C structures (include/swraster.h)
# ---------------------------------------------------------------------------
class sr_texture_desc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("type", C.c_int32),
                ("texels", C.POINTER(C.c_uint32))]


class sr_texref(C.Structure):
    _fields_ = [("texture", C.c_int32), ("wrap_s", C.c_int32), ("wrap_t", C.c_int32)]


class sr_material_desc(C.Structure):
    _fields_ = [("base_color_factor", C.c_float * 4), ("metallic_factor", C.c_float),
                ("roughness_factor", C.c_float), ("emissive_factor", C.c_float * 3),
                ("occlusion_strength", C.c_float), ("alpha_cutoff", C.c_float),
                ("transmission", C.c_float), ("alpha_mode", C.c_int32), ("is_translucent", C.c_int32),
                ("base_color", sr_texref), ("metallic_roughness", sr_texref), ("normal", sr_texref),
                ("emissive", sr_texref), ("occlusion", sr_texref), ("transmission_texture", sr_texref)]


class sr_primitive_desc(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("texcoords", C.POINTER(C.c_float)), ("tangents", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("num_vertices", C.c_uint32),
                ("num_indices", C.c_uint32), ("material", C.c_int32)]


class sr_mesh_desc(C.Structure):
    _fields_ = [("first_primitive", C.c_int32), ("num_primitives", C.c_int32)]


class sr_node_desc(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("mesh", C.c_int32)]


class sr_environment(C.Structure):
    _fields_ = [("cubemap_size", C.c_uint32), ("cubemap", C.POINTER(C.c_uint32)),
                ("specular_size", C.c_uint32), ("specular_mips", C.c_uint32),
                ("specular", C.POINTER(C.c_uint32)), ("brdf_lut_size", C.c_uint32),
                ("brdf_lut", C.POINTER(C.c_uint32)), ("irradiance_sh", (C.c_float * 3) * 4)]


class sr_voxel_grid(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("world_min", C.c_float * 3), ("world_max", C.c_float * 3),
                ("light_intensity", C.POINTER(C.c_float)), ("sh_scale", C.c_float),
                ("sky_visibility", C.c_float), ("sh4", C.POINTER(C.c_float))]


class sr_light(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("color", C.c_float * 3)]


class sr_scene_desc(C.Structure):
    _fields_ = [("primitives", C.POINTER(sr_primitive_desc)), ("num_primitives", C.c_int32),
                ("meshes", C.POINTER(sr_mesh_desc)), ("num_meshes", C.c_int32),
                ("nodes", C.POINTER(sr_node_desc)), ("num_nodes", C.c_int32),
                ("materials", C.POINTER(sr_material_desc)), ("num_materials", C.c_int32),
                ("textures", C.POINTER(sr_texture_desc)), ("num_textures", C.c_int32),
                ("environment", C.POINTER(sr_environment)), ("voxels", C.POINTER(sr_voxel_grid)),
                ("light", sr_light)]


class sr_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("fov", C.c_float), ("width", C.c_float),
                ("height", C.c_float), ("one_over_width", C.c_float), ("one_over_height", C.c_float),
                ("near_plane", C.c_float), ("far_plane", C.c_float), ("yaw", C.c_float),
                ("pitch", C.c_float), ("roll", C.c_float), ("view_matrix", C.c_float * 16),
                ("projection_matrix", C.c_float * 16), ("view_project_matrix", C.c_float * 16),
                ("inverse_view_project_matrix", C.c_float * 16),
                ("skybox_matrix_transposed", C.c_float * 16), ("view_clip_planes", (C.c_float * 4) * 6),
                ("view_normal", C.c_float * 3)]


class sr_renderer_opts(C.Structure):
    _fields_ = [("tile_size", C.c_int32), ("raster_mode", C.c_int32), ("device", C.c_int32),
                ("reserved", C.c_int32 * 5)]


class sr_frame_stats(C.Structure):
    _fields_ = [("input_triangles", C.c_uint64), ("setup_triangles", C.c_uint64),
                ("binned_pairs", C.c_uint64), ("tiles", C.c_uint32), ("draws", C.c_uint32),
                ("ms_setup", C.c_float), ("ms_bin", C.c_float), ("ms_raster_shade", C.c_float),
                ("ms_blit", C.c_float), ("ms_total", C.c_float), ("ms_raster", C.c_float),
                ("ms_shade", C.c_float), ("depth_rejection", C.c_uint32)]


class sr_view(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", sr_camera), ("width", C.c_int32),
                ("height", C.c_int32), ("exposure", C.c_float)]


class sr_batch_opts(C.Structure):
    _fields_ = [("tile_size", C.c_int32), ("use_graph", C.c_int32), ("persistent_blocks", C.c_int32),
                ("device", C.c_int32), ("reserved", C.c_int32 * 4)]


class sr_batch_stats(C.Structure):
    _fields_ = [("input_triangles", C.c_uint64), ("setup_triangles", C.c_uint64),
                ("binned_pairs", C.c_uint64), ("pixels", C.c_uint64), ("views", C.c_uint32),
                ("work_items", C.c_uint32), ("ms_total", C.c_float)]


class sr_server_opts(C.Structure):
    _fields_ = [("admission_bytes", C.c_uint64), ("max_batch_views", C.c_int32), ("streams", C.c_int32),
                ("tile_size", C.c_int32), ("device", C.c_int32), ("reserved", C.c_int32 * 4)]


class sr_request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("camera", sr_camera), ("width", C.c_int32), ("height", C.c_int32),
                ("exposure", C.c_float), ("user_tag", C.c_uint64)]


class sr_completion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("user_tag", C.c_uint64), ("status", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32), ("submit_time_s", C.c_double),
                ("start_time_s", C.c_double), ("complete_time_s", C.c_double)]


# Every entry point declared in include/swraster.h, with its argument types.
_P = C.c_void_p
_SIGS = {
    "sr_last_error": (C.c_char_p, []),
    "sr_version": (C.c_int, [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "sr_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "sr_prepare_environment": (C.c_int, [C.POINTER(C.c_uint32), C.c_uint32, C.c_int32, C.POINTER(sr_environment)]),
    "sr_environment_free": (C.c_int, [C.POINTER(sr_environment)]),
    "sr_prepare_environment_device": (C.c_int, [C.POINTER(C.c_uint32), C.c_uint32, C.c_int32,
                                                C.POINTER(sr_environment)]),
    "sr_prepare_voxel_grid": (C.c_int, [C.POINTER(sr_scene_desc), C.c_int32, C.c_int32, C.POINTER(sr_voxel_grid)]),
    "sr_voxel_grid_free": (C.c_int, [C.POINTER(sr_voxel_grid)]),
    "sr_prepare_gi": (C.c_int, [C.POINTER(sr_scene_desc), C.c_int32, C.c_int32, C.c_int32,
                                C.POINTER(sr_voxel_grid)]),
    "sr_gi_cache_save": (C.c_int, [C.c_char_p, C.POINTER(sr_voxel_grid)]),
    "sr_gi_cache_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "sr_specular_cache_save": (C.c_int, [C.c_char_p, C.POINTER(sr_environment)]),
    "sr_specular_cache_load": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "sr_prepare_voxel_grid_device": (C.c_int, [C.POINTER(sr_scene_desc), C.c_int32, C.c_int32,
                                               C.POINTER(sr_voxel_grid)]),
    "sr_scene_desc_bounds": (C.c_int, [C.POINTER(sr_scene_desc), C.c_float * 3, C.c_float * 3, C.c_float * 3,
                                       C.POINTER(C.c_float)]),
    "sr_texture_mip_chain": (C.c_int, [C.POINTER(sr_texture_desc), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32), C.c_uint32 * 32, C.c_uint32 * 32, C.c_uint32 * 32]),
    "sr_gltf_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "sr_gltf_scene_desc": (C.c_int, [_P, C.POINTER(sr_scene_desc)]),
    "sr_gltf_destroy": (C.c_int, [_P]),
    "sr_png_decode": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint32), C.c_size_t]),
    "sr_jpeg_decode": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint32), C.c_size_t]),
    "sr_camera_new": (C.c_int, [C.c_float * 3, C.c_float * 3, C.c_float, C.c_float, C.c_float, C.c_float,
                                C.POINTER(sr_camera)]),
    "sr_camera_from_pose": (C.c_int, [C.c_float * 3, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_float, C.c_float, C.POINTER(sr_camera)]),
    "sr_camera_update_matrices": (C.c_int, [C.POINTER(sr_camera)]),
    "sr_scene_create": (C.c_int, [C.POINTER(sr_scene_desc), C.c_int32, C.POINTER(_P)]),
    "sr_scene_destroy": (C.c_int, [_P]),
    "sr_scene_info": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "sr_renderer_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(sr_renderer_opts), C.POINTER(_P)]),
    "sr_renderer_destroy": (C.c_int, [_P]),
    "sr_renderer_render_scene": (C.c_int, [_P, _P, C.POINTER(sr_camera)]),
    "sr_render_frame_noalloc": (C.c_int, [_P, C.POINTER(sr_camera), C.c_int32, C.c_int32, C.c_float, _P, _P, _P, _P]),
    "sr_renderer_update_auto_exposure": (C.c_int, [_P, C.c_float]),
"sr_renderer_update_auto_exposure_device": (C.c_int, [_P, C.c_float]),
    "sr_renderer_blit_to_buffer": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_size_t, C.c_size_t]),
    "sr_renderer_blit_to_device": (C.c_int, [_P, _P]),
    "sr_renderer_present": (C.c_int, [_P, C.POINTER(C.c_uint8), C.c_size_t, C.c_size_t]),
    "sr_renderer_synchronize": (C.c_int, [_P]),
    "sr_renderer_present_to_device": (C.c_int, [_P, _P]),
    "sr_renderer_exposure": (C.c_float, [_P]),
    "sr_renderer_set_exposure": (C.c_int, [_P, C.c_float]),
    "sr_renderer_read_vbuffer": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "sr_renderer_read_color": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "sr_renderer_read_tile_luminance": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "sr_renderer_get_stats": (C.c_int, [_P, C.POINTER(sr_frame_stats)]),
    "sr_renderer_set_stage_timing": (C.c_int, [_P, C.c_int32]),
    "sr_renderer_device_buffers": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "sr_batch_create": (C.c_int, [C.POINTER(sr_batch_opts), C.POINTER(_P)]),
    "sr_batch_destroy": (C.c_int, [_P]),
    "sr_batch_configure": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "sr_batch_render": (C.c_int, [_P, C.POINTER(sr_view), C.c_int32]),
    "sr_batch_read_vbuffer": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "sr_batch_read_pixels": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint32)]),
    "sr_batch_get_stats": (C.c_int, [_P, C.POINTER(sr_batch_stats)]),
    "sr_server_create": (C.c_int, [C.POINTER(sr_server_opts), C.POINTER(_P)]),
    "sr_server_destroy": (C.c_int, [_P]),
    "sr_server_add_scene": (C.c_int, [_P, _P, C.POINTER(C.c_int32)]),
    "sr_server_submit": (C.c_int, [_P, C.POINTER(sr_request), C.POINTER(C.c_uint64)]),
    "sr_server_poll": (C.c_int, [_P, C.POINTER(sr_completion), C.c_int32, C.POINTER(C.c_int32)]),
    "sr_server_read_pixels": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_uint32)]),
    "sr_server_read_vbuffer": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "sr_server_release": (C.c_int, [_P, C.c_uint64]),
    "sr_renderer_read_bins": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.c_uint64, C.POINTER(C.c_uint32)]),
    "sr_server_drain": (C.c_int, [_P]),
    "sr_server_clock": (C.c_double, [_P]),
}

_lib = None


def lib():
    """The loaded product library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SwrasterError(-1, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != SR_OK:
        raise SwrasterError(code, lib().sr_last_error().decode(errors="replace"))
    return code


def jpeg_decode(data):
    """Decode JPEG bytes with the library's decoder (the one the glTF loader
    uses): (height, width) array of packed r<<24|g<<16|b<<8|0xFF texels."""
    buf = bytes(data)
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().sr_jpeg_decode(buf, len(buf), C.byref(w), C.byref(h), None, 0))
    out = np.empty(w.value * h.value, np.uint32)
    check(lib().sr_jpeg_decode(buf, len(buf), C.byref(w), C.byref(h), out.ctypes.data_as(C.POINTER(C.c_uint32)),
                               out.size))
    return out.reshape(h.value, w.value)


def png_decode(data):
    """Decode PNG bytes with the library's decoder (the one the glTF loader
    uses): (height, width) array of packed r<<24|g<<16|b<<8|a texels."""
    buf = bytes(data)
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().sr_png_decode(buf, len(buf), C.byref(w), C.byref(h), None, 0))
    out = np.empty(w.value * h.value, np.uint32)
    check(lib().sr_png_decode(buf, len(buf), C.byref(w), C.byref(h), out.ctypes.data_as(C.POINTER(C.c_uint32)),
                              out.size))
    return out.reshape(h.value, w.value)


def device_count():
    n = C.c_int32(0)
    check(lib().sr_device_count(C.byref(n)))
    return n.value


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32)) if a is not None else None


def pack_rgba(rgba):
    """(..., 4) uint8 -> packed r<<24 | g<<16 | b<<8 | a (src/util.rs:98-100)."""
    rgba = np.asarray(rgba, dtype=np.uint32)
    return ((rgba[..., 0] << 24) | (rgba[..., 1] << 16) | (rgba[..., 2] << 8) | rgba[..., 3]).astype(np.uint32)


def unpack_rgba(px):
    px = np.asarray(px, dtype=np.uint32)
    return np.stack([(px >> 24) & 255, (px >> 16) & 255, (px >> 8) & 255, px & 255], -1).astype(np.uint8)


# ---------------------------------------------------------------------------
# Camera (RenderCamera, src/rendercamera.rs)
# ---------------------------------------------------------------------------
class RenderCamera:
    """Mirror of RenderCamera: ``new`` derives yaw/pitch from a look-at point
    (with the reference's pitch convention), ``from_pose`` takes them
    directly; ``update_matrices`` fills view/projection/skybox matrices and
    the view-space cull planes."""

    def __init__(self, cam):
        self.c = cam

    @classmethod
    def new(cls, position, look_at, fov, width, height, far_plane):
        c = sr_camera()
        check(lib().sr_camera_new((C.c_float * 3)(*position), (C.c_float * 3)(*look_at), fov, width, height,
                                  far_plane, C.byref(c)))
        return cls(c)

    @classmethod
    def from_pose(cls, position, yaw, pitch, fov, width, height, far_plane, roll=0.0):
        c = sr_camera()
        check(lib().sr_camera_from_pose((C.c_float * 3)(*position), yaw, pitch, roll, fov, width, height,
                                        far_plane, C.byref(c)))
        return cls(c)

    def update_matrices(self):
        check(lib().sr_camera_update_matrices(C.byref(self.c)))
        return self

    def matrix(self, name):
        return np.array(getattr(self.c, name), dtype=np.float32).reshape(4, 4).T  # row-major view


# ---------------------------------------------------------------------------
# Environment and voxel inputs
# ---------------------------------------------------------------------------
CACHE_DIR = os.path.join(REPO_DIR, "assets", ".cache")


def cubemap_faces():
    """Skybox faces (+X,-X,+Y,-Y,+Z,-Z) packed RGBA8 with alpha 255."""
    faces = np.load(os.path.join(REPO_DIR, "assets", "cubemap_faces.npz"))["faces"]
    rgba = np.concatenate([faces, np.full(faces.shape[:-1] + (1,), 255, np.uint8)], -1)
    return np.ascontiguousarray(pack_rgba(rgba).reshape(6, faces.shape[1], faces.shape[2]))


class Environment:
    """Image-based lighting inputs (src/scene.rs:152-228), cached on disk."""

    def __init__(self, cubemap, specular, specular_mips, lut, sh):
        self.cubemap = np.ascontiguousarray(cubemap, np.uint32)
        self.specular = np.ascontiguousarray(specular, np.uint32)
        self.specular_mips = int(specular_mips)
        self.lut = np.ascontiguousarray(lut, np.uint32)
        self.sh = np.ascontiguousarray(sh, np.float32).reshape(4, 3)
        self.desc = sr_environment()
        self.desc.cubemap_size = self.cubemap.shape[-1]
        self.desc.cubemap = _uptr(self.cubemap)
        self.desc.specular_size = self.cubemap.shape[-1]
        self.desc.specular_mips = self.specular_mips
        self.desc.specular = _uptr(self.specular)
        self.desc.brdf_lut_size = int(round(np.sqrt(self.lut.size)))
        self.desc.brdf_lut = _uptr(self.lut)
        for i in range(4):
            for k in range(3):
                self.desc.irradiance_sh[i][k] = float(self.sh[i, k])

    @classmethod
    def load(cls, threads=0, cache=True, device=None):
        """The skybox's IBL inputs, computed on the host (threads) or on a
        HIP device (device ordinal), cached on disk per variant."""
        path = os.path.join(CACHE_DIR, "environment_v1.npz" if device is None else "environment_device_v1.npz")
        faces = cubemap_faces()
        if cache and os.path.exists(path):
            z = np.load(path)
            if np.array_equal(z["cubemap"], faces):
                return cls(z["cubemap"], z["specular"], int(z["mips"]), z["lut"], z["sh"])
        env = sr_environment()
        if device is None:
            check(lib().sr_prepare_environment(_uptr(faces), faces.shape[-1], threads, C.byref(env)))
        else:
            check(lib().sr_prepare_environment_device(_uptr(faces), faces.shape[-1], int(device), C.byref(env)))
        size = env.cubemap_size
        spec = np.ctypeslib.as_array(env.specular, shape=(env.specular_mips * 6 * size * size,)).copy()
        lut = np.ctypeslib.as_array(env.brdf_lut, shape=(env.brdf_lut_size ** 2,)).copy()
        sh = np.array([[env.irradiance_sh[i][k] for k in range(3)] for i in range(4)], np.float32)
        mips = env.specular_mips
        lib().sr_environment_free(C.byref(env))
        if cache:
            os.makedirs(CACHE_DIR, exist_ok=True)
            np.savez(path, cubemap=faces, specular=spec, mips=mips, lut=lut, sh=sh)
        return cls(faces, spec, mips, lut, sh)


def save_specular_cache(path, env):
    """Write env's GGX-prefiltered cubemap as the reference's .ggx file
    (src/texture.rs:520-552)."""
    check(lib().sr_specular_cache_save(os.fsencode(path), C.byref(env.desc)))


def load_specular_cache(path, face_size):
    """The specular mip chain from a .ggx file (src/texture.rs:426-518), or
    None when the file is missing or does not fit face_size."""
    mips = 1 + int(face_size).bit_length() - 1
    spec = np.empty(mips * 6 * face_size * face_size, np.uint32)
    if lib().sr_specular_cache_load(os.fsencode(path), face_size, _uptr(spec)) != 0:
        return None
    return spec, mips


class VoxelGrid:
    """Voxel lighting grid: per-voxel sun visibility and, after a GI bake,
    16 SH floats per voxel (sh4, x fastest); without sh4 every voxel uses
    sh_scale times the scene SH."""

    def __init__(self, dims, wmin, wmax, light, sh_scale=0.25, sky=1.0, sh4=None):
        self.light = np.ascontiguousarray(light, np.float32).ravel()
        self.sh4 = None if sh4 is None else np.ascontiguousarray(sh4, np.float32).ravel()
        self.desc = sr_voxel_grid()
        for k in range(3):
            self.desc.dims[k] = int(dims[k])
            self.desc.world_min[k] = float(wmin[k])
            self.desc.world_max[k] = float(wmax[k])
        self.desc.light_intensity = _fptr(self.light)
        self.desc.sh_scale = sh_scale
        self.desc.sky_visibility = sky
        self.desc.sh4 = None if self.sh4 is None else _fptr(self.sh4)
        self.dims = tuple(int(d) for d in dims)
        self.wmin = np.array(wmin, np.float32)
        self.wmax = np.array(wmax, np.float32)


# ---------------------------------------------------------------------------
# Scene description (host arrays) and the HBM-resident scene
# ---------------------------------------------------------------------------
def default_light():
    d = np.array([-0.2, 1.0, 0.5], np.float32)
    # Vec3A::normalize: v * (1 / sqrt(dot)) in f32 (src/scene.rs:237)
    dot = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2])
    r = np.float32(1.0) / np.float32(np.sqrt(dot))
    col = np.array([1.0, 1.0, 0.95], np.float32) * np.float32(5.0)
    return (d * r).astype(np.float32), col.astype(np.float32)


class SceneData:
    """Host-side scene: primitives, meshes, nodes, materials, textures.

    primitives: list of dicts with positions (N,3) f32, normals (N,3) f32,
    texcoords (N,2) f32 or None, tangents (N,4) f32 or None, indices (M,) u32,
    material int.  meshes: list of (first_primitive, num_primitives).
    nodes: list of (transform (4,4) row-major f32, mesh index or -1).
    materials: list of dicts (see material()).  textures: list of
    (width, height, type, packed u32 texels)."""

    def __init__(self, primitives, meshes, nodes, materials, textures, name="scene"):
        self.primitives = primitives
        self.meshes = meshes
        self.nodes = nodes
        self.materials = materials
        self.textures = textures
        self.name = name
        self.light_dir, self.light_color = default_light()
        self.environment = None
        self.voxels = None
        self._build_desc()

    def rebuild(self):
        """Rebuild the C descriptor after editing primitives, materials or
        textures in place (before prepare/attach)."""
        self._build_desc()

    @property
    def triangles(self):
        return sum(len(p["indices"]) // 3 for p in self.primitives)

    def _build_desc(self):
        keep = []
        prims = (sr_primitive_desc * max(1, len(self.primitives)))()
        for i, p in enumerate(self.primitives):
            for k in ("positions", "normals", "texcoords", "tangents"):
                if p.get(k) is not None:
                    p[k] = np.ascontiguousarray(p[k], np.float32)
            p["indices"] = np.ascontiguousarray(p["indices"], np.uint32).ravel()
            prims[i].positions = _fptr(p["positions"])
            prims[i].normals = _fptr(p["normals"])
            prims[i].texcoords = _fptr(p.get("texcoords"))
            prims[i].tangents = _fptr(p.get("tangents"))
            prims[i].indices = _uptr(p["indices"])
            prims[i].num_vertices = len(p["positions"])
            prims[i].num_indices = len(p["indices"])
            prims[i].material = int(p.get("material", 0))
        meshes = (sr_mesh_desc * max(1, len(self.meshes)))()
        for i, (f, n) in enumerate(self.meshes):
            meshes[i].first_primitive = f
            meshes[i].num_primitives = n
        nodes = (sr_node_desc * max(1, len(self.nodes)))()
        for i, (t, m) in enumerate(self.nodes):
            col = np.asarray(t, np.float32).reshape(4, 4).T.ravel()  # column-major
            for k in range(16):
                nodes[i].transform[k] = float(col[k])
            nodes[i].mesh = int(m)
        mats = (sr_material_desc * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            d = mats[i]
            for k in range(4):
                d.base_color_factor[k] = float(m["base_color"][k])
            d.metallic_factor = float(m.get("metallic", 1.0))
            d.roughness_factor = float(m.get("roughness", 1.0))
            for k in range(3):
                d.emissive_factor[k] = float(m.get("emissive", (0, 0, 0))[k])
            d.occlusion_strength = float(m.get("occlusion_strength", 1.0))
            d.alpha_cutoff = float(m.get("alpha_cutoff", 0.5))
            d.transmission = float(m.get("transmission", 0.0))
            d.alpha_mode = int(m.get("alpha_mode", SR_ALPHA_OPAQUE))
            d.is_translucent = int(m.get("is_translucent", 0))
            for slot in ("base_color", "metallic_roughness", "normal", "emissive", "occlusion",
                         "transmission_texture"):
                ref = m.get(slot + "_tex")
                r = getattr(d, slot)
                if ref is None:
                    r.texture, r.wrap_s, r.wrap_t = -1, SR_WRAP_REPEAT, SR_WRAP_REPEAT
                else:
                    r.texture, r.wrap_s, r.wrap_t = ref[0], ref[1], ref[2]
        texs = (sr_texture_desc * max(1, len(self.textures)))()
        for i, (w, h, ty, data) in enumerate(self.textures):
            data = np.ascontiguousarray(data, np.uint32).ravel()
            self.textures[i] = (w, h, ty, data)
            texs[i].width, texs[i].height, texs[i].type = w, h, ty
            texs[i].texels = _uptr(data)
        d = sr_scene_desc()
        d.primitives, d.num_primitives = prims, len(self.primitives)
        d.meshes, d.num_meshes = meshes, len(self.meshes)
        d.nodes, d.num_nodes = nodes, len(self.nodes)
        d.materials, d.num_materials = mats, len(self.materials)
        d.textures, d.num_textures = texs, len(self.textures)
        for k in range(3):
            d.light.direction[k] = float(self.light_dir[k])
            d.light.color[k] = float(self.light_color[k])
        keep += [prims, meshes, nodes, mats, texs]
        self._keep = keep
        self.desc = d

    _TEX_SLOTS = (("base_color", "base_color"), ("metallic_roughness", "metallic_roughness"),
                  ("normal", "normal"), ("emissive", "emissive"), ("occlusion", "occlusion"),
                  ("transmission_texture", "transmission_texture"))

    @classmethod
    def from_gltf(cls, path, name=None):
        """Load a .gltf or .glb file with the library's C++ loader
        (sr_gltf_load: load_gltf + Scene::from_gltf, src/main.rs:81-97,
        src/scene.rs:145-815) and copy its arrays into a SceneData."""
        h = _P()
        check(lib().sr_gltf_load(os.fsencode(str(path)), C.byref(h)))
        try:
            d = sr_scene_desc()
            check(lib().sr_gltf_scene_desc(h, C.byref(d)))

            def arr(ptr, n, width):
                if not ptr or n == 0:
                    return None
                return np.ctypeslib.as_array(ptr, (n * width,)).copy().reshape(n, width) if width > 1 else \
                    np.ctypeslib.as_array(ptr, (n,)).copy()

            prims = []
            for i in range(d.num_primitives):
                p = d.primitives[i]
                nv = p.num_vertices
                pos = arr(p.positions, nv, 3)
                prims.append({"positions": pos if pos is not None else np.zeros((0, 3), np.float32),
                              "normals": arr(p.normals, nv, 3) if nv else np.zeros((0, 3), np.float32),
                              "texcoords": arr(p.texcoords, nv, 2), "tangents": arr(p.tangents, nv, 4),
                              "indices": arr(p.indices, p.num_indices, 1) if p.num_indices else
                              np.zeros(0, np.uint32), "material": int(p.material)})
            meshes = [(d.meshes[i].first_primitive, d.meshes[i].num_primitives) for i in range(d.num_meshes)]
            nodes = [(np.array(d.nodes[i].transform, np.float32).reshape(4, 4).T.copy(), int(d.nodes[i].mesh))
                     for i in range(d.num_nodes)]
            mats = []
            for i in range(d.num_materials):
                m = d.materials[i]
                md = {"base_color": tuple(float(x) for x in m.base_color_factor), "metallic": m.metallic_factor,
                      "roughness": m.roughness_factor, "emissive": tuple(float(x) for x in m.emissive_factor),
                      "occlusion_strength": m.occlusion_strength, "alpha_cutoff": m.alpha_cutoff,
                      "transmission": m.transmission, "alpha_mode": int(m.alpha_mode),
                      "is_translucent": int(m.is_translucent)}
                for key, field in cls._TEX_SLOTS:
                    r = getattr(m, field)
                    if r.texture >= 0:
                        md[key + "_tex"] = (int(r.texture), int(r.wrap_s), int(r.wrap_t))
                mats.append(md)
            texs = []
            for i in range(d.num_textures):
                t = d.textures[i]
                texs.append((int(t.width), int(t.height), int(t.type),
                             np.ctypeslib.as_array(t.texels, (t.width * t.height,)).copy()))
        finally:
            lib().sr_gltf_destroy(h)
        base = os.path.splitext(os.path.basename(str(path)))[0]
        return cls(prims, meshes, nodes, mats, texs, name=name or base)

    def bounds(self):
        mn, mx, ce = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
        dg = C.c_float()
        check(lib().sr_scene_desc_bounds(C.byref(self.desc), mn, mx, ce, C.byref(dg)))
        return np.array(mn[:], np.float32), np.array(mx[:], np.float32), np.array(ce[:], np.float32), dg.value

    def attach(self, environment, voxels):
        self.environment = environment
        self.voxels = voxels
        self.desc.environment = C.pointer(environment.desc)
        self.desc.voxels = C.pointer(voxels.desc)
        return self

    def prepare(self, environment=None, grid=128, threads=0, cache_key=None, device=None, gi=False,
                gi_cache=None):
        """Attach the IBL environment and build (or load) the voxel grid; with
        a device ordinal the bakes run on that GPU (identical result).  gi
        adds the voxel GI bake (the reference's --gi, src/main.rs:238-282);
        gi_cache names its .gi file (the reference uses the glTF path with
        the extension .gi): loaded when it fits, else written after the bake."""
        environment = environment or Environment.load(threads=threads)
        path = None
        if cache_key:
            path = os.path.join(CACHE_DIR, f"voxels_{cache_key}_{grid}{'_gi' if gi else ''}.npz")
            if os.path.exists(path):
                z = np.load(path)
                sh4 = z["sh4"] if gi else None
                return self.attach(environment, VoxelGrid(z["dims"], z["wmin"], z["wmax"], z["light"],
                                                          sh_scale=1.0 if gi else 0.25, sh4=sh4))
        g = sr_voxel_grid()
        if gi and gi_cache and os.path.exists(gi_cache):
            # the reference's --gi load path (src/main.rs:247-276): sun
            # visibility first, then the .gi cache if it fits the grid
            sun = self.prepare(environment, grid=grid, threads=threads, device=device).voxels
            sh4 = np.empty(int(np.prod(sun.dims)) * 16, np.float32)
            dims = (C.c_int32 * 3)(*sun.dims)
            if lib().sr_gi_cache_load(os.fsencode(gi_cache), dims, _fptr(sh4)) == 0:
                return self.attach(environment, VoxelGrid(sun.dims, sun.wmin, sun.wmax, sun.light,
                                                          sh_scale=1.0, sh4=sh4))
        if gi:
            self.environment = environment
            self.desc.environment = C.pointer(environment.desc)
            check(lib().sr_prepare_gi(C.byref(self.desc), grid, -1 if device is None else int(device),
                                      threads, C.byref(g)))
            n = g.dims[0] * g.dims[1] * g.dims[2]
            light = np.ctypeslib.as_array(g.light_intensity, shape=(n,)).copy()
            sh4 = np.ctypeslib.as_array(g.sh4, shape=(n * 16,)).copy()
            dims, wmin, wmax = list(g.dims), list(g.world_min), list(g.world_max)
            if gi_cache and lib().sr_gi_cache_save(os.fsencode(gi_cache), C.byref(g)) != 0:
                # as the reference (src/main.rs:258-262): report and go on
                print(f"Failed to write GI cache {gi_cache}: {lib().sr_last_error().decode(errors='replace')}",
                      file=sys.stderr)
            lib().sr_voxel_grid_free(C.byref(g))
            if path:
                os.makedirs(CACHE_DIR, exist_ok=True)
                np.savez(path, dims=dims, wmin=wmin, wmax=wmax, light=light, sh4=sh4)
            return self.attach(environment, VoxelGrid(dims, wmin, wmax, light, sh_scale=1.0, sh4=sh4))
        if device is None:
            check(lib().sr_prepare_voxel_grid(C.byref(self.desc), grid, threads, C.byref(g)))
        else:
            check(lib().sr_prepare_voxel_grid_device(C.byref(self.desc), grid, int(device), C.byref(g)))
        n = g.dims[0] * g.dims[1] * g.dims[2]
        light = np.ctypeslib.as_array(g.light_intensity, shape=(n,)).copy()
        dims, wmin, wmax = list(g.dims), list(g.world_min), list(g.world_max)
        lib().sr_voxel_grid_free(C.byref(g))
        if path:
            os.makedirs(CACHE_DIR, exist_ok=True)
            np.savez(path, dims=dims, wmin=wmin, wmax=wmax, light=light)
        return self.attach(environment, VoxelGrid(dims, wmin, wmax, light))


class Scene:
    """A scene resident in HBM (sr_scene_create)."""

    def __init__(self, data, device=0):
        if data.environment is None or data.voxels is None:
            data.prepare()
        self.data = data
        self.handle = C.c_void_p()
        check(lib().sr_scene_create(C.byref(data.desc), device, C.byref(self.handle)))

    def info(self):
        t, v, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().sr_scene_info(self.handle, C.byref(t), C.byref(v), C.byref(b)))
        return {"triangles": t.value, "vertices": v.value, "device_bytes": b.value}

    def close(self):
        """Frees the scene; raises SwrasterError (SR_ERR_BUSY) while a
        renderer's queued frame still uses it."""
        if self.handle:
            check(lib().sr_scene_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """Mirror of Renderer (src/renderer.rs:145-355) on one GPU."""

    def __init__(self, width, height, tile_size=0, device=0):
        self.width, self.height = int(width), int(height)
        o = sr_renderer_opts()
        o.tile_size = tile_size
        o.device = device
        self.handle = C.c_void_p()
        check(lib().sr_renderer_create(self.width, self.height, C.byref(o), C.byref(self.handle)))

    def render_scene(self, scene, camera):
        check(lib().sr_renderer_render_scene(self.handle, scene.handle, C.byref(camera.c)))
        self._scene = scene      # the queued frame uses it until the frame is read

    def synchronize(self):
        """Waits for the queued frame and checks it (render_scene queues it)."""
        check(lib().sr_renderer_synchronize(self.handle))

    def update_auto_exposure(self, delta_time):
        check(lib().sr_renderer_update_auto_exposure(self.handle, float(delta_time)))

    def update_auto_exposure_device(self, delta_time):
        """The exposure step on the device (no host round trip)."""
        check(lib().sr_renderer_update_auto_exposure_device(self.handle, float(delta_time)))

    @property
    def exposure(self):
        return lib().sr_renderer_exposure(self.handle)

    def set_exposure(self, exposure):
        check(lib().sr_renderer_set_exposure(self.handle, float(exposure)))

    def blit_to_buffer(self, buffer=None):
        if buffer is None:
            buffer = np.zeros((self.height, self.width), np.uint32)
        check(lib().sr_renderer_blit_to_buffer(self.handle, _uptr(buffer), self.width, self.height))
        return buffer

    def present(self, frame=None):
        """The presented frame as (height, width, 4) bytes r, g, b, a: the
        blit and App::render_frame's unpack into the pixels frame
        (src/main.rs:574-590) in one device pass.  Rows come in pairs as in
        the blit, so an odd last row keeps what the frame held."""
        if frame is None:
            frame = np.zeros((self.height, self.width, 4), np.uint8)
        if frame.shape != (self.height, self.width, 4) or frame.dtype != np.uint8 or not frame.flags.c_contiguous:
            raise ValueError(f"frame must be a contiguous ({self.height}, {self.width}, 4) uint8 array")
        check(lib().sr_renderer_present(self.handle, frame.ctypes.data_as(C.POINTER(C.c_uint8)),
                                        self.width, self.height))
        return frame

    def present_to_device(self, device_ptr):
        """The presented frame into device memory (width*height*4 bytes)."""
        check(lib().sr_renderer_present_to_device(self.handle, C.c_void_p(int(device_ptr))))

    def read_vbuffer(self):
        ids = np.empty((self.height, self.width), np.uint32)
        depth = np.empty((self.height, self.width), np.float32)
        check(lib().sr_renderer_read_vbuffer(self.handle, _uptr(ids), _fptr(depth)))
        return ids, depth

    def read_color(self):
        rgb = np.empty((self.height, self.width, 3), np.float32)
        check(lib().sr_renderer_read_color(self.handle, _fptr(rgb)))
        return rgb

    def read_tile_luminance(self):
        n = C.c_int32()
        check(lib().sr_renderer_read_tile_luminance(self.handle, None, C.byref(n)))
        out = np.empty(n.value, np.float32)
        check(lib().sr_renderer_read_tile_luminance(self.handle, _fptr(out), C.byref(n)))
        return out

    def read_bins(self):
        """The last frame's per-tile lists: (start, end, list) arrays."""
        n = C.c_uint32()
        check(lib().sr_renderer_read_bins(self.handle, None, None, None, 0, C.byref(n)))
        start = np.empty(n.value, np.uint32)
        end = np.empty(n.value, np.uint32)
        check(lib().sr_renderer_read_bins(self.handle, _uptr(start), _uptr(end), None, 0, C.byref(n)))
        total = int(self.stats()["binned_pairs"])
        lst = np.empty(max(total, 1), np.uint32)
        check(lib().sr_renderer_read_bins(self.handle, None, None, _uptr(lst), total, C.byref(n)))
        return start, end, lst[:total]

    def stats(self):
        s = sr_frame_stats()
        check(lib().sr_renderer_get_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in sr_frame_stats._fields_}

    def set_stage_timing(self, mode):
        """How later frames time their stages: "clock" (default; the kernels'
        wall-clock stamps), "events" (HIP events on the stage launches) or
        "off"."""
        check(lib().sr_renderer_set_stage_timing(self.handle, STAGE_TIMING[mode]))

    def close(self):
        if self.handle:
            lib().sr_renderer_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many views (scene, camera, resolution) in one launch sequence with a
    persistent raster kernel draining (view, tile) items."""

    def __init__(self, tile_size=0, use_graph=False, persistent_blocks=0, device=0):
        o = sr_batch_opts()
        o.tile_size, o.use_graph, o.persistent_blocks, o.device = tile_size, int(use_graph), persistent_blocks, device
        self.handle = C.c_void_p()
        check(lib().sr_batch_create(C.byref(o), C.byref(self.handle)))
        self.sizes = []

    def configure(self, persistent_blocks=0, use_graph=False):
        """Persistent raster workgroup count (0 keeps it) and hipGraph use."""
        check(lib().sr_batch_configure(self.handle, int(persistent_blocks), int(use_graph)))

    def render(self, views):
        arr = (sr_view * len(views))()
        for i, (scene, cam, w, h, exposure) in enumerate(views):
            arr[i].scene = scene.handle.value
            arr[i].camera = cam.c
            arr[i].width, arr[i].height, arr[i].exposure = w, h, exposure
        check(lib().sr_batch_render(self.handle, arr, len(views)))
        self.sizes = [(w, h) for (_, _, w, h, _) in views]

    def read_vbuffer(self, i):
        w, h = self.sizes[i]
        ids = np.empty((h, w), np.uint32)
        depth = np.empty((h, w), np.float32)
        check(lib().sr_batch_read_vbuffer(self.handle, i, _uptr(ids), _fptr(depth)))
        return ids, depth

    def read_pixels(self, i):
        w, h = self.sizes[i]
        px = np.zeros((h, w), np.uint32)
        check(lib().sr_batch_read_pixels(self.handle, i, _uptr(px)))
        return px

    def stats(self):
        s = sr_batch_stats()
        check(lib().sr_batch_get_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in sr_batch_stats._fields_}

    def close(self):
        if self.handle:
            lib().sr_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Render server (sr_server_*): requests are batched continuously as they
    arrive, rendered on `streams` overlapping pipeline slots, and admitted by
    framebuffer-plus-bin memory."""

    def __init__(self, admission_bytes=0, max_batch_views=0, streams=0, tile_size=0, device=0):
        o = sr_server_opts()
        o.admission_bytes, o.max_batch_views, o.streams = admission_bytes, max_batch_views, streams
        o.tile_size, o.device = tile_size, device
        self.handle = C.c_void_p()
        check(lib().sr_server_create(C.byref(o), C.byref(self.handle)))
        self.sizes = {}
        self._scenes = []

    def add_scene(self, scene):
        idx = C.c_int32()
        check(lib().sr_server_add_scene(self.handle, scene.handle, C.byref(idx)))
        self._scenes.append(scene)
        return idx.value

    def submit(self, scene_index, camera, width, height, exposure=0.0, tag=0):
        r = sr_request()
        r.scene, r.camera, r.width, r.height, r.exposure, r.user_tag = scene_index, camera.c, width, height, exposure, tag
        rid = C.c_uint64()
        check(lib().sr_server_submit(self.handle, C.byref(r), C.byref(rid)))
        self.sizes[rid.value] = (width, height)
        return rid.value

    def poll(self, max_items=256):
        buf = (sr_completion * max_items)()
        n = C.c_int32()
        check(lib().sr_server_poll(self.handle, buf, max_items, C.byref(n)))
        return [{k: getattr(buf[i], k) for k, _ in sr_completion._fields_} for i in range(n.value)]

    def read_pixels(self, rid):
        w, h = self.sizes[rid]
        px = np.zeros((h, w), np.uint32)
        check(lib().sr_server_read_pixels(self.handle, rid, _uptr(px)))
        return px

    def read_vbuffer(self, rid):
        w, h = self.sizes[rid]
        ids = np.empty((h, w), np.uint32)
        depth = np.empty((h, w), np.float32)
        check(lib().sr_server_read_vbuffer(self.handle, rid, _uptr(ids), _fptr(depth)))
        return ids, depth

    def release(self, rid):
        check(lib().sr_server_release(self.handle, rid))
        self.sizes.pop(rid, None)

    def drain(self):
        check(lib().sr_server_drain(self.handle))

    def clock(self):
        return lib().sr_server_clock(self.handle)

    def close(self):
        if self.handle:
            lib().sr_server_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
