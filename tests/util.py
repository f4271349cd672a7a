"""Helpers for building small test scenes (host arrays only)."""
import numpy as np


def dummy_environment(sr, color=(128, 160, 200)):
    """A tiny uniform environment: 4x4 cubemap faces, 3 specular mips, 8x8 LUT."""
    c = sr.pack_rgba(np.array([[color[0], color[1], color[2], 255]], np.uint8))[0]
    cube = np.full((6, 4, 4), c, np.uint32)
    spec = np.full(3 * 6 * 16, c, np.uint32)
    lut = np.full(64, sr.pack_rgba(np.array([[200, 40, 0, 255]], np.uint8))[0], np.uint32)
    sh = np.array([[0.3, 0.3, 0.3], [0.1, 0.1, 0.1], [0.05, 0.05, 0.05], [0.0, 0.0, 0.0]], np.float32)
    return sr.Environment(cube, spec, 3, lut, sh)


def dummy_voxels(sr, scene):
    mn, mx, _, _ = scene.bounds()
    return sr.VoxelGrid((2, 2, 2), mn - 1, mx + 1, np.ones(8, np.float32))


def unproject(sx, sy, depth, width, height, fov):
    """Screen position (pixels, y down) at view depth -> view/world point for
    a camera at the origin looking down -z (yaw = pitch = 0)."""
    aspect = width / height
    ty = np.tan(fov / 2)
    ndx = 2 * np.asarray(sx, np.float64) / width - 1
    ndy = 1 - 2 * np.asarray(sy, np.float64) / height
    d = np.asarray(depth, np.float64)
    return np.stack([ndx * ty * aspect * d, ndy * ty * d, -d], -1)


def screen_triangle_scene(sr, tris_sub, depths, width, height, fov=np.pi / 3, far=150.0, colors=None):
    """One primitive whose triangles have the given subpixel screen vertices
    (n, 3, 2) and per-triangle view depths (n,) or per-vertex (n, 3).

    Vertices are placed so that the snapped subpixel coordinates are exactly
    tris_sub (the unprojection error is far below half a subpixel)."""
    tris_sub = np.asarray(tris_sub, np.float64)
    n = len(tris_sub)
    d = np.asarray(depths, np.float64)
    if d.ndim == 1:
        d = np.repeat(d[:, None], 3, 1)
    pts = unproject(tris_sub[..., 0] / 16.0, tris_sub[..., 1] / 16.0, d, width, height, fov).astype(np.float32)
    pos = pts.reshape(-1, 3)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (len(pos), 1))
    prim = {"positions": pos, "normals": nrm, "texcoords": None, "tangents": None,
            "indices": np.arange(3 * n, dtype=np.uint32), "material": 0}
    mat = {"base_color": (0.8, 0.5, 0.3, 1.0), "metallic": 0.0, "roughness": 0.5}
    scene = sr.SceneData([prim], [(0, 1)], [(np.eye(4, dtype=np.float32), 0)], [mat], [], name="test")
    cam = {"position": [0.0, 0.0, 0.0], "yaw": 0.0, "pitch": 0.0, "fov": float(fov), "width": width,
           "height": height, "far": float(far)}
    return scene, cam


def attach_dummy(sr, scene):
    return scene.attach(dummy_environment(sr), dummy_voxels(sr, scene))


def exact_coverage(tri, width, height):
    """Exact integer coverage of one snapped triangle (3, 2) at pixel centres:
    inside = every edge function >= 0 where edges with a < 0 or (a == 0 and
    b > 0) include their boundary and the others exclude it."""
    (x0, y0), (x1, y1), (x2, y2) = [(int(p[0]), int(p[1])) for p in tri]
    X = np.arange(width, dtype=np.int64) * 16 + 8
    Y = np.arange(height, dtype=np.int64) * 16 + 8
    X, Y = np.meshgrid(X, Y)
    inside = np.ones((height, width), bool)
    for (ax, ay), (bx, by) in (((x1, y1), (x2, y2)), ((x2, y2), (x0, y0)), ((x0, y0), (x1, y1))):
        a = by - ay
        b = ax - bx
        e = a * X + b * Y + (bx * ay - ax * by)
        inclusive = a < 0 or (a == 0 and b > 0)
        inside &= (e >= 0) if inclusive else (e > 0)
    return inside


def signed_area(tri):
    (x0, y0), (x1, y1), (x2, y2) = tri
    return (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)


def quad_layers_scene(sr, layers, width, height, fov=np.pi / 3, far=150.0):
    """Camera-facing screen rectangles at given view depths, one primitive,
    mesh and node each.  layers: list of ((x0, y0, x1, y1) in pixels,
    depth, material dict).  Texcoords run 0..1 across each rectangle; the
    camera sits at the origin looking down -z."""
    prims, meshes, nodes, mats = [], [], [], []
    for i, ((x0, y0, x1, y1), depth, mat) in enumerate(layers):
        corners = unproject(np.array([x0, x1, x1, x0], np.float64), np.array([y0, y0, y1, y1], np.float64),
                            np.full(4, depth), width, height, fov).astype(np.float32)
        uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
        # corners TL, TR, BR, BL; (TL, BR, TR) and (TL, BL, BR) face the camera
        idx = np.array([0, 2, 1, 0, 3, 2], np.uint32)
        prims.append({"positions": corners, "normals": np.tile(np.array([0, 0, 1], np.float32), (4, 1)),
                      "texcoords": uv, "tangents": None, "indices": idx, "material": i})
        meshes.append((i, 1))
        nodes.append((np.eye(4, dtype=np.float32), i))
        mats.append(mat)
    scene = sr.SceneData(prims, meshes, nodes, mats, [], name="quad_layers")
    cam = {"position": [0.0, 0.0, 0.0], "yaw": 0.0, "pitch": 0.0, "fov": float(fov), "width": width,
           "height": height, "far": float(far)}
    return scene, cam


def dark_environment(sr, sky=(128, 160, 200)):
    """Uniform sky, black specular cubemap, zero SH irradiance: shading then
    reduces to the emissive term (and the translucent blend over it)."""
    c = sr.pack_rgba(np.array([[sky[0], sky[1], sky[2], 255]], np.uint8))[0]
    black = sr.pack_rgba(np.array([[0, 0, 0, 255]], np.uint8))[0]
    cube = np.full((6, 4, 4), c, np.uint32)
    spec = np.full(3 * 6 * 16, black, np.uint32)
    lut = np.full(64, sr.pack_rgba(np.array([[200, 40, 0, 255]], np.uint8))[0], np.uint32)
    return sr.Environment(cube, spec, 3, lut, np.zeros((4, 3), np.float32))
