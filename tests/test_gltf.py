"""CPU tests of the glTF loader and the PNG decoder behind it (sr_gltf_load,
sr_png_decode; src/main.rs:81-97, src/scene.rs:145-815, src/texture.rs:866-1015).

The PNG decoder is checked against PNGs encoded here from numpy arrays
(zlib + explicit scanline filters, every colour type and bit depth, stored,
fixed and dynamic Huffman streams) and against PIL for 8-bit files; the
expected RGBA8 follows the image crate's to_rgba8 expansion.  The loader is
checked by round trips of generated scenes (arrays bit-identical), .glb and
data-URI buffers, node hierarchies with TRS against float64 closed forms,
computed smooth normals, normalized texture coordinates, materials and error
paths."""
import base64
import io
import json
import os
import struct
import zlib

import numpy as np
import pytest


# ---------------------------------------------------------------------------
# PNG encoding (independent of the decoder under test)
# ---------------------------------------------------------------------------
def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def _filter_line(raw, prev, bpp, ftype):
    out = bytearray(len(raw))
    for i in range(len(raw)):
        a = raw[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        if ftype == 0:
            p = 0
        elif ftype == 1:
            p = a
        elif ftype == 2:
            p = b
        elif ftype == 3:
            p = (a + b) >> 1
        else:
            pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
            p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
        out[i] = (raw[i] - p) & 255
    return bytes(out)


def encode_png(samples, ctype, depth, plte=None, trns=None, filters=(0, 1, 2, 3, 4), level=6, strategy=None):
    """samples: (h, w, channels) integer array at the given bit depth."""
    h, w, ch = samples.shape
    rows = []
    for y in range(h):
        if depth == 16:
            rows.append(samples[y].astype(">u2").tobytes())
        elif depth == 8:
            rows.append(samples[y].astype(np.uint8).tobytes())
        else:
            bits = np.unpackbits(samples[y].reshape(-1, 1).astype(np.uint8), axis=1)[:, 8 - depth:].ravel()
            rows.append(np.packbits(bits).tobytes())
    bpp = max(1, ch * depth // 8)
    prev = bytes(len(rows[0]))
    data = bytearray()
    for y, raw in enumerate(rows):
        f = filters[y % len(filters)]
        data.append(f)
        data += _filter_line(raw, prev, bpp, f)
        prev = raw
    if strategy is None:
        comp = zlib.compress(bytes(data), level)
    else:
        co = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
        comp = co.compress(bytes(data)) + co.flush()
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if plte is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(plte, np.uint8).ravel()))
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    # split the stream over two IDAT chunks
    half = len(comp) // 2
    return out + _chunk(b"IDAT", comp[:half]) + _chunk(b"IDAT", comp[half:]) + _chunk(b"IEND", b"")


def to8(v, depth):
    v = np.asarray(v, np.int64)
    if depth == 16:
        return ((v + 128) // 257).astype(np.uint8)
    if depth == 8:
        return v.astype(np.uint8)
    return (v * 255 // ((1 << depth) - 1)).astype(np.uint8)


def pack(r, g, b, a):
    return (r.astype(np.uint32) << 24) | (g.astype(np.uint32) << 16) | (b.astype(np.uint32) << 8) | a.astype(np.uint32)


@pytest.mark.parametrize("ctype,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (4, 8), (4, 16),
                                         (6, 8), (6, 16), (3, 1), (3, 2), (3, 4), (3, 8)])
def test_png_colour_types_and_depths(sr, ctype, depth):
    rng = np.random.Generator(np.random.PCG64(ctype * 100 + depth))
    h, w = 13, 37
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    s = rng.integers(0, 1 << depth, (h, w, ch))
    plte = trns = None
    if ctype == 3:
        n = 1 << depth
        plte = rng.integers(0, 256, (n, 3))
        trns = bytes(rng.integers(0, 256, max(n // 2, 1)).astype(np.uint8))
    png = encode_png(s, ctype, depth, plte, trns)
    got = sr.png_decode(png)
    if ctype == 3:
        idx = s[..., 0]
        al = np.full(1 << depth, 255, np.uint8)
        al[:len(trns)] = np.frombuffer(trns, np.uint8)
        want = pack(plte[idx, 0], plte[idx, 1], plte[idx, 2], al[idx])
    elif ctype in (0, 4):
        g = to8(s[..., 0], depth)
        a = to8(s[..., 1], depth) if ctype == 4 else np.full(g.shape, 255, np.uint8)
        want = pack(g, g, g, a)
    else:
        c = [to8(s[..., k], depth) for k in range(3)]
        a = to8(s[..., 3], depth) if ctype == 6 else np.full(c[0].shape, 255, np.uint8)
        want = pack(c[0], c[1], c[2], a)
    assert got.shape == (h, w)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("level,strategy", [(0, None), (1, None), (9, None), (6, zlib.Z_FIXED),
                                            (6, zlib.Z_HUFFMAN_ONLY), (6, zlib.Z_RLE)])
def test_png_inflate_block_types(sr, level, strategy):
    rng = np.random.Generator(np.random.PCG64(level + 7))
    # smooth plus noisy regions so every block kind sees matches and literals
    y, x = np.mgrid[0:64, 0:96]
    s = np.stack([(x * 3) % 256, (y * 5) % 256, rng.integers(0, 256, x.shape), np.full(x.shape, 200)], -1)
    s[20:30] = rng.integers(0, 256, (10, 96, 4))
    png = encode_png(s, 6, 8, level=level, strategy=strategy)
    got = sr.png_decode(png)
    assert np.array_equal(got, pack(s[..., 0], s[..., 1], s[..., 2], s[..., 3]))


def test_png_colour_key_transparency(sr):
    s = np.array([[[10, 20, 30], [1, 2, 3]], [[10, 20, 30], [10, 20, 31]]], np.int64)
    png = encode_png(s, 2, 8, trns=struct.pack(">HHH", 10, 20, 30))
    got = sr.png_decode(png)
    assert [int(v & 255) for v in got.ravel()] == [0, 255, 0, 255]
    g = np.array([[[5], [6]]], np.int64)
    got = sr.png_decode(encode_png(g, 0, 8, trns=struct.pack(">H", 6)))
    assert [int(v & 255) for v in got.ravel()] == [255, 0]


@pytest.mark.parametrize("mode", ["RGBA", "RGB", "L", "LA", "P"])
def test_png_matches_pil(sr, mode):
    from PIL import Image
    rng = np.random.Generator(np.random.PCG64(len(mode)))
    rgba = rng.integers(0, 256, (40, 50, 4), dtype=np.uint8)
    img = Image.fromarray(rgba, "RGBA")
    img = img.convert(mode) if mode != "P" else img.convert("RGB").quantize(64)
    buf = io.BytesIO()
    img.save(buf, "PNG")
    got = sr.png_decode(buf.getvalue())
    ref = np.asarray(img.convert("RGBA"), np.uint32)
    assert np.array_equal(got, pack(ref[..., 0], ref[..., 1], ref[..., 2], ref[..., 3]))


def test_png_rejects_bad_input(sr):
    with pytest.raises(sr.SwrasterError):
        sr.png_decode(b"not a png at all")
    s = np.zeros((4, 4, 4), np.int64)
    png = bytearray(encode_png(s, 6, 8))
    png[40:44] = b"\xff\xff\xff\xff"          # damage the compressed stream
    with pytest.raises(sr.SwrasterError):
        sr.png_decode(bytes(png))


# ---------------------------------------------------------------------------
# glTF
# ---------------------------------------------------------------------------
def _same_scene(a, b):
    assert a.meshes == b.meshes
    assert len(a.primitives) == len(b.primitives)
    for pa, pb in zip(a.primitives, b.primitives):
        for k in ("positions", "normals", "indices"):
            assert np.array_equal(np.asarray(pa[k]), np.asarray(pb[k])), k
        if pa.get("texcoords") is not None:
            assert np.array_equal(np.asarray(pa["texcoords"], np.float32), pb["texcoords"])
        else:
            assert not np.any(pb["texcoords"])
        assert pa["material"] == pb["material"]
    for (ta, ma), (tb, mb) in zip(a.nodes, b.nodes):
        assert np.array_equal(np.asarray(ta, np.float32), tb) and ma == mb
    for ma, mb in zip(a.materials, b.materials):
        assert np.array_equal(np.float32(ma["base_color"]), np.float32(mb["base_color"]))
        assert np.float32(ma.get("metallic", 1)) == np.float32(mb["metallic"])
        assert np.float32(ma.get("roughness", 1)) == np.float32(mb["roughness"])
        assert (ma.get("base_color_tex") is None) == (mb.get("base_color_tex") is None)
    for (w, h, _t, d), (w2, h2, _t2, d2) in zip(a.textures, b.textures):
        assert (w, h) == (w2, h2) and np.array_equal(d, d2)


def test_gltf_round_trip_c1(sr, scenes, tmp_path):
    scene, _ = scenes.make_c1(width=320, height=180)
    path = scenes.write_gltf(scene, str(tmp_path / "c1.gltf"))
    _same_scene(scene, sr.SceneData.from_gltf(path))


def test_gltf_round_trip_c3_untextured(sr, scenes, tmp_path):
    scene, _ = scenes.make_c3(triangles=2000, primitives=16)
    path = scenes.write_gltf(scene, str(tmp_path / "c3.gltf"))
    _same_scene(scene, sr.SceneData.from_gltf(path))


def _doc_and_bin(scenes, scene, tmp_path):
    path = scenes.write_gltf(scene, str(tmp_path / "s.gltf"))
    doc = json.load(open(path))
    blob = open(str(tmp_path / "s.bin"), "rb").read()
    return doc, blob


def test_glb_and_data_uri_buffers(sr, scenes, tmp_path):
    scene, _ = scenes.make_c1(width=160, height=90, spheres=3)
    doc, blob = _doc_and_bin(scenes, scene, tmp_path)
    ref = sr.SceneData.from_gltf(str(tmp_path / "s.gltf"))
    # data URI buffer
    d2 = json.loads(json.dumps(doc))
    d2["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
    p2 = tmp_path / "inline.gltf"
    p2.write_text(json.dumps(d2))
    _same_scene(ref, sr.SceneData.from_gltf(str(p2)))
    # GLB: JSON chunk + BIN chunk, buffer 0 without a uri
    d3 = json.loads(json.dumps(doc))
    del d3["buffers"][0]["uri"]
    js = json.dumps(d3).encode()
    js += b" " * (-len(js) % 4)
    bn = blob + b"\0" * (-len(blob) % 4)
    glb = struct.pack("<4sII", b"glTF", 2, 12 + 8 + len(js) + 8 + len(bn))
    glb += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bn), 0x004E4942) + bn
    p3 = tmp_path / "s.glb"
    p3.write_bytes(glb)
    _same_scene(ref, sr.SceneData.from_gltf(str(p3)))


def _write(tmp_path, doc, arrays):
    """A .gltf whose single buffer holds the given arrays, one bufferView
    each; returns the path."""
    blob = bytearray()
    views = []
    for a in arrays:
        a = np.ascontiguousarray(a)
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": a.nbytes})
        blob += a.tobytes()
    doc = dict(doc)
    doc["asset"] = {"version": "2.0"}
    doc["bufferViews"] = views
    doc["buffers"] = [{"uri": "b.bin", "byteLength": len(blob)}]
    (tmp_path / "b.bin").write_bytes(bytes(blob))
    p = tmp_path / "t.gltf"
    p.write_text(json.dumps(doc))
    return str(p)


def _quad():
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    return pos, idx


def _acc(view, comp, count, typ, **kw):
    a = {"bufferView": view, "componentType": comp, "count": count, "type": typ}
    a.update(kw)
    return a


def test_gltf_hierarchy_and_trs(sr, tmp_path):
    pos, idx = _quad()
    q = [0.0, 0.0, np.sin(np.pi / 8), np.cos(np.pi / 8)]      # 45 degrees about z
    doc = {"accessors": [_acc(0, 5126, 4, "VEC3"), _acc(1, 5123, 6, "SCALAR")],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
           "nodes": [{"mesh": 0, "children": [1]},
                     {"mesh": 0, "translation": [1, 2, 3], "rotation": q, "scale": [2, 2, 2], "children": [2]},
                     {"mesh": 0, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 5, 0, 0, 1]},
                     {"translation": [7, 0, 0]}]}
    doc["nodes"][0]["translation"] = [0, 0, -10]
    g = sr.SceneData.from_gltf(_write(tmp_path, doc, [pos, idx]))
    assert [m for _, m in g.nodes] == [0, 0, 0, -1]

    def trs(t, r, s):
        x, y, z, w = r
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        M = np.eye(4)
        M[:3, :3] = R * np.asarray(s, np.float64)
        M[:3, 3] = t
        return M

    m0 = trs([0, 0, -10], [0, 0, 0, 1], [1, 1, 1])
    m1 = m0 @ trs([1, 2, 3], q, [2, 2, 2])
    m2 = m1 @ trs([5, 0, 0], [0, 0, 0, 1], [1, 1, 1])
    m3 = trs([7, 0, 0], [0, 0, 0, 1], [1, 1, 1])
    for (t, _), want in zip(g.nodes, [m0, m1, m2, m3]):
        assert np.allclose(t, want, atol=1e-5), (t, want)
    # no normals in the file: smooth normals of a flat quad are +z
    assert np.array_equal(g.primitives[0]["normals"], np.tile(np.float32([0, 0, 1]), (4, 1)))
    assert not np.any(g.primitives[0]["texcoords"])
    # no materials: one glTF-default material
    assert len(g.materials) == 1 and g.materials[0]["metallic"] == 1.0 and g.materials[0]["roughness"] == 1.0


def test_gltf_smooth_normals_and_normalized_texcoords(sr, tmp_path):
    # a unit cube with 8 shared corners and one unused vertex
    corners = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)] + [[5, 5, 5]], np.float32)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    idx = np.array([[a, b, c, a, c, d] for a, b, c, d in faces], np.uint8).ravel()
    uv8 = np.arange(18, dtype=np.uint8).reshape(9, 2) * 13
    uv16 = (np.arange(18, dtype=np.uint16).reshape(9, 2) * 3001).astype(np.uint16)
    doc = {"accessors": [_acc(0, 5126, 9, "VEC3"), _acc(1, 5121, 36, "SCALAR"),
                         _acc(2, 5121, 9, "VEC2", normalized=True), _acc(3, 5123, 9, "VEC2", normalized=True)],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 2}, "indices": 1},
                                      {"attributes": {"POSITION": 0, "TEXCOORD_0": 3}, "indices": 1}]}],
           "nodes": [{"mesh": 0}]}
    g = sr.SceneData.from_gltf(_write(tmp_path, doc, [corners, idx, uv8, uv16]))
    p0, p1 = g.primitives
    assert np.array_equal(p0["indices"], idx.astype(np.uint32))
    # each corner gets face normals summed with weights of twice the triangle
    # areas: one triangle or two per face, so compare the direction in float64
    n = np.zeros((9, 3))
    for t in idx.reshape(-1, 3):
        e1, e2 = corners[t[1]] - corners[t[0]], corners[t[2]] - corners[t[0]]
        n[t] += np.cross(e1, e2)
    want = n[:8] / np.linalg.norm(n[:8], axis=1, keepdims=True)
    assert np.allclose(p0["normals"][:8], want, atol=1e-6)
    assert np.array_equal(p0["normals"][8], np.float32([0, 0, 1]))
    assert np.array_equal(p0["texcoords"], uv8.astype(np.float32) / np.float32(255))
    assert np.array_equal(p1["texcoords"], uv16.astype(np.float32) / np.float32(65535))


def test_gltf_materials(sr, tmp_path):
    pos, idx = _quad()
    from PIL import Image
    Image.fromarray(np.full((4, 4, 4), 200, np.uint8), "RGBA").save(str(tmp_path / "t.png"))
    doc = {"accessors": [_acc(0, 5126, 4, "VEC3"), _acc(1, 5123, 6, "SCALAR")],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1, "material": 1}]}],
           "nodes": [{"mesh": 0}],
           "images": [{"uri": "t.png"}], "samplers": [{"wrapS": 33071, "wrapT": 33648}],
           "textures": [{"source": 0, "sampler": 0}, {"source": 0}],
           "materials": [
               {"pbrMetallicRoughness": {"baseColorTexture": {"index": 1}}, "alphaMode": "MASK", "alphaCutoff": 0.3},
               {"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.25, 1, 0.75], "metallicFactor": 0.1,
                                         "roughnessFactor": 0.2,
                                         "metallicRoughnessTexture": {"index": 0}},
                "emissiveFactor": [1, 2, 3], "occlusionTexture": {"index": 0, "strength": 0.4},
                "alphaMode": "BLEND",
                "extensions": {"KHR_materials_transmission": {"transmissionFactor": 0.6,
                                                              "transmissionTexture": {"index": 0}}}}]}
    g = sr.SceneData.from_gltf(_write(tmp_path, doc, [pos, idx]))
    m0, m1 = g.materials
    assert m0["alpha_mode"] == sr.SR_ALPHA_MASK and np.float32(m0["alpha_cutoff"]) == np.float32(0.3)
    assert m0["base_color_tex"] == (0, sr.SR_WRAP_REPEAT, sr.SR_WRAP_REPEAT)
    assert m0["base_color"] == (1.0, 1.0, 1.0, 1.0) and m0["is_translucent"] == 0
    assert m1["base_color"] == (0.5, 0.25, 1.0, 0.75)
    assert np.float32(m1["metallic"]) == np.float32(0.1) and np.float32(m1["roughness"]) == np.float32(0.2)
    assert m1["emissive"] == (1.0, 2.0, 3.0) and np.float32(m1["occlusion_strength"]) == np.float32(0.4)
    assert m1["is_translucent"] == 1 and np.float32(m1["transmission"]) == np.float32(0.6)
    assert m1["metallic_roughness_tex"] == (0, sr.SR_WRAP_CLAMP_TO_EDGE, sr.SR_WRAP_MIRRORED_REPEAT)
    assert m1["transmission_texture_tex"][0] == 0 and m1["occlusion_tex"][0] == 0
    assert m1["alpha_mode"] == sr.SR_ALPHA_BLEND
    assert g.primitives[0]["material"] == 1
    # one texture per image, typed by its first use (base colour: sRGB)
    assert len(g.textures) == 1 and g.textures[0][2] == sr.SR_TEX_SRGB
    assert np.all(g.textures[0][3] == np.uint32(0xC8C8C8C8))


def test_gltf_errors(sr, tmp_path):
    pos, idx = _quad()
    with pytest.raises(sr.SwrasterError):
        sr.SceneData.from_gltf(str(tmp_path / "missing.gltf"))
    doc = {"accessors": [_acc(0, 5126, 4, "VEC3")],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}], "nodes": [{"mesh": 0}]}
    with pytest.raises(sr.SwrasterError, match="without indices"):
        sr.SceneData.from_gltf(_write(tmp_path, doc, [pos]))
    (tmp_path / "j.jpg").write_bytes(b"\xff\xd8\xff\xe0" + bytes(64))
    doc = {"accessors": [_acc(0, 5126, 4, "VEC3"), _acc(1, 5123, 6, "SCALAR")],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}], "nodes": [{"mesh": 0}],
           "images": [{"uri": "j.jpg"}], "textures": [{"source": 0}],
           "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}]}
    with pytest.raises(sr.SwrasterError, match="JPEG"):
        sr.SceneData.from_gltf(_write(tmp_path, doc, [pos, idx]))
    (tmp_path / "bad.gltf").write_text("{\"asset\": {\"version\": \"2.0\"}, \"nodes\": [")
    with pytest.raises(sr.SwrasterError, match="JSON"):
        sr.SceneData.from_gltf(str(tmp_path / "bad.gltf"))


@pytest.mark.parametrize("field,value", [("view_offset", -48.0), ("view_offset", 1.8e19), ("count", 1e15),
                                         ("count", -1.0), ("accessor_offset", -12.0), ("stride", -4.0),
                                         ("view_length", 1.8e19), ("buffer_view", -1.0)])
def test_gltf_rejects_out_of_range_numbers(sr, tmp_path, field, value):
    """Negative or huge offsets, lengths, strides, counts and indices are
    refused before any read (a mutation fuzz of the loader under
    AddressSanitizer found a negative byteOffset wrapping the bounds
    check)."""
    pos, idx = _quad()
    acc = [_acc(0, 5126, 4, "VEC3"), _acc(1, 5123, 6, "SCALAR")]
    doc = {"accessors": acc, "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
           "nodes": [{"mesh": 0}]}
    path = _write(tmp_path, doc, [pos, idx])
    d = json.loads(open(path).read())
    if field == "view_offset":
        d["bufferViews"][0]["byteOffset"] = value
    elif field == "view_length":
        d["bufferViews"][0]["byteLength"] = value
    elif field == "count":
        d["accessors"][0]["count"] = value
    elif field == "accessor_offset":
        d["accessors"][0]["byteOffset"] = value
    elif field == "stride":
        d["bufferViews"][0]["byteStride"] = value
    else:
        d["accessors"][0]["bufferView"] = value
    open(path, "w").write(json.dumps(d))
    with pytest.raises(sr.SwrasterError):
        sr.SceneData.from_gltf(path)


def test_gltf_load_time_c2(sr, scenes, tmp_path):
    """The 1,045,458-triangle heightfield with its 2048^2 texture: loaded
    arrays identical, load time reported."""
    import time
    scene, _ = scenes.make_c2()
    path = scenes.write_gltf(scene, str(tmp_path / "c2.gltf"))
    nbytes = sum(os.path.getsize(str(tmp_path / f)) for f in os.listdir(str(tmp_path)))
    t0 = time.perf_counter()
    g = sr.SceneData.from_gltf(path)
    dt = time.perf_counter() - t0
    print(f"[gltf] c2: {nbytes / 1e6:.1f} MB of files, {scene.triangles} triangles, loaded in {dt * 1e3:.0f} ms "
          f"({nbytes / 1e6 / dt:.0f} MB/s)")
    _same_scene(scene, g)

