"""CPU tests of the JPEG decoder behind the glTF loader (sr_jpeg_decode; the
reference decodes textures with the image crate, src/texture.rs:897-1015).

Files are encoded by PIL (libjpeg-turbo) -- baseline, progressive, optimized
Huffman tables, 4:4:4 / 4:2:2 / 4:2:0 chroma, odd sizes, grey, restart
intervals -- and decoded by both.  Agreement is within 3 LSB per channel,
mean within 0.3: the two decoders differ in the inverse DCT (libjpeg's
integer transform against a double-precision one here) and in the rounding
bias of chroma upsampling and colour conversion.  The reference's own decoder
(zune-jpeg inside the image crate) could not be run here, so agreement with
its output is parity unpinned; the reference's skybox image is checked
against PIL when the reference tree is present."""
import io
import json
import os

import numpy as np
import pytest

Image = pytest.importorskip("PIL.Image")


def picture(w, h, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    y, x = np.mgrid[0:h, 0:w]
    r = (128 + 100 * np.sin(x / 7.0) * np.cos(y / 11.0)).astype(np.uint8)
    g = ((x * 255) // max(w - 1, 1)).astype(np.uint8)
    b = ((y * 255) // max(h - 1, 1)).astype(np.uint8)
    a = np.stack([r, g, b], -1)
    a[h // 3: h // 2, w // 4: w // 2] = rng.integers(0, 256, (h // 2 - h // 3, w // 2 - w // 4, 3), dtype=np.uint8)
    return a


def encode(a, **kw):
    buf = io.BytesIO()
    Image.fromarray(a).save(buf, "JPEG", **kw)
    return buf.getvalue()


def compare(sr, data, name):
    ref = np.asarray(Image.open(io.BytesIO(data)).convert("RGB")).astype(int)
    got = sr.unpack_rgba(sr.jpeg_decode(data)).astype(int)
    assert got.shape[:2] == ref.shape[:2], name
    assert np.all(got[..., 3] == 255)
    d = np.abs(got[..., :3] - ref)
    assert d.max() <= 3 and d.mean() <= 0.3, f"{name}: max {d.max()} mean {d.mean():.3f}"
    return d


@pytest.mark.parametrize("size", [(64, 48), (37, 23), (257, 131)])
@pytest.mark.parametrize("subsampling", [0, 1, 2])
@pytest.mark.parametrize("mode", ["baseline", "progressive", "optimized"])
def test_jpeg_matches_pil(sr, size, subsampling, mode):
    a = picture(*size)
    kw = {"quality": 88, "subsampling": subsampling}
    if mode == "progressive":
        kw["progressive"] = True
    if mode == "optimized":
        kw["optimize"] = True
    compare(sr, encode(a, **kw), f"{size} sub{subsampling} {mode}")


@pytest.mark.parametrize("progressive", [False, True])
def test_jpeg_grey(sr, progressive):
    a = picture(53, 41)
    buf = io.BytesIO()
    Image.fromarray(a).convert("L").save(buf, "JPEG", quality=85, progressive=progressive)
    d = compare(sr, buf.getvalue(), "grey")
    assert d.max() <= 1


@pytest.mark.parametrize("kw", [{"restart_marker_blocks": 3}, {"restart_marker_rows": 1},
                                {"restart_marker_blocks": 5, "progressive": True}])
def test_jpeg_restart_intervals(sr, kw):
    a = picture(129, 77, seed=2)
    data = encode(a, quality=85, **kw)
    assert b"\xff\xdd" in data                  # a DRI segment is present
    compare(sr, data, f"restarts {kw}")


def test_jpeg_quality_extremes(sr):
    a = picture(96, 64, seed=3)
    for q in (5, 50, 100):
        compare(sr, encode(a, quality=q), f"quality {q}")


def test_jpeg_rejects_bad_input(sr):
    for bad in (b"", b"\xff\xd8", b"\xff\xd8\xff\xe0" + bytes(64), b"PNG not jpeg"):
        with pytest.raises(sr.SwrasterError):
            sr.jpeg_decode(bad)
    data = bytearray(encode(picture(32, 32), quality=80))
    sof = data.index(b"\xff\xc0")
    data[sof + 1] = 0xC3                        # lossless SOF: not decoded
    with pytest.raises(sr.SwrasterError, match="lossless"):
        sr.jpeg_decode(bytes(data))


def test_jpeg_rejects_oversubscribed_huffman_lengths(sr):
    """A DHT whose code-length counts ask for more codes of a length than
    that many bits hold (three 1-bit codes) is refused: the table is never
    built past its 9-bit lookup (a mutation fuzz of the decoder under
    AddressSanitizer found the lookup written out of bounds here)."""
    data = bytearray(encode(picture(16, 16), quality=80))
    i = data.index(b"\xff\xc4")
    counts = data[i + 5:i + 21]
    nv = sum(counts)
    assert nv >= 3
    data[i + 5:i + 21] = bytes([3] + [0] * 14 + [nv - 3])
    with pytest.raises(sr.SwrasterError, match="Huffman"):
        sr.jpeg_decode(bytes(data))


def test_jpeg_refuses_huge_frames_before_allocating(sr):
    """A frame header claiming 65535 x 65535 samples is refused at once
    (the header alone would size about 25 GB of coefficient planes); at
    most 2^27 pixels, a 512 MiB RGBA8 image, are decoded."""
    data = bytearray(encode(picture(16, 16), quality=80))
    i = data.index(b"\xff\xc0")
    data[i + 5:i + 9] = b"\xff\xff\xff\xff"
    with pytest.raises(sr.SwrasterError, match="2\\^27"):
        sr.jpeg_decode(bytes(data))


REF_CUBEMAP = "/root/reference/assets/cubemap.jpg"


@pytest.mark.skipif(not os.path.exists(REF_CUBEMAP), reason="reference tree not present")
def test_reference_skybox_jpeg(sr):
    """The reference's skybox image (baseline, 4:4:4, 1024 x 768)."""
    with open(REF_CUBEMAP, "rb") as fh:
        d = compare(sr, fh.read(), "cubemap.jpg")
    assert d.mean() < 0.05


def test_gltf_with_jpeg_texture(sr, tmp_path):
    """A glTF written by hand here (not by the scene generator's writer):
    one textured quad whose base colour image is a JPEG next to the file,
    its buffer inline as a data URI.  The loaded texture is the decoder's
    image."""
    import base64
    a = picture(40, 24, seed=5)
    jpg = encode(a, quality=90, subsampling=2)
    (tmp_path / "albedo.jpg").write_bytes(jpg)
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    uv = np.array([[0, 1], [1, 1], [1, 0], [0, 0]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    blob = pos.tobytes() + uv.tobytes() + idx.tobytes()
    doc = {
        "asset": {"version": "2.0", "generator": "hand-written test file"},
        "scene": 0, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": 0}]}],
        "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicFactor": 0.0}}],
        "textures": [{"source": 0}], "images": [{"uri": "albedo.jpg"}],
        "buffers": [{"byteLength": len(blob),
                     "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 48},
                        {"buffer": 0, "byteOffset": 48, "byteLength": 32},
                        {"buffer": 0, "byteOffset": 80, "byteLength": 12}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3",
                       "min": [-1, -1, 0], "max": [1, 1, 0]},
                      {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC2"},
                      {"bufferView": 2, "componentType": 5123, "count": 6, "type": "SCALAR"}]}
    path = tmp_path / "quad.gltf"
    path.write_text(json.dumps(doc))
    g = sr.SceneData.from_gltf(str(path))
    assert len(g.textures) == 1
    w, h, ty, tex = g.textures[0]
    assert (w, h, ty) == (40, 24, sr.SR_TEX_SRGB)
    assert np.array_equal(np.asarray(tex, np.uint32).reshape(h, w), sr.jpeg_decode(jpg))
    assert np.array_equal(g.primitives[0]["positions"], pos)
    assert np.array_equal(g.primitives[0]["texcoords"], uv)
    assert np.array_equal(g.primitives[0]["indices"], idx.astype(np.uint32))
