"""CPU tests of the reference's cache files: .gi (VGI0, src/gi.rs:30-117)
and .ggx (GGX0, src/texture.rs:426-552).  The payload layout is pinned by
round trips and by header bytes; the brotli reader by a stream written here
by hand from RFC 7932 (uncompressed meta-blocks), so it does not rest on the
library's own encoder."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from test_gi import BOUNDS, ROOM, _box, _gi, _mat, _scene
from util import dark_environment, dummy_environment

NOT_FOUND, INVALID = -7, -1


def _rfc7932_stored(payload):
    """A brotli stream of uncompressed meta-blocks (RFC 7932 sections 9.1-9.2):
    WBITS = 16 (one 0 bit); per block ISLAST = 0, MNIBBLES = 4 (code 0),
    MLEN - 1 in 16 bits, ISUNCOMPRESSED = 1, zero pad to a byte boundary,
    MLEN raw bytes; finally ISLAST = 1, ISLASTEMPTY = 1."""
    out = bytearray()
    acc, nbits = 0, 0

    def put(v, n):
        nonlocal acc, nbits
        acc |= v << nbits
        nbits += n

    def flush_bytes(pad):
        nonlocal acc, nbits
        if pad and nbits % 8:
            nbits += 8 - nbits % 8
        while nbits >= 8:
            out.append(acc & 0xFF)
            acc >>= 8
            nbits -= 8

    put(0, 1)
    for i in range(0, len(payload), 65536):
        chunk = payload[i:i + 65536]
        put(0, 1)
        put(0, 2)
        put(len(chunk) - 1, 16)
        put(1, 1)
        flush_bytes(True)
        out += chunk
    put(1, 1)
    put(1, 1)
    flush_bytes(True)
    return bytes(out)


def _grid_desc(sr, dims, sh4):
    g = sr.sr_voxel_grid()
    for k in range(3):
        g.dims[k] = dims[k]
    g.sh4 = sh4.ctypes.data_as(C.POINTER(C.c_float))
    return g


def _load_gi(sr, path, dims):
    out = np.zeros(int(np.prod(dims)) * 16, np.float32)
    rc = sr.lib().sr_gi_cache_load(os.fsencode(str(path)), (C.c_int32 * 3)(*dims),
                                   out.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, out


def test_gi_cache_round_trip_and_header(sr, tmp_path):
    env = dark_environment(sr, sky=(0, 0, 0))
    _, sh4 = _gi(sr, _scene(sr, _box(*ROOM, [0] * 6), [_mat(emissive=(0.5, 0.25, 0.125))], BOUNDS), env, 12)
    sh4 = np.ascontiguousarray(sh4.ravel())
    p = tmp_path / "scene.gi"
    sr.check(sr.lib().sr_gi_cache_save(os.fsencode(str(p)), C.byref(_grid_desc(sr, (12, 12, 12), sh4))))
    raw = p.read_bytes()
    assert raw[:4] == b"VGI0" and struct.unpack("<4I", raw[4:20]) == (8, 12, 12, 12)
    assert len(raw) < 20 + sh4.nbytes             # compressed
    rc, back = _load_gi(sr, p, (12, 12, 12))
    assert rc == 0 and np.array_equal(back.view(np.uint32), sh4.view(np.uint32))
    assert _load_gi(sr, p, (12, 12, 13))[0] == NOT_FOUND
    assert _load_gi(sr, tmp_path / "missing.gi", (12, 12, 12))[0] == NOT_FOUND


def test_gi_cache_reads_a_hand_written_stream(sr, tmp_path):
    """Payload order: per voxel (x fastest), per coefficient r, g, b, w as
    little-endian f32 (src/gi.rs:95-101)."""
    dims = (5, 3, 7)
    vals = np.arange(np.prod(dims) * 16, dtype=np.float32) * np.float32(0.25) - np.float32(7.0)
    hdr = b"VGI0" + struct.pack("<4I", 8, *dims)
    p = tmp_path / "hand.gi"
    p.write_bytes(hdr + _rfc7932_stored(vals.astype("<f4").tobytes()))
    rc, back = _load_gi(sr, p, dims)
    assert rc == 0 and np.array_equal(back, vals)
    # a payload one float short, a wrong version, a wrong magic, a cut stream
    p.write_bytes(hdr + _rfc7932_stored(vals[:-1].astype("<f4").tobytes()))
    assert _load_gi(sr, p, dims)[0] == NOT_FOUND
    p.write_bytes(b"VGI0" + struct.pack("<4I", 7, *dims) + _rfc7932_stored(vals.tobytes()))
    assert _load_gi(sr, p, dims)[0] == NOT_FOUND
    p.write_bytes(b"VGJ0" + struct.pack("<4I", 8, *dims) + _rfc7932_stored(vals.tobytes()))
    assert _load_gi(sr, p, dims)[0] == NOT_FOUND
    p.write_bytes(hdr + _rfc7932_stored(vals.tobytes())[:-100])
    assert _load_gi(sr, p, dims)[0] in (NOT_FOUND, INVALID)
    p.write_bytes(hdr[:12])
    assert _load_gi(sr, p, dims)[0] == NOT_FOUND


def test_specular_cache_round_trip(sr, tmp_path):
    env = dummy_environment(sr)
    env.specular[:] = np.arange(env.specular.size, dtype=np.uint32) * np.uint32(2654435761)
    p = tmp_path / "cubemap.ggx"
    sr.save_specular_cache(str(p), env)
    raw = p.read_bytes()
    assert raw[:4] == b"GGX0" and struct.unpack("<4I", raw[4:20]) == (1, 4, 4, 3)
    spec, mips = sr.load_specular_cache(str(p), 4)
    assert mips == 3 and np.array_equal(spec, env.specular)
    assert sr.load_specular_cache(str(p), 8) is None
    # mips at full face size, little-endian u32 (src/texture.rs:467-476)
    p.write_bytes(raw[:20] + _rfc7932_stored(env.specular.astype("<u4").tobytes()))
    spec, _ = sr.load_specular_cache(str(p), 4)
    assert np.array_equal(spec, env.specular)


def test_prepare_gi_writes_then_reads_the_cache(sr, tmp_path):
    """As load_scene with --gi (src/main.rs:247-276): the first prepare bakes
    and writes <scene>.gi, the second loads it instead of baking."""
    env = dark_environment(sr, sky=(0, 0, 0))
    scene = _scene(sr, _box(*ROOM, [0] * 6), [_mat(emissive=(0.5, 0.25, 0.125))], BOUNDS)
    path = str(tmp_path / "room.gi")
    scene.prepare(env, grid=12, threads=4, gi=True, gi_cache=path)
    first = scene.voxels.sh4.copy()
    assert os.path.exists(path)
    raw = bytearray(open(path, "rb").read())
    # replace the payload: the next prepare must take the file's values
    vals = first * np.float32(2.0)
    open(path, "wb").write(bytes(raw[:20]) + _rfc7932_stored(vals.astype("<f4").tobytes()))
    scene.prepare(env, grid=12, threads=4, gi=True, gi_cache=path)
    assert np.array_equal(scene.voxels.sh4, vals)
    assert scene.voxels.desc.sh_scale == 1.0
