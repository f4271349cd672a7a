"""CPU tests pinning the oracle's translucent path.

- The per-tile sort against a restatement of the reference's
  BumpQueue::sort_by (recursive Lomuto quicksort, rightmost pivot) with the
  comparator b.avg_z.cmp(&a.avg_z) under OrderedFloat, on keys full of ties,
  NaNs and signed zeros; and, for distinct keys, against any descending sort.
- The forward blend by closed form: with no direct light, a black specular
  cubemap and zero irradiance, a translucent fragment turns the colour c
  beneath it into c * base * transmission + emissive, in that f32 order, and
  layers apply back to front over the opaque frame; the depth test is against
  opaque depth only and translucent fragments write no depth.
"""
import math

import numpy as np
import pytest

from util import dark_environment, dummy_voxels, quad_layers_scene

NO_ID = 0xFFFFFFFF


def _of_less(a, b):
    """OrderedFloat a < b: NaN is the greatest value and equals itself."""
    if math.isnan(a):
        return False
    if math.isnan(b):
        return True
    return a < b


def reference_sort(keys):
    """BumpQueue::sort_by (src/bumpqueue.rs:155-200) with the comparator of
    src/tilerasterizer.rs:93-94, written from the reference source: partition
    moves an element left when compare(element, pivot) == Less, i.e. when the
    pivot's avg_z is below the element's."""
    v = list(range(len(keys)))

    def partition(lo, hi):
        pivot = keys[v[hi]]
        i = lo
        for j in range(lo, hi):
            if _of_less(pivot, keys[v[j]]):
                v[i], v[j] = v[j], v[i]
                i += 1
        v[i], v[hi] = v[hi], v[i]
        return i

    def quicksort(lo, hi):
        if lo < hi:
            p = partition(lo, hi)
            if p > 0:
                quicksort(lo, p - 1)
            if p + 1 < hi:
                quicksort(p + 1, hi)

    if len(v) > 1:
        quicksort(0, len(v) - 1)
    return np.array(v, np.uint32)


@pytest.mark.parametrize("seed,n,levels", [(1, 2, 2), (2, 3, 2), (3, 17, 3), (4, 200, 5), (5, 513, 40),
                                           (6, 64, 1), (7, 900, 900)])
def test_sort_matches_reference_quicksort(orc, seed, n, levels):
    rng = np.random.Generator(np.random.PCG64(seed))
    pool = np.concatenate([rng.uniform(-1, 1, max(levels - 3, 1)).astype(np.float32),
                           np.array([np.nan, 0.0, -0.0], np.float32)])[:levels]
    keys = pool[rng.integers(0, len(pool), n)]
    got = orc.sort_translucent(keys)
    want = reference_sort([float(k) for k in keys])
    assert np.array_equal(got, want)
    assert np.array_equal(np.sort(got), np.arange(n))
    # back to front: no element compares above the one before it
    s = keys[got]
    for a, b in zip(s[:-1], s[1:]):
        assert not _of_less(float(a), float(b))


def test_sort_distinct_keys_is_descending(orc):
    rng = np.random.Generator(np.random.PCG64(11))
    keys = rng.permutation(np.linspace(-1, 1, 1000).astype(np.float32))
    keys[17] = np.nan
    got = orc.sort_translucent(keys)
    want = np.concatenate([[17], np.argsort(-np.where(np.isnan(keys), np.inf, keys), kind="stable")[1:]])
    assert np.array_equal(got, want)


def _render(orc, scenes, scene, cam):
    o = orc.OracleRenderer(cam["width"], cam["height"])
    o.render_scene(orc.OracleScene(scene), scenes.make_camera(cam))
    return o.read_frame()


def _blend_layers(sr, W, H):
    """An opaque emissive slab over the right half, two translucent layers in
    front of it and one behind it."""
    opaque = {"base_color": (0.3, 0.3, 0.3, 1.0), "metallic": 0.0, "roughness": 1.0, "emissive": (0.25, 0.5, 0.75)}

    def tr(base, t, em):
        return {"base_color": base + (1.0,), "metallic": 0.0, "roughness": 0.5, "is_translucent": 1,
                "transmission": t, "emissive": em}

    layers = [((0.5 * W, -4, W + 4, H + 4), 10.0, opaque),
              ((0.2 * W, 0.2 * H, 0.8 * W, 0.8 * H), 8.0, tr((0.8, 0.6, 0.4), 0.7, (0.05, 0.0, 0.1))),
              ((0.4 * W, 0.3 * H, 0.9 * W, 0.9 * H), 6.0, tr((0.3, 0.9, 0.5), 0.45, (0.0, 0.2, 0.02))),
              ((0.6 * W, 0.1 * H, 1.0 * W, 0.5 * H), 12.0, tr((0.1, 0.1, 0.1), 0.1, (3.0, 3.0, 3.0)))]
    return layers


def test_translucent_blend_closed_form(sr, orc, scenes):
    W, H = 160, 120
    layers = _blend_layers(sr, W, H)
    full, cam = quad_layers_scene(sr, layers, W, H)
    base, _ = quad_layers_scene(sr, layers[:1], W, H)
    env = dark_environment(sr)
    for s in (full, base):
        s.light_color = (0.0, 0.0, 0.0)
        s.rebuild()
        s.attach(env, dummy_voxels(sr, full))
    after = _render(orc, scenes, full, cam)
    before = _render(orc, scenes, base, cam)
    # translucent fragments write no depth and take no visibility-buffer slot
    assert np.array_equal(after["depth"].view(np.uint32), before["depth"].view(np.uint32))
    assert np.array_equal(after["ids"] != NO_ID, before["ids"] != NO_ID)

    ys, xs = np.mgrid[0:H, 0:W]
    cx, cy = xs + 0.5, ys + 0.5
    margin = 1.5
    interior = np.ones((H, W), bool)
    regions = []
    for (x0, y0, x1, y1), depth, mat in layers[1:]:
        inside = (cx > x0 + margin) & (cx < x1 - margin) & (cy > y0 + margin) & (cy < y1 - margin)
        near = (cx > x0 - margin) & (cx < x1 + margin) & (cy > y0 - margin) & (cy < y1 + margin)
        interior &= inside | ~near
        # the opaque slab (depth 10) hides what lies behind it
        on = inside & ~((before["ids"] != NO_ID) & (depth > 10.0))
        regions.append((depth, on, mat))

    def blend(order):
        c = before["rgb"].astype(np.float32).copy()
        for depth, on, mat in order:
            b = np.array(mat["base_color"][:3], np.float32)
            c[on] = (c[on] * b) * np.float32(mat["transmission"]) + np.array(mat["emissive"], np.float32)
        return c

    got = after["rgb"].astype(np.float32)
    back_to_front = blend(sorted(regions, key=lambda r: -r[0]))
    front_to_back = blend(sorted(regions, key=lambda r: r[0]))
    covered = interior & np.any([r[1] for r in regions], 0)
    assert covered.sum() > 3000
    assert np.array_equal(got[interior].view(np.uint32), back_to_front[interior].view(np.uint32))
    overlap = interior & regions[0][1] & regions[1][1]
    assert overlap.sum() > 500
    assert not np.array_equal(got[overlap], front_to_back[overlap])
    # the layer behind the slab is hidden everywhere
    assert not regions[2][1].any()
    x0, y0, x1, y1 = layers[3][0]
    behind = interior & (cx > x0) & (cx < x1) & (cy > y0) & (cy < y1) & ~regions[0][1] & ~regions[1][1]
    assert behind.sum() > 50
    assert np.array_equal(got[behind].view(np.uint32), before["rgb"][behind].view(np.uint32))


def test_translucent_transmission_texture(sr, orc, scenes):
    """The transmission factor times the transmission texture's red channel
    (nearest sample), clamped to [0, 1]."""
    W, H = 96, 64
    n = 8
    red = np.arange(n * n, dtype=np.uint8).reshape(n, n) * 4
    rgba = np.stack([red, np.zeros_like(red), np.zeros_like(red), np.full_like(red, 255)], -1).reshape(-1, 4)
    tex = (n, n, sr.SR_TEX_LINEAR, sr.pack_rgba(rgba).ravel())
    mat = {"base_color": (1.0, 1.0, 1.0, 1.0), "metallic": 0.0, "roughness": 0.5, "is_translucent": 1,
           "transmission": 1.5, "emissive": (0.0, 0.0, 0.0),
           "transmission_texture_tex": (0, sr.SR_WRAP_CLAMP_TO_EDGE, sr.SR_WRAP_CLAMP_TO_EDGE)}
    scene, cam = quad_layers_scene(sr, [((-4, -4, W + 4, H + 4), 5.0, mat)], W, H)
    scene.textures = [tex]
    plain = {"base_color": (1.0, 1.0, 1.0, 1.0), "metallic": 0.0, "roughness": 0.5}
    empty, _ = quad_layers_scene(sr, [((0, 0, 1, 1), 5.0, plain)], W, H)
    env = dark_environment(sr)
    for s in (scene, empty):
        s.light_color = (0.0, 0.0, 0.0)
        s.rebuild()
        s.attach(env, dummy_voxels(sr, scene))
    after = _render(orc, scenes, scene, cam)["rgb"].astype(np.float32)
    sky = _render(orc, scenes, empty, cam)["rgb"].astype(np.float32)
    # with base 1 and no emission the colour is sky * clamp(1.5 * r), r in {0, 4, ..., 252} / 255
    ratio = after[8:-8, 8:-8, 0] / sky[8:-8, 8:-8, 0]
    allowed = np.minimum(np.float32(1.5) * (np.arange(64, dtype=np.float32) * 4 / np.float32(255.0)), 1.0)
    d = np.abs(ratio.reshape(-1, 1) - allowed.reshape(1, -1)).min(1)
    assert d.max() < 1e-6
    assert len(np.unique(np.round(ratio, 5))) > 10
    assert ratio.max() == 1.0
