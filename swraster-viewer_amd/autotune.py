"""Autotuner for the batched path.

Picks the screen-tile size (8, 16 or 32 px) and the number of persistent
raster workgroups (or -1: one raster workgroup per (view, tile), which wins
for few large views) for a set of views by timing the real batch, and keeps the
choice in a JSON cache keyed by device and a coarse workload signature (view
count, log2 of total pixels and of total triangles).  Numerics do not depend
on either knob; only time does.  The bin lists hold one 4-byte entry per
(triangle, tile), so the per-workgroup triangle load follows from the tile
size and is not a separate knob here.
"""
import json
import math
import os
import time

CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "assets", ".cache", "autotune.json")
TILES = (8, 16, 32)
WORKGROUPS = (-1, 1024, 2048, 4096)     # grid launch; 4, 8 and 16 per CU on a 256-CU MI355X


def signature(views):
    px = sum(int(w) * int(h) for (_, _, w, h, _) in views)
    tris = sum(int(scene.data.triangles) for (scene, _, _, _, _) in views)
    return f"v{len(views)}_p{round(math.log2(max(px, 1)))}_t{round(math.log2(max(tris, 1)))}"


def _time(batch, views, reps):
    batch.render(views)                                 # warm (and, the first time, allocate)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        batch.render(views)
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts[len(ts) // 2]


def tune(sr, views, tiles=TILES, workgroups=WORKGROUPS, reps=2, use_cache=True, device=0, cache_path=None):
    """Returns {"tile_size", "persistent_blocks", "ms", "table"} for `views`
    (a list of (Scene, RenderCamera, width, height, exposure))."""
    path = cache_path or CACHE_PATH
    key = f"dev{device}:{signature(views)}"
    cache = {}
    if use_cache and os.path.exists(path):
        with open(path) as fh:
            cache = json.load(fh)
        if key in cache:
            return dict(cache[key], cached=True)
    table = []
    for t in tiles:
        b = sr.Batch(tile_size=t, device=device)
        try:
            for wg in workgroups:
                b.configure(persistent_blocks=wg)
                table.append({"tile_size": t, "persistent_blocks": wg, "ms": round(_time(b, views, reps) * 1e3, 3)})
        finally:
            b.close()
    best = min(table, key=lambda r: r["ms"])
    out = {"tile_size": best["tile_size"], "persistent_blocks": best["persistent_blocks"], "ms": best["ms"],
           "table": table}
    if use_cache:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        cache[key] = out
        with open(path, "w") as fh:
            json.dump(cache, fh, indent=1)
    return dict(out, cached=False)
