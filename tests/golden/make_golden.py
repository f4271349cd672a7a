"""Regenerate the committed fixtures under tests/golden/.

  c1_160x90.npz       oracle visibility buffer (canonical ids, depth), blit
                      (exposure 2.0) and 64px-tile centre luminance of the
                      16-icosphere scene rendered at 160x90
  scene_digests.json  SHA-256 of the generator's arrays for small instances
                      of each configuration, plus the decoded skybox faces

They freeze the oracle and the scene generator against regressions; they do
not by themselves pin agreement with the reference binary.
Usage: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import __graft_entry__ as entry  # noqa: E402

SMALL = {"c1": {}, "c2": {"grid": 64, "tex_size": 256}, "c3": {"triangles": 2000, "width": 640, "height": 360}}


def digests(sr, sc):
    out = {}
    for name, kw in SMALL.items():
        scene, spec = sc.CONFIGS[name](**kw)
        out[name] = {"kwargs": kw, "sha256": sc.array_digest(scene), "triangles": scene.triangles,
                     "camera": spec}
    faces = sr.cubemap_faces()
    out["cubemap_faces_sha256"] = hashlib.sha256(faces.tobytes()).hexdigest()
    return out


def main():
    sr = entry.load_package()
    orc = entry.load_oracle()
    import swraster_viewer_amd.scenes as sc
    scene, spec = sc.make_c1(width=160, height=90)
    scene.prepare(sr.Environment.load(), cache_key=scene.name)
    cam = sc.make_camera(spec)
    o = orc.OracleRenderer(160, 90)
    o.render_scene(orc.OracleScene(scene), cam)
    f = o.read_frame()
    lum = o.tile_luminance()
    o.update_auto_exposure(0.0)
    np.savez_compressed(os.path.join(HERE, "c1_160x90.npz"), ids=f["ids"], depth=f["depth"], pixels=o.blit(),
                        tile_luminance=lum)
    with open(os.path.join(HERE, "scene_digests.json"), "w") as fh:
        json.dump(digests(sr, sc), fh, indent=1)
    print("wrote fixtures")


if __name__ == "__main__":
    main()
