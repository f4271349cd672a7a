"""Decode the skybox cross image into six packed faces.

The reference loads ``assets/cubemap.jpg`` (a 4x3 cross of 256^2 faces) and
re-packs it face by face in the order +X, -X, +Y, -Y, +Z, -Z, taking face
(col, row) cells (2,1), (0,1), (1,0), (1,2), (1,1), (3,1)
(src/texture.rs:922-971).  This script does the same decode once with PIL
(libjpeg-turbo) and stores the RGB bytes, so nothing at run time needs the
JPEG or the reference tree.  The image crate's decoder used by the reference
(zune-jpeg) may differ from libjpeg-turbo by small amounts per texel.

Usage: python assets/make_cubemap_faces.py /path/to/cubemap.jpg
"""
import sys
import numpy as np
from PIL import Image

FACE_CELLS = [(2, 1), (0, 1), (1, 0), (1, 2), (1, 1), (3, 1)]


def main(src):
    img = np.asarray(Image.open(src).convert("RGB"))
    h, w, _ = img.shape
    fw, fh = w // 4, h // 3
    faces = np.stack([img[cy * fh:(cy + 1) * fh, cx * fw:(cx + 1) * fw]
                      for cx, cy in FACE_CELLS])
    out = __file__.replace("make_cubemap_faces.py", "cubemap_faces.npz")
    np.savez_compressed(out, faces=faces)
    print(out, faces.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/assets/cubemap.jpg")
