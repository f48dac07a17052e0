"""Write tests/golden/tiff_goldens.json (and copy electrode.tiff next to it):
the decoded content of the reference's TIFF data files, decoded with Pillow
(an implementation independent of taufactor_amd.io), as SHA-256 of the page
bytes and label counts.

Run from the repository root:  python tests/golden/make_tiff_goldens.py

  electrode.tiff  docs/notebooks/electrode.tiff of the reference (256 pages of
                  256x256 uint8, Adobe Deflate): copied as a fixture; the
                  reference's overview notebook solves Solver(img == 0) on it
  TF-500          docs/paper/TF-500x500x500-1x2x6-v3.tiff.zip (500 uncompressed
                  pages of 500x500 uint8): only its digest is kept
"""
import hashlib
import io
import json
import os
import shutil
import zipfile

import numpy as np
from PIL import Image, ImageSequence

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = "/root/reference"
OUT = os.path.join(ROOT, "tests", "golden")


def _pages(fp):
    return np.stack([np.array(p) for p in ImageSequence.Iterator(Image.open(fp))])


def _entry(vol):
    vals, counts = np.unique(vol, return_counts=True)
    return {"shape": list(vol.shape), "dtype": str(vol.dtype),
            "sha256": hashlib.sha256(np.ascontiguousarray(vol).tobytes()).hexdigest(),
            "labels": {str(int(v)): int(c) for v, c in zip(vals, counts)}}


def main():
    src = os.path.join(REF, "docs", "notebooks", "electrode.tiff")
    shutil.copyfile(src, os.path.join(OUT, "electrode.tiff"))
    out = {"electrode.tiff": _entry(_pages(src))}
    z = zipfile.ZipFile(os.path.join(REF, "docs", "paper", "TF-500x500x500-1x2x6-v3.tiff.zip"))
    name = [n for n in z.namelist() if n.lower().endswith((".tif", ".tiff"))][0]
    out["TF-500x500x500-1x2x6-v3.tiff.zip"] = _entry(_pages(io.BytesIO(z.read(name))))
    with open(os.path.join(OUT, "tiff_goldens.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
