"""CPU: the TIFF reader (taufactor_amd.io) against Pillow, an independent
decoder, on files Pillow writes (every compression and sample type the
reader supports, several pages) and on the reference's own data files
(digests in tests/golden/tiff_goldens.json, written by
tests/golden/make_tiff_goldens.py with Pillow)."""
import hashlib
import io
import json
import os

import numpy as np
import pytest

from taufactor_amd import io as tio

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "tiff_goldens.json")))
PIL = pytest.importorskip("PIL.Image")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("compression", ["raw", "tiff_deflate", "tiff_adobe_deflate", "packbits"])
@pytest.mark.parametrize("dtype", ["uint8", "uint16", "float32"])
def test_reads_what_pillow_writes(compression, dtype):
    rng = np.random.default_rng(0)
    pages = [(rng.random((37, 53)) * 200).astype(dtype) for _ in range(4)]
    if compression == "packbits" and dtype != "uint8":
        pytest.skip("Pillow writes PackBits for 8-bit pages only")
    ims = [PIL.fromarray(p) for p in pages]
    buf = io.BytesIO()
    ims[0].save(buf, format="TIFF", save_all=True, append_images=ims[1:], compression=compression)
    got = tio.imread_bytes(buf.getvalue())
    assert got.shape == (4, 37, 53)
    assert np.array_equal(got, np.stack(pages))


def test_single_page_and_errors():
    a = np.arange(30, dtype=np.uint8).reshape(5, 6)
    buf = io.BytesIO()
    PIL.fromarray(a).save(buf, format="TIFF")
    assert np.array_equal(tio.imread_bytes(buf.getvalue()), a)
    with pytest.raises(tio.TiffError):
        tio.imread_bytes(b"not a tiff at all")


def test_electrode_fixture_matches_pillow_digest():
    g = GOLD["electrode.tiff"]
    vol = tio.imread(os.path.join(HERE, "golden", "electrode.tiff"))
    assert list(vol.shape) == g["shape"] and str(vol.dtype) == g["dtype"]
    assert _sha(vol) == g["sha256"]
    vals, counts = np.unique(vol, return_counts=True)
    assert {str(int(v)): int(c) for v, c in zip(vals, counts)} == g["labels"]


ZIP = "/root/reference/docs/paper/TF-500x500x500-1x2x6-v3.tiff.zip"


@pytest.mark.skipif(not os.path.exists(ZIP), reason="reference data file not present")
def test_paper_volume_zip_matches_pillow_digest():
    g = GOLD["TF-500x500x500-1x2x6-v3.tiff.zip"]
    vol = tio.imread(ZIP)
    assert list(vol.shape) == g["shape"]
    assert _sha(vol) == g["sha256"]
