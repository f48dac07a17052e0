"""Reading segmented volumes from TIFF stacks.

The reference's examples load microstructures with ``tifffile.imread``
(docs/notebooks/00-overview.ipynb, cell 2: ``electrode.tiff``; the paper's
``TF-500x500x500-1x2x6-v3.tiff.zip``).  tifffile is not a dependency here,
so this module reads the baseline TIFF forms such stacks use, following the
TIFF 6.0 specification:

* little- or big-endian classic TIFF (not BigTIFF), one IFD per page;
* one sample per pixel of 8, 16 or 32 bits (unsigned, signed or float);
* strips or tiles, compression none (1), Deflate (8 and the older code
  32946) or PackBits (32773), with or without horizontal differencing
  (predictor 2).

``imread(path)`` returns a numpy array [pages, rows, cols] (a single page
gives [rows, cols]); a ``.zip`` holding one TIFF is read in place.
"""
from __future__ import annotations

import struct
import zipfile
import zlib

import numpy as np

_TYPE_SIZES = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8}
_TYPE_CODES = {1: "B", 3: "H", 4: "I", 6: "b", 8: "h", 9: "i", 11: "f", 12: "d"}


class TiffError(ValueError):
    """The file is not a TIFF this reader supports."""


def _ifd(buf: bytes, off: int, bo: str) -> tuple[dict, int]:
    """Tags of the IFD at `off` (tag -> tuple of values) and the next IFD offset."""
    (n,) = struct.unpack_from(bo + "H", buf, off)
    tags = {}
    for i in range(n):
        tag, typ, count, raw = struct.unpack_from(bo + "HHI4s", buf, off + 2 + 12 * i)
        size = _TYPE_SIZES.get(typ)
        if size is None:
            continue
        nbytes = size * count
        if nbytes <= 4:
            data = raw[:nbytes]
        else:
            (at,) = struct.unpack(bo + "I", raw)
            data = buf[at:at + nbytes]
        if typ == 2:
            tags[tag] = (data.rstrip(b"\0").decode("latin-1"),)
        elif typ in (5, 10):
            vals = struct.unpack(bo + ("I" if typ == 5 else "i") * (2 * count), data)
            tags[tag] = tuple(vals[k] / vals[k + 1] if vals[k + 1] else 0.0 for k in range(0, len(vals), 2))
        elif typ == 7:
            tags[tag] = (data,)
        else:
            tags[tag] = struct.unpack(bo + _TYPE_CODES[typ] * count, data)
    (nxt,) = struct.unpack_from(bo + "I", buf, off + 2 + 12 * n)
    return tags, nxt


def _packbits(data: bytes, size: int) -> bytes:
    out = bytearray()
    i = 0
    while i < len(data) and len(out) < size:
        n = data[i] - 256 if data[i] > 127 else data[i]
        i += 1
        if n >= 0:
            out += data[i:i + n + 1]
            i += n + 1
        elif n != -128:
            out += bytes([data[i]]) * (1 - n)
            i += 1
    return bytes(out)


def _decode(chunk: bytes, compression: int, size: int) -> bytes:
    if compression == 1:
        return chunk
    if compression in (8, 32946):
        return zlib.decompress(chunk)
    if compression == 32773:
        return _packbits(chunk, size)
    raise TiffError(f"TIFF compression {compression} is not supported (1, 8, 32946, 32773 are)")


def _dtype(bits: int, fmt: int, bo: str) -> np.dtype:
    kind = {1: "u", 2: "i", 3: "f"}.get(fmt)
    if kind is None or bits not in (8, 16, 32) or (kind == "f" and bits != 32):
        raise TiffError(f"unsupported sample type: {bits} bits, format {fmt}")
    return np.dtype(("<" if bo == "<" else ">") + kind + str(bits // 8))


def _page(buf: bytes, tags: dict, bo: str) -> np.ndarray:
    width, height = tags[256][0], tags[257][0]
    spp = tags.get(277, (1,))[0]
    if spp != 1:
        raise TiffError(f"{spp} samples per pixel: only single-channel stacks are supported")
    bits = tags.get(258, (1,))[0]
    dt = _dtype(bits, tags.get(339, (1,))[0], bo)
    compression = tags.get(259, (1,))[0]
    predictor = tags.get(317, (1,))[0]
    item = dt.itemsize
    if 324 in tags:  # tiles
        tw, tl = tags[322][0], tags[323][0]
        offsets, counts = tags[324], tags[325]
        out = np.empty((height, width), dtype=dt)
        across = (width + tw - 1) // tw
        for t, (o, c) in enumerate(zip(offsets, counts)):
            tile = np.frombuffer(_decode(buf[o:o + c], compression, tw * tl * item), dtype=dt)
            tile = tile[:tw * tl].reshape(tl, tw)
            if predictor == 2:
                tile = np.cumsum(tile, axis=1, dtype=dt)
            r0, c0 = (t // across) * tl, (t % across) * tw
            h, w = min(tl, height - r0), min(tw, width - c0)
            out[r0:r0 + h, c0:c0 + w] = tile[:h, :w]
        return out
    rows = tags.get(278, (height,))[0]
    parts = []
    for s, (o, c) in enumerate(zip(tags[273], tags[279])):
        nrows = min(rows, height - s * rows)
        raw = np.frombuffer(_decode(buf[o:o + c], compression, nrows * width * item), dtype=dt)
        strip = raw[:nrows * width].reshape(nrows, width)
        if predictor == 2:
            strip = np.cumsum(strip, axis=1, dtype=dt)
        parts.append(strip)
    img = np.concatenate(parts, axis=0)
    if img.shape != (height, width):
        raise TiffError("strip data does not cover the page")
    return img


def imread(path) -> np.ndarray:
    """Read every page of a TIFF (or of the single TIFF inside a .zip)."""
    path = str(path)
    if path.lower().endswith(".zip"):
        with zipfile.ZipFile(path) as z:
            # skip macOS metadata entries (__MACOSX/._name)
            names = [n for n in z.namelist() if n.lower().endswith((".tif", ".tiff"))
                     and not n.startswith("__MACOSX/") and not n.rsplit("/", 1)[-1].startswith("._")]
            if len(names) != 1:
                raise TiffError(f"{path} must hold exactly one .tif/.tiff (found {names})")
            buf = z.read(names[0])
    else:
        with open(path, "rb") as f:
            buf = f.read()
    return imread_bytes(buf)


def imread_bytes(buf: bytes) -> np.ndarray:
    if buf[:2] == b"II":
        bo = "<"
    elif buf[:2] == b"MM":
        bo = ">"
    else:
        raise TiffError("not a TIFF file")
    (magic, off) = struct.unpack_from(bo + "HI", buf, 2)
    if magic != 42:
        raise TiffError("BigTIFF and other variants are not supported")
    pages = []
    seen = set()
    while off and off not in seen:
        seen.add(off)
        tags, off = _ifd(buf, off, bo)
        if tags.get(254, (0,))[0] & 1:  # reduced-resolution copy
            continue
        pages.append(_page(buf, tags, bo))
    if not pages:
        raise TiffError("no image pages")
    vol = np.stack(pages) if len(pages) > 1 else pages[0]
    return vol.astype(vol.dtype.newbyteorder("="), copy=False)


__all__ = ["imread", "imread_bytes", "TiffError"]
