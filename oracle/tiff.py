"""TIFF files as the reference's vtkTIFFWriter writes them (test infrastructure).

The reference dumps volumes and projections with vtkTIFFWriter
(Reconstructors.h:49-52,78-81,98-101,127-130,166-169,195-198,218-221,246-249;
mainReconstruct.cpp:127-130,145-148,163-166,188-191), VTK 9.0.1
(CMakeLists.txt:53), which drives libtiff 4.0.x.  Neither is vendored in the
reference, so this module restates their published behaviour:

* vtkTIFFWriter::WriteFileHeader / WriteFile (one z slice): rows written from
  the highest y down (the VTK y axis points up, ORIENTATION_TOPLEFT), tags
  ImageWidth, ImageLength, Orientation, SamplesPerPixel, BitsPerSample,
  PlanarConfig contiguous, SampleFormat IEEEFP for float, ExtraSamples
  (associated alpha) beyond three components, Compression (PackBits by
  default), Photometric (min-is-black for one component, else RGB),
  RowsPerStrip = TIFFDefaultStripSize, X/YResolution = 10 / spacing in
  pixels per centimetre.
* vtkTIFFWriter::WriteVolume (more than one z slice): one page per slice,
  rows in memory order, SubfileType = page, PageNumber (page, pages), the same
  size, sample, compression and resolution tags, no Orientation.
* libtiff: the PackBits encoder (tif_packbits.c: PackBitsEncode, one call per
  scanline), TIFFDefaultStripSize (8192 bytes / scanline bytes, at least 1),
  strips appended at the end of the file, each directory written at the
  even-rounded end of the file with its entries sorted by tag, out-of-line
  values after it in the order TIFFWriteDirectorySec emits them (resolution,
  bits per sample, strip byte counts, strip offsets, extra samples, sample
  format) each at an even offset, strip arrays as LONG, integral rationals
  as n/1.

Pinned by the reference's own resources/testImage.tiff (a 460x360 RGBA
PackBits file of this writer): re-encoding its decoded pixels reproduces the
file byte for byte.  Pages and float samples are checked by decoding with an
independent reader (Pillow).
"""
import struct

import numpy as np

STRIPSIZE_DEFAULT = 8192

COMPRESSION_NONE = 1
COMPRESSION_PACKBITS = 32773


def packbits_encode_row(row):
    """libtiff's PackBitsEncode over one scanline (bytes -> bytes)."""
    out = bytearray()
    state = "base"
    lastlit = -1
    i, n_all = 0, len(row)
    while i < n_all:
        b = row[i]
        n = 1
        i += 1
        while i < n_all and row[i] == b:
            n += 1
            i += 1
        while True:  # the encoder's "again" label
            if state == "base":
                if n > 1:
                    state = "run"
                    if n > 128:
                        out += bytes([0x81, b])  # -127: 128 copies
                        n -= 128
                        continue
                    out += bytes([(1 - n) & 0xFF, b])
                else:
                    lastlit = len(out)
                    out += bytes([0, b])
                    state = "literal"
                break
            if state == "literal":
                if n > 1:
                    state = "literal_run"
                    if n > 128:
                        out += bytes([0x81, b])
                        n -= 128
                        continue
                    out += bytes([(1 - n) & 0xFF, b])
                else:
                    out[lastlit] += 1
                    if out[lastlit] == 127:
                        state = "base"
                    out.append(b)
                break
            if state == "run":
                if n > 1:
                    if n > 128:
                        out += bytes([0x81, b])
                        n -= 128
                        continue
                    out += bytes([(1 - n) & 0xFF, b])
                else:
                    lastlit = len(out)
                    out += bytes([0, b])
                    state = "literal"
                break
            # literal_run: a 2-byte run between literals joins the literal
            if n == 1 and out[-2] == 0xFF and out[lastlit] < 126:
                out[lastlit] += 2
                state = "base" if out[lastlit] == 127 else "literal"
                out[-2] = out[-1]
            else:
                state = "run"
            # re-examine the current byte in the new state
    return bytes(out)


def packbits_decode(data, expected):
    """PackBits decoding (TIFF 6.0, section 9)."""
    out = bytearray()
    i = 0
    while len(out) < expected and i < len(data):
        c = data[i]
        i += 1
        if c < 128:
            out += data[i:i + c + 1]
            i += c + 1
        elif c > 128:
            out += bytes([data[i]]) * (257 - c)
            i += 1
    return bytes(out[:expected])


def _rational(v):
    """libtiff 4.0.x TIFFWriteDirectoryTagCheckedRational of a float value."""
    v = float(np.float32(v))
    if v <= 0.0:
        return 0, 1
    if v == float(int(v)) and v < 2**32:
        return int(v), 1
    if v < 1.0:
        return int(v * 0xFFFFFFFF), 0xFFFFFFFF
    return 0xFFFFFFFF, int(0xFFFFFFFF / v)


SHORT, LONG, RATIONAL = 3, 4, 5


class _File:
    def __init__(self):
        self.buf = bytearray(b"II*\x00" + b"\x00" * 4)
        self.last_next = 4  # where the next directory's offset goes

    def append(self, data):
        off = len(self.buf)
        self.buf += data
        return off

    def directory(self, entries, order):
        """entries: {tag: (type, values)}; order: tags in libtiff's data order."""
        diroff = (len(self.buf) + 1) & ~1
        if diroff > len(self.buf):
            self.buf += b"\x00" * (diroff - len(self.buf))
        ndir = len(entries)
        dataoff = diroff + 2 + 12 * ndir + 4
        payload = {}
        blobs = bytearray()
        for tag in order:
            if tag not in entries:
                continue
            typ, vals = entries[tag]
            raw = _pack(typ, vals)
            if len(raw) <= 4:
                payload[tag] = raw.ljust(4, b"\x00")
            else:
                payload[tag] = struct.pack("<I", dataoff + len(blobs))
                blobs += raw
                if len(blobs) & 1:
                    blobs += b"\x00"
        d = bytearray(struct.pack("<H", ndir))
        for tag in sorted(entries):
            typ, vals = entries[tag]
            d += struct.pack("<HHI", tag, typ, len(vals) if typ != RATIONAL else len(vals) // 2) + payload[tag]
        d += b"\x00" * 4
        self.buf += d
        self.buf += blobs  # every value emitted here has even length
        struct.pack_into("<I", self.buf, self.last_next, diroff)
        self.last_next = diroff + 2 + 12 * ndir


def _pack(typ, vals):
    if typ == SHORT:
        return struct.pack("<%dH" % len(vals), *vals)
    if typ == LONG:
        return struct.pack("<%dI" % len(vals), *vals)
    return struct.pack("<%dI" % len(vals), *vals)


_DATA_ORDER = [256, 257, 282, 283, 254, 258, 259, 262, 274, 277, 278, 284, 296, 297, 279, 273, 338, 339]


def _page(f, rows, width, height, spp, bps, float_samples, resolution, compression, extra):
    scanline = width * spp * (bps // 8)
    rps = max(1, STRIPSIZE_DEFAULT // max(scanline, 1))
    nstrips = (height + rps - 1) // rps
    offsets, counts = [], []
    for s in range(nstrips):
        chunk = b"".join(packbits_encode_row(r) if compression == COMPRESSION_PACKBITS else r
                         for r in rows[s * rps:(s + 1) * rps])
        offsets.append(f.append(chunk))
        counts.append(len(chunk))
    e = {256: (SHORT if width < 65536 else LONG, [width]),
         257: (SHORT if height < 65536 else LONG, [height]),
         258: (SHORT, [bps] * spp),
         259: (SHORT, [compression]),
         262: (SHORT, [1 if spp == 1 else 2]),
         273: (LONG, offsets),
         277: (SHORT, [spp]),
         278: (SHORT if rps < 65536 else LONG, [rps]),
         279: (LONG, counts),
         284: (SHORT, [1])}
    if resolution is not None:
        e[282] = (RATIONAL, list(_rational(resolution[0])))
        e[283] = (RATIONAL, list(_rational(resolution[1])))
        e[296] = (SHORT, [3])
    if float_samples:
        e[339] = (SHORT, [3] * spp)
    e.update(extra)
    f.directory(e, _DATA_ORDER)


def _resolution(spacing):
    if spacing is None:
        return None
    rx, ry = 10.0 / spacing[0], 10.0 / spacing[1]
    return (rx, ry) if rx > 0 and ry > 0 else None


def write_vtk_tiff(img, spacing=(1.0, 1.0), compression=COMPRESSION_PACKBITS):
    """Bytes of the file vtkTIFFWriter writes for a VTK image.

    img: array [z][y][x] or [z][y][x][c] in VTK memory order (x fastest),
    dtype uint8, uint16 or float32; spacing: the image's (x, y) spacing.
    """
    a = np.ascontiguousarray(img)
    if a.ndim == 3:
        a = a[..., None]
    nz, ny, nx, spp = a.shape
    if a.dtype not in (np.uint8, np.uint16, np.float32):
        raise ValueError("unsigned char, unsigned short or float scalars only")
    bps = a.dtype.itemsize * 8
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    res = _resolution(spacing)
    f = _File()
    if nz == 1:
        rows = [a[0, y].tobytes() for y in range(ny - 1, -1, -1)]
        extra = {274: (SHORT, [1])}
        if spp > 3:
            extra[338] = (SHORT, [1] + [0] * (spp - 4))  # associated alpha, then unspecified
        _page(f, rows, nx, ny, spp, bps, a.dtype == np.float32, res, compression, extra)
    else:
        if spp != 1:
            raise ValueError("a multi-page file holds one component per pixel")
        for z in range(nz):
            rows = [a[z, y].tobytes() for y in range(ny)]
            extra = {254: (LONG, [2]), 297: (SHORT, [z, nz])}
            _page(f, rows, nx, ny, 1, bps, bps == 32, res, compression, extra)
    return bytes(f.buf)


def read_tiff(data):
    """Minimal reader of the files above: list of (tags, page array [y][x][c])
    in file row order."""
    bo = "<" if data[:2] == b"II" else ">"
    off = struct.unpack(bo + "I", data[4:8])[0]
    sizes = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 11: 4, 12: 8}
    fmts = {1: "B", 2: "c", 3: "H", 4: "I", 5: "I", 11: "f", 12: "d"}
    pages = []
    while off:
        n = struct.unpack(bo + "H", data[off:off + 2])[0]
        tags = {}
        for i in range(n):
            e = data[off + 2 + 12 * i:off + 14 + 12 * i]
            tag, typ, cnt = struct.unpack(bo + "HHI", e[:8])
            cnt2 = cnt * (2 if typ == 5 else 1)
            size = sizes[typ] * cnt
            raw = e[8:8 + size] if size <= 4 else data[struct.unpack(bo + "I", e[8:12])[0]:][:size]
            tags[tag] = list(struct.unpack(bo + fmts[typ] * cnt2, raw[:struct.calcsize(fmts[typ]) * cnt2]))
        w, h = tags[256][0], tags[257][0]
        spp = tags.get(277, [1])[0]
        bps = tags[258][0]
        rps = tags.get(278, [h])[0]
        dt = {8: np.uint8, 16: np.uint16, 32: np.float32 if tags.get(339, [1])[0] == 3 else np.uint32}[bps]
        line = w * spp * bps // 8
        raw = bytearray()
        for s, (so, sc) in enumerate(zip(tags[273], tags[279])):
            rows = min(rps, h - s * rps)
            chunk = data[so:so + sc]
            raw += packbits_decode(chunk, rows * line) if tags[259][0] == COMPRESSION_PACKBITS else chunk
        arr = np.frombuffer(bytes(raw), dtype=np.dtype(dt).newbyteorder(bo)).reshape(h, w, spp)
        pages.append((tags, arr))
        off = struct.unpack(bo + "I", data[off + 2 + 12 * n:off + 6 + 12 * n])[0]
    return pages
