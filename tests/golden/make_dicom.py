"""Writes the DICOM fixtures of tests/test_dicom.py (tests/golden/dicom/).

The files are encoded here from the standard (PS3.10 file format: preamble,
"DICM", explicit VR little endian file meta group; PS3.5 data encoding:
explicit / implicit VR, little / big endian, deflate, sequences of undefined
length, pixel cells of BitsAllocated with the BitsStored bits ending at
HighBit), with pixel values chosen here, so the reader under test is checked
against known values.  expected.npz holds, per file, the stored values a
reader must return ([frame][row][column], first row of the file first) and
cases.json the attributes and the read outcome.

Run from the repository root: python tests/golden/make_dicom.py
"""
import json
import os
import struct
import zlib

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dicom")
LONG_VR = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ", "SV", "UC", "UN", "UR", "UT", "UV"}

TS_IMPLICIT = "1.2.840.10008.1.2"
TS_EXPLICIT = "1.2.840.10008.1.2.1"
TS_DEFLATED = "1.2.840.10008.1.2.1.99"
TS_BIG = "1.2.840.10008.1.2.2"
TS_JPEG = "1.2.840.10008.1.2.4.50"


def pad(vr, b):
    if len(b) % 2:
        b += b"\x00" if vr in ("UI", "OB", "UN") else b" "
    return b


class Enc:
    def __init__(self, explicit=True, big=False):
        self.explicit, self.big = explicit, big
        self.o = ">" if big else "<"

    def tag(self, g, e):
        return struct.pack(self.o + "HH", g, e)

    def el(self, g, e, vr, value, undefined=False):
        value = pad(vr, value)
        n = 0xFFFFFFFF if undefined else len(value)
        if not self.explicit:
            return self.tag(g, e) + struct.pack(self.o + "I", n) + value
        if vr in LONG_VR:
            return self.tag(g, e) + vr.encode() + b"\x00\x00" + struct.pack(self.o + "I", n) + value
        return self.tag(g, e) + vr.encode() + struct.pack(self.o + "H", n) + value

    def us(self, g, e, v):
        return self.el(g, e, "US", struct.pack(self.o + "H", v))

    def s(self, g, e, vr, text):
        return self.el(g, e, vr, text.encode())

    def item(self, body, undefined):
        if undefined:
            return self.tag(0xFFFE, 0xE000) + struct.pack(self.o + "I", 0xFFFFFFFF) + body + \
                self.tag(0xFFFE, 0xE00D) + struct.pack(self.o + "I", 0)
        return self.tag(0xFFFE, 0xE000) + struct.pack(self.o + "I", len(body)) + body

    def sequence(self, g, e, items):
        head = self.el(g, e, "SQ", b"", undefined=True)
        return head + b"".join(items) + self.tag(0xFFFE, 0xE0DD) + struct.pack(self.o + "I", 0)


def meta(ts):
    m = Enc()
    body = m.el(0x0002, 0x0001, "OB", b"\x00\x01") + m.s(0x0002, 0x0002, "UI", "1.2.840.10008.5.1.4.1.1.13.1.3") + \
        m.s(0x0002, 0x0003, "UI", "1.2.3.4.5.6.7") + m.s(0x0002, 0x0010, "UI", ts) + \
        m.s(0x0002, 0x0012, "UI", "1.2.3.4")
    return b"\x00" * 128 + b"DICM" + m.el(0x0002, 0x0000, "UL", struct.pack("<I", len(body))) + body


def cells(values, alloc, stored, high, big):
    """Pixel cells: each stored value (two's complement in `stored` bits)
    placed at bits [high - stored + 1, high]; the other bits set to junk."""
    v = np.asarray(values, np.int64) & ((1 << stored) - 1)
    shift = high + 1 - stored
    junk = np.int64(((1 << alloc) - 1) ^ (((1 << stored) - 1) << shift))
    raw = (v << shift) | (junk & 0x5A5A5A5A)
    dt = {8: np.uint8, 16: np.uint16, 32: np.uint32}[alloc]
    a = raw.astype(dt)
    return a.astype(a.dtype.newbyteorder(">" if big else "<")).tobytes()


def dataset(enc, frames, alloc=16, stored=None, high=None, signed=False, patient=True, time="101500.000",
            spacing="0.14\\0.15", nframes_tag=True, seq=True, spp=1, photometric="MONOCHROME2", truncate=0):
    nf, rows, cols = frames.shape
    stored = alloc if stored is None else stored
    high = stored - 1 if high is None else high
    d = enc.s(0x0008, 0x0016, "UI", "1.2.840.10008.5.1.4.1.1.13.1.3")
    if time is not None:
        d += enc.s(0x0008, 0x0032, "TM", time)
    if seq:
        inner = enc.s(0x0008, 0x1150, "UI", "1.2.3") + enc.s(0x0008, 0x1155, "UI", "1.2.3.4.5")
        nested = enc.sequence(0x0040, 0xA730, [enc.item(enc.s(0x0040, 0xA040, "CS", "TEXT"), True)])
        d += enc.sequence(0x0008, 0x1115, [enc.item(inner + nested, True), enc.item(inner, False)])
    if patient:
        d += enc.s(0x0010, 0x0020, "LO", "KV-0001")
    d += enc.s(0x0018, 0x1164, "DS", "0.2\\0.2")  # ImagerPixelSpacing: used without PixelSpacing
    d += enc.us(0x0028, 0x0002, spp) + enc.s(0x0028, 0x0004, "CS", photometric)
    if nframes_tag:
        d += enc.s(0x0028, 0x0008, "IS", str(nf))
    d += enc.us(0x0028, 0x0010, rows) + enc.us(0x0028, 0x0011, cols)
    if spacing is not None:
        d += enc.s(0x0028, 0x0030, "DS", spacing)
    d += enc.us(0x0028, 0x0100, alloc) + enc.us(0x0028, 0x0101, stored) + enc.us(0x0028, 0x0102, high) + \
        enc.us(0x0028, 0x0103, 1 if signed else 0)
    px = cells(frames.ravel(), alloc, stored, high, enc.big)
    if truncate:
        px = px[:-truncate]
    d += enc.el(0x7FE0, 0x0010, "OB" if alloc == 8 else "OW", px)
    return d


def write(name, data):
    with open(os.path.join(OUT, name), "wb") as f:
        f.write(data)


def main():
    os.makedirs(os.path.join(OUT, "series"), exist_ok=True)
    rng = np.random.default_rng(7)
    cases, expected = {}, {}

    def add(name, data, frames, attrs):
        write(name, data)
        expected[name] = frames
        cases[name] = attrs

    f = rng.integers(0, 4096, (2, 6, 5))
    add("explicit_le_u16_12bit.dcm", meta(TS_EXPLICIT) + dataset(Enc(), f, 16, 12, 11),
        f, {"ok": True, "spacing": [0.15, 0.14], "patient": True, "time": "101500.000"})
    f = rng.integers(-32768, 32768, (1, 4, 7))
    add("implicit_le_s16.dcm", meta(TS_IMPLICIT) + dataset(Enc(explicit=False), f, 16, signed=True, time="090000",
                                                           nframes_tag=False),
        f, {"ok": True, "spacing": [0.15, 0.14], "patient": True, "time": "090000"})
    f = rng.integers(0, 256, (3, 5, 4))
    add("explicit_be_u8.dcm", meta(TS_BIG) + dataset(Enc(big=True), f, 8, spacing=None),
        f, {"ok": True, "spacing": [0.2, 0.2], "patient": True, "time": "101500.000"})
    f = rng.integers(0, 65536, (2, 3, 9))
    add("explicit_be_u16.dcm", meta(TS_BIG) + dataset(Enc(big=True), f, 16, time=None),
        f, {"ok": True, "spacing": [0.15, 0.14], "patient": True, "time": ""})
    f = rng.integers(0, 1 << 14, (2, 8, 6))
    comp = zlib.compressobj(9, zlib.DEFLATED, -15)
    body = dataset(Enc(), f, 16, 14, 13)
    add("deflated_u16_14bit.dcm", meta(TS_DEFLATED) + comp.compress(body) + comp.flush(),
        f, {"ok": True, "spacing": [0.15, 0.14], "patient": True, "time": "101500.000"})
    f = rng.integers(0, 1000, (1, 3, 3))
    add("bare_implicit.dcm", dataset(Enc(explicit=False), f, 16, patient=False, seq=False),
        f, {"ok": True, "spacing": [0.15, 0.14], "patient": False, "time": "101500.000"})
    f = rng.integers(-2048, 2048, (2, 5, 5))
    add("signed_12bit_high_15.dcm", meta(TS_EXPLICIT) + dataset(Enc(), f, 16, 12, 15, signed=True),
        f, {"ok": True, "spacing": [0.15, 0.14], "patient": True, "time": "101500.000"})
    f = rng.integers(0, 1 << 31, (1, 2, 3)) * 2 + 1
    add("explicit_le_u32.dcm", meta(TS_EXPLICIT) + dataset(Enc(), f, 32),
        f, {"ok": True, "spacing": [0.15, 0.14], "patient": True, "time": "101500.000"})
    # rejected
    f = rng.integers(0, 100, (1, 2, 2))
    jpeg = Enc().el(0x7FE0, 0x0010, "OB", b"", undefined=True)
    bad = dataset(Enc(), f, 16)
    add("reject_jpeg.dcm", meta(TS_JPEG) + bad, f, {"ok": False, "why": "transfer syntax"})
    add("reject_encapsulated.dcm", meta(TS_EXPLICIT) + bad[:bad.index(b"\xe0\x7f\x10\x00")] + jpeg +
        Enc().tag(0xFFFE, 0xE0DD) + b"\x00" * 4, f, {"ok": False, "why": "encapsulated"})
    add("reject_rgb.dcm", meta(TS_EXPLICIT) + dataset(Enc(), np.zeros((1, 2, 6), int), 8, spp=3,
                                                      photometric="RGB"),
        f, {"ok": False, "why": "one sample per pixel"})
    add("reject_short_pixels.dcm", meta(TS_EXPLICIT) + dataset(Enc(), f, 16, truncate=2), f,
        {"ok": False, "why": "shorter"})
    add("reject_not_dicom.dcm", bytes(rng.integers(0, 256, 300, dtype=np.uint8)) + b"junk", f,
        {"ok": False, "why": "not a DICOM file"})

    # a series: order by acquisition time (read only where a PatientID is
    # present), ties in file-name order, other extensions ignored
    series = {}
    for name, (nf, patient, time) in {"b.dcm": (2, True, "120000"), "a.dcm": (1, True, "110000"),
                                      "c.DCM": (1, False, "130000"), "d.dcm": (1, True, "110000"),
                                      "e.Dcm": (1, True, "000000")}.items():
        fr = rng.integers(0, 4096, (nf, 7, 9))
        write(os.path.join("series", name), meta(TS_EXPLICIT) + dataset(Enc(), fr, 16, 12, 11, patient=patient,
                                                                        time=time))
        expected["series/" + name] = fr
        series[name] = {"patient": patient, "time": time, "frames": nf}
    write(os.path.join("series", "notes.txt"), b"not an image\n")
    cases["series"] = {"files": series, "order": ["c.DCM", "a.dcm", "d.dcm", "b.dcm"]}
    np.savez_compressed(os.path.join(OUT, "expected.npz"),
                        **{k.replace("/", "__"): np.asarray(v, np.int64) for k, v in expected.items()})
    with open(os.path.join(OUT, "cases.json"), "w") as fh:
        json.dump(cases, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
