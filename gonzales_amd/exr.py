"""OpenEXR read and write for the film's images (the reference writes EXR through OpenImageIO,
Image/OpenImageIOWriter.swift:6-97, and keeps its regression image as a tiled, ZIP-compressed
float EXR in Tests/reference).

Written: scanline images, FLOAT channels B, G, R (and A when given; channel names sorted as the
format requires), NO or ZIP compression (16 scanlines per block; zlib over the byte-interleaved,
delta-coded block). Read: scanline or single-level tiled images with HALF or FLOAT channels and
NO, ZIPS or ZIP compression, which covers the files the reference writes. The format follows the
OpenEXR file layout documentation (magic 20000630, version 2, attribute header, chunk offset
table, chunks).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

MAGIC = 20000630
_COMP_NONE, _COMP_RLE, _COMP_ZIPS, _COMP_ZIP = 0, 1, 2, 3
_LINES = {_COMP_NONE: 1, _COMP_ZIPS: 1, _COMP_ZIP: 16}
_PIXEL = {0: (np.uint32, 4), 1: (np.float16, 2), 2: (np.float32, 4)}  # UINT, HALF, FLOAT


def _attr(name: str, typ: str, data: bytes) -> bytes:
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def _predictor_encode(raw: bytes) -> bytes:
    """Byte interleave (even bytes, then odd bytes) then delta coding + 128, as the ZIP codec does."""
    a = np.frombuffer(raw, np.uint8)
    t = np.concatenate([a[0::2], a[1::2]])
    d = t.astype(np.int16)
    d[1:] = (t[1:].astype(np.int16) - t[:-1].astype(np.int16) + 128) & 255
    return d.astype(np.uint8).tobytes()


def _predictor_decode(buf: bytes) -> bytes:
    """Undo the ZIP codec's delta coding (t[i] = t[i-1] + d[i] - 128) and byte interleave."""
    d = np.frombuffer(buf, np.uint8).astype(np.int64)
    d[1:] -= 128
    t = (np.cumsum(d) & 255).astype(np.uint8)
    half = (len(t) + 1) // 2
    out = np.empty(len(t), np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def write_exr(path: str, rgb: np.ndarray, alpha: np.ndarray | None = None, compression: str = "zip") -> None:
    """A (H, W, 3) float image (and optional (H, W) alpha) as a float scanline EXR."""
    rgb = np.asarray(rgb, np.float32)
    h, w, _ = rgb.shape
    chans = {"B": rgb[..., 2], "G": rgb[..., 1], "R": rgb[..., 0]}
    if alpha is not None:
        chans["A"] = np.asarray(alpha, np.float32)
    names = sorted(chans)
    comp = {"none": _COMP_NONE, "zip": _COMP_ZIP}[compression]
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for n in names) + b"\0"
    hdr = struct.pack("<ii", MAGIC, 2)
    hdr += _attr("channels", "chlist", chlist)
    hdr += _attr("compression", "compression", bytes([comp]))
    hdr += _attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    hdr += _attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    hdr += _attr("lineOrder", "lineOrder", b"\0")
    hdr += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    hdr += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    hdr += b"\0"
    lines = _LINES[comp]
    chunks = []
    for y0 in range(0, h, lines):
        y1 = min(h, y0 + lines)
        raw = b"".join(np.ascontiguousarray(chans[n][y, :]).astype("<f4").tobytes() for y in range(y0, y1) for n in names)
        data = raw
        if comp == _COMP_ZIP:
            z = zlib.compress(_predictor_encode(raw))
            data = z if len(z) < len(raw) else raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    table_at = len(hdr)
    first = table_at + 8 * len(chunks)
    offsets, pos = [], first
    for c in chunks:
        offsets.append(pos)
        pos += len(c)
    with open(path, "wb") as f:
        f.write(hdr + struct.pack(f"<{len(offsets)}Q", *offsets) + b"".join(chunks))


def _parse_header(data: bytes):
    magic, version = struct.unpack_from("<ii", data, 0)
    if magic != MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0xFF != 2:
        raise ValueError(f"EXR version {version & 0xFF}")
    flags = version >> 8
    if flags & ~0x2:
        raise ValueError("deep, multi-part or long-name EXR files are not supported")
    tiled = bool(flags & 0x2)
    i, attrs = 8, {}
    while True:
        j = data.index(b"\0", i)
        name = data[i:j].decode()
        i = j + 1
        if not name:
            break
        j = data.index(b"\0", i)
        typ = data[i:j].decode()
        i = j + 1
        (size,) = struct.unpack_from("<i", data, i)
        i += 4
        attrs[name] = (typ, data[i:i + size])
        i += size
    return tiled, attrs, i


def read_exr(path: str) -> dict[str, np.ndarray]:
    """Channels of a single-part scanline or single-level tiled EXR as float32 (H, W) arrays."""
    with open(path, "rb") as f:
        data = f.read()
    tiled, attrs, at = _parse_header(data)
    chl = attrs["channels"][1]
    chans, k = [], 0
    while chl[k] != 0:
        e = chl.index(b"\0", k)
        name = chl[k:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", chl, e + 1)
        if (xs, ys) != (1, 1):
            raise ValueError("subsampled channels are not supported")
        chans.append((name, ptype))
        k = e + 1 + 16
    comp = attrs["compression"][1][0]
    if comp not in (_COMP_NONE, _COMP_ZIPS, _COMP_ZIP):
        raise ValueError(f"EXR compression {comp} is not supported (none, ZIPS and ZIP are)")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    out = {n: np.zeros((h, w), np.float32) for n, _ in chans}

    def decode(buf: bytes, raw_size: int) -> bytes:
        if len(buf) == raw_size or comp == _COMP_NONE:
            return buf
        return _predictor_decode(zlib.decompress(buf))

    def fill(raw: bytes, ty0: int, tx0: int, th: int, tw: int):
        pos = 0
        for yy in range(th):
            for name, ptype in chans:
                dt, size = _PIXEL[ptype]
                row = np.frombuffer(raw, "<" + np.dtype(dt).str[1:], tw, pos)
                out[name][ty0 + yy, tx0:tx0 + tw] = row.astype(np.float32)
                pos += tw * size

    bpp = sum(_PIXEL[p][1] for _, p in chans)
    if tiled:
        tx, ty, mode = struct.unpack("<IIB", attrs["tiles"][1])
        if mode & 0xF != 0:
            raise ValueError("mip-mapped or rip-mapped EXR files are not supported")
        nx, ny = (w + tx - 1) // tx, (h + ty - 1) // ty
        offsets = struct.unpack_from(f"<{nx * ny}Q", data, at)
        for off in offsets:
            cx, cy, lx, ly, size = struct.unpack_from("<iiiii", data, off)
            tw, th = min(tx, w - cx * tx), min(ty, h - cy * ty)
            raw = decode(data[off + 20:off + 20 + size], tw * th * bpp)
            fill(raw, cy * ty, cx * tx, th, tw)
    else:
        lines = _LINES[comp]
        n_chunks = (h + lines - 1) // lines
        offsets = struct.unpack_from(f"<{n_chunks}Q", data, at)
        for off in offsets:
            y, size = struct.unpack_from("<ii", data, off)
            th = min(lines, y1 - y + 1)
            raw = decode(data[off + 8:off + 8 + size], th * w * bpp)
            fill(raw, y - y0, 0, th, w)
    return out


def read_rgb(path: str) -> np.ndarray:
    c = read_exr(path)
    return np.stack([c["R"], c["G"], c["B"]], -1)


__all__ = ["write_exr", "read_exr", "read_rgb"]
