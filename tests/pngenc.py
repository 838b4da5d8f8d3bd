"""A small PNG encoder for tests, written from the PNG specification
(ISO/IEC 15948): signature, IHDR/IDAT/IEND chunks with CRC-32, zlib stream,
the five row filters and Adam7 interlacing.  It is the independent half of
the decoder's round-trip tests and writes the depth images of test datasets.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _pack_rows(samples: np.ndarray, bit_depth: int) -> list[bytes]:
    """samples: h x (w*channels) integers -> packed row bytes."""
    rows = []
    for r in samples:
        if bit_depth == 16:
            rows.append(r.astype(">u2").tobytes())
        elif bit_depth == 8:
            rows.append(r.astype(np.uint8).tobytes())
        else:
            per = 8 // bit_depth
            n = (len(r) + per - 1) // per
            out = bytearray(n)
            for i, v in enumerate(r):
                out[i // per] |= int(v) << (8 - bit_depth * (i % per + 1))
            rows.append(bytes(out))
    return rows


def _paeth(a: int, b: int, c: int) -> int:
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter(rows: list[bytes], bpp: int, choose) -> bytes:
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for y, row in enumerate(rows):
        ft = choose(y)
        f = bytearray(len(row))
        for i in range(len(row)):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = (0, a, b, (a + b) >> 1, _paeth(a, b, c))[ft]
            f[i] = (row[i] - pred) & 0xFF
        out.append(ft)
        out += f
        prev = row
    return bytes(out)


def encode(samples: np.ndarray, bit_depth: int = 16, color_type: int = 0, interlace: bool = False,
           filters=None, chunk_split: int = 0) -> bytes:
    """samples: h x w (gray) or h x w x 2 (gray + alpha) integers.
    filters: None (cycle 0..4 by row), an int, or a function of the row."""
    s = np.asarray(samples)
    h, w = s.shape[:2]
    channels = 1 if s.ndim == 2 else s.shape[2]
    flat = s.reshape(h, w * channels)
    bpp = max(1, bit_depth * channels // 8)
    choose = (lambda y: y % 5) if filters is None else (filters if callable(filters) else (lambda y: filters))
    if interlace:
        data = b""
        for x0, y0, dx, dy in ADAM7:
            sub = s[y0::dy, x0::dx]
            if sub.shape[0] == 0 or sub.shape[1] == 0:
                continue
            data += _filter(_pack_rows(sub.reshape(sub.shape[0], -1), bit_depth), bpp, choose)
    else:
        data = _filter(_pack_rows(flat, bit_depth), bpp, choose)
    z = zlib.compress(data, 6)
    ihdr = struct.pack(">IIBBBBB", w, h, bit_depth, color_type, 0, 0, 1 if interlace else 0)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)
    if chunk_split:
        for i in range(0, len(z), chunk_split):
            out += _chunk(b"IDAT", z[i:i + chunk_split])
    else:
        out += _chunk(b"IDAT", z)
    return out + _chunk(b"IEND", b"")


def write_depth_png(path, raw: np.ndarray):
    """16-bit grayscale PNG of a uint16 h x w array."""
    with open(path, "wb") as f:
        f.write(encode(np.asarray(raw, np.uint16), 16, 0, filters=lambda y: (y * 3) % 5))
