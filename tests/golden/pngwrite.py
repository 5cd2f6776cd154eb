"""Minimal PNG encoder for the texture fixtures (PNG specification, ISO/IEC
15948): non-interlaced grey / RGB / palette / grey+alpha / RGBA images at
8 or 16 bits, each row filtered with a chosen filter type (0 None, 1 Sub,
2 Up, 3 Average, 4 Paeth) so a decoder's unfiltering of every type is
exercised.  Written independently of the decoder it checks."""
import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter_row(row, prev, bpp, ftype):
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = (0, a, b, (a + b) // 2, _paeth(a, b, c))[ftype]
        out[i] = (x - pred) & 0xFF
    return bytes(out)


def png_bytes(samples, color_type, depth=8, palette=None, filters=(0, 1, 2, 3, 4)):
    """samples: (h, w, channels) integer array, rows top first (palette
    images: (h, w) or (h, w, 1) indices).  Row y uses filters[y % len]."""
    s = np.asarray(samples)
    if s.ndim == 2:
        s = s[:, :, None]
    h, w, c = s.shape
    assert c == CHANNELS[color_type] and depth in (8, 16)
    if depth == 16:
        raw = s.astype(">u2").tobytes()
    else:
        raw = s.astype(np.uint8).tobytes()
    stride = w * c * depth // 8
    bpp = c * depth // 8
    body, prev = b"", bytes(stride)
    for y in range(h):
        row = raw[y * stride:(y + 1) * stride]
        f = filters[y % len(filters)]
        body += bytes([f]) + _filter_row(row, prev, bpp, f)
        prev = row
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, 0))
    if color_type == 3:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    # split the stream over two IDAT chunks: a decoder must concatenate them
    z = zlib.compress(body, 9)
    half = len(z) // 2
    out += _chunk(b"IDAT", z[:half]) + _chunk(b"IDAT", z[half:]) + _chunk(b"IEND", b"")
    return out
