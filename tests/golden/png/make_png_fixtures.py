"""Writes the PNG fixtures of tests/test_png.py and their expected samples.

expected.npz maps each file name to the channels CImg receives for it in
the reference build (ImageMagick's PNM output: gray and gray+alpha as one
channel, colour and palette as three, alpha dropped, samples as stored),
shape (height, width, channels).

* cornell_box_default.png, mitsuba_default.png: the reference's own
  textures (scene/cornell-box/default.png, 8-bit RGB; scene/mitsuba/
  default.png, 8-bit RGBA with an ICC profile), copied as data; expected
  samples from PIL's decoder.
* synthetic files, seeded: written by PIL where it can write the form, else
  by the small encoder below (16-bit colour, gray+alpha 16, Adam7
  interlacing, each filter type, fixed-Huffman and stored deflate blocks);
  every file is read back with PIL and must give its source samples (PIL
  keeps 8-bit samples and 16-bit gray; for 16-bit colour it keeps the high
  bytes, which are compared instead).

Run from the repository root in the build container (needs PIL and
/root/reference):  python tests/golden/png/make_png_fixtures.py
"""
import io
import os
import shutil
import struct
import zlib

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/scene"


def chunk(kind, data):
    c = struct.pack(">I", len(data)) + kind + data
    return c + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def filter_row(f, row, prev, bpp):
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        if f == 0:
            p = 0
        elif f == 1:
            p = a
        elif f == 2:
            p = b
        elif f == 3:
            p = (a + b) >> 1
        else:
            pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
            p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
        out[i] = (x - p) & 0xFF
    return bytes(out)


def pack_rows(samples, depth):
    """(h, w, s) integer samples -> list of row byte strings."""
    h, w, s = samples.shape
    rows = []
    for y in range(h):
        vals = samples[y].reshape(-1)
        if depth == 16:
            rows.append(vals.astype(">u2").tobytes())
        elif depth == 8:
            rows.append(vals.astype(np.uint8).tobytes())
        else:
            bits = np.zeros(len(vals) * depth, np.uint8)
            for k in range(depth):
                bits[k::depth] = (vals >> (depth - 1 - k)) & 1
            pad = (-len(bits)) % 8
            rows.append(np.packbits(np.concatenate([bits, np.zeros(pad, np.uint8)])).tobytes())
    return rows


def encode(samples, depth, ctype, palette=None, interlace=False, filters=(0, 1, 2, 3, 4),
           strategy=zlib.Z_DEFAULT_STRATEGY, level=6):
    h, w, s = samples.shape
    bpp = max(1, s * depth // 8)
    if interlace:
        passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4),
                  (1, 0, 2, 2), (0, 1, 1, 2)]
    else:
        passes = [(0, 0, 1, 1)]
    raw = bytearray()
    k = 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = pack_rows(sub, depth)
        prev = bytes(len(rows[0]))
        for r in rows:
            f = filters[k % len(filters)]
            k += 1
            raw += bytes([f]) + filter_row(f, r, prev, bpp)
            prev = r
    co = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
    z = co.compress(bytes(raw)) + co.flush()
    ihdr = struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr)
    if palette is not None:
        out += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    out += chunk(b"tEXt", b"Comment\x00ancillary chunk to skip")
    # split the image data over two IDAT chunks
    out += chunk(b"IDAT", z[: len(z) // 2]) + chunk(b"IDAT", z[len(z) // 2:])
    return out + chunk(b"IEND", b"")


def main():
    rng = np.random.default_rng(2024)
    expected = {}

    def put(name, data, expect, pil_expect=None):
        with open(os.path.join(HERE, name), "wb") as f:
            f.write(data)
        got = np.asarray(Image.open(io.BytesIO(data)))
        want = expect if pil_expect is None else pil_expect
        if got.ndim == 2:
            got = got[:, :, None]
        got = got[:, :, : want.shape[2]] if got.shape[2] >= want.shape[2] else got
        assert np.array_equal(got.astype(np.int64), want.astype(np.int64)), name
        expected[name] = expect.astype(np.uint16)

    # the reference's own textures
    for src, name in [("cornell-box/default.png", "cornell_box_default.png"),
                      ("mitsuba/default.png", "mitsuba_default.png")]:
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, name))
        a = np.asarray(Image.open(os.path.join(HERE, name)))
        expected[name] = a[:, :, :3].astype(np.uint16)

    # PIL-written 8-bit forms
    img = rng.integers(0, 256, (37, 53, 4)).astype(np.uint8)
    img[10:20] = img[10:11]                         # repeats: long matches
    for mode, ch in [("RGB", 3), ("RGBA", 4), ("L", 1), ("LA", 2)]:
        src = img[:, :, :ch] if ch > 1 else img[:, :, :1]
        b = io.BytesIO()
        Image.fromarray(src if ch > 1 else src[:, :, 0], mode).save(b, "PNG", compress_level=9)
        keep = 1 if mode in ("L", "LA") else 3
        put(f"pil_{mode.lower()}.png", b.getvalue(), src[:, :, :keep])
    b = io.BytesIO()
    Image.fromarray(img[:, :, :3], "RGB").save(b, "PNG", compress_level=0)   # stored blocks
    put("pil_rgb_stored.png", b.getvalue(), img[:, :, :3])
    pal_img = Image.fromarray(img[:, :, :3], "RGB").quantize(colors=200)
    b = io.BytesIO()
    pal_img.save(b, "PNG")
    pal = np.asarray(pal_img.getpalette(), np.uint8).reshape(-1, 3)
    put("pil_palette8.png", b.getvalue(), pal[np.asarray(pal_img)],
        pil_expect=np.asarray(pal_img)[:, :, None])
    g16 = rng.integers(0, 65536, (29, 31)).astype(np.uint16)
    b = io.BytesIO()
    Image.fromarray(g16.astype(np.int32), "I").convert("I;16").save(b, "PNG")
    put("pil_gray16.png", b.getvalue(), g16[:, :, None])

    # encoder-written forms
    rgb16 = rng.integers(0, 65536, (23, 41, 3)).astype(np.int64)
    put("rgb16_filters.png", encode(rgb16, 16, 2), rgb16, pil_expect=rgb16 >> 8)
    rgba16 = rng.integers(0, 65536, (19, 17, 4)).astype(np.int64)
    put("rgba16_interlaced.png", encode(rgba16, 16, 6, interlace=True), rgba16[:, :, :3],
        pil_expect=rgba16[:, :, :3] >> 8)
    ga16 = rng.integers(0, 65536, (13, 27, 2)).astype(np.int64)
    put("gray_alpha16.png", encode(ga16, 16, 4), ga16[:, :, :1], pil_expect=ga16[:, :, :1] >> 8)
    # smooth fields with a little noise compress, so deflate codes them
    # with Huffman blocks rather than storing them
    yy, xx = np.indices((33, 35))
    rgb8 = np.stack([(3 * xx + 2 * yy) % 256, (5 * yy) % 256, (xx * yy) % 256], axis=2)
    rgb8 = (rgb8 + rng.integers(0, 3, rgb8.shape)) % 256
    for f in range(5):
        put(f"rgb8_filter{f}.png", encode(rgb8, 8, 2, filters=(f,)), rgb8)
    put("rgb8_fixed_huffman.png", encode(rgb8, 8, 2, strategy=zlib.Z_FIXED), rgb8)
    put("rgb8_interlaced.png", encode(rgb8, 8, 2, interlace=True), rgb8)
    rgba8 = rng.integers(0, 256, (9, 11, 4)).astype(np.int64)
    put("rgba8_interlaced_tiny.png", encode(rgba8, 8, 6, interlace=True), rgba8[:, :, :3])
    for depth in (1, 2, 4):
        g = rng.integers(0, 1 << depth, (21, 37, 1)).astype(np.int64)
        scale = 255 // ((1 << depth) - 1)
        put(f"gray{depth}.png", encode(g, depth, 0), g, pil_expect=g * (scale if depth > 1 else 1))
        n = 1 << depth
        palette = rng.integers(0, 256, (n, 3))
        idx = rng.integers(0, n, (21, 37, 1)).astype(np.int64)
        put(f"palette{depth}_interlaced.png", encode(idx, depth, 3, palette=palette,
                                                     interlace=True),
            palette[idx[:, :, 0]], pil_expect=idx)
    big = (np.indices((300, 400)).sum(axis=0) % 251).astype(np.int64)[:, :, None]
    big = np.concatenate([big, (big * 7) % 256, (big * 13) % 256], axis=2)
    put("rgb8_large_patterns.png", encode(big, 8, 2), big)

    np.savez_compressed(os.path.join(HERE, "expected.npz"), **expected)
    print(f"wrote {len(expected)} PNG fixtures and expected.npz to {HERE}")


if __name__ == "__main__":
    main()
