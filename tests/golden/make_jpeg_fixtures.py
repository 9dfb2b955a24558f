"""Writes tests/golden/jpeg: small JPEG files and, in expected.npz, the
samples libjpeg's default decode gives for each (read through PIL, which
decodes with libjpeg-turbo's defaults: islow inverse DCT, fancy
upsampling -- the decode ImageMagick runs when the reference's CImg loads a
.jpg texture).

Files written by PIL cover its encoder's forms (baseline 4:4:4, 4:2:2,
4:2:0, gray, progressive, restart markers, optimized Huffman tables, a
coarse quality, Adobe RGB, widths of 3 to 5 samples).  The rest come from
`encode_baseline` below, a plain sequential encoder for what PIL cannot
write: other sampling factors (4:4:0, 4:1:1, mixed), one scan per
component, and 16-bit quantization tables.

Run from the repository root: python tests/golden/make_jpeg_fixtures.py
"""
from __future__ import annotations

import io
import os

import numpy as np
from PIL import Image

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jpeg")


def test_image(w, h, seed):
    """Smooth gradients, an edge and some noise, as uint8 RGB."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    r = 128 + 100 * np.sin(x / max(w, 1) * 5.0 + seed) * np.cos(y / max(h, 1) * 3.0)
    g = 255 * (x + y) / max(w + h - 2, 1)
    b = np.where((x - w / 2) ** 2 + (y - h / 2) ** 2 < (min(w, h) / 3) ** 2, 230.0, 25.0)
    img = np.stack([r, g, b], axis=-1) + rng.normal(0, 12, (h, w, 3))
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


# ----------------------------------------------------------- own encoder
def _zigzag():
    order = []
    for s in range(15):
        lo, hi = (0, s) if s < 8 else (s - 7, 7)
        rows = range(lo, hi + 1) if s & 1 else range(hi, lo - 1, -1)
        order += [r * 8 + (s - r) for r in rows]
    return order


ZIGZAG = _zigzag()
_u = np.arange(8)
DCT = np.array([[(np.sqrt(0.5) if k == 0 else 1.0) * np.cos((2 * _u + 1) * k * np.pi / 16)
                 for k in range(8)]]).reshape(8, 8) / 2.0  # rows k, columns x


class BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, value, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc = 0
                self.n = 0

    def flush(self):
        while self.n:
            self.put(1, 1)
        return bytes(self.out)


def _category(v):
    a = abs(int(v))
    s = a.bit_length()
    return s, (v if v >= 0 else v + (1 << s) - 1)


def encode_baseline(rgb, factors, quality_scale=1.0, qbits16=False, interleaved=True,
                    restart=0):
    """A sequential Huffman JPEG of `rgb` (uint8 HxWx3) with per-component
    (h, v) `factors`; fixed-length Huffman codes (DC 4 bits, AC 8 bits) keep
    the encoder small.  qbits16 writes 16-bit tables (SOF1)."""
    h_img, w_img = rgb.shape[:2]
    f = rgb.astype(np.float64)
    ycc = np.stack([0.299 * f[..., 0] + 0.587 * f[..., 1] + 0.114 * f[..., 2],
                    -0.168736 * f[..., 0] - 0.331264 * f[..., 1] + 0.5 * f[..., 2] + 128,
                    0.5 * f[..., 0] - 0.418688 * f[..., 1] - 0.081312 * f[..., 2] + 128], -1)
    hmax = max(a for a, _ in factors)
    vmax = max(b for _, b in factors)
    mcux = -(-w_img // (8 * hmax))
    mcuy = -(-h_img // (8 * vmax))
    base = np.array([16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55] * 4, float)
    qt = [np.clip(np.rint((base + 4 * np.arange(64) % 7) * quality_scale * (1 + c)), 1,
                  65535 if qbits16 else 255).astype(int) for c in range(2)]
    comps = []
    for ci, (hs, vs) in enumerate(factors):
        dw, dh = -(-w_img * hs // hmax), -(-h_img * vs // vmax)
        sx, sy = hmax // hs, vmax // vs
        full = np.pad(ycc[..., ci], ((0, dh * sy - h_img), (0, dw * sx - w_img)), mode="edge")
        plane = full.reshape(dh, sy, dw, sx).mean(axis=(1, 3))
        bw, bh = mcux * hs, mcuy * vs
        padded = np.pad(plane, ((0, bh * 8 - dh), (0, bw * 8 - dw)), mode="edge")
        q = qt[0 if ci == 0 else 1].reshape(8, 8)
        blocks = np.zeros((bh, bw, 64), int)
        for r in range(bh):
            for c in range(bw):
                px = np.rint(padded[r * 8:r * 8 + 8, c * 8:c * 8 + 8]) - 128
                coef = DCT @ px @ DCT.T
                blocks[r, c] = np.rint(coef / q).astype(int).reshape(64)[ZIGZAG]
        comps.append(dict(id=ci + 1, h=hs, v=vs, tq=0 if ci == 0 else 1, blocks=blocks,
                          bcols=-(-dw // 8), brows=-(-dh // 8)))

    def seg(marker, payload):
        return bytes([0xFF, marker]) + (len(payload) + 2).to_bytes(2, "big") + payload

    out = bytearray(b"\xff\xd8")
    out += seg(0xE0, b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00")
    for t in range(2):
        if qbits16:
            out += seg(0xDB, bytes([0x10 | t]) + b"".join(int(v).to_bytes(2, "big")
                                                          for v in qt[t][ZIGZAG]))
        else:
            out += seg(0xDB, bytes([t]) + bytes(int(v) for v in qt[t][ZIGZAG]))
    sof = bytes([8]) + h_img.to_bytes(2, "big") + w_img.to_bytes(2, "big") + bytes([len(comps)])
    for c in comps:
        sof += bytes([c["id"], (c["h"] << 4) | c["v"], c["tq"]])
    out += seg(0xC1 if qbits16 else 0xC0, sof)
    dc_syms = list(range(12))
    ac_syms = [0x00, 0xF0] + [(r << 4) | s for r in range(16) for s in range(1, 11)]
    for cls, syms, length in ((0, dc_syms, 4), (1, ac_syms, 8)):
        counts = [0] * 16
        counts[length - 1] = len(syms)
        out += seg(0xC4, bytes([cls << 4]) + bytes(counts) + bytes(syms))
    dc_code = {s: i for i, s in enumerate(dc_syms)}
    ac_code = {s: i for i, s in enumerate(ac_syms)}
    if restart:
        out += seg(0xDD, restart.to_bytes(2, "big"))

    def write_block(bw_, blk, pred):
        diff = int(blk[0]) - pred
        s, bits = _category(diff)
        bw_.put(dc_code[s], 4)
        if s:
            bw_.put(bits, s)
        run = 0
        last = max([k for k in range(1, 64) if blk[k] != 0], default=0)
        for k in range(1, last + 1):
            v = int(blk[k])
            if v == 0:
                run += 1
                continue
            while run > 15:
                bw_.put(ac_code[0xF0], 8)
                run -= 16
            s, bits = _category(v)
            bw_.put(ac_code[(run << 4) | s], 8)
            bw_.put(bits, s)
            run = 0
        if last < 63:
            bw_.put(ac_code[0x00], 8)
        return int(blk[0])

    scans = [comps] if interleaved else [[c] for c in comps]
    for sc in scans:
        sos = bytes([len(sc)])
        for c in sc:
            sos += bytes([c["id"], 0x00])
        out += seg(0xDA, sos + bytes([0, 63, 0]))
        units = []
        if len(sc) == 1:
            c = sc[0]
            units = [[(c, r, col)] for r in range(c["brows"]) for col in range(c["bcols"])]
        else:
            for my in range(mcuy):
                for mx in range(mcux):
                    units.append([(c, my * c["v"] + y, mx * c["h"] + x)
                                  for c in sc for y in range(c["v"]) for x in range(c["h"])])
        bw_ = BitWriter()
        preds = {c["id"]: 0 for c in sc}
        rst = 0
        for m, unit in enumerate(units):
            if restart and m and m % restart == 0:
                out += bw_.flush() + bytes([0xFF, 0xD0 + rst])
                rst = (rst + 1) & 7
                bw_ = BitWriter()
                preds = {c["id"]: 0 for c in sc}
            for c, r, col in unit:
                preds[c["id"]] = write_block(bw_, c["blocks"][r, col], preds[c["id"]])
        out += bw_.flush()
    out += b"\xff\xd9"
    return bytes(out)


# --------------------------------------------------------------- fixtures
def pil_jpeg(img, **kw):
    b = io.BytesIO()
    img.save(b, "JPEG", **kw)
    return b.getvalue()


def fixtures():
    rgb = lambda w, h, s: Image.fromarray(test_image(w, h, s))  # noqa: E731
    gray = lambda w, h, s: Image.fromarray(test_image(w, h, s)[..., 1])  # noqa: E731
    yield "pil_420.jpg", pil_jpeg(rgb(37, 29, 1), quality=90, subsampling=2)
    yield "pil_422.jpg", pil_jpeg(rgb(41, 23, 2), quality=85, subsampling=1)
    yield "pil_444.jpg", pil_jpeg(rgb(19, 17, 3), quality=95, subsampling=0)
    yield "pil_gray.jpg", pil_jpeg(gray(33, 21, 4), quality=80)
    yield "pil_progressive_420.jpg", pil_jpeg(rgb(45, 31, 5), quality=88, subsampling=2,
                                              progressive=True)
    yield "pil_progressive_444.jpg", pil_jpeg(rgb(26, 35, 6), quality=92, subsampling=0,
                                              progressive=True)
    yield "pil_progressive_gray.jpg", pil_jpeg(gray(40, 24, 7), quality=75, progressive=True)
    yield "pil_restart.jpg", pil_jpeg(rgb(53, 37, 8), quality=90, subsampling=2,
                                      restart_marker_blocks=3)
    yield "pil_optimized.jpg", pil_jpeg(rgb(30, 30, 9), quality=90, subsampling=1, optimize=True)
    yield "pil_coarse.jpg", pil_jpeg(rgb(48, 40, 10), quality=3, subsampling=2)
    yield "pil_adobe_rgb.jpg", pil_jpeg(rgb(24, 20, 11), quality=90, subsampling=0, keep_rgb=True)
    yield "pil_width3_420.jpg", pil_jpeg(rgb(3, 5, 12), quality=90, subsampling=2)
    yield "pil_width4_422.jpg", pil_jpeg(rgb(4, 4, 13), quality=90, subsampling=1)
    yield "pil_width5_420.jpg", pil_jpeg(rgb(5, 3, 14), quality=90, subsampling=2)
    src = test_image(36, 28, 15)
    yield "own_440.jpg", encode_baseline(src, [(1, 2), (1, 1), (1, 1)])
    yield "own_411.jpg", encode_baseline(src, [(4, 1), (1, 1), (1, 1)])
    yield "own_mixed.jpg", encode_baseline(src, [(2, 2), (2, 1), (1, 1)])
    yield "own_noninterleaved_420.jpg", encode_baseline(src, [(2, 2), (1, 1), (1, 1)],
                                                        interleaved=False)
    yield "own_qtable16.jpg", encode_baseline(src, [(2, 1), (1, 1), (1, 1)], quality_scale=40.0,
                                              qbits16=True)
    yield "own_restart_444.jpg", encode_baseline(src, [(1, 1), (1, 1), (1, 1)], restart=5)


def main():
    os.makedirs(OUT, exist_ok=True)
    expected = {}
    for name, data in fixtures():
        with open(os.path.join(OUT, name), "wb") as fh:
            fh.write(data)
        im = Image.open(io.BytesIO(data))
        im.load()
        arr = np.asarray(im)
        if arr.ndim == 2:
            arr = arr[:, :, None]
        expected[name] = arr
        print(f"{name}: {im.mode} {im.size} {len(data)} bytes")
    np.savez_compressed(os.path.join(OUT, "expected.npz"), **expected)


if __name__ == "__main__":
    main()
