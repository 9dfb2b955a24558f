"""Malformed texture files are refused with an error, never a crash: seeded
mutations of the PNG and JPEG fixtures read through hpt_host_read_image.
The PNG mutations keep every chunk CRC and the zlib Adler-32 valid, so they
reach the header checks, inflate, the filters, interlacing and palettes;
JPEG has no checksums, so plain byte edits reach every marker and the
entropy decoder.  The files are decoded in a child process, so that a crash
shows up as a failed test naming the seed."""
from __future__ import annotations

import glob
import os
import random
import struct
import subprocess
import sys
import zlib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

DECODE_ALL = r"""
import sys
sys.path.insert(0, sys.argv[1])
from hptracing_amd import _native
from hptracing_amd.scene import read_image
ok = bad = 0
for line in open(sys.argv[2]):
    path = line.strip()
    print(path, flush=True)            # the last path printed names a crash
    try:
        img = read_image(path)
        assert img.ndim == 3 and img.size > 0
        ok += 1
    except _native.HptError:
        bad += 1
print(f"decoded {ok} refused {bad}", flush=True)
"""


def _edit(d, rnd, k):
    d = bytearray(d)
    for _ in range(k):
        if not d:
            break
        a = rnd.randrange(len(d))
        op = rnd.randrange(5)
        if op == 0:
            d[a] ^= 1 << rnd.randrange(8)
        elif op == 1:
            d[a] = rnd.randrange(256)
        elif op == 2:
            del d[a:a + rnd.randint(1, 8)]
        elif op == 3:
            d[a:a] = bytes(rnd.randrange(256) for _ in range(rnd.randint(1, 8)))
        else:
            d[a] = rnd.choice([0, 255, 127, 128])
    return d


def _chunks(b):
    i, out = 8, []
    while i + 8 <= len(b):
        ln, = struct.unpack(">I", b[i:i + 4])
        out.append([b[i + 4:i + 8], bytearray(b[i + 8:i + 8 + ln])])
        i += 12 + ln
    return out


def _png(chunks):
    o = bytearray(b"\x89PNG\r\n\x1a\n")
    for t, d in chunks:
        o += struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + bytes(d)))
    return bytes(o)


def _mutate_png(src, rnd):
    cs = _chunks(src)
    raw = zlib.decompress(b"".join(bytes(c[1]) for c in cs if c[0] == b"IDAT"))
    which = rnd.randrange(4)
    if which == 0:                                   # header fields
        for c in cs:
            if c[0] == b"IHDR":
                c[1] = _edit(c[1], rnd, 1)
    elif which == 1:                                 # filter bytes and samples
        cs = [c for c in cs if c[0] != b"IDAT"]
        cs.insert(-1, [b"IDAT", bytearray(zlib.compress(bytes(_edit(raw, rnd, 4))))])
    elif which == 2:                                 # the deflate stream, Adler-32 kept
        z = zlib.compress(raw, rnd.choice([0, 1, 9]))
        body = _edit(z[2:-4], rnd, rnd.randint(1, 3))
        cs = [c for c in cs if c[0] != b"IDAT"]
        cs.insert(-1, [b"IDAT", bytearray(z[:2] + body + struct.pack(">I", zlib.adler32(raw)))])
    else:                                            # palette and ancillary chunks
        for c in cs:
            if c[0] in (b"PLTE", b"tRNS"):
                c[1] = _edit(c[1], rnd, 2)
        if rnd.random() < 0.3:
            cs = [c for c in cs if c[0] != b"PLTE"]
    return _png(cs)


def _run(tmp_path, files):
    lst = tmp_path / "files.txt"
    lst.write_text("\n".join(files))
    p = subprocess.run([sys.executable, "-c", DECODE_ALL, ROOT, str(lst)],
                       capture_output=True, text=True, timeout=600)
    last = p.stdout.strip().splitlines()[-1] if p.stdout.strip() else ""
    assert p.returncode == 0, f"decoder crashed (rc {p.returncode}) on {last}: {p.stderr[-500:]}"
    assert last.startswith("decoded"), last
    ok, bad = int(last.split()[1]), int(last.split()[3])
    assert ok + bad == len(files)
    return ok, bad


def test_mutated_pngs_are_decoded_or_refused(tmp_path):
    rnd = random.Random(2026)
    seeds = [open(p, "rb").read() for p in sorted(glob.glob(os.path.join(GOLDEN, "png", "*.png")))]
    files = []
    for i in range(600):
        f = tmp_path / f"m{i:04d}.png"
        f.write_bytes(_mutate_png(rnd.choice(seeds), rnd))
        files.append(str(f))
    ok, bad = _run(tmp_path, files)
    assert ok > 50 and bad > 50        # both outcomes are exercised


def test_mutated_jpegs_are_decoded_or_refused(tmp_path):
    rnd = random.Random(2027)
    seeds = [open(p, "rb").read() for p in sorted(glob.glob(os.path.join(GOLDEN, "jpeg", "*.jpg")))]
    files = []
    for i in range(600):
        f = tmp_path / f"m{i:04d}.jpg"
        f.write_bytes(bytes(_edit(rnd.choice(seeds), rnd, rnd.randint(1, 8))))
        files.append(str(f))
    ok, bad = _run(tmp_path, files)
    assert ok > 50 and bad > 50
