"""JPEG textures (host/jpeg.cpp) through the texture pipeline of
scene/base.cpp:98-110.  The reference's CImg hands a .jpg to ImageMagick,
whose JPEG coder decodes with libjpeg's default settings; each fixture of
tests/golden/jpeg (written by make_jpeg_fixtures.py, expected samples from
libjpeg-turbo's default decode through PIL) must decode to exactly those
samples, and its 512x512 layer must be CImg's normalize(0,255) and nearest
resize of them.  The reference's own JPEG textures (the sponza scene's
map_Kd files, baseline 4:2:0 and progressive 4:4:4) are checked against the
same decode when the reference tree is present."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

from conftest import GOLDEN
from hptracing_amd import _native as N
from hptracing_amd.scene import read_image
from textures import cimg_layer, load_textured

JPEG_DIR = os.path.join(GOLDEN, "jpeg")
EXPECTED = np.load(os.path.join(JPEG_DIR, "expected.npz"))
SPONZA = "/root/reference/scene/dabrovic-sponza"


@pytest.mark.parametrize("name", sorted(EXPECTED.files))
def test_jpeg_fixture_samples(name):
    got = read_image(os.path.join(JPEG_DIR, name))
    want = EXPECTED[name]
    assert got.shape == want.shape
    assert np.array_equal(got, want.astype(np.float32))


@pytest.mark.parametrize("name", ["pil_420.jpg", "pil_progressive_444.jpg", "pil_gray.jpg",
                                  "own_mixed.jpg", "pil_adobe_rgb.jpg"])
def test_jpeg_fixture_through_the_texture_pipeline(tmp_path, name):
    kd = load_textured(tmp_path, name, os.path.join(JPEG_DIR, name))
    assert kd.n_textures == 1
    layer = kd.texture_data.reshape(512, 512, 4)
    assert np.array_equal(layer[:, :, :3], cimg_layer(EXPECTED[name]))
    assert (layer[:, :, 3] == 0).all()


def _sponza_textures():
    if not os.path.isdir(SPONZA):
        return []
    names = set()
    for line in open(os.path.join(SPONZA, "sponza.mtl"), errors="replace"):
        m = re.match(r"\s*map_Kd\s+(\S+)", line)
        if m:
            names.add(m.group(1))
    files = {f.lower(): f for f in os.listdir(SPONZA)}
    return sorted(files[n.lower()] for n in names if n.lower() in files)


@pytest.mark.skipif(not os.path.isdir(SPONZA), reason="reference scenes absent")
@pytest.mark.parametrize("name", _sponza_textures())
def test_reference_jpeg_textures(name):
    """The sponza scene's diffuse textures (its MTL names them in upper case,
    the files are mixed case; the scene ships without its OBJ), sample for
    sample against libjpeg's default decode."""
    PIL = pytest.importorskip("PIL.Image")
    path = os.path.join(SPONZA, name)
    with PIL.open(path) as im:
        want = np.asarray(im)
    want = want[:, :, None] if want.ndim == 2 else want
    got = read_image(path)
    assert got.shape == want.shape
    assert np.array_equal(got, want.astype(np.float32))


def _edited(tmp_path, edit, src="pil_420.jpg"):
    data = bytearray(open(os.path.join(JPEG_DIR, src), "rb").read())
    edit(data)
    p = tmp_path / "x.jpg"
    p.write_bytes(bytes(data))
    return p


def _sof(data):
    for i in range(2, len(data) - 1):
        if data[i] == 0xFF and data[i + 1] in (0xC0, 0xC1, 0xC2):
            return i
    raise AssertionError("no frame header")


@pytest.mark.parametrize("case,edit,message", [
    ("arithmetic", lambda d: d.__setitem__(_sof(d) + 1, 0xC9), "arithmetic"),
    ("twelve_bit", lambda d: d.__setitem__(_sof(d) + 4, 12), "8-bit"),
    ("no_scan", lambda d: d.__delitem__(slice(_sof(d) + 20, None)), "jpeg"),
])
def test_jpeg_errors(tmp_path, case, edit, message):
    p = _edited(tmp_path, edit)
    with pytest.raises(N.HptError) as e:
        read_image(str(p))
    assert e.value.code == N.HPT_ERR_IO
    assert message in str(e.value)


def test_cmyk_jpeg_is_a_clear_error(tmp_path):
    PIL = pytest.importorskip("PIL.Image")
    p = tmp_path / "cmyk.jpg"
    PIL.fromarray(np.full((8, 8, 4), 100, np.uint8), mode="CMYK").save(p, "JPEG")
    with pytest.raises(N.HptError) as e:
        read_image(str(p))
    assert "CMYK" in str(e.value)


def test_truncated_jpeg_keeps_what_was_decoded(tmp_path):
    """A file cut inside its image data decodes as libjpeg decodes it (the
    missing data read as zero bits, no EOI needed): same size, and the rows
    decoded before the cut equal the full file's."""
    full = read_image(os.path.join(JPEG_DIR, "pil_restart.jpg"))
    data = open(os.path.join(JPEG_DIR, "pil_restart.jpg"), "rb").read()
    p = tmp_path / "cut.jpg"
    p.write_bytes(data[: len(data) - 400])
    cut = read_image(str(p))
    assert cut.shape == full.shape
    assert np.array_equal(cut[:8], full[:8])
