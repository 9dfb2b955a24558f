"""PNG textures (host/png.cpp) through the texture pipeline of
scene/base.cpp:98-110: each fixture of tests/golden/png (written by
make_png_fixtures.py; the reference's own default.png textures among them)
must decode to the samples CImg receives from ImageMagick in the reference
build -- gray as one channel, colour and palette as three, alpha dropped --
and the resulting 512x512 layer must be CImg's normalize(0,255) and nearest
resize of them."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN
from hptracing_amd import _native as N
from hptracing_amd.scene import KDTree, read_image
from textures import cimg_layer, load_textured

PNG_DIR = os.path.join(GOLDEN, "png")
EXPECTED = np.load(os.path.join(PNG_DIR, "expected.npz"))
REFERENCE_SCENES = "/root/reference/scene"


@pytest.mark.parametrize("name", sorted(EXPECTED.files))
def test_png_fixture_through_the_texture_pipeline(tmp_path, name):
    kd = load_textured(tmp_path, name, os.path.join(PNG_DIR, name))
    assert kd.n_textures == 1
    layer = kd.texture_data.reshape(512, 512, 4)
    assert np.array_equal(layer[:, :, :3], cimg_layer(EXPECTED[name]))
    assert (layer[:, :, 3] == 0).all()


@pytest.mark.parametrize("name", sorted(EXPECTED.files))
def test_png_fixture_samples(name):
    """The whole decoded image, sample for sample (the 512x512 layer above
    samples only part of a larger image)."""
    got = read_image(os.path.join(PNG_DIR, name))
    want = EXPECTED[name]
    want = want[:, :, None] if want.ndim == 2 else want
    assert got.shape == want.shape
    assert np.array_equal(got, want.astype(np.float32))


def _corrupt(tmp_path, name, edit):
    data = bytearray(open(os.path.join(PNG_DIR, "pil_rgb.png"), "rb").read())
    edit(data)
    p = tmp_path / "src.bin"
    p.write_bytes(bytes(data))
    return p


@pytest.mark.parametrize("case,edit,message", [
    ("crc", lambda d: d.__setitem__(60, d[60] ^ 0x40), b"CRC"),
    ("truncated", lambda d: d.__delitem__(slice(len(d) // 2, None)), b"png"),
    ("signature", lambda d: d.__setitem__(1, ord("Q")), b"only PNG, JPEG and PNM"),
])
def test_png_errors(tmp_path, case, edit, message):
    src = _corrupt(tmp_path, "x.png", edit)
    with pytest.raises(N.HptError) as e:
        load_textured(tmp_path, "x.png", src)
    assert e.value.code == N.HPT_ERR_IO
    assert message in str(e.value).encode()


@pytest.mark.skipif(not os.path.isdir(REFERENCE_SCENES), reason="reference scenes absent")
@pytest.mark.parametrize("scene,obj,fixture", [
    ("cornell-box", "untitled.obj", "cornell_box_default.png"),
    ("vr-scene", "final.obj", "cornell_box_default.png"),
])
def test_reference_textured_scenes_load(scene, obj, fixture):
    """The reference's textured scenes load with their PNG texture, and the
    layer is the one the fixture of the same file gives.  (scene/mitsuba
    holds its RGBA texture but no OBJ; the fixture covers that texture.)"""
    d = os.path.join(REFERENCE_SCENES, scene)
    if not os.path.exists(os.path.join(d, obj)):
        pytest.skip(f"{scene}/{obj} absent")
    kd = KDTree(os.path.join(d, obj), d + "/")
    assert kd.n_textures >= 1
    tex = kd.texture_data.reshape(-1, 512, 512, 4)
    same = open(os.path.join(d, "default.png"), "rb").read() == \
        open(os.path.join(PNG_DIR, fixture), "rb").read()
    if same:
        assert np.array_equal(tex[0, :, :, :3], cimg_layer(EXPECTED[fixture]))
