"""Image writers (raygun_amd/imageio.py): PNG decoded by an independent
reader (Pillow, when installed), PPM and PFM parsed back, PFM rows stored
bottom to top as the format defines. CPU only."""
import numpy as np
import pytest

from raygun_amd import imageio


def _img(h, w, ch, seed):
    return np.random.default_rng(seed).integers(0, 256, size=(h, w, ch), dtype=np.uint8)


@pytest.mark.parametrize("ch", [3, 4])
def test_png_decodes_to_the_same_pixels(tmp_path, ch):
    Image = pytest.importorskip("PIL.Image")
    a = _img(37, 53, ch, seed=ch)
    p = tmp_path / "x.png"
    imageio.write_png(p, a)
    with Image.open(p) as im:
        assert im.mode == ("RGBA" if ch == 4 else "RGB")
        b = np.asarray(im)
    assert np.array_equal(a, b)


def test_ppm_header_and_pixels(tmp_path):
    a = _img(5, 7, 4, seed=1)
    p = tmp_path / "x.ppm"
    imageio.write_ppm(p, a)
    data = p.read_bytes()
    assert data.startswith(b"P6\n7 5\n255\n")
    px = np.frombuffer(data[len(b"P6\n7 5\n255\n"):], np.uint8).reshape(5, 7, 3)
    assert np.array_equal(px, a[:, :, :3])


def test_pfm_round_trip_and_row_order(tmp_path):
    a = np.random.default_rng(2).normal(size=(4, 6, 4)).astype(np.float32)
    p = tmp_path / "x.pfm"
    imageio.write_pfm(p, a)
    data = p.read_bytes()
    header = b"PF\n6 4\n-1.0\n"
    assert data.startswith(header)
    first_row = np.frombuffer(data[len(header):len(header) + 6 * 3 * 4], "<f4").reshape(6, 3)
    assert np.array_equal(first_row, a[3, :, :3])  # bottom row first
    assert np.array_equal(imageio.read_pfm(p), a[:, :, :3])
    g = a[:, :, 0]
    imageio.write_pfm(p, g)
    assert np.array_equal(imageio.read_pfm(p), g)


def test_png_rejects_bad_shapes(tmp_path):
    with pytest.raises(ValueError):
        imageio.write_png(tmp_path / "x.png", np.zeros((4, 4), np.uint8))
