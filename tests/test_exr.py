"""EXR I/O (gonzales_amd/exr.py): round trips of our writer through our reader, and the
reference's regression image (tiled, ZIP-compressed float RGBA, tests/golden/regression-test.exr)
read back with the values its own tests hold (128 x 128, alpha 1, a Cornell image with the warm
light of radiance (17, 12, 4): red > green > blue on average)."""
from pathlib import Path

import numpy as np
import pytest

from gonzales_amd.exr import read_exr, read_rgb, write_exr

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("compression", ["none", "zip"])
@pytest.mark.parametrize("shape", [(1, 1), (16, 16), (37, 53), (100, 3)])
def test_round_trip(tmp_path, compression, shape):
    g = np.random.default_rng(shape[0] * 7 + shape[1])
    img = (g.random(shape + (3,)) * 10 - 1).astype(np.float32)
    img[0, 0] = [np.inf, -0.0, 1e-30]
    alpha = g.random(shape).astype(np.float32)
    write_exr(str(tmp_path / "a.exr"), img, alpha, compression=compression)
    c = read_exr(str(tmp_path / "a.exr"))
    assert np.array_equal(np.stack([c["R"], c["G"], c["B"]], -1).view(np.uint32), img.view(np.uint32))
    assert np.array_equal(c["A"], alpha)
    data = (tmp_path / "a.exr").read_bytes()
    assert data[:4] == bytes([0x76, 0x2F, 0x31, 0x01])


def test_zip_compresses_smooth_images(tmp_path):
    img = np.tile(np.linspace(0, 1, 64, dtype=np.float32)[None, :, None], (64, 1, 3))
    write_exr(str(tmp_path / "z.exr"), img, compression="zip")
    write_exr(str(tmp_path / "n.exr"), img, compression="none")
    assert (tmp_path / "z.exr").stat().st_size < 0.5 * (tmp_path / "n.exr").stat().st_size
    assert np.array_equal(read_rgb(str(tmp_path / "z.exr")), img)


def test_reference_regression_image_reads():
    c = read_exr(str(GOLDEN / "regression-test.exr"))
    assert sorted(c) == ["A", "B", "G", "R"]
    assert all(v.shape == (128, 128) for v in c.values())
    assert np.all(c["A"] == 1.0)
    means = [float(c[k].mean()) for k in "RGB"]
    assert means[0] > means[1] > means[2] > 0
    assert all(np.isfinite(c[k]).all() and (c[k] >= 0).all() for k in "RGB")


def test_bad_files_are_rejected(tmp_path):
    (tmp_path / "x.exr").write_bytes(b"not an exr file at all")
    with pytest.raises(ValueError, match="OpenEXR"):
        read_exr(str(tmp_path / "x.exr"))
