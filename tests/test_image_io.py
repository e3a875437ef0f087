"""Frame ingest (the reference's image loaders, loaders/*.cpp): PNM P5 and
baseline TIFF read by the engine's own loaders, pixels exactly as written,
errors with the reference's texts.  Host destinations: no GPU needed."""
import numpy as np
import pytest

import openpiv_amd as oa
from conftest import load_npz
from dropin_util import write_pgm, write_tiff


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("comment", [False, True])
def test_pnm_roundtrip(tmp_path, dtype, comment):
    rng = np.random.default_rng(1)
    a = rng.integers(0, np.iinfo(dtype).max + 1, (37, 53)).astype(dtype)
    write_pgm(tmp_path / "a.pgm", a, comment=comment)
    assert oa.image_info(tmp_path / "a.pgm") == (53, 37, np.dtype(dtype), 1)
    assert np.array_equal(oa.read_image(tmp_path / "a.pgm"), a)


def test_pnm_reference_example_crop(tmp_path):
    """the 512^2 crop of the reference's examples/process/piv1.pgm"""
    g = load_npz("piv_crop512_u8.npz")
    write_pgm(tmp_path / "piv.pgm", g["a"])
    assert np.array_equal(oa.read_image(tmp_path / "piv.pgm"), g["a"])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("big_endian", [False, True])
@pytest.mark.parametrize("rps", [0, 7])
def test_tiff_pages(tmp_path, dtype, big_endian, rps):
    rng = np.random.default_rng(2)
    pages = [rng.integers(0, np.iinfo(dtype).max + 1, (45, 61)).astype(dtype) for _ in range(3)]
    write_tiff(tmp_path / "a.tif", pages, big_endian=big_endian, rows_per_strip=rps)
    assert oa.image_info(tmp_path / "a.tif") == (61, 45, np.dtype(dtype), 3)
    for i, p in enumerate(pages):
        assert np.array_equal(oa.read_image(tmp_path / "a.tif", i), p)


def test_loader_errors(tmp_path):
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "missing.pgm")
    assert "failed to open" in str(e.value)
    (tmp_path / "x.bin").write_bytes(b"not an image at all")
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "x.bin")
    assert "failed to find loader" in str(e.value)
    write_tiff(tmp_path / "one.tif", [np.zeros((4, 4), np.uint8)])
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "one.tif", 1)
    assert "image index out of range" in str(e.value)
    (tmp_path / "c.ppm").write_bytes(b"P6\n2 2\n255\n" + bytes(12))
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "c.ppm")
    assert "colour" in str(e.value)
    (tmp_path / "short.pgm").write_bytes(b"P5\n10 10\n255\n" + bytes(20))
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "short.pgm")
    assert "short pixel data" in str(e.value)
    # a TIFF cut after its header: the directory lies past the end
    write_tiff(tmp_path / "t.tif", [np.ones((8, 8), np.uint16)])
    (tmp_path / "t3.tif").write_bytes((tmp_path / "t.tif").read_bytes()[:12])
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "t3.tif")
    assert "TIFF" in str(e.value)
