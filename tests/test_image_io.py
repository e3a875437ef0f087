"""Frame ingest (the reference's image loaders, loaders/*.cpp): PNM P5/P6 and
TIFF (mono/RGB, 8/16 bits, uncompressed, Deflate, LZW, PackBits, predictor)
read by the engine's own decoder into greyscale pixels of their own depth,
errors with the reference's texts.  Host destinations: no GPU needed.

Pinned by: the reference's own loader (PNM goldens written through
oracle/_ref/ref_piv, tests/golden/make_loader_golden.py); the reference's
TIFF fixtures (tests/corr_a.tiff = test-rgb.tiff, test-mono.tiff,
examples/process/exp1_001_b.tiff) at the 511 x 369 the reference's loader
test asserts (image_loader_test.cpp:21-48), their pixels decoded here
independently with Python's zlib; TIFFs written by PIL's libtiff in every
codec, decoded back by PIL.  RGB -> grey is the integer rule of
pixel_types.h:397-400."""
import struct
import zlib

import numpy as np
import pytest

import openpiv_amd as oa
from conftest import GOLDEN, load_npz
from dropin_util import write_pgm, write_tiff


def grey(rgb: np.ndarray) -> np.ndarray:
    """(218 r + 732 g + 74 b) >> 10 on integral samples (pixel_types.h:397-400)"""
    c = rgb.astype(np.uint64)
    return (c[..., 0] * 218 + c[..., 1] * 732 + c[..., 2] * 74) >> 10


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


@pytest.mark.parametrize("case", ["p5_u8", "p5_u16", "p6_u8", "p6_u16"])
def test_pnm_against_reference_loader(tmp_path, case):
    """pixels as the reference's PNM loader delivers them to a gf_image,
    including its signed-char reading of 8-bit P6 samples"""
    g = load_npz("loader_cases.npz")
    f = tmp_path / "x.pnm"
    f.write_bytes(g[f"{case}_file"].tobytes())
    got = oa.read_image(f)
    ref = g[f"{case}_ref"]
    assert got.shape == ref.shape
    assert np.array_equal(got.astype(np.float64), ref)
    # 8-bit P6 comes out 16-bit: bytes >= 128 leave the 8-bit range
    assert got.dtype == (np.uint8 if case == "p5_u8" else np.uint16)


def test_p6_rule_closed_form(tmp_path):
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 128, (6, 9, 3)).astype(np.uint8)   # below 128: no sign extension
    (tmp_path / "c.ppm").write_bytes(b"P6\n9 6\n255\n" + rgb.tobytes())
    assert np.array_equal(oa.read_image(tmp_path / "c.ppm"), grey(rgb))


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


def tiff_tags(path):
    """the first IFD of a little-endian TIFF as {tag: values}"""
    d = path.read_bytes()
    assert d[:4] == b"II*\x00"
    ifd = struct.unpack("<I", d[4:8])[0]
    n = struct.unpack("<H", d[ifd:ifd + 2])[0]
    tags = {}
    for k in range(n):
        t, ty, c, v = struct.unpack("<HHII", d[ifd + 2 + 12 * k:ifd + 14 + 12 * k])
        size = {1: 1, 3: 2, 4: 4}.get(ty, 0)
        if size * c <= 4:
            raw = d[ifd + 10 + 12 * k:ifd + 10 + 12 * k + size * c]
        else:
            raw = d[v:v + size * c]
        fmt = {1: "B", 3: "H", 4: "I"}.get(ty)
        tags[t] = list(struct.unpack("<" + fmt * c, raw)) if fmt else None
    return d, tags


@pytest.mark.parametrize("name", ["corr_a.tiff", "exp1_001_b.tiff"])
def test_reference_rgb_deflate_fixture(name):
    """the reference's RGB, Deflate-compressed fixtures (corr_a.tiff is
    also tests/test-rgb.tiff and examples/process/exp1_001_a.tiff)"""
    path = GOLDEN / name
    assert oa.image_info(path) == (511, 369, np.dtype(np.uint8), 1)   # image_loader_test.cpp:36-48
    d, tags = tiff_tags(path)
    assert tags[259] == [8] and tags[277] == [3] and tags[258] == [8, 8, 8]
    raw = b"".join(zlib.decompress(d[o:o + c]) for o, c in zip(tags[273], tags[279]))
    rgb = np.frombuffer(raw, np.uint8)[:369 * 511 * 3].reshape(369, 511, 3)
    assert np.array_equal(oa.read_image(path), grey(rgb).astype(np.uint8))


def test_reference_mono_fixture():
    """tests/test-mono.tiff: uncompressed 8-bit, 16-row strips"""
    path = GOLDEN / "test-mono.tiff"
    assert oa.image_info(path) == (511, 369, np.dtype(np.uint8), 1)   # image_loader_test.cpp:21-34
    d, tags = tiff_tags(path)
    raw = b"".join(d[o:o + c] for o, c in zip(tags[273], tags[279]))
    assert np.array_equal(oa.read_image(path), np.frombuffer(raw, np.uint8)[:369 * 511].reshape(369, 511))


PIL_CODECS = [None, "tiff_lzw", "packbits", "tiff_adobe_deflate", "tiff_deflate"]


@pytest.mark.parametrize("codec", PIL_CODECS)
@pytest.mark.parametrize("mode", ["L", "I;16", "RGB"])
@pytest.mark.parametrize("predictor", [False, True])
def test_pil_written_tiffs(tmp_path, codec, mode, predictor):
    """every codec through PIL's libtiff writer, checked against PIL's own
    reading of the file (the predictor tag is honoured by libtiff for LZW and
    Deflate only, and so here)"""
    Image = pytest.importorskip("PIL.Image")
    if predictor and codec in (None, "packbits"):
        pytest.skip("libtiff defines the predictor for LZW and Deflate only (PIL treats the tag its own way)")
    rng = np.random.default_rng(7)
    shape = (53, 67)
    if mode == "L":
        arr = (rng.integers(0, 40, shape).cumsum(1) % 256).astype(np.uint8)
    elif mode == "I;16":
        arr = (rng.integers(0, 9000, shape).cumsum(1) % 65536).astype(np.uint16)
    else:
        arr = rng.integers(0, 256, shape + (3,)).astype(np.uint8)
    kw = {"tiffinfo": {317: 2}} if predictor else {}
    path = tmp_path / "p.tif"
    Image.fromarray(arr).save(path, format="TIFF", compression=codec, **kw)
    back = np.array(Image.open(path))
    assert np.array_equal(back, arr)
    want = grey(back).astype(np.uint8) if mode == "RGB" else back
    assert oa.image_info(path)[:3] == (67, 53, want.dtype)
    assert np.array_equal(oa.read_image(path), want)


def write_rgb16_tiff(path, rgb: np.ndarray, deflate: bool, big_endian: bool):
    """16-bit RGB, one strip, uncompressed or zlib (Deflate, compression 8)"""
    e = ">" if big_endian else "<"
    h, w, _ = rgb.shape
    data = rgb.astype(e + "u2").tobytes()
    if deflate:
        data = zlib.compress(data)
    head = (b"MM\x00\x2a" if big_endian else b"II\x2a\x00") + struct.pack(e + "I", 8)
    bps_at = 8 + 2 + 9 * 12 + 4
    strip_at = bps_at + 6
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 3, bps_at), (259, 3, 1, 8 if deflate else 1),
               (262, 3, 1, 2), (273, 4, 1, strip_at), (277, 3, 1, 3), (278, 4, 1, h), (279, 4, 1, len(data))]
    ifd = struct.pack(e + "H", len(entries))
    for t, ty, c, v in entries:
        if ty == 3 and c == 1:
            ifd += struct.pack(e + "HHIHH", t, ty, c, v, 0)
        else:
            ifd += struct.pack(e + "HHII", t, ty, c, v)
    ifd += struct.pack(e + "I", 0)
    path.write_bytes(head + ifd + struct.pack(e + "HHH", 16, 16, 16) + data)


@pytest.mark.parametrize("deflate", [False, True])
@pytest.mark.parametrize("big_endian", [False, True])
def test_rgb16_tiff(tmp_path, deflate, big_endian):
    rng = np.random.default_rng(9)
    rgb = rng.integers(0, 65536, (21, 34, 3)).astype(np.uint16)
    write_rgb16_tiff(tmp_path / "c.tif", rgb, deflate, big_endian)
    assert oa.image_info(tmp_path / "c.tif") == (34, 21, np.dtype(np.uint16), 1)
    assert np.array_equal(oa.read_image(tmp_path / "c.tif"), grey(rgb).astype(np.uint16))


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


def test_crafted_sizes_are_rejected(tmp_path):
    """header sizes whose products wrap 64 bits, or exceed the file, fail
    cleanly instead of sizing a buffer from a wrapped product"""
    h = 2 ** 31 + 2 ** 15
    w = (2 ** 63 + 2 ** 15) // h
    for kind in (5, 6):
        (tmp_path / "big.pnm").write_bytes(f"P{kind}\n{w} {h}\n65535\n".encode() + bytes(64))
        with pytest.raises(oa.OpivError) as e:
            oa.read_image(tmp_path / "big.pnm")
        assert e.value.code == 1
    (tmp_path / "big2.pgm").write_bytes(b"P5\n4294967295 4294967295\n255\n" + bytes(16))
    with pytest.raises(oa.OpivError):
        oa.read_image(tmp_path / "big2.pgm")
    # TIFF: 65535 x 4294967295 16-bit pixels in a 100-byte file
    write_tiff(tmp_path / "t.tif", [np.ones((8, 8), np.uint16)])
    d = bytearray((tmp_path / "t.tif").read_bytes())
    ifd = struct.unpack("<I", d[4:8])[0]
    n = struct.unpack("<H", d[ifd:ifd + 2])[0]
    for k in range(n):
        e0 = ifd + 2 + 12 * k
        tag = struct.unpack("<H", d[e0:e0 + 2])[0]
        if tag in (256, 257):
            d[e0 + 2:e0 + 12] = struct.pack("<HII", 4, 1, 0xffffffff if tag == 257 else 65535)
    (tmp_path / "t4.tif").write_bytes(bytes(d))
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "t4.tif")
    assert e.value.code in (1, 6)


def test_unsupported_tiffs(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    rgba = np.zeros((8, 8, 4), np.uint8)
    Image.fromarray(rgba).save(tmp_path / "a.tif", format="TIFF")
    with pytest.raises(oa.OpivError) as e:
        oa.read_image(tmp_path / "a.tif")
    assert "spp != 1 or 3" in str(e.value)
    Image.fromarray(np.zeros((64, 64), np.uint8)).save(tmp_path / "t.tif", format="TIFF",
                                                        tiffinfo={322: 16, 323: 16})
    _, tags = tiff_tags(tmp_path / "t.tif")
    if 322 in tags:
        with pytest.raises(oa.OpivError) as e:
            oa.read_image(tmp_path / "t.tif")
        assert "tiled" in str(e.value)


def test_corrupt_compressed_strip(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    arr = np.random.default_rng(3).integers(0, 256, (40, 40)).astype(np.uint8)
    for codec in ("tiff_lzw", "tiff_adobe_deflate", "packbits"):
        Image.fromarray(arr).save(tmp_path / "c.tif", format="TIFF", compression=codec)
        d, tags = tiff_tags(tmp_path / "c.tif")
        off, cnt = tags[273][0], tags[279][0]
        cut = bytearray(d)
        cut[off + cnt // 2:off + cnt] = bytes(cnt - cnt // 2)   # zero the back half of the strip
        (tmp_path / "x.tif").write_bytes(bytes(cut))
        try:
            got = oa.read_image(tmp_path / "x.tif")
        except oa.OpivError as e:
            assert "corrupt TIFF strip" in str(e)
        else:   # a stream that still decodes to the full strip is accepted
            assert got.shape == arr.shape
