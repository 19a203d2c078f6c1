"""Image files (csrc/image_encode.cpp), the save step of the reference's CLI
(backend/artefact-cli/main.rs:156): PNG, BMP, TIFF and lossless WebP written here, read back with
Pillow (libpng, its BMP reader, libtiff, libwebp) and compared byte for byte with the RGB input;
the CLI's output path rules (main.rs:57-98)."""
import ctypes
import io
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

import artefact_amd
from artefact_amd import image, synth


def images():
    rng = np.random.default_rng(3)
    out = [rng.integers(0, 256, (h, w, 3), dtype=np.uint8) for h, w in ((1, 1), (5, 3), (9, 17), (33, 64))]
    out.append(np.stack(synth.synth_rgb(96, 130, 4), -1))  # smooth areas, edges and noise
    out.append(np.zeros((7, 11, 3), np.uint8))
    out.append(np.full((7, 11, 3), 255, np.uint8))
    return out


@pytest.mark.parametrize("fmt", ["png", "bmp", "tiff", "webp"])
def test_round_trip_through_pillow(fmt):
    for rgb in images():
        data = artefact_amd.encode_image(rgb, fmt)
        with Image.open(io.BytesIO(data)) as im:
            assert im.format == {"png": "PNG", "bmp": "BMP", "tiff": "TIFF", "webp": "WEBP"}[fmt]
            assert im.mode == "RGB" and im.size == (rgb.shape[1], rgb.shape[0])
            back = np.asarray(im.convert("RGB"))
        assert np.array_equal(back, rgb), (fmt, rgb.shape)


def png_chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, out = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        out.append((kind, body))
        pos += 12 + n
    return out


def unfilter(raw, w, h):
    """PNG scanline filters undone in numpy (ISO/IEC 15948 section 9), independent of the encoder."""
    stride = 3 * w
    out = np.zeros((h, stride), np.int64)
    for y in range(h):
        t = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        prev = out[y - 1] if y else np.zeros(stride, np.int64)
        cur = np.zeros(stride, np.int64)
        for i in range(stride):
            a = cur[i - 3] if i >= 3 else 0
            b = prev[i]
            c = prev[i - 3] if i >= 3 else 0
            p = a + b - c
            pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
            pred = [0, a, b, (a + b) // 2, a if pa <= pb and pa <= pc else (b if pb <= pc else c)][t]
            cur[i] = (line[i] + pred) & 255
        out[y] = cur
    return out.astype(np.uint8).reshape(h, w, 3)


def test_png_structure_and_filters():
    """Chunk CRCs, IHDR (8-bit RGB, no interlace), one zlib stream over the IDAT chunks, every
    filter type used somewhere on a mixed image, and the unfiltered rows equal the input."""
    rgb = np.stack(synth.synth_rgb(64, 80, 9), -1)
    rgb[10:20] = 0
    data = artefact_amd.encode_image(rgb, "png")
    chunks = png_chunks(data)
    kinds = [k for k, _ in chunks]
    assert kinds[0] == b"IHDR" and kinds[-1] == b"IEND" and set(kinds[1:-1]) == {b"IDAT"}
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (w, h, depth, ctype, comp, filt, inter) == (80, 64, 8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for k, b in chunks if k == b"IDAT"))
    assert len(raw) == h * (1 + 3 * w)
    types = {raw[y * (1 + 3 * w)] for y in range(h)}
    assert types <= {0, 1, 2, 3, 4} and len(types) >= 2
    assert np.array_equal(unfilter(raw, w, h), rgb)


def test_bmp_layout():
    rgb = np.random.default_rng(1).integers(0, 256, (3, 5, 3), dtype=np.uint8)
    data = artefact_amd.encode_image(rgb, "bmp")
    assert data[:2] == b"BM"
    size, _, _, off = struct.unpack("<IHHI", data[2:14])
    assert size == len(data) and off == 54
    hdr = struct.unpack("<IiiHHIIiiII", data[14:54])
    assert hdr[:5] == (40, 5, 3, 1, 24)
    row = (3 * 5 + 3) & ~3
    bottom = np.frombuffer(data, np.uint8, 15, 54).reshape(5, 3)
    assert np.array_equal(bottom[:, ::-1], rgb[2])  # first stored row is the last image row, BGR
    assert len(data) == 54 + row * 3


def webp_images():
    """Images that take every prefix-code form of the WebP writer: one and two symbols per
    channel (simple codes), every byte value in a channel (a code-length code of one length
    and a run of zeros), skewed histograms (long codes), and a size at the format's limit."""
    rng = np.random.default_rng(11)
    out = [np.full((3, 4, 3), (7, 200, 7), np.uint8)]
    two = np.zeros((5, 9, 3), np.uint8)
    two[::2] = (1, 0, 255)
    out.append(two)
    full = np.arange(256 * 3, dtype=np.uint32).reshape(16, 16, 3).astype(np.uint8)
    out.append(full)
    # red and blue take every value exactly once (with green 0): all 256 code lengths are 8,
    # so the code-length sequence is one token repeated
    flat = np.zeros((16, 16, 3), np.uint8)
    flat[..., 0] = np.arange(256, dtype=np.uint8).reshape(16, 16)
    flat[..., 2] = 255 - flat[..., 0]
    out.append(flat)
    skew = rng.geometric(0.35, (64, 200, 3)).clip(0, 255).astype(np.uint8)
    out.append(skew)
    out.append(rng.integers(0, 256, (2, 16384, 3), dtype=np.uint8))
    return out


def test_webp_is_lossless_vp8l_and_every_code_form_decodes():
    for rgb in webp_images():
        data = artefact_amd.encode_image(rgb, "webp")
        assert data[:4] == b"RIFF" and data[8:16] == b"WEBPVP8L"
        riff, = struct.unpack("<I", data[4:8])
        chunk, = struct.unpack("<I", data[16:20])
        assert riff == len(data) - 8 and chunk + (chunk & 1) == len(data) - 20
        assert data[20] == 0x2F
        bits = int.from_bytes(data[21:25], "little")
        assert (bits & 0x3FFF) + 1 == rgb.shape[1] and ((bits >> 14) & 0x3FFF) + 1 == rgb.shape[0]
        with Image.open(io.BytesIO(data)) as im:
            assert im.format == "WEBP" and im.size == (rgb.shape[1], rgb.shape[0])
            back = np.asarray(im.convert("RGB"))
        assert np.array_equal(back, rgb), rgb.shape
    lib = artefact_amd._lib.load()
    assert lib.artefact_image_encode_bound(16385, 2, artefact_amd._lib.ARTEFACT_IMAGE_WEBP) == 0


def test_bound_and_errors():
    lib = artefact_amd._lib.load()
    assert lib.artefact_image_encode_bound(0, 5, 0) == 0
    assert lib.artefact_image_encode_bound(5, 5, 9) == 0
    rgb = np.zeros((4, 4, 3), np.uint8)
    small = np.empty(10, np.uint8)
    n = ctypes.c_size_t()
    U8 = artefact_amd._lib.U8_P
    for fmt in (0, 1, 2, 3):
        rc = lib.artefact_image_encode(rgb.ctypes.data_as(U8), 4, 4, fmt, small.ctypes.data_as(U8), small.size,
                                       ctypes.byref(n))
        assert rc == -1 and n.value == 0
    with pytest.raises(artefact_amd.ImageFormatError):
        artefact_amd.encode_image(rgb, "gif")


@pytest.mark.parametrize("inp,out,fmt,want", [
    ("in.jpg", None, "auto", "in.png"),
    ("in.jpg", "out.bmp", "auto", "out.bmp"),
    ("in.jpg", "out", "auto", "out.png"),
    ("in.jpg", "out", "tiff", "out.tiff"),
    ("dir/in.jpeg", None, "bmp", "dir/in.bmp"),
    ("in.jpg", "out.tiff", "tiff", "out.tiff"),
])
def test_cli_output_path(inp, out, fmt, want):
    assert image.output_path(inp, out, fmt) == want


@pytest.mark.parametrize("out,fmt", [("out.png", "bmp"), (None, "jpg")])
def test_cli_output_path_errors(out, fmt):
    with pytest.raises(artefact_amd.ImageFormatError):
        image.output_path("in.jpg", out, fmt)


def test_save_picks_the_format_from_the_extension(tmp_path):
    rgb = np.stack(synth.synth_rgb(16, 24, 2), -1)
    for name, kind in (("a.png", "PNG"), ("a.BMP", "BMP"), ("a.tif", "TIFF"), ("a.tiff", "TIFF"), ("a.webp", "WEBP")):
        p = tmp_path / name
        artefact_amd.save_image(rgb, str(p))
        with Image.open(p) as im:
            assert im.format == kind
            assert np.array_equal(np.asarray(im.convert("RGB")), rgb)
    with pytest.raises(artefact_amd.ImageFormatError):
        artefact_amd.save_image(rgb, str(tmp_path / "a.jpg"))
