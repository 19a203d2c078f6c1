"""JPEG input stage (host C++ decoder in the library) against independent references.

* libjpeg through Pillow: luma (and 4:4:4 chroma) rebuilt from the decoded coefficients
  (dequantise, orthonormal IDCT in f64, +128, round) is within 1 of libjpeg's own decoded
  samples, for baseline and progressive files at several qualities and subsamplings, with
  and without restart markers; quantisation tables equal the IJG tables scaled to the
  quality, in natural order.
* progressive == baseline: libjpeg quantises once, so the progressive and the sequential
  file of the same image and quality hold identical coefficients.
* a T.81 writer (tests/jpeg_writer.py) over known coefficients: the decoder returns them
  exactly, including the fork's block placement for a (h=1, v=2) component.
No GPU is needed: decoding is host code.
"""
import io

import numpy as np
import pytest
from PIL import Image

import artefact_amd
from artefact_amd import synth
from jpeg_writer import ZIGZAG, ScanBits, write_baseline, write_grey_scans

D = synth.dct_matrix()


def rebuild(c):
    b = c.dct_coefs.reshape(c.block_h, c.block_w, 8, 8).astype(np.float64) * c.quant_table.reshape(8, 8)
    px = np.einsum("ux,abuv,vy->abxy", D, b, D) + 128.0
    return px.transpose(0, 2, 1, 3).reshape(c.rounded_px_h, c.rounded_px_w)


def pil_jpeg(img, **kw):
    f = io.BytesIO()
    Image.fromarray(img).save(f, "JPEG", **kw)
    return f.getvalue()


def pil_ycc(data):
    im = Image.open(io.BytesIO(data))
    im.draft("YCbCr" if im.mode != "L" else "L", im.size)
    im.load()
    a = np.asarray(im).astype(np.float64)
    return a[..., None] if a.ndim == 2 else a


CASES = [(sub, prog, q, rst) for sub in (0, 1, 2) for prog in (False, True) for q in (30, 75, 95)
         for rst in (None,)] + [(2, False, 75, 2), (0, False, 50, 5), (1, True, 75, 3)]


@pytest.mark.parametrize("sub,prog,quality,rst", CASES)
def test_matches_libjpeg_samples(sub, prog, quality, rst):
    h, w = 61, 93
    rgb = np.stack(synth.synth_rgb(h, w, quality + sub), -1)
    kw = dict(quality=quality, subsampling=sub, progressive=prog)
    if rst:
        kw["restart_marker_blocks"] = rst
    data = pil_jpeg(rgb, **kw)
    j = artefact_amd.decode_jpeg(data)
    assert (j.nchannel, j.real_px_w, j.real_px_h) == (3, w, h)
    hs, vs = {0: (1, 1), 1: (2, 1), 2: (2, 2)}[sub]
    assert [(c.horizontal_samp_factor, c.vertical_samp_factor) for c in j.coefs] == [(1, 1), (hs, vs), (hs, vs)]
    W = -(-w // (8 * hs)) * 8 * hs
    H = -(-h // (8 * vs)) * 8 * vs
    assert (j.coefs[0].rounded_px_w, j.coefs[0].rounded_px_h) == (W, H)
    assert (j.coefs[1].rounded_px_w, j.coefs[1].rounded_px_h) == (W // hs, H // vs)
    assert np.array_equal(j.coefs[0].quant_table, synth.quant_table(quality))
    assert np.array_equal(j.coefs[1].quant_table, synth.quant_table(quality, chroma=True))
    ycc = pil_ycc(data)
    planes = [0, 1, 2] if sub == 0 else [0]
    for p in planes:
        mine = np.clip(np.round(rebuild(j.coefs[p])[:h, :w]), 0, 255)
        assert np.abs(mine - ycc[..., p]).max() <= 1, f"plane {p}"


@pytest.mark.parametrize("sub", [0, 1, 2])
def test_progressive_equals_sequential(sub):
    rgb = np.stack(synth.synth_rgb(80, 72, 11), -1)
    a = artefact_amd.decode_jpeg(pil_jpeg(rgb, quality=80, subsampling=sub))
    b = artefact_amd.decode_jpeg(pil_jpeg(rgb, quality=80, subsampling=sub, progressive=True))
    for x, y in zip(a.coefs, b.coefs):
        assert np.array_equal(x.dct_coefs, y.dct_coefs)


@pytest.mark.parametrize("prog", [False, True])
def test_grey(prog):
    g = synth.synth_plane(37, 50, 4)
    data = pil_jpeg(g, quality=60, progressive=prog)
    j = artefact_amd.decode_jpeg(data)
    assert (j.nchannel, j.real_px_w, j.real_px_h) == (1, 50, 37)
    assert (j.coefs[0].rounded_px_w, j.coefs[0].rounded_px_h) == (56, 40)
    mine = np.clip(np.round(rebuild(j.coefs[0])[:37, :50]), 0, 255)
    assert np.abs(mine - pil_ycc(data)[..., 0]).max() <= 1


def _std_components(coefs, jpeg_hv):
    out = []
    for c, (h, v) in zip(coefs, jpeg_hv):
        out.append((h, v, c.quant_table, c.dct_coefs, c.block_w))
    return out


@pytest.mark.parametrize("sub,rst", [("444", 0), ("420", 0), ("422", 0), ("420", 3), ("444", 7)])
def test_writer_round_trip_exact(sub, rst):
    w, h = 88, 56
    coefs = synth.encode(synth.synth_rgb(h, w, 5), subsampling=sub, quality=70)
    hv = {"444": [(1, 1)] * 3, "420": [(2, 2), (1, 1), (1, 1)], "422": [(2, 1), (1, 1), (1, 1)]}[sub]
    data = write_baseline(w, h, _std_components(coefs, hv), restart_interval=rst)
    j = artefact_amd.decode_jpeg(data)
    for a, b in zip(j.coefs, coefs):
        assert (a.rounded_px_w, a.rounded_px_h, a.horizontal_samp_factor, a.vertical_samp_factor) == \
               (b.rounded_px_w, b.rounded_px_h, b.horizontal_samp_factor, b.vertical_samp_factor)
        assert np.array_equal(a.dct_coefs, b.dct_coefs)
        assert np.array_equal(a.quant_table, b.quant_table)


def test_one_by_two_component_keeps_the_fork_placement():
    """A (h=1, v=2) luma: the fork stores an MCU's two blocks at 2m and 2m+1 (mcu.rs:235-249)."""
    w, h = 40, 32
    coefs = synth.encode(synth.synth_rgb(h, w, 6), subsampling="440", quality=70)
    data = write_baseline(w, h, _std_components(coefs, [(1, 2), (1, 1), (1, 1)]))
    j = artefact_amd.decode_jpeg(data)
    y = coefs[0]
    mw, mh = y.block_w, y.block_h // 2
    std = y.dct_coefs.reshape(-1, 64)
    want = np.empty_like(std)
    for r in range(mh):
        for c in range(mw):
            m = r * mw + c
            want[2 * m] = std[2 * r * mw + c]
            want[2 * m + 1] = std[(2 * r + 1) * mw + c]
    assert np.array_equal(j.coefs[0].dct_coefs.reshape(-1, 64), want)
    for a, b in zip(j.coefs[1:], coefs[1:]):
        assert np.array_equal(a.dct_coefs, b.dct_coefs)


def test_truncated_entropy_data_keeps_zeros():
    rgb = np.stack(synth.synth_rgb(128, 128, 8), -1)
    data = pil_jpeg(rgb, quality=75, subsampling=2)
    full = artefact_amd.decode_jpeg(data)
    cut = artefact_amd.decode_jpeg(data[: len(data) * 2 // 3])
    # The first block rows agree; the missing tail decodes to zeros or stops early.
    assert np.array_equal(cut.coefs[0].dct_coefs[: 64 * 32], full.coefs[0].dct_coefs[: 64 * 32])
    assert not np.array_equal(cut.coefs[0].dct_coefs, full.coefs[0].dct_coefs)


@pytest.mark.parametrize("data", [b"", b"not a jpeg at all", b"\xff\xd8\xff\xd9", b"\xff\xd8\xff\xdb\x00"])
def test_malformed_data_raises(data):
    with pytest.raises(artefact_amd.JpegDecodeError):
        artefact_amd.decode_jpeg(data)


# The reference decoder's own invalid-input cases (zune-jpeg/tests/invalid_images.rs): each
# must be rejected.  The messages are this decoder's own, so only the rejection is compared.
@pytest.mark.parametrize("data", [
    bytes([0xFF, 0xD8, 0xA4]),                                        # eof, :13-17
    bytes([0xFF, 0xD8, 0xFF, 0x00, 0x00, 0x00]),                      # bad_ff_marker_size, :20-24
    bytes([255, 216, 255, 218, 232, 197, 255]),                       # bad_number_of_scans, :27-35
    bytes([255, 216, 255, 196, 0, 0]),                                # huffman_length_subtraction_overflow, :38-46
    bytes([255, 216, 255, 218, 0, 8, 1, 0, 8, 1]),                    # index_oob, :49-53
    bytes([255, 216, 255, 192, 255, 1, 8, 9, 119, 48, 255, 192]),     # mul_with_overflow, :56-66
], ids=["eof", "bad_ff_marker_size", "bad_number_of_scans", "huffman_length_subtraction_overflow",
        "index_oob", "mul_with_overflow"])
def test_reference_invalid_images_rejected(data):
    with pytest.raises(artefact_amd.JpegDecodeError):
        artefact_amd.decode_jpeg(data)


def _with_frame_size(data, w, h):
    data = bytearray(data)
    i = data.index(b"\xff\xc0")
    data[i + 5:i + 7] = h.to_bytes(2, "big")
    data[i + 7:i + 9] = w.to_bytes(2, "big")
    return bytes(data)


def test_frame_size_limit():
    """zune-core's default decoder options (zune.rs:12) allow 1 << 14 pixels per side
    (zune-jpeg/src/headers.rs:216-230).  A 16384-wide frame decodes (its missing entropy data
    keeps zeros); one pixel more on either side is rejected before any plane is allocated."""
    data = pil_jpeg(np.full((8, 16), 128, np.uint8), quality=90)
    j = artefact_amd.decode_jpeg(_with_frame_size(data, 16384, 8))
    assert (j.coefs[0].rounded_px_w, j.coefs[0].rounded_px_h) == (16384, 8)
    with pytest.raises(artefact_amd.JpegDecodeError, match="width 16385 greater than width limit 16384"):
        artefact_amd.decode_jpeg(_with_frame_size(data, 16385, 8))
    with pytest.raises(artefact_amd.JpegDecodeError, match="height 65535 greater than height limit 16384"):
        artefact_amd.decode_jpeg(_with_frame_size(data, 16, 65535))


def test_mutated_files_decode_or_raise():
    """Seeded byte flips, overwrites, insertions, deletions and truncations of valid files
    (baseline and progressive, 4:2:0 / 4:4:4 / grey): each either decodes into planes of the
    size it reports or raises JpegDecodeError."""
    rng = np.random.default_rng(11)
    rgb = np.stack(synth.synth_rgb(40, 24, 3), -1).astype(np.uint8)
    seeds = [pil_jpeg(rgb, quality=q, subsampling=s, progressive=p)
             for q, s, p in [(30, 2, False), (90, 0, True), (60, 2, True)]]
    seeds.append(pil_jpeg(rgb[..., 0], quality=75, progressive=True))
    decoded = rejected = 0
    for _ in range(1500):
        b = bytearray(seeds[rng.integers(len(seeds))])
        for _ in range(rng.integers(1, 5)):
            if not b:
                break
            at = int(rng.integers(len(b)))
            op = rng.integers(6)
            if op == 0:
                b[at] ^= 1 << int(rng.integers(8))
            elif op == 1:
                b[at] = int(rng.integers(256))
            elif op == 2:
                b[at] = 0xFF
            elif op == 3:
                del b[at:]
            elif op == 4:
                b.insert(at, int(rng.integers(256)))
            else:
                del b[at:at + int(rng.integers(1, 16))]
        try:
            j = artefact_amd.decode_jpeg(bytes(b))
        except artefact_amd.JpegDecodeError:
            rejected += 1
            continue
        decoded += 1
        for c in j.coefs:
            assert c.dct_coefs.size == c.rounded_px_w * c.rounded_px_h
            assert c.quant_table.size == 64
    assert decoded > 100 and rejected > 100, (decoded, rejected)


def test_unsupported_process_raises():
    data = bytearray(pil_jpeg(np.zeros((16, 16), np.uint8), quality=90))
    i = data.index(b"\xff\xc0")
    data[i + 1] = 0xC3  # lossless
    with pytest.raises(artefact_amd.JpegDecodeError, match="unsupported"):
        artefact_amd.decode_jpeg(bytes(data))


def test_source_errors_use_the_reference_messages(tmp_path):
    with pytest.raises(artefact_amd.ArtefactProcessError, match="Failed to read JPEG: Failed to read JPEG file"):
        artefact_amd.Artefact().source(artefact_amd.JpegSource.File(str(tmp_path / "missing.jpg"))).process()
    with pytest.raises(artefact_amd.ArtefactProcessError, match="Failed to read JPEG: Failed to decode JPEG"):
        artefact_amd.Artefact().source(artefact_amd.JpegSource.Buffer(b"garbage")).process()


def _grey_block(data):
    j = artefact_amd.decode_jpeg(data)
    assert j.nchannel == 1
    return j.coefs[0].dct_coefs.reshape(-1, 64)[0]


def test_refinement_run_past_the_band_places_the_new_coefficient_at_its_end():
    """A refinement symbol whose zero run reaches past the band (zune-jpeg/src/bitstream.rs:579-616):
    the fork stops the run at the band's last position and puts the new coefficient there."""
    q = np.full(64, 4)
    scans = [
        (0, 0, 0, 0, ScanBits().dc(0)),
        (1, 5, 0, 1, ScanBits().ac(0, 1).ac(0, 0)),  # zigzag 1 = 1 << 1, then EOB
        # run 4 with a new +1: zigzag 2..5 hold only four zeros, so the run ends at zigzag 5;
        # the correction bit of zigzag 1 (passed on the way) is 0.
        (1, 5, 1, 0, ScanBits().ac(4, 1).raw(0, 1)),
    ]
    blk = _grey_block(write_grey_scans(8, 8, q, scans))
    want = np.zeros(64, np.int16)
    want[ZIGZAG[1]] = 2
    want[ZIGZAG[5]] = 1
    assert np.array_equal(blk, want)


@pytest.mark.parametrize("progressive", [False, True])
@pytest.mark.parametrize("size,where", [(2, "wrap"), (1, "clamp")])
def test_ac_run_past_63_lands_where_the_fork_puts_it(progressive, size, where):
    """A run that carries the zigzag position past 63 (corrupt data).  The fork's combined
    run/value table path clamps the position to 63; its other path wraps it modulo 64
    (bitstream.rs:359-378 sequential, :487-503 progressive).  With this writer's 8-bit AC codes,
    a 1-bit value takes the table path (8 + 1 <= 9 bits) and a 2-bit value the other one."""
    q = np.full(64, 4)
    ac = ScanBits()
    if not progressive:
        ac.dc(0)
    ac.ac(15, 0).ac(15, 0).ac(15, 0)   # zigzag 1 -> 49
    ac.ac(12, 1)                       # value 1 at zigzag 61
    ac.ac(15, size)                    # run 15 from zigzag 62 -> 77
    if progressive:
        scans = [(0, 0, 0, 0, ScanBits().dc(0)), (1, 63, 0, 0, ac)]
    else:
        scans = [(0, 63, 0, 0, ac)]
    blk = _grey_block(write_grey_scans(8, 8, q, scans, progressive=progressive))
    want = np.zeros(64, np.int16)
    want[ZIGZAG[61]] = 1
    want[ZIGZAG[77 & 63] if where == "wrap" else ZIGZAG[63]] = size  # value 2 (2 bits) or 1 (1 bit)
    assert np.array_equal(blk, want)


def test_missing_restart_marker_follows_the_fork():
    """A sequential file whose restart interval is one MCU loses one RSTn marker that follows an
    interval ending on a byte boundary (no padding bits), so the next interval's bits follow
    seamlessly.  The fork restarts only at an MCU end where its countdown is 0 and the reader
    has met an RSTn marker (mcu.rs:288-330): the next interval then decodes with the DC
    predictions carried over (each component's DC values offset by its prediction at the missing
    marker), and the interval after it, whose marker is present, decodes exactly again."""
    w, h = 64, 48
    coefs = synth.encode(synth.synth_rgb(h, w, 21), subsampling="444", quality=90)
    comps = _std_components(coefs, [(1, 1)] * 3)
    log = []
    data = write_baseline(w, h, comps, restart_interval=1, rst_log=log)
    k = next(i for i, (_, pad) in enumerate(log) if pad == 0 and i + 1 < len(log))
    at = log[k][0]
    assert data[at] == 0xFF and 0xD0 <= data[at + 1] <= 0xD7
    cut = data[:at] + data[at + 2:]
    full = artefact_amd.decode_jpeg(data)
    got = artefact_amd.decode_jpeg(cut)
    # With a one-MCU interval, MCU k + 1 is the interval after the missing marker.
    for c in range(3):
        want = full.coefs[c].dct_coefs.reshape(-1, 64).astype(np.int32).copy()
        want[k + 1, 0] += want[k, 0]  # the prediction at the missing marker: the last DC before it
        assert np.array_equal(got.coefs[c].dct_coefs.reshape(-1, 64).astype(np.int32), want), c
