"""GPU output stage and Artefact::process against the oracle, through the C ABI.

The loop is bit-exact against the oracle in norm_mode="f64", and the output stage
is integer-valued, so every byte must match.
"""
import numpy as np
import pytest

import artefact_amd
import oracle
from artefact_amd import synth

pytestmark = pytest.mark.gpu


def jpeg_of(wl):
    return artefact_amd.Jpeg(len(wl["coefs"]), wl["width"], wl["height"], wl["coefs"])


@pytest.mark.parametrize("name,iters", [("tgv_420_93x61", 6), ("tgv_444_64x40", 3), ("tv_luma_37x21", 4),
                                        ("tgv_420_128x64", 2)])
def test_download_rgb_matches_oracle(name, iters):
    wl = synth.workload(name, seed=21)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case {name} iters={iters}", flush=True)
    with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"]) as plan:
        plan.init()
        plan.iterate()
        planes = plan.download()
        rgb = plan.download_rgb(wl["width"], wl["height"])
    assert np.array_equal(rgb, oracle.output_rgb(planes, W, wl["width"], wl["height"]))


@pytest.mark.parametrize("real_w", [1, 2, 3, 4, 5, 61, 64])
def test_download_rgb_special_values_and_ragged_rows(real_w):
    W, H = 64, 8
    rng = np.random.default_rng(real_w)
    y = rng.uniform(-300, 300, (H, W)).astype(np.float32)
    cb = rng.uniform(-300, 300, (H, W)).astype(np.float32)
    cr = rng.uniform(-300, 300, (H, W)).astype(np.float32)
    y[0, :8] = [np.nan, -0.01, 126.99, 127.0, -128.0, np.inf, -np.inf, 0.0]
    cb[1, :3] = [np.nan, np.inf, -np.inf]
    wl = synth.workload("tgv_444_64x8", seed=1)
    print(f"case special real_w={real_w}", flush=True)
    with artefact_amd.Plan(wl["coefs"], 0.3, wl["pweight"], 0, W, H, init_override=[y, cb, cr]) as plan:
        plan.init()
        plan.iterate()
        rgb = plan.download_rgb(real_w, 7)
    assert np.array_equal(rgb, oracle.output_rgb([y, cb, cr], W, real_w, 7))


@pytest.mark.parametrize("name,sep", [("tgv_420_93x61", False), ("tgv_420_93x61", True), ("tgv_444_40x24", True),
                                      ("tv_luma_37x21", False)])
def test_process_matches_oracle(name, sep):
    wl = synth.workload(name, seed=8)
    weight, pweight, iters = [0.3, 0.2, 0.1], [0.001, 0.01, 0.002], [5, 3, 2]
    print(f"case process {name} separate={sep}", flush=True)
    img = artefact_amd.process_coefficients(jpeg_of(wl), weight, pweight, iters, sep)
    ref = oracle.process(wl["coefs"], wl["width"], wl["height"], weight, pweight, iters, sep, norm_mode="f64")
    assert np.array_equal(img, ref)


def test_artefact_builder_defaults_and_benchmark():
    wl = synth.workload("tgv_420_48x32", seed=5)
    img = artefact_amd.Artefact().source(jpeg_of(wl)).process()
    ref = oracle.process(wl["coefs"], 48, 32, [0.3] * 3, [0.001] * 3, [50] * 3, norm_mode="f64")
    assert np.array_equal(img, ref)
    a = (artefact_amd.Artefact().source(jpeg_of(wl)).weight(artefact_amd.ForEach((0.3, 0.3, 0.3)))
         .iterations(artefact_amd.ForAll(2)).benchmark(True))
    with pytest.raises(artefact_amd.ArtefactProcessError, match="BENCHMARK"):
        a.process()
    with pytest.raises(artefact_amd.ArtefactProcessError, match="Source is not set"):
        artefact_amd.Artefact().process()


@pytest.mark.parametrize("sub,prog,sep", [(2, False, False), (0, True, True), (1, False, False)])
def test_process_from_jpeg_bytes(sub, prog, sep):
    """Artefact::process end to end: JPEG bytes (written by libjpeg through Pillow) ->
    host decoder -> GPU loop(s) -> GPU output stage, byte-exact against the oracle run on
    the decoded coefficients."""
    import io
    from PIL import Image
    rgb = np.stack(synth.synth_rgb(45, 70, 12), -1)
    f = io.BytesIO()
    Image.fromarray(rgb).save(f, "JPEG", quality=60, subsampling=sub, progressive=prog)
    data = f.getvalue()
    print(f"case jpeg sub={sub} progressive={prog} separate={sep}", flush=True)
    img = (artefact_amd.Artefact().source(artefact_amd.JpegSource.Buffer(data))
           .iterations(artefact_amd.ForAll(4)).separate_components(sep).process())
    j = artefact_amd.decode_jpeg(data)
    ref = oracle.process(j.coefs, 70, 45, [0.3] * 3, [0.001] * 3, [4] * 3, sep, norm_mode="f64")
    assert img.shape == (45, 70, 3)
    assert np.array_equal(img, ref)


def test_process_largest_jpeg():
    """Artefact::process on the largest image the decoder accepts, 16384 x 16384 4:2:0 (3 GiB
    of planes: the 64-bit-addressed gradient sweep).  The oracle is too slow here, so the
    check is the shape and closeness to libjpeg's decode after 2 iterations, which move the
    decoded image by little."""
    import io
    from PIL import Image
    n = 16384
    yy = np.arange(n, dtype=np.uint32)[:, None]
    xx = np.arange(n, dtype=np.uint32)[None, :]
    rgb = np.empty((n, n, 3), np.uint8)
    rgb[..., 0] = ((xx * 3 + yy) >> 7) & 255
    rgb[..., 1] = ((xx + yy * 5) >> 8) & 255
    rgb[..., 2] = 128 + (((xx >> 9) ^ (yy >> 9)) & 63)
    f = io.BytesIO()
    Image.fromarray(rgb).save(f, "JPEG", quality=75, subsampling=2)
    del rgb
    data = f.getvalue()
    print(f"case process {n}x{n} 4:2:0, {len(data)} bytes", flush=True)
    img = (artefact_amd.Artefact().source(artefact_amd.JpegSource.Buffer(data))
           .iterations(artefact_amd.ForAll(2)).process())
    assert img.shape == (n, n, 3) and img.dtype == np.uint8
    limit, Image.MAX_IMAGE_PIXELS = Image.MAX_IMAGE_PIXELS, None  # PIL's guard stops at 179 M pixels
    try:
        ref = np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))
    finally:
        Image.MAX_IMAGE_PIXELS = limit
    diff = np.abs(img.astype(np.int16) - ref.astype(np.int16))
    rows = diff.mean(axis=(1, 2))
    cols = diff.mean(axis=(0, 2))
    tail_r, tail_c = diff[-64:].mean(), diff[:, -64:].mean()
    print(f"mean abs diff {diff.mean():.3f}, max {diff.max()}; worst row {rows.max():.3f} (row {rows.argmax()}), "
          f"worst column {cols.max():.3f} (column {cols.argmax()}); last 64 rows {tail_r:.3f}, columns {tail_c:.3f}",
          flush=True)
    assert diff.mean() < 3.0
    # An addressing error confined to some rows or columns (a 32-bit wrap past 2^32 bytes of
    # planes) leaves the global mean small but not these.
    assert rows.max() < 6.0 and cols.max() < 6.0
    assert tail_r < 3.0 and tail_c < 3.0
