"""Warm-start CNN kernels (csrc/cnn.hip) against the CPU reference (oracle/cnn.py), and the
streaming inference kernel against the layered one.

BF16 activations and weights with f32 accumulation are not bit-identical to the float64
reference, so the checks are relative: the predicted residual within 3 % (relative L2), each
layer's weight and bias gradients with cosine similarity >= 0.99 and relative L2 error
<= 8 %, and Adam steps on a fixed batch lower the loss.
"""
import numpy as np
import pytest

from artefact_amd import warmstart
from oracle import cnn

pytestmark = pytest.mark.gpu


features = cnn.features


def params_with_bias(seed):
    p = warmstart.init_params(seed).astype(np.float64)
    ws, bs = warmstart.split_params(p)
    rng = np.random.default_rng(seed + 7)
    for b in bs:
        b[:] = rng.normal(scale=0.05, size=b.shape)
    flat = np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in zip(ws, bs)])
    return flat.astype(np.float32), ws, bs


def batch(seed, B=2, H=24, W=48):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    x0 = np.stack([60 * np.sin(xx / (5 + i)) + 40 * np.cos(yy / 7) + rng.normal(0, 8, (H, W)) for i in range(B)])
    q = rng.uniform(2, 80, size=(B, 64))  # a quantisation table per plane
    t = x0 + rng.normal(0, 4, x0.shape)
    return x0.astype(np.float32), q.astype(np.float32), t.astype(np.float32)


def test_forward_matches_oracle():
    flat, ws, bs = params_with_bias(1)
    x0, q, _ = batch(2)
    print("case cnn forward", flush=True)
    net = warmstart.WarmStartNet(flat)
    got = net.forward(x0, q)
    net.close()
    ref, _, _ = cnn.forward(features(x0, q), ws, bs)
    ref = ref * 64.0
    r_got, r_ref = got - x0, ref - x0
    assert np.linalg.norm(r_got - r_ref) <= 0.03 * np.linalg.norm(r_ref)


def test_gradients_match_oracle():
    flat, ws, bs = params_with_bias(3)
    x0, q, t = batch(4)
    print("case cnn gradients", flush=True)
    net = warmstart.WarmStartNet(flat)
    loss, g = net.gradients(x0, q, t)
    net.close()
    rloss, gws, gbs = cnn.loss_and_grads(features(x0, q), t / 64.0, ws, bs)
    assert abs(loss - rloss) <= 0.03 * rloss
    gw, gb = warmstart.split_params(g)
    for l in range(len(ws)):
        for a, b in ((gw[l], gws[l]), (gb[l], gbs[l])):
            a, b = a.reshape(-1), b.reshape(-1)
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
            rel = float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))
            print(f"layer {l}: cos {cos:.5f} rel {rel:.4f}")
            assert cos >= 0.99 and rel <= 0.08, (l, cos, rel)


def test_adam_lowers_the_loss():
    """A learnable target (a pointwise nonlinearity of the input): 150 Adam steps on one batch
    cut the loss at least fourfold."""
    x0, q, _ = batch(5, B=4, H=32, W=32)
    t = (x0 + 10.0 * np.tanh(x0 / 30.0)).astype(np.float32)
    net = warmstart.WarmStartNet(warmstart.init_params(0))
    net.set_dataset(x0, q, t)
    losses = [net.train_step(np.arange(4), 2e-3) for _ in range(150)]
    net.close()
    print(f"loss {losses[0]:.5g} -> {losses[-1]:.5g}")
    assert losses[-1] < 0.25 * losses[0], (losses[0], losses[-1])


def test_out_of_range_index_takes_no_step():
    """A batch with an index past the dataset is rejected before anything runs: the weights
    stay as they were, and the next valid step still trains."""
    from artefact_amd import _lib

    x0, q, t = batch(3, B=2, H=16, W=32)
    net = warmstart.WarmStartNet(warmstart.init_params(1))
    net.set_dataset(x0, q, t)
    before = net.params()
    with pytest.raises(_lib.ArtefactError) as err:
        net.train_step([0, 2], 1e-3)
    assert err.value.code == -1
    np.testing.assert_array_equal(net.params(), before)
    net.train_step([0, 1], 1e-3)
    assert not np.array_equal(net.params(), before)
    net.close()


@pytest.mark.parametrize("H,W", [(24, 48), (70, 100), (200, 136), (9, 17), (8, 8)])
def test_streaming_kernel_matches_layered(H, W, monkeypatch):
    """Inference runs all five layers in one streaming kernel (v_mfma_f32_16x16x32_bf16, each
    tap over all 32 channels at once).  The layered kernels (16x16x16, two channel halves per
    tap) stay behind ARTEFACT_CNN_LAYERED=1.  Both round to bf16 between layers after
    different MFMA groupings of the same sums, so they agree to bf16 precision, not bit for
    bit: the residuals within 1 % (relative L2) and, at every pixel, within 2 % of the
    residual's range.  The sizes give partial bands and tiles, several row segments, and
    planes smaller than one segment and one tile."""
    flat, _, _ = params_with_bias(11)
    x0, q, _ = batch(12, B=2, H=H, W=W)
    print(f"case cnn streaming vs layered {H}x{W}", flush=True)
    net = warmstart.WarmStartNet(flat)
    got = net.forward(x0, q)
    monkeypatch.setenv("ARTEFACT_CNN_LAYERED", "1")
    ref = net.forward(x0, q)
    net.close()
    r_got, r_ref = got - x0, ref - x0
    rel = np.linalg.norm(r_got - r_ref) / np.linalg.norm(r_ref)
    mx = np.max(np.abs(r_got - r_ref)) / (np.max(r_ref) - np.min(r_ref))
    print(f"relative L2 {rel:.3g}, max {mx:.3g} of the range")
    assert rel <= 0.01
    assert mx <= 0.02


def test_failed_workspace_allocation_leaves_the_net_usable():
    """A forward too large to allocate returns ARTEFACT_E_HIP, and the net's workspace pointers are
    cleared as they are released: a normal forward afterwards succeeds and destroy frees nothing
    twice.  The input buffers are never read: the allocation fails first."""
    import ctypes

    from artefact_amd import _lib

    flat, _, _ = params_with_bias(2)
    net = warmstart.WarmStartNet(flat)
    x0, q, _ = batch(6, B=1, H=24, W=32)
    net.forward(x0, q)  # allocates a small workspace
    lib = _lib.load()
    F = ctypes.POINTER(ctypes.c_float)
    tiny = np.zeros(64 * 64, np.float32)
    rc = lib.artefact_cnn_forward(net._h, tiny.ctypes.data_as(F), tiny.ctypes.data_as(F), 64, 65535, 65535,
                                  tiny.ctypes.data_as(F))
    assert rc == -4
    again = net.forward(x0, q)
    assert np.isfinite(again).all()
    net.close()
