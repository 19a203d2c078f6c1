"""The CPU reference of the warm-start CNN (oracle/cnn.py) pinned by central finite
differences in float64, and its convolution by a brute-force loop."""
import numpy as np

from oracle import cnn


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 6, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    out = cnn.conv3x3(x, w, b)
    ref = np.zeros_like(out)
    for n in range(2):
        for i in range(5):
            for j in range(6):
                for co in range(4):
                    s = b[co]
                    for ci in range(3):
                        for dy in range(3):
                            for dx in range(3):
                                y, xx = i + dy - 1, j + dx - 1
                                if 0 <= y < 5 and 0 <= xx < 6:
                                    s += w[co, ci, dy, dx] * x[n, y, xx, ci]
                    ref[n, i, j, co] = s
    assert np.allclose(out, ref, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    widths = (2, 4, 3, 1)
    ws, bs = cnn.init_params(2, widths)
    for b in bs:
        b[:] = rng.normal(scale=0.1, size=b.shape)
    x = rng.normal(size=(2, 6, 5, 2))
    t = rng.normal(size=(2, 6, 5))
    _, gws, gbs = cnn.loss_and_grads(x, t, ws, bs)
    eps = 1e-6
    for li in range(len(ws)):
        for idx in [(0, 0, 0, 0), (1, 1, 2, 1), (0, 1, 1, 2)]:
            if any(i >= s for i, s in zip(idx, ws[li].shape)):
                continue
            ws[li][idx] += eps
            lp, _, _ = cnn.loss_and_grads(x, t, ws, bs)
            ws[li][idx] -= 2 * eps
            lm, _, _ = cnn.loss_and_grads(x, t, ws, bs)
            ws[li][idx] += eps
            assert abs((lp - lm) / (2 * eps) - gws[li][idx]) < 1e-6 * max(1.0, abs(gws[li][idx]))
        bs[li][0] += eps
        lp, _, _ = cnn.loss_and_grads(x, t, ws, bs)
        bs[li][0] -= 2 * eps
        lm, _, _ = cnn.loss_and_grads(x, t, ws, bs)
        bs[li][0] += eps
        assert abs((lp - lm) / (2 * eps) - gbs[li][0]) < 1e-6 * max(1.0, abs(gbs[li][0]))


def test_adam_first_step_moves_by_lr():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -3.0])
    m, v = np.zeros(2), np.zeros(2)
    cnn.adam_step(p, g, m, v, 1, 0.01)
    assert np.allclose(p, [1.0 - 0.01, -2.0 + 0.01], atol=1e-7)  # |mhat/sqrt(vhat)| = 1 at t = 1


def test_features_tile_the_table_over_the_blocks():
    """Feature 1 at pixel (y, x) is the plane's table entry (y mod 8, x mod 8) / 64; feature 0 is
    the plane / 64."""
    rng = np.random.default_rng(4)
    x0 = rng.normal(size=(2, 13, 21))
    q = rng.uniform(1, 99, size=(2, 64))
    f = cnn.features(x0, q)
    assert f.shape == (2, 13, 21, 2)
    assert np.array_equal(f[..., 0], x0 / 64.0)
    for b in range(2):
        for y in range(13):
            for x in range(21):
                assert f[b, y, x, 1] == q[b, (y % 8) * 8 + x % 8] / 64.0
