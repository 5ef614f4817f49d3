"""Pins the screener's FP64 oracle (CPU only): chain-rule gradients of the
network equal central finite differences, and conv1d equals a direct loop."""
import numpy as np

from oracle import screener_ref as ref


def _params(rng, layers=((3, 4), (4, 4), (4, 1))):
    return [(rng.standard_normal((co, ci, ref.K)) * 0.4, rng.standard_normal(co) * 0.1)
            for ci, co in layers]


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2, 3, 17))
    w = rng.standard_normal((4, 3, ref.K))
    b = rng.standard_normal(4)
    y = ref.conv(a, w, b)
    for bb in range(2):
        for co in range(4):
            for x in range(17):
                s = b[co]
                for ci in range(3):
                    for k in range(ref.K):
                        xx = x + k - ref.PAD
                        if 0 <= xx < 17:
                            s += w[co, ci, k] * a[bb, ci, xx]
                assert abs(y[bb, co, x] - s) < 1e-12


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 23))
    params = _params(rng)
    target = np.abs(rng.standard_normal((2, 23)))
    pred, acts = ref.forward(x, params)
    gy = ref.loss_grad(pred, target, 1.5)
    grads = ref.backward(acts, params, gy)
    h = 1e-6
    for li in range(len(params)):
        w, b = params[li]
        for idx in [(0, 0, 0), (w.shape[0] - 1, w.shape[1] - 1, ref.K - 1), (0, w.shape[1] - 1, 4)]:
            old = w[idx]
            w[idx] = old + h
            lp = ref.loss(ref.forward(x, params)[0], target, 1.5)
            w[idx] = old - h
            lm = ref.loss(ref.forward(x, params)[0], target, 1.5)
            w[idx] = old
            assert abs((lp - lm) / (2 * h) - grads[li][0][idx]) < 1e-6
        old = b[0]
        b[0] = old + h
        lp = ref.loss(ref.forward(x, params)[0], target, 1.5)
        b[0] = old - h
        lm = ref.loss(ref.forward(x, params)[0], target, 1.5)
        b[0] = old
        assert abs((lp - lm) / (2 * h) - grads[li][1][0]) < 1e-6
