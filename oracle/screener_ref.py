"""ORACLE — test infrastructure only.

FP64 NumPy restatement of the screener network's layers (no reference
counterpart: the screener is an addition of this project) and of its feature
construction, to check the HIP kernels against:

* conv1d with kernel 9 and 'same' zero padding:
  y[b, co, x] = bias[co] + sum_ci sum_k w[co, ci, k] * a[b, ci, x + k - 4],
  a = relu(input) for every layer but the first;
* its gradients by the chain rule (``backward``), themselves pinned against
  central finite differences in tests/test_screener_oracle.py;
* the peak-weighted squared-error loss.
"""
import numpy as np

K = 9
PAD = 4


def conv(a, w, b):
    """a (B, Cin, X), w (Cout, Cin, K), b (Cout) -> (B, Cout, X)."""
    B, Cin, X = a.shape
    ap = np.pad(a, ((0, 0), (0, 0), (PAD, PAD)))
    cols = np.stack([ap[:, :, k:k + X] for k in range(K)], axis=-1)  # B,Cin,X,K
    return np.einsum("bixk,oik->box", cols, w) + b[None, :, None]


def forward(x, params):
    """params: list of (w, b); returns prediction (B, X) and the list of
    pre-activations [x, z1, z2, z3, z4]."""
    acts = [x]
    a = x
    for i, (w, b) in enumerate(params):
        inp = a if i == 0 else np.maximum(a, 0.0)
        z = conv(inp, w, b)
        acts.append(z)
        a = z
    return acts[-1][:, 0, :], acts


def backward(acts, params, gy):
    """Gradients [(gw, gb), ...] of sum(gy * prediction)."""
    grads = [None] * len(params)
    gz = gy[:, None, :]
    for i in reversed(range(len(params))):
        w, b = params[i]
        ain = acts[i]
        a = ain if i == 0 else np.maximum(ain, 0.0)
        B, Cin, X = a.shape
        ap = np.pad(a, ((0, 0), (0, 0), (PAD, PAD)))
        cols = np.stack([ap[:, :, k:k + X] for k in range(K)], axis=-1)
        gw = np.einsum("box,bixk->oik", gz, cols)
        gb = gz.sum(axis=(0, 2))
        grads[i] = (gw, gb)
        if i > 0:
            gzp = np.pad(gz, ((0, 0), (0, 0), (PAD, PAD)))
            ga = np.zeros_like(a)
            for k in range(K):
                # a[x'] receives w[.., k] * gz[x' - k + 4]
                ga += np.einsum("oi,box->bix", w[:, :, k], gzp[:, :, 2 * PAD - k:2 * PAD - k + X])
            gz = ga * (ain > 0)
    return grads


def loss(pred, target, boost):
    m = 1.0 + boost * target
    return float(np.sum(m * (pred - target) ** 2) / np.sum(m))


def loss_grad(pred, target, boost):
    m = 1.0 + boost * target
    return 2.0 * m * (pred - target) / np.sum(m)
