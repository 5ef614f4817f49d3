"""Screening variant of the MFMA kernel (FMF_KERNEL_BF16; no reference
counterpart): one bf16 product per MAC, a 1/sqrt(energy) table, no 1e-6
threshold.  Bars against the FP64 oracle (oracle/fp64.py):

  * |dcc| <= 2^-8 * sum |w| per template row: rounding template and data to
    bf16 changes each product T_k d_k by at most 2^-8 |T_k d_k| (plus 2^-18),
    and sum |T_k d_k| <= sqrt(sum T^2 sum d^2) (Cauchy-Schwarz), so each
    channel's cc moves by at most 2^-8;
  * RMS |dcc| <= 1e-3 (the roundings are unbiased);
  * exactly 0 where every channel's window is zero and past the last lag;
  * the per-channel layout within 2^-8 of the oracle's per-channel cc;
  * windows that start before the data (negative moveouts) stay within the
    bound;
  * steps other than 1 and 2 and 'full' normalization are refused.
"""
import numpy as np
import pytest

from tests.golden import cases

pytestmark = pytest.mark.gpu

BOUND = 2.0 ** -8


def _expand(x, T, S, C):
    x = np.asarray(x)
    if x.ndim == 2:  # (T, S) -> (T, S, C), as the reference wrapper does
        x = np.repeat(x[:, :, None], C, axis=2)
    return x.reshape(T, S, C)


def _screen(inp, step=None, network_sum=True, normalize=0):
    import torch
    from fast_matched_filter_amd import _lib
    from fast_matched_filter_amd.device import matched_filter_device
    tp = np.asarray(inp["templates"], np.float32)
    T, S, C, N = tp.shape
    data = np.asarray(inp["data"], np.float32).reshape(S, C, -1)
    L = data.shape[-1]
    step = step or inp["step"]
    mo = _expand(inp["moveouts"], T, S, C).astype(np.int32)
    w = _expand(inp["weights"], T, S, C).astype(np.float32)
    sst = np.sum(tp * tp, axis=-1, dtype=np.float32)
    n_corr = (L - N) // step + 1
    print("screening kernel: T=%d S=%d C=%d N=%d L=%d step=%d sum=%s normalize=%d" %
          (T, S, C, N, L, step, network_sum, normalize), flush=True)
    shape = (T, n_corr) if network_sum else (T, n_corr, S * C)
    out = torch.zeros(shape, dtype=torch.float32, device="cuda")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    matched_filter_device(dev(tp), dev(sst), dev(mo), dev(w), dev(data), step, N, L, T, S, C,
                          n_corr, out, normalize=normalize, network_sum=network_sum,
                          kernel=_lib.KERNEL_BF16)
    got = out.cpu().numpy()
    return got, mo, w


def _oracle(inp, mo, w, step, network_sum=True, normalize="short"):
    from oracle import fp64
    return fp64.matched_filter(inp["templates"], mo, w, inp["data"], step, normalize=normalize,
                               network_sum=network_sum)


@pytest.fixture(scope="module")
def inputs():
    return {name: cases.CASES[name]() for name in ("small", "noisy", "selfdet", "edge")}


@pytest.mark.parametrize("name,step", [("small", 1), ("small", 2), ("noisy", 1),
                                       ("selfdet", 2)])
def test_within_bf16_bound_of_fp64(inputs, name, step):
    inp = inputs[name]
    got, mo, w = _screen(inp, step=step)
    ref = _oracle(inp, mo, w, step)
    assert got.shape == ref.shape
    err = np.abs(got - ref)
    bound = BOUND * np.abs(w).reshape(w.shape[0], -1).sum(axis=1) + 1e-6
    print("%s step %d: max |dcc| %.3g, rms %.3g, bound %.3g" %
          (name, step, err.max(), np.sqrt((err ** 2).mean()), bound.max()), flush=True)
    assert np.all(err.max(axis=1) <= bound)
    assert np.sqrt((err ** 2).mean()) <= 1e-3
    # same zero tail as the reference: nothing past each row's last lag
    from oracle import fp64
    tp = np.asarray(inp["templates"])
    T, N, L = tp.shape[0], tp.shape[-1], np.asarray(inp["data"]).shape[-1]
    for t in range(T):
        start, stop = fp64.lag_range(mo[t].reshape(-1), step, N, L)
        hi = -(-stop // step) if stop > 0 else 0
        assert not got[t, hi:].any()


def test_per_channel_layout(inputs):
    inp = inputs["small"]
    got, mo, w = _screen(inp, network_sum=False)
    ref = _oracle(inp, mo, w, 1, network_sum=False)
    ref = ref.reshape(got.shape)
    assert np.abs(got - ref).max() <= BOUND + 1e-6


def test_zero_gap_gives_exact_zeros(inputs):
    inp = dict(inputs["small"])
    data = np.array(inp["data"], np.float32, copy=True)
    L = data.shape[-1]
    g0, g1 = L // 3, L // 3 + 6000
    data[..., g0:g1] = 0.0
    inp["data"] = data
    got, mo, w = _screen(inp)
    N = np.asarray(inp["templates"]).shape[-1]
    mo2 = mo.reshape(mo.shape[0], -1)
    checked = 0
    for t in range(got.shape[0]):
        # lags whose windows lie inside the gap on every channel
        lo = g0 - mo2[t].min()
        hi = g1 - N - mo2[t].max()
        if hi > lo:
            assert not got[t, lo:hi + 1].any()
            checked += hi + 1 - lo
    assert checked > 1000
    ref = _oracle(inp, mo, w, 1)
    err = np.abs(got - ref)
    assert np.all(err.max(axis=1) <= BOUND * np.abs(w).reshape(w.shape[0], -1).sum(axis=1) + 1e-6)


def test_other_steps_and_full_normalization_refused(inputs):
    with pytest.raises(RuntimeError):
        _screen(inputs["edge"])  # step 3
    with pytest.raises(RuntimeError):
        _screen(inputs["small"], normalize=1)



@pytest.mark.parametrize("step", [1, 2])
def test_negative_moveouts_within_bound(step):
    """Windows that start before the data (moveouts down to -70 in the first
    block): data and energies there take the out-of-range offset and read 0;
    the result stays within the bound of the FP64 oracle, which applies the
    reference's lag rule (lags whose windows leave the data are 0)."""
    from tests.test_gpu_mfma32 import _negative_moveouts
    inp = _negative_moveouts(step, 50 + step)
    got, mo, w = _screen(inp)
    ref = _oracle(inp, mo, w, step)
    err = np.abs(got - ref)
    bound = BOUND * np.abs(w).reshape(w.shape[0], -1).sum(axis=1) + 1e-6
    print("negative moveouts step %d: max |dcc| %.3g, bound %.3g" % (step, err.max(), bound.max()),
          flush=True)
    assert np.all(err.max(axis=1) <= bound)
    assert np.count_nonzero(ref) > 0.9 * ref.size
