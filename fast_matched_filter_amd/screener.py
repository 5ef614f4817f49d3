"""Learned screener in front of the exact matched filter.

A 4-layer 1D convolutional network (kernel 9, 32 channels, ReLU, FP16 weights,
FP32 accumulation) predicts, per template, the decimated peak |cc| (max |cc|
over each run of 8 lags, in units of the template's cc median absolute
deviation, log-compressed).  Its input is a coarse correlation: the matched
filter's screening kernel (one bf16 product per MAC) run on data and templates
decimated by COARSE = 8 (mean of 8 samples, moveouts rounded to the coarse
grid), at 1/64 of the exact scan's MACs, turned into 8 features per screener
position (coarse |cc|
and its signed extremes at the position, maximum, mean and rms over +-3
positions, all in units of the coarse cc's MAD, and the template's coarse
noise level).  The exact bf16x3 kernel then runs only on the output tiles
whose prediction exceeds a calibrated gate, plus a random sample of tiles
from which each template's cc MAD is estimated; detections are cc > 6 x MAD.
When the gate keeps too many tiles the full exact scan runs instead
(fallback).

The envelope features of the first version (moveout-aligned data energies,
``features``) stay available for comparison: they cannot see events whose
energy is a few percent of the noise's.

Everything numeric runs in libfmf_hip.so kernels: decimation, the coarse and
exact correlations, the exact MAD (radix select), the features, the network
forward (matrix cores for inference, FP32 vector kernels for training) and
backward, the loss, Adam and the gate.  torch provides device memory, random
numbers for the synthetic days and small index operations.
"""
import ctypes as ct
import math

import numpy as np

from . import _lib

DEC = 8                      # lags per screener sample
COARSE = 8                   # decimation of the coarse scan
# MFMA tile shape of the screened exact scan: 212 (2048 lags per item) keeps
# 2.74 % of a day's lags where 412 (4096) keeps 3.39 % and the full scan's
# 422 (8192) 4.70 % at the same event recall; an event needs only the lags
# around it.  On one box: 23.0 ms per day against 24.4 ms with 412.
ITEM_SHAPE = 212
# item shapes whose bf16x3 and screening blocks cover the same lags (the MAD
# sample then runs on the screening kernel)
MIXED_SHAPES = (212, 412)
LAYERS = ((8, 32), (32, 32), (32, 32), (32, 1))
K = 9
N_FEATURES = 8


def _p(t):
    return ct.c_void_p(t.data_ptr())


def _s(stream):
    return ct.c_void_p(stream.cuda_stream)


def _bind(lib):
    if getattr(lib, "_screener_bound", False):
        return lib
    V, I, F, L = ct.c_void_p, ct.c_int, ct.c_float, ct.c_longlong
    sig = {
        "fmf_screen_envelope": [V, I, I, I, V, V],
        "fmf_screen_box_mean": [V, I, I, I, V, V],
        "fmf_screen_features": [V, V, V, V, V, V, I, I, I, I, I, F, V, V],
        "fmf_conv1d_forward": [V, I, I, I, I, V, V, I, V, V],
        "fmf_conv1d_backward": [V, V, V, I, I, I, I, V, I, V, V, V, V],
        "fmf_screen_loss": [V, V, L, F, F, V, V, I, V],
        "fmf_adam_step": [V, V, V, V, I, F, F, F, F, I, V, V, V],
        "fmf_screen_gate": [V, I, V, I, I, I, I, V, V],
        "fmf_screen_peak_target": [V, I, I, I, V, V],
        "fmf_matched_filter_items_device": [V] * 5 + [ct.c_size_t] * 7 + [V, I, I, V, L, I, V,
                                                                          ct.c_size_t, V],
        "fmf_matched_filter_items_mixed_device": [V] * 5 + [ct.c_size_t] * 7 + [V, V, L, V, L, I,
                                                                                V, ct.c_size_t,
                                                                                V],
        "fmf_mfma_item_outputs": [I],
        "fmf_mfma_shape_outputs": [I],
        "fmf_screen_decimate": [V, L, L, I, L, V, V],
        "fmf_screen_coarse_features": [V, I, I, V, I, I, V, V],
        "fmf_screen_coarse_features_f16": [V, I, I, V, I, I, V, V],
        "fmf_conv1d_forward_fast": [V, I, I, I, I, V, V, I, V, V, ct.c_size_t, V],
        "fmf_screen_net_forward": [V, I, I] + [V] * 8 + [V, ct.c_size_t, V, V],
        "fmf_screen_net_forward_f16": [V, I, I] + [V] * 8 + [V, ct.c_size_t, V, V],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = I
    lib.fmf_conv1d_forward_fast_scratch_bytes.argtypes = [I]
    lib.fmf_conv1d_forward_fast_scratch_bytes.restype = ct.c_size_t
    lib.fmf_screen_net_scratch_bytes.argtypes = []
    lib.fmf_screen_net_scratch_bytes.restype = ct.c_size_t
    lib._screener_bound = True
    return lib


def lib():
    return _bind(_lib.require())


def _check(rc, what):
    _lib.check(rc, what)


# ---- features -----------------------------------------------------------------

def features(data, moveouts, weights, n_tp, n_corr, stream=None):
    """Screener input (T, 8, n_out) float32 for one day.  data (S*C, L),
    moveouts int32 / weights float32 (T, S*C), all device tensors."""
    import torch
    dev = data.device
    stream = stream or torch.cuda.current_stream(dev)
    SC, L = data.shape
    T = moveouts.shape[0]
    n_dec = (L + DEC - 1) // DEC
    n_out = (n_corr + DEC - 1) // DEC
    env = torch.empty((SC, n_dec), dtype=torch.float32, device=dev)
    lb = lib()
    _check(lb.fmf_screen_envelope(_p(data), SC, L, n_dec, _p(env), _s(stream)), "envelope")
    nw = max(1, n_tp // DEC)
    win = torch.empty_like(env)
    winh = torch.empty_like(env)
    _check(lb.fmf_screen_box_mean(_p(env), SC, n_dec, nw, _p(win), _s(stream)), "box")
    _check(lb.fmf_screen_box_mean(_p(env), SC, n_dec, max(1, nw // 2), _p(winh), _s(stream)), "box")
    chinv = 1.0 / env.mean(dim=1).clamp_min(1e-30)
    feat = torch.empty((T, N_FEATURES, n_out), dtype=torch.float32, device=dev)
    _check(lb.fmf_screen_features(_p(win), _p(winh), _p(env), _p(chinv.contiguous()),
                                  _p(moveouts), _p(weights), T, SC, n_dec, n_out,
                                  max(1, nw // 2), float(n_tp) / 1000.0, _p(feat),
                                  _s(stream)), "features")
    return feat


def decimate(x, factor, stream=None):
    """(rows, n) device float32 -> (rows, ceil(n / factor)), mean of `factor`
    consecutive samples (past the end: 0)."""
    import torch
    stream = stream or torch.cuda.current_stream(x.device)
    x = x.contiguous()
    rows, n = x.shape
    n_out = (n + factor - 1) // factor
    out = torch.empty((rows, n_out), dtype=torch.float32, device=x.device)
    _check(lib().fmf_screen_decimate(_p(x), rows, n, int(factor), n_out, _p(out), _s(stream)),
           "decimate")
    return out


def coarse_scan(day, factor=COARSE, workspace=None, stream=None):
    """Coarse cc of a day: the screening MFMA scan (one bf16 product per MAC,
    |dcc| <= 2^-8 per unit weight) of data and templates decimated by
    `factor`, moveouts rounded to the coarse grid.  Returns the coarse cc
    (T, n_coarse) and its exact per-template median absolute deviation."""
    import torch
    from .detect import median_mad
    from .device import matched_filter_device
    tp, data = day["templates"], day["data"]
    T, SC, N = tp.shape
    stream = stream or torch.cuda.current_stream(data.device)
    d_c = decimate(data, factor, stream)
    t_c = decimate(tp.reshape(T * SC, N), factor, stream).reshape(T, SC, -1).contiguous()
    N_c, L_c = t_c.shape[-1], d_c.shape[1]
    mo_c = torch.div(day["moveouts"] + factor // 2, factor, rounding_mode="floor").to(torch.int32)
    sst = t_c.pow(2).sum(dim=-1).contiguous()
    n_coarse = L_c - N_c + 1
    cc = torch.empty((T, n_coarse), dtype=torch.float32, device=data.device)
    matched_filter_device(t_c.reshape(-1), sst.reshape(-1), mo_c.contiguous().reshape(-1),
                          day["weights"].reshape(-1), d_c.reshape(-1), 1, N_c, L_c, T, SC, 1,
                          n_coarse, cc, kernel=_lib.KERNEL_BF16, workspace=workspace,
                          stream=stream)
    _, mad_c = median_mad(cc)
    return cc, mad_c


def coarse_features(cc_c, mad_c, n_out, factor=COARSE, stream=None, half=False):
    """Screener input (T, 8, n_out) from the coarse cc and its MAD: float32,
    or float16 with half=True (the FP16 rounding of the same values, what
    the fused inference kernel reads either way)."""
    import torch
    stream = stream or torch.cuda.current_stream(cc_c.device)
    T, n_coarse = cc_c.shape
    feat = torch.empty((T, N_FEATURES, n_out), dtype=torch.float16 if half else torch.float32,
                       device=cc_c.device)
    fn = lib().fmf_screen_coarse_features_f16 if half else lib().fmf_screen_coarse_features
    _check(fn(_p(cc_c.contiguous()), T, n_coarse, _p(mad_c.contiguous()), int(factor), n_out,
              _p(feat), _s(stream)), "coarse features")
    return feat


def day_features(day, n_corr, factor=COARSE, stream=None):
    """Coarse scan + features of a day: (feat (T, 8, n_out), coarse cc, MAD)."""
    cc_c, mad_c = coarse_scan(day, factor=factor, stream=stream)
    n_out = (n_corr + DEC - 1) // DEC
    return coarse_features(cc_c, mad_c, n_out, factor=factor, stream=stream), cc_c, mad_c


def peak_target(cc, stream=None):
    """(T, n_out) max |cc| over each run of 8 lags."""
    import torch
    stream = stream or torch.cuda.current_stream(cc.device)
    T, n_corr = cc.shape
    n_out = (n_corr + DEC - 1) // DEC
    out = torch.empty((T, n_out), dtype=torch.float32, device=cc.device)
    _check(lib().fmf_screen_peak_target(_p(cc), n_corr, n_out, T, _p(out), _s(stream)), "peak")
    return out


def mad(x, dim=-1):
    """Median absolute deviation along `dim` (torch, on device)."""
    med = x.median(dim=dim, keepdim=True).values
    return (x - med).abs().median(dim=dim).values


def mad_rows(x, n_sample=1 << 20, seed=0, valid=None):
    """Per-row MAD of a (T, n) device tensor from the same `n_sample` random
    columns of every row (restricted to columns where `valid` (T, n) holds,
    when given, by resampling rows independently)."""
    import torch
    T, n = x.shape
    g = torch.Generator(device=x.device).manual_seed(int(seed))
    if valid is None:
        cols = torch.randint(0, n, (min(n_sample, n),), generator=g, device=x.device)
        return mad(x[:, cols], dim=1)
    out = torch.empty(T, device=x.device)
    for t in range(T):
        idx = valid[t].nonzero().view(-1)
        if idx.numel() < 100:
            out[t] = mad(x[t])
            continue
        pick = idx[torch.randint(0, idx.numel(), (min(n_sample, idx.numel()),), generator=g,
                                 device=x.device)]
        out[t] = mad(x[t, pick])
    return out


# ---- network --------------------------------------------------------------------

class Network:
    """4 x conv1d(k=9, same) with ReLU between layers; parameters in one flat
    float32 master vector, its FP16 rounding (what the network computes with)
    and a float32 copy of that rounding read by the kernels."""

    def __init__(self, device, seed=0):
        import torch
        self.device = device
        self.offsets = []
        off = 0
        for cin, cout in LAYERS:
            self.offsets.append((off, off + cout * cin * K))
            off += cout * cin * K + cout
        self.n_params = off
        g = torch.Generator(device="cpu").manual_seed(seed)
        master = torch.zeros(off, dtype=torch.float32)
        for (cin, cout), (w0, b0) in zip(LAYERS, self.offsets):
            std = math.sqrt(2.0 / (cin * K))
            master[w0:b0] = torch.randn(cout * cin * K, generator=g) * std
        self.master = master.to(device)
        self.w16 = self.master.half()
        self.w32 = self.w16.float()
        self.m = torch.zeros_like(self.master)
        self.v = torch.zeros_like(self.master)
        self.t = 0

    def _w(self, i):
        w0, b0 = self.offsets[i]
        cin, cout = LAYERS[i]
        return self.w32[w0:b0], self.w32[b0:b0 + cout]

    def forward(self, x, keep=False, stream=None, fast=False):
        """x (B, 8, X) -> prediction (B, X); with keep, also the layer inputs
        (pre-activation) for backward.  fast: the 32-channel layers on the
        matrix cores (FP16 activations; inference)."""
        import torch
        stream = stream or torch.cuda.current_stream(x.device)
        B, _, X = x.shape
        acts = [x]
        a = x
        lb = lib()
        scratch = None
        if fast:
            scratch = torch.empty(int(lb.fmf_conv1d_forward_fast_scratch_bytes(32)),
                                  dtype=torch.uint8, device=x.device)
        for i, (cin, cout) in enumerate(LAYERS):
            w, b = self._w(i)
            z = torch.empty((B, cout, X), dtype=torch.float32, device=x.device)
            if fast and cout == 32:
                _check(lb.fmf_conv1d_forward_fast(_p(a), B, cin, X, int(i > 0), _p(w), _p(b),
                                                  cout, _p(z), _p(scratch), scratch.numel(),
                                                  _s(stream)), "conv forward (matrix cores)")
            else:
                _check(lb.fmf_conv1d_forward(_p(a), B, cin, X, int(i > 0), _p(w), _p(b),
                                             cout, _p(z), _s(stream)), "conv forward")
            acts.append(z)
            a = z
        y = acts[-1][:, 0, :]
        return (y, acts) if keep else y

    def infer(self, x, stream=None):
        """Prediction (B, X) for x (B, 8, X), float32 or float16, with the
        fused inference kernel (input rounded to FP16, hidden activations
        FP16 in LDS, matrix cores for every layer)."""
        import torch
        stream = stream or torch.cuda.current_stream(x.device)
        x = x.contiguous()
        B, _, X = x.shape
        lb = lib()
        scratch = torch.empty(int(lb.fmf_screen_net_scratch_bytes()), dtype=torch.uint8,
                              device=x.device)
        pred = torch.empty((B, X), dtype=torch.float32, device=x.device)
        ptrs = []
        for i in range(len(LAYERS)):
            w, b = self._w(i)
            ptrs += [_p(w), _p(b)]
        if x.dtype not in (torch.float16, torch.float32):
            raise TypeError("screener input must be float32 or float16, not %s" % x.dtype)
        fn = lb.fmf_screen_net_forward_f16 if x.dtype == torch.float16 else lb.fmf_screen_net_forward
        _check(fn(_p(x), B, X, *ptrs, _p(scratch), scratch.numel(), _p(pred), _s(stream)),
               "screener network")
        return pred

    def backward(self, acts, gy, stream=None):
        """Gradient of the flat parameter vector for d loss / d prediction = gy (B, X)."""
        import torch
        stream = stream or torch.cuda.current_stream(gy.device)
        grad = torch.zeros(self.n_params, dtype=torch.float32, device=gy.device)
        gz = gy.reshape(gy.shape[0], 1, gy.shape[1]).contiguous()
        B, X = gy.shape
        for i in reversed(range(len(LAYERS))):
            cin, cout = LAYERS[i]
            w, _ = self._w(i)
            w0, b0 = self.offsets[i]
            ain = acts[i]
            ga = torch.empty((B, cin, X), dtype=torch.float32, device=gy.device) if i > 0 else None
            gw = grad[w0:b0]
            gb = grad[b0:b0 + cout]
            _check(lib().fmf_conv1d_backward(_p(gz), _p(ain), _p(ain), B, cin, X, int(i > 0),
                                             _p(w), cout, _p(ga) if ga is not None else None,
                                             _p(gw), _p(gb), _s(stream)), "conv backward")
            gz = ga
        return grad

    def adam(self, grad, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        import torch
        stream = stream or torch.cuda.current_stream(grad.device)
        self.t += 1
        _check(lib().fmf_adam_step(_p(self.master), _p(grad), _p(self.m), _p(self.v),
                                   self.n_params, lr, b1, b2, eps, self.t, _p(self.w16),
                                   _p(self.w32), _s(stream)), "adam")

    def loss_and_grad(self, pred, target, boost=1.0, stream=None):
        import torch
        stream = stream or torch.cuda.current_stream(pred.device)
        n = pred.numel()
        grad = torch.empty_like(pred)
        blocks = 256
        partial = torch.empty(blocks, dtype=torch.float32, device=pred.device)
        norm = float((1.0 + boost * target).sum().item())
        _check(lib().fmf_screen_loss(_p(pred.contiguous()), _p(target.contiguous()), n,
                                     float(boost), 1.0 / norm, _p(grad), _p(partial), blocks,
                                     _s(stream)), "loss")
        return partial.sum(), grad

    def state(self):
        return {"w16": self.w16.cpu().numpy()}

    def load(self, w16):
        import torch
        self.w16 = torch.as_tensor(np.asarray(w16, dtype=np.float16)).to(self.device)
        self.master = self.w16.float()
        self.w32 = self.w16.float()


# ---- synthetic days (device) -------------------------------------------------------

def synthetic_day(seed, device, n_templates=100, n_stations=20, n_components=3,
                  n_tp=800, n_data=8_640_000, snr=(0.1, 3.0), n_events=20,
                  max_moveout=2000, gap_fraction=0.05, gap_window=6000):
    """One day at the 'day' workload's shapes generated on the device: N(0,1)
    noise with 5 % of 6000-sample windows zeroed per channel, Hann-tapered
    Gaussian templates (unit RMS), moveouts U[0, max_moveout) per station,
    equal weights, `n_events` injections per template at amplitude
    log-uniform in `snr`.  Returns device tensors and the host event list."""
    import torch
    g = torch.Generator(device=device).manual_seed(int(seed))
    T, S, C, N, L = n_templates, n_stations, n_components, n_tp, n_data
    SC = S * C
    data = torch.randn((SC, L), generator=g, device=device, dtype=torch.float32)
    n_win = L // gap_window
    gaps = torch.rand((SC, n_win), generator=g, device=device) < gap_fraction
    mask = gaps.repeat_interleave(gap_window, dim=1)
    data[:, :n_win * gap_window].masked_fill_(mask, 0.0)
    hann = 0.5 - 0.5 * torch.cos(2 * math.pi * torch.arange(N, device=device) / (N - 1))
    tp = torch.randn((T, SC, N), generator=g, device=device) * hann
    tp = tp / tp.pow(2).mean(dim=-1, keepdim=True).sqrt()
    mo_s = torch.randint(0, max_moveout, (T, S), generator=g, device=device, dtype=torch.int32)
    mo = mo_s.repeat_interleave(C, dim=1).contiguous()
    w = torch.full((T, SC), 1.0 / SC, device=device)
    span = N + max_moveout
    lags = torch.randint(0, L - span - 1, (T, n_events), generator=g, device=device)
    amp = torch.exp(torch.empty((T, n_events), device=device).uniform_(
        math.log(snr[0]), math.log(snr[1]), generator=g))
    flat = data.view(-1)
    k = torch.arange(N, device=device)
    ch_off = (torch.arange(SC, device=device) * L)
    for t in range(T):
        start = lags[t][:, None] + mo[t][None, :].long()                 # (E, SC)
        idx = (ch_off[None, :, None] + start[:, :, None] + k[None, None, :]).reshape(-1)
        val = (amp[t][:, None, None] * tp[t][None, :, :]).reshape(-1)
        flat.index_add_(0, idx, val)
    events = [(t, int(l)) for t in range(T) for l in lags[t].tolist()]
    return dict(templates=tp.contiguous(), moveouts=mo, weights=w, data=data,
                events=events, n_tp=N)


def _sum_squares(day):
    """Templates' sums of squares (T, SC): the day's own 'sst' when it
    carries one (a caller that scanned the day with its own values gets the
    same bits), else computed from the templates."""
    if day.get("sst") is not None:
        return day["sst"]
    return day["templates"].pow(2).sum(dim=-1).contiguous()


def exact_scan(day, out=None, workspace=None, stream=None):
    """Full bf16x3 scan of a synthetic day: (T, n_corr) on the device."""
    import torch
    from .device import matched_filter_device
    tp, data = day["templates"], day["data"]
    T, SC, N = tp.shape
    L = data.shape[1]
    n_corr = L - N + 1
    sst = _sum_squares(day)
    if out is None:
        out = torch.empty((T, n_corr), dtype=torch.float32, device=data.device)
    matched_filter_device(tp.reshape(-1), sst.reshape(-1), day["moveouts"].reshape(-1),
                          day["weights"].reshape(-1), data.reshape(-1), 1, N, L, T, SC, 1,
                          n_corr, out, kernel=_lib.KERNEL_BF16X3, workspace=workspace,
                          stream=stream)
    return out


def screened_scan(day, net, gate, sample_fraction=0.02, margin=16, max_keep=0.5,
                  seed=0, stream=None, workspace=None, factor=COARSE, item_shape=ITEM_SHAPE):
    """Screened scan of a day: coarse scan -> features -> network -> gate ->
    exact bf16x3 on the kept tiles, and a stratified random sample of tiles
    per template for the MAD estimate (on the screening kernel where the
    gate did not keep them, item shapes 212 and 412).  Returns (cc: the
    exact cc on the kept tiles, zeros elsewhere; kept mask (T, n_tiles); MAD
    estimate (T); prediction; info with per-stage times and the kept and
    sample-only tile fractions).  Falls back to the full scan when more than
    `max_keep` of the tiles are kept or a row is shorter than one item."""
    import time
    import torch
    from .detect import median_mad
    tp, data = day["templates"], day["data"]
    T, SC, N = tp.shape
    L = data.shape[1]
    n_corr = L - N + 1
    dev = data.device
    stream = stream or torch.cuda.current_stream(dev)
    times = {}

    def mark(name, t0):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        times[name] = t1 - t0
        return t1

    t0 = time.perf_counter()
    cc_c, mad_c = coarse_scan(day, factor=factor, stream=stream)
    t0 = mark("coarse", t0)
    n_out = (n_corr + DEC - 1) // DEC
    feat = coarse_features(cc_c, mad_c, n_out, factor=factor, stream=stream, half=True)
    t0 = mark("features", t0)
    pred = net.infer(feat, stream=stream)                          # (T, n_out)
    t0 = mark("network", t0)
    lb = lib()
    tile_out = (lb.fmf_mfma_shape_outputs(int(item_shape)) if item_shape
                else lb.fmf_mfma_item_outputs(1))
    n_tiles = (n_corr + tile_out - 1) // tile_out
    keep = torch.empty((T, n_tiles), dtype=torch.int32, device=dev)
    thr = torch.full((T,), float(gate), dtype=torch.float32, device=dev)
    _check(lb.fmf_screen_gate(_p(pred), n_out, _p(thr), n_tiles, tile_out // DEC, margin, T,
                              _p(keep), _s(stream)), "gate")
    # the same number of whole tiles per template for the MAD: one drawn at
    # random from each of k equal strata of the row (no sort; the sample
    # covers the whole day)
    full_tiles = max(1, n_corr // tile_out)
    k = max(1, int(round(sample_fraction * full_tiles)))
    g = torch.Generator(device=dev).manual_seed(int(seed))
    i = torch.arange(k, device=dev)
    s_lo, s_hi = (i * full_tiles) // k, ((i + 1) * full_tiles) // k
    u = torch.rand((T, k), generator=g, device=dev)
    pick = torch.minimum(s_lo + (u * (s_hi - s_lo)).long(), s_hi - 1)
    sample = torch.zeros((T, n_tiles), dtype=torch.bool, device=dev)
    sample.scatter_(1, pick, True)
    kept = keep > 0
    sample_only = sample & ~kept
    frac = float(kept.float().mean().item())
    frac_sample = float(sample_only.float().mean().item())
    t0 = mark("gate", t0)
    out = torch.zeros((T, n_corr), dtype=torch.float32, device=dev)
    # rows shorter than one whole item leave no tile for the MAD sample: the
    # full scan (a day this short gains nothing from screening)
    if frac > max_keep or n_corr < tile_out:
        exact_scan(day, out=out, workspace=workspace, stream=stream)
        kept[:] = True
        fallback = True
    else:
        # items are tile * T + template, ascending (tile-major): the gate's
        # on the bf16x3 kernel; the sampled tiles the gate did not keep only
        # feed the MAD, on the screening kernel where the tile shape allows
        # (one bf16 product per MAC: its rounding, ~1e-5 RMS of cc, moves the
        # MAD by a relative ~1e-6), in the same call and energy pass
        to_items = lambda m: m.t().contiguous().view(-1).nonzero().view(-1).to(torch.int32)
        sst = _sum_squares(day)
        from .device import workspace_bytes
        if workspace is None:
            workspace = torch.empty(workspace_bytes(N, L, T, SC, 1, 0, _lib.KERNEL_BF16X3, 0),
                                    dtype=torch.uint8, device=dev)
        args = (_p(tp), _p(sst), _p(day["moveouts"]), _p(data), _p(day["weights"]), 1, N, L, T,
                SC, 1, n_corr, _p(out))
        if item_shape in MIXED_SHAPES:
            items, sitems = to_items(kept), to_items(sample_only)
            _check(lb.fmf_matched_filter_items_mixed_device(
                *args, _p(items), int(items.numel()), _p(sitems), int(sitems.numel()),
                int(item_shape), _p(workspace), workspace.numel(), _s(stream)), "items")
        else:
            items = to_items(kept | sample_only)
            _check(lb.fmf_matched_filter_items_device(
                *args, 0, 0, _p(items), int(items.numel()), int(item_shape or 0),
                _p(workspace), workspace.numel(), _s(stream)), "items")
        fallback = False
    t0 = mark("exact", t0)
    # MAD of cc per template: the whole rows after a fallback, else the
    # sampled tiles; then the output keeps the gate's tiles only
    if fallback:
        _, mads = median_mad(out)
    else:
        view = out[:, :full_tiles * tile_out].view(T, full_tiles, tile_out)
        samp = torch.gather(view, 1, pick[:, :, None].expand(-1, -1, tile_out)).reshape(T, -1)
        _, mads = median_mad(samp)
        # zero only those tiles: integer indices (a boolean mask with a scalar
        # becomes a masked fill that passes over the whole output)
        zt = sample_only[:, :full_tiles].nonzero()
        view[zt[:, 0], zt[:, 1]] = 0.0
    mark("mad", t0)
    info = dict(kept_fraction=1.0 if fallback else frac,
                sampled_fraction=0.0 if fallback else frac_sample, fallback=fallback,
                n_tiles=n_tiles, tile_out=tile_out, times=times, coarse_factor=factor)
    return out, kept, mads, pred, info
