"""Screener kernels on the GPU (no reference counterpart; pinned by the FP64
oracle oracle/screener_ref.py, itself checked against finite differences):

  * conv1d forward and backward of the HIP kernels equal the oracle's
    (FP16-rounded weights, FP32 accumulation: <= 1e-4 relative);
  * the decimated features equal a NumPy restatement;
  * the exact scan restricted to a tile list equals the full scan on those
    tiles bit for bit and leaves the others untouched;
  * a few Adam steps of the HIP training loop lower the loss.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _net(seed=0):
    import torch
    from fast_matched_filter_amd import screener
    return screener.Network(torch.device("cuda:0"), seed=seed)


def _lib_check(rc):
    from fast_matched_filter_amd import _lib
    _lib.check(rc, "fmf_set_mfma_tiles")


def _params(net):
    from fast_matched_filter_amd import screener
    flat = net.w32.double().cpu().numpy()
    out = []
    for (cin, cout), (w0, b0) in zip(screener.LAYERS, net.offsets):
        out.append((flat[w0:b0].reshape(cout, cin, screener.K), flat[b0:b0 + cout]))
    return out


def test_conv_forward_backward_vs_oracle():
    import torch
    from oracle import screener_ref as ref
    net = _net(3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3, 8, 700)).astype(np.float32)
    target = np.abs(rng.standard_normal((3, 700))).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    pred, acts = net.forward(xd, keep=True)
    params = _params(net)
    p_ref, a_ref = ref.forward(x.astype(np.float64), params)
    got = pred.cpu().numpy()
    assert np.abs(got - p_ref).max() <= 1e-4 * max(1.0, np.abs(p_ref).max())
    loss, gy = net.loss_and_grad(pred, torch.from_numpy(target).cuda(), boost=1.0)
    assert abs(float(loss) - ref.loss(p_ref, target, 1.0)) <= 1e-4 * ref.loss(p_ref, target, 1.0)
    grad = net.backward(acts, gy).cpu().numpy()
    g_ref = ref.backward(a_ref, params, ref.loss_grad(p_ref, target.astype(np.float64), 1.0))
    for (w0, b0), (gw, gb), (cin, cout) in zip(net.offsets, g_ref, [(8, 32), (32, 32), (32, 32), (32, 1)]):
        gw_got = grad[w0:b0].reshape(gw.shape)
        gb_got = grad[b0:b0 + cout]
        assert np.abs(gw_got - gw).max() <= 2e-4 * max(1e-3, np.abs(gw).max())
        assert np.abs(gb_got - gb).max() <= 2e-4 * max(1e-3, np.abs(gb).max())


def test_features_match_numpy():
    import torch
    from fast_matched_filter_amd import screener
    day = screener.synthetic_day(5, torch.device("cuda:0"), n_templates=3, n_stations=4,
                                 n_tp=160, n_data=40_000, max_moveout=300, n_events=2)
    N = day["n_tp"]
    L = day["data"].shape[1]
    n_corr = L - N + 1
    feat = screener.features(day["data"], day["moveouts"], day["weights"], N, n_corr).cpu().numpy()
    d = day["data"].cpu().numpy().astype(np.float64)
    mo = day["moveouts"].cpu().numpy()
    w = day["weights"].cpu().numpy()
    n_dec = (L + 7) // 8
    dpad = np.zeros((d.shape[0], n_dec * 8))
    dpad[:, :L] = d
    env = (dpad ** 2).reshape(d.shape[0], n_dec, 8).mean(axis=2)
    nw = N // 8

    def box(e, width):
        c = np.concatenate([np.zeros((e.shape[0], 1)), np.cumsum(e, axis=1)], axis=1)
        q = np.arange(n_dec)
        hi = np.minimum(q + width, n_dec)
        return (c[:, hi] - c[:, q]) / (hi - q)

    win, winh = box(env, nw), box(env, nw // 2)
    chinv = 1.0 / env.mean(axis=1)
    n_out = (n_corr + 7) // 8
    j = np.arange(n_out)
    for t in range(3):
        q = np.clip(j[None, :] + mo[t][:, None] // 8, 0, n_dec - 1)
        r = win[np.arange(d.shape[0])[:, None], q] * chinv[:, None]
        f0 = (w[t][:, None] * r).sum(0) / w[t].sum()
        np.testing.assert_allclose(feat[t, 0], f0, rtol=1e-4, atol=1e-6)
        np.testing.assert_allclose(feat[t, 1], r.max(0), rtol=1e-4, atol=1e-6)
        assert np.all(feat[t, 7] == 1.0)


@pytest.mark.parametrize("junk,shape", [(False, 0), (True, 0), (True, 43)])
def test_items_equal_full_scan_on_listed_tiles(junk, shape):
    """Listed items equal the full scan bit for bit and nothing else is
    written; with `junk`, entries outside [0, n_tiles * T) (negative, one
    past the end, far past it) are mixed into the list and skipped, on the
    32x32 tiles (the default shape) and on the 16x16 tiles (shape 43)."""
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd.device import workspace_bytes
    from fast_matched_filter_amd import _lib
    dev = torch.device("cuda:0")
    day = screener.synthetic_day(6, dev, n_templates=5, n_stations=4, n_tp=400,
                                 n_data=120_000, max_moveout=500, n_events=3)
    lb = screener.lib()
    before = int(lb.fmf_get_mfma_tiles(1))
    if shape:
        _lib.check(lb.fmf_set_mfma_tiles(1, shape), "tiles")
    try:
        _items_case(day, junk, dev, lb, screener, workspace_bytes, _lib)
    finally:
        _lib.check(lb.fmf_set_mfma_tiles(1, before), "tiles")


def _items_case(day, junk, dev, lb, screener, workspace_bytes, _lib):
    import torch
    full = screener.exact_scan(day)
    T, n_corr = full.shape
    tile = lb.fmf_mfma_item_outputs(1)
    n_tiles = (n_corr + tile - 1) // tile
    rng = np.random.default_rng(0)
    keep = rng.random((n_tiles, T)) < 0.4
    listed = np.nonzero(keep.reshape(-1))[0].astype(np.int32)
    if junk:
        bad = np.array([-2**31, -1, n_tiles * T, n_tiles * T + 7, 2**31 - 1], dtype=np.int32)
        listed = np.concatenate([bad[:2], listed, bad[2:]])
    print("case: listed items, junk entries %s" % junk, flush=True)
    items = torch.from_numpy(listed).to(dev)
    out = torch.full((T, n_corr), 7.0, device=dev)
    tp = day["templates"]
    N = tp.shape[2]
    SC = tp.shape[1]
    L = day["data"].shape[1]
    sst = tp.pow(2).sum(-1).contiguous()
    ws = torch.empty(workspace_bytes(N, L, T, SC, 1, 0, _lib.KERNEL_BF16X3, 0), dtype=torch.uint8, device=dev)
    p = screener._p
    rc = lb.fmf_matched_filter_items_device(p(tp), p(sst), p(day["moveouts"]), p(day["data"]),
                                            p(day["weights"]), 1, N, L, T, SC, 1, n_corr, p(out), 0, 0,
                                            p(items), int(items.numel()), 0, p(ws), ws.numel(),
                                            screener._s(torch.cuda.current_stream()))
    _lib.check(rc, "items")
    o, f = out.cpu().numpy(), full.cpu().numpy()
    for k in range(n_tiles):
        sl = slice(k * tile, min((k + 1) * tile, n_corr))
        for t in range(T):
            if keep[k, t]:
                assert np.array_equal(o[t, sl], f[t, sl])
            else:
                assert np.all(o[t, sl] == 7.0)


def test_training_steps_reduce_loss():
    import torch
    net = _net(7)
    rng = np.random.default_rng(8)
    x = torch.from_numpy(rng.standard_normal((16, 8, 512)).astype(np.float32)).cuda()
    # a learnable target: a smoothed, rectified mix of two input channels
    tgt = torch.relu(x[:, 0] + 0.5 * x[:, 3]).contiguous()
    losses = []
    for _ in range(60):
        pred, acts = net.forward(x, keep=True)
        loss, gy = net.loss_and_grad(pred, tgt, boost=0.5)
        losses.append(float(loss))
        net.adam(net.backward(acts, gy), lr=3e-3)
    assert losses[-1] < 0.5 * losses[0]
    assert torch.equal(net.w16, net.master.half())


# ---- coarse-correlation front end and matrix-core forward --------------------

def test_decimate_matches_numpy():
    import torch
    from fast_matched_filter_amd import screener
    rng = np.random.default_rng(11)
    x = rng.standard_normal((5, 1003)).astype(np.float32)
    for D in (1, 2, 4, 7):
        got = screener.decimate(torch.from_numpy(x).cuda(), D).cpu().numpy()
        n_out = (1003 + D - 1) // D
        pad = np.zeros((5, n_out * D), np.float32)
        pad[:, :1003] = x
        want = pad.reshape(5, n_out, D).astype(np.float64).mean(axis=2)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


def test_decimate_vector_path_matches_numpy():
    """Rows whose length is a multiple of 4 with D a multiple of 4 take the
    16-byte-load path (a partial last group falls back to the scalar loop)."""
    import torch
    from fast_matched_filter_amd import screener
    rng = np.random.default_rng(13)
    n = 1000
    x = rng.standard_normal((4, n)).astype(np.float32)
    for D in (4, 8, 12):
        got = screener.decimate(torch.from_numpy(x).cuda(), D).cpu().numpy()
        n_out = (n + D - 1) // D
        pad = np.zeros((4, n_out * D), np.float32)
        pad[:, :n] = x
        want = pad.reshape(4, n_out, D).astype(np.float64).mean(axis=2)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("D", [4, 8, 16])
def test_coarse_features_match_numpy(D):
    import torch
    from fast_matched_filter_amd import screener
    rng = np.random.default_rng(12)
    T, nc = 3, 5000
    cc = (rng.standard_normal((T, nc)) * 0.01).astype(np.float32)
    cc[1, 2000:2003] = 0.3
    mad = np.array([0.006, 0.007, 0.008], np.float32)
    n_out = (nc * D + 7) // 8
    f = screener.coarse_features(torch.from_numpy(cc).cuda(), torch.from_numpy(mad).cuda(),
                                 n_out, factor=D).cpu().numpy()
    ext = np.concatenate([np.zeros((T, 64)), cc.astype(np.float64), np.zeros((T, 64))], axis=1)
    for t in range(T):
        for j in sorted({0, 1, 999, 1000, 1001, n_out // 2, n_out - 1}):
            q0, q1 = 8 * j // D, (8 * j + 7) // D
            span = (8 + D - 1) // D
            loc = ext[t, 64 + q0:64 + q1 + 1]
            win = ext[t, 64 + q0 - 3 * span:64 + q1 + 3 * span + 1]
            want = [np.abs(loc).max(), max(loc.max(), 0), max(-loc.min(), 0), np.abs(win).max(),
                    np.abs(win).mean(), np.sqrt((win ** 2).mean())]
            np.testing.assert_allclose(f[t, :6, j], np.log1p(np.array(want) / mad[t]),
                                       rtol=2e-5, atol=2e-6)
            assert abs(f[t, 6, j] - (np.log(mad[t]) + 5)) < 1e-5 and f[t, 7, j] == 1.0
    # the FP16 variant stores the round-to-nearest FP16 of the same values
    h = screener.coarse_features(torch.from_numpy(cc).cuda(), torch.from_numpy(mad).cuda(),
                                 n_out, factor=D, half=True)
    assert h.dtype == torch.float16
    assert np.array_equal(h.cpu().numpy().view(np.uint16), f.astype(np.float16).view(np.uint16))


def test_fast_forward_matches_fp32_forward():
    """Matrix-core layers (FP16 activations) against the FP32 kernels and the
    FP64 oracle on FP16-rounded inputs: the rounding of activations to FP16
    bounds the difference (<= 5e-3 of the output scale)."""
    import torch
    from oracle import screener_ref as ref
    net = _net(13)
    rng = np.random.default_rng(14)
    x = rng.standard_normal((4, 8, 3000)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    slow = net.forward(xd).cpu().numpy()
    fast = net.forward(xd, fast=True).cpu().numpy()
    scale = np.abs(slow).max()
    assert np.abs(fast - slow).max() <= 5e-3 * scale
    p_ref, _ = ref.forward(x.astype(np.float16).astype(np.float64), _params(net))
    assert np.abs(fast - p_ref).max() <= 5e-3 * scale


def test_fast_conv_layer_exact_on_fp16_inputs():
    """One matrix-core layer on inputs that are exact in FP16: equals the FP64
    convolution of the same values to FP32 accumulation error."""
    import torch
    from fast_matched_filter_amd import screener
    from oracle import screener_ref as ref
    rng = np.random.default_rng(15)
    lb = screener.lib()
    for cin in (8, 32):
        x = (rng.integers(-64, 65, (3, cin, 700)) / 32.0).astype(np.float32)
        w = (rng.integers(-32, 33, (32, cin, 9)) / 64.0).astype(np.float32)
        b = rng.standard_normal(32).astype(np.float32)
        out = torch.empty((3, 32, 700), dtype=torch.float32, device="cuda")
        sc = torch.empty(int(lb.fmf_conv1d_forward_fast_scratch_bytes(cin)), dtype=torch.uint8,
                         device="cuda")
        xd, wd, bd = (torch.from_numpy(a).cuda() for a in (x, w, b))
        screener._check(lb.fmf_conv1d_forward_fast(screener._p(xd), 3, cin, 700, 0,
                                                   screener._p(wd), screener._p(bd), 32,
                                                   screener._p(out), screener._p(sc), sc.numel(),
                                                   screener._s(torch.cuda.current_stream())),
                        "fast conv")
        want = ref.conv(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64))
        assert np.abs(out.cpu().numpy() - want).max() <= 1e-4 * max(1.0, np.abs(want).max())


def test_coarse_scan_is_matched_filter_of_decimated_inputs():
    """The coarse scan equals the bf16x3 matched filter of the decimated
    arrays within the screening kernel's bound (2^-8 per unit weight; the
    weights sum to 1), and its MAD is the exact median of |cc - median|."""
    import torch
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd import screener
    dev = torch.device("cuda:0")
    day = screener.synthetic_day(21, dev, n_templates=3, n_stations=4, n_tp=160,
                                 n_data=60_000, max_moveout=300, n_events=2)
    cc_c, mad_c = screener.coarse_scan(day)
    D = screener.COARSE
    tp = day["templates"].cpu().numpy()
    T, SC, N = tp.shape
    data = day["data"].cpu().numpy()
    dec = lambda a: np.pad(a, [(0, 0)] * (a.ndim - 1) + [(0, (-a.shape[-1]) % D)]).reshape(
        a.shape[:-1] + (-1, D)).astype(np.float64).mean(-1).astype(np.float32)
    mo = (day["moveouts"].cpu().numpy() + D // 2) // D
    ref = fmf.matched_filter(dec(tp)[:, :, None, :].reshape(T, SC, 1, -1), mo.reshape(T, SC, 1),
                             day["weights"].cpu().numpy().reshape(T, SC, 1),
                             dec(data).reshape(SC, 1, -1), 1, arch="hip_mfma", check_zeros=False)
    got = cc_c.cpu().numpy()
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 2.0 ** -8 + 1e-5
    for t in range(T):
        m = np.float32(np.median(got[t]))
        assert mad_c[t].item() == np.float32(np.median(np.abs(got[t] - m)))


def test_screened_scan_paths():
    """gate -inf: the full scan (fallback) and its exact MAD; gate +inf:
    nothing kept, a zero output, and the MAD of the sampled tiles, computed
    on the screening kernel, within 2e-3 of the exact MAD of the same tiles
    (the sample's own sampling error is separate); a middle gate: the kept
    tiles equal the full scan bit for bit (the items' tile shape 212 and the
    full scan's 422 sum every output in the same order; the test runs on 422
    whatever an earlier test left) and every other tile is zero."""
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd.detect import median_mad
    dev = torch.device("cuda:0")
    day = screener.synthetic_day(22, dev, n_templates=4, n_stations=4, n_tp=200,
                                 n_data=400_000, max_moveout=300, n_events=3)
    net = _net(23)
    lib = screener.lib()

    def body():
        full = screener.exact_scan(day)
        out, computed, mads, pred, info = screener.screened_scan(day, net, -1e30)
        assert info["fallback"] and torch.equal(out, full)
        # the items run on their own tile shape (2048 lags per item), the
        # process-wide shape untouched
        assert info["tile_out"] == 2048 and int(lib.fmf_get_mfma_tiles(1)) == 422
        assert torch.equal(mads, median_mad(full)[1])
        out, computed, mads, pred, info = screener.screened_scan(day, net, 1e30,
                                                                 sample_fraction=0.2,
                                                                 max_keep=0.9)
        assert not info["fallback"] and not bool(computed.any()) and not bool(out.any())
        assert info["kept_fraction"] == 0.0 and 0.1 < info["sampled_fraction"] < 0.3
        # the same tiles from the full scan: which ones the seed-0 draw picked
        tile = info["tile_out"]
        T, n_corr = full.shape
        full_tiles = n_corr // tile
        k = max(1, int(round(0.2 * full_tiles)))
        g = torch.Generator(device=dev).manual_seed(0)
        i = torch.arange(k, device=dev)
        s_lo, s_hi = (i * full_tiles) // k, ((i + 1) * full_tiles) // k
        u = torch.rand((T, k), generator=g, device=dev)
        pick = torch.minimum(s_lo + (u * (s_hi - s_lo)).long(), s_hi - 1)
        view = full[:, :full_tiles * tile].view(T, full_tiles, tile)
        samp = torch.gather(view, 1, pick[:, :, None].expand(-1, -1, tile)).reshape(T, -1)
        exact_mads = median_mad(samp)[1]
        assert float(((mads - exact_mads).abs() / exact_mads).max()) <= 2e-3
        # a middle gate: the median of the tiles' maxima keeps about half
        from fast_matched_filter_amd import screener_train as st
        n_tiles = (n_corr + tile - 1) // tile
        thr = float(st._tile_max(pred, tile // screener.DEC, n_tiles, 16).median())
        out, computed, mads, pred, info = screener.screened_scan(day, net, thr,
                                                                 sample_fraction=0.2,
                                                                 max_keep=0.9)
        assert not info["fallback"]
        o, f, c = out.cpu().numpy(), full.cpu().numpy(), computed.cpu().numpy()
        assert 0 < c.mean() < 0.9
        for t in range(o.shape[0]):
            for kk in range(c.shape[1]):
                sl = slice(kk * tile, (kk + 1) * tile)
                if c[t, kk]:
                    assert np.array_equal(o[t, sl], f[t, sl])
                else:
                    assert not o[t, sl].any()

    old = int(lib.fmf_get_mfma_tiles(1))
    _lib_check(lib.fmf_set_mfma_tiles(1, 422))
    try:
        body()
    finally:
        _lib_check(lib.fmf_set_mfma_tiles(1, old))


def test_screened_scan_uses_the_days_sums_of_squares():
    """A day that carries its own template sums of squares (here summed on
    the host in FP64, then rounded) is scanned with them: the kept items
    equal the device entry point's full scan with the same values bit for
    bit, on the (stations, components) layout of that call."""
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd.device import matched_filter_device
    dev = torch.device("cuda:0")
    day = screener.synthetic_day(27, dev, n_templates=4, n_stations=4, n_components=3, n_tp=200,
                                 n_data=400_000, max_moveout=300, n_events=3)
    T, SC, N = day["templates"].shape
    L = day["data"].shape[1]
    n_corr = L - N + 1
    tp64 = day["templates"].double().cpu().numpy()
    day["sst"] = torch.from_numpy((tp64 ** 2).sum(axis=-1).astype(np.float32)).to(dev)
    lib = screener.lib()
    old = int(lib.fmf_get_mfma_tiles(1))
    _lib_check(lib.fmf_set_mfma_tiles(1, 422))
    try:
        full = torch.empty((T, n_corr), dtype=torch.float32, device=dev)
        matched_filter_device(day["templates"].reshape(-1), day["sst"].reshape(-1),
                              day["moveouts"].reshape(-1), day["weights"].reshape(-1),
                              day["data"].reshape(-1), 1, N, L, T, SC // 3, 3, n_corr, full,
                              kernel=1)
        net = _net(28)
        _, _, _, pred, info = screener.screened_scan(day, net, 1e30, sample_fraction=0.2,
                                                     max_keep=0.9)
        # a gate at the median of the tiles' maxima keeps about half of them
        from fast_matched_filter_amd import screener_train as st
        thr = float(st._tile_max(pred, info["tile_out"] // screener.DEC, info["n_tiles"],
                                 16).median())
        out, computed, _, _, info = screener.screened_scan(day, net, thr, sample_fraction=0.2,
                                                           max_keep=0.9)
    finally:
        _lib_check(lib.fmf_set_mfma_tiles(1, old))
    assert not info["fallback"]
    mask = computed.repeat_interleave(info["tile_out"], dim=1)[:, :n_corr]
    assert 0 < float(mask.float().mean()) < 0.9
    assert torch.equal(out[mask], full[mask])
    assert not out[~mask].any()


def test_screened_scan_short_day_falls_back():
    """Rows shorter than one item (2048 lags): the full scan, its exact MAD."""
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd.detect import median_mad
    dev = torch.device("cuda:0")
    day = screener.synthetic_day(25, dev, n_templates=3, n_stations=2, n_tp=200,
                                 n_data=1_800, max_moveout=100, n_events=1)
    net = _net(26)
    out, computed, mads, pred, info = screener.screened_scan(day, net, 1e30)
    full = screener.exact_scan(day)
    assert info["fallback"] and bool(computed.all())
    assert torch.allclose(out, full, atol=1e-6) and out.shape == full.shape
    assert torch.allclose(mads, median_mad(full)[1], atol=1e-7)


def test_fused_inference_matches_layer_by_layer():
    """The fused kernel against the layer-by-layer matrix-core path (same FP16
    rounding of hidden activations except the last layer's input, which the
    layer-by-layer path keeps in FP32) and the FP64 oracle, including the
    sequence ends (zero padding of every layer) and lengths that are not a
    multiple of the block."""
    import torch
    from oracle import screener_ref as ref
    net = _net(31)
    rng = np.random.default_rng(32)
    for X in (1, 223, 224, 1000, 4099):
        x = rng.standard_normal((3, 8, X)).astype(np.float32)
        xd = torch.from_numpy(x).cuda()
        fused = net.infer(xd).cpu().numpy()
        fast = net.forward(xd, fast=True).cpu().numpy()
        p_ref, _ = ref.forward(x.astype(np.float16).astype(np.float64), _params(net))
        scale = max(1.0, np.abs(p_ref).max())
        assert fused.shape == (3, X)
        assert np.abs(fused - fast).max() <= 5e-3 * scale, X
        assert np.abs(fused - p_ref).max() <= 5e-3 * scale, X
        # FP16 input: the kernel rounds float input to FP16 when it stages it,
        # so the FP16 rounding of the same features gives the same bits
        half = net.infer(xd.half()).cpu().numpy()
        assert np.array_equal(half.view(np.uint32), fused.view(np.uint32)), X


def test_trained_network_gain_over_coarse_gate():
    """What the trained network gains over the method without it: on a
    held-out full-size day (seed 5000), the committed network and the coarse
    |cc| feature alone, each at its own gate calibrated for 99.7 % of the
    detected events on training days.  Both keep >= 99 % of the events the
    exact scan detects; the network needs at most 3/4 of the coarse-only
    gate's tiles (it reads the coarse cc's context: neighbouring positions
    and signed extremes)."""
    import os
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd import screener_train as st
    from fast_matched_filter_amd.detect import median_mad
    path = os.path.join(os.path.dirname(screener.__file__), "screener_weights.npz")
    w = np.load(path)
    dev = torch.device("cuda:0")
    net = screener.Network(dev)
    net.load(w["w16"])
    print("held-out day 5000, full size", flush=True)
    day = screener.synthetic_day(5000, dev)
    cc = screener.exact_scan(day)
    _, mads = median_mad(cc)
    T, n_corr = cc.shape
    tile_out = screener.lib().fmf_mfma_item_outputs(1)
    n_tiles = (n_corr + tile_out - 1) // tile_out
    tile_dec = tile_out // screener.DEC
    evs = st._event_windows(cc, mads, day["events"], tile_out)
    assert len(evs) > 1500
    feat, _, _ = screener.day_features(day, n_corr, factor=int(w["coarse"]))
    pred = net.infer(feat)
    keep_net = (st._tile_max(pred, tile_dec, n_tiles, 16) > float(w["gate"])).cpu().numpy()
    keep_co = (st._tile_max(feat[:, 0].contiguous(), tile_dec, n_tiles, 16) >
               float(w["gate_coarse_only"])).cpu().numpy()
    rec_net = np.mean([keep_net[t, k] for t, _, k in evs])
    rec_co = np.mean([keep_co[t, k] for t, _, k in evs])
    print("network: recall %.4f, tiles %.4f; coarse only: recall %.4f, tiles %.4f" %
          (rec_net, keep_net.mean(), rec_co, keep_co.mean()), flush=True)
    assert rec_net >= 0.99 and rec_co >= 0.99
    assert keep_net.mean() <= 0.75 * keep_co.mean()


@pytest.mark.parametrize("n_out,tile_dec,margin", [(10_000, 512, 16), (4_097, 256, 0),
                                                   (300, 512, 16), (70_001, 512, 3)])
def test_gate_kernel_against_numpy(n_out, tile_dec, margin):
    """fmf_screen_gate keeps tile (t, k) iff some prediction in
    [k * tile_dec - margin, (k + 1) * tile_dec + margin) (clipped to the row)
    exceeds thr[t] strictly: rows ending inside a tile, margin 0, values equal
    to the threshold."""
    import torch
    from fast_matched_filter_amd import screener
    rng = np.random.default_rng(n_out + tile_dec + margin)
    T = 7
    pred = rng.standard_normal((T, n_out)).astype(np.float32) - 4.0
    n_tiles = (n_out + tile_dec - 1) // tile_dec
    # sparse peaks, some exactly at the threshold
    thr = rng.uniform(-0.5, 0.5, T).astype(np.float32)
    for t in range(T):
        for x in rng.integers(0, n_out, size=max(1, n_tiles // 3)):
            pred[t, x] = thr[t] + (0.0 if rng.random() < 0.3 else 0.25)
    dev = torch.device("cuda:0")
    d_pred = torch.from_numpy(pred).to(dev)
    d_thr = torch.from_numpy(thr).to(dev)
    keep = torch.full((T, n_tiles), -1, dtype=torch.int32, device=dev)
    lib = screener.lib()
    print("gate kernel: T=%d n_out=%d tile_dec=%d margin=%d" % (T, n_out, tile_dec, margin),
          flush=True)
    screener._check(lib.fmf_screen_gate(screener._p(d_pred), n_out, screener._p(d_thr), n_tiles,
                                        tile_dec, margin, T, screener._p(keep),
                                        screener._s(torch.cuda.current_stream(dev))), "gate")
    got = keep.cpu().numpy()
    exp = np.zeros((T, n_tiles), dtype=np.int32)
    for t in range(T):
        for k in range(n_tiles):
            lo, hi = max(k * tile_dec - margin, 0), min((k + 1) * tile_dec + margin, n_out)
            exp[t, k] = int(np.any(pred[t, lo:hi] > thr[t]))
    assert np.array_equal(got, exp)
    assert 0 < exp.mean() < 1


def test_mixed_items_entry_point():
    """fmf_matched_filter_items_mixed_device: the bf16x3 list equals the full
    scan bit for bit, the screening list is within 2^-8 * sum |w| of it (one
    bf16 product per MAC, 1/sqrt of the energies taken in the kernel), other
    tiles keep the caller's values; both tile shapes; shapes whose screening
    blocks cover other lags are refused."""
    import torch
    from fast_matched_filter_amd import screener, _lib
    from fast_matched_filter_amd.device import workspace_bytes
    dev = torch.device("cuda:0")
    day = screener.synthetic_day(29, dev, n_templates=5, n_stations=4, n_tp=300,
                                 n_data=200_000, max_moveout=400, n_events=3)
    tp, data, mo, w = day["templates"], day["data"], day["moveouts"], day["weights"]
    T, SC, N = tp.shape
    L = data.shape[1]
    n_corr = L - N + 1
    sst = tp.pow(2).sum(-1).contiguous()
    lib = screener.lib()
    old = int(lib.fmf_get_mfma_tiles(1))
    _lib_check(lib.fmf_set_mfma_tiles(1, 422))
    try:
        full = screener.exact_scan(day)
    finally:
        _lib_check(lib.fmf_set_mfma_tiles(1, old))
    ws = torch.empty(workspace_bytes(N, L, T, SC, 1, 0, _lib.KERNEL_BF16X3, 0), dtype=torch.uint8,
                     device=dev)
    p = screener._p
    bound = 2.0 ** -8 * float(w.abs().sum(dim=1).max()) + 1e-6
    for shape in (212, 412):
        tile = lib.fmf_mfma_shape_outputs(shape)
        n_tiles = (n_corr + tile - 1) // tile
        rng = np.random.default_rng(shape)
        kind = rng.integers(0, 3, size=(n_tiles, T))  # 0 untouched, 1 bf16x3, 2 screening
        ex = torch.from_numpy(np.nonzero(kind.reshape(-1) == 1)[0].astype(np.int32)).to(dev)
        sc = torch.from_numpy(np.nonzero(kind.reshape(-1) == 2)[0].astype(np.int32)).to(dev)
        out = torch.full((T, n_corr), 7.0, device=dev)
        print("mixed items: shape %d, %d + %d items" % (shape, ex.numel(), sc.numel()), flush=True)
        _lib.check(lib.fmf_matched_filter_items_mixed_device(
            p(tp), p(sst), p(mo), p(data), p(w), 1, N, L, T, SC, 1, n_corr, p(out),
            p(ex), int(ex.numel()), p(sc), int(sc.numel()), shape, p(ws), ws.numel(),
            screener._s(torch.cuda.current_stream())), "mixed items")
        o, f = out.cpu().numpy(), full.cpu().numpy()
        for k in range(n_tiles):
            sl = slice(k * tile, min((k + 1) * tile, n_corr))
            for t in range(T):
                if kind[k, t] == 1:
                    assert np.array_equal(o[t, sl], f[t, sl]), (shape, k, t)
                elif kind[k, t] == 2:
                    assert np.abs(o[t, sl] - f[t, sl]).max() <= bound, (shape, k, t)
                else:
                    assert np.all(o[t, sl] == 7.0), (shape, k, t)
    rc = lib.fmf_matched_filter_items_mixed_device(
        p(tp), p(sst), p(mo), p(data), p(w), 1, N, L, T, SC, 1, n_corr, p(out),
        p(ex), int(ex.numel()), p(sc), int(sc.numel()), 422, p(ws), ws.numel(),
        screener._s(torch.cuda.current_stream()))
    assert rc != 0
