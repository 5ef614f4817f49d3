"""Screener kernels on the GPU (no reference counterpart; pinned by the FP64
oracle oracle/screener_ref.py, itself checked against finite differences):

  * conv1d forward and backward of the HIP kernels equal the oracle's
    (FP16-rounded weights, FP32 accumulation: <= 1e-4 relative);
  * the decimated features equal a NumPy restatement;
  * the exact scan restricted to a tile list equals the full scan on those
    tiles bit for bit and leaves the others untouched;
  * a few Adam steps of the HIP training loop lower the loss.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _net(seed=0):
    import torch
    from fast_matched_filter_amd import screener
    return screener.Network(torch.device("cuda:0"), seed=seed)


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


def test_items_equal_full_scan_on_listed_tiles():
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd.device import workspace_bytes
    from fast_matched_filter_amd import _lib
    dev = torch.device("cuda:0")
    day = screener.synthetic_day(6, dev, n_templates=5, n_stations=4, n_tp=400,
                                 n_data=120_000, max_moveout=500, n_events=3)
    full = screener.exact_scan(day)
    T, n_corr = full.shape
    lb = screener.lib()
    tile = lb.fmf_mfma_item_outputs(1)
    n_tiles = (n_corr + tile - 1) // tile
    rng = np.random.default_rng(0)
    keep = rng.random((n_tiles, T)) < 0.4
    items = torch.from_numpy(np.nonzero(keep.reshape(-1))[0].astype(np.int32)).to(dev)
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
                                            p(items), int(items.numel()), p(ws), ws.numel(),
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
