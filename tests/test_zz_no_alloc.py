"""The allocation-free entry point (fmf_matched_filter_device_noalloc) against
the workspace entry point (fmf_matched_filter_device), through the C ABI:
two calls in a row — the second on a smaller input, so it runs on what the
first left in the library's scratch array — each bit-identical to the
workspace call on the same inputs (same kernels, same order of operations);
a call whose workspace exceeds the scratch size is refused before anything
runs."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(name, **over):
    from fast_matched_filter_amd.synthetic import make_workload
    w = make_workload(name, **over)
    tp = w["templates"]
    return dict(tp=tp, sst=np.sum(tp ** 2, axis=-1).astype(np.float32), mo=w["moveouts"],
                w=w["weights"], data=w["data"])


def _both(inp, step, kernel, network_sum):
    import torch
    from fast_matched_filter_amd import _lib
    from fast_matched_filter_amd.device import matched_filter_device, workspace_bytes
    lib = _lib.require()
    dev = torch.device("cuda:0")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).to(dev)
    T, S, C, N = inp["tp"].shape
    L = inp["data"].shape[-1]
    n_corr = (L - N) // step + 1
    shape = (T, n_corr) if network_sum else (T, n_corr, S * C)
    args = [d(inp["tp"]), d(inp["sst"]), d(inp["mo"]), d(inp["data"]), d(inp["w"])]
    need = workspace_bytes(N, L, T, S, C, 0, kernel, 0)
    assert 0 < need <= int(lib.fmf_noalloc_scratch_bytes())
    ref = torch.full(shape, 7.0, device=dev)
    matched_filter_device(args[0], args[1], args[2], args[4], args[3], step, N, L, T, S, C,
                          n_corr, ref, network_sum=network_sum, kernel=kernel)
    out = torch.full(shape, -3.0, device=dev)
    stream = torch.cuda.current_stream(dev)
    rc = lib.fmf_matched_filter_device_noalloc(
        *[a.data_ptr() for a in args], step, N, L, T, S, C, n_corr, out.data_ptr(), 0,
        int(network_sum), kernel, 0, ct.c_void_p(stream.cuda_stream))
    assert rc == 0, lib.fmf_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_equals_workspace_call_twice():
    from fast_matched_filter_amd import _lib
    print("case: noalloc, small workload T=10 L=100000 bf16x3 step 1", flush=True)
    big = _inputs("small")
    out, ref = _both(big, 1, _lib.KERNEL_BF16X3, True)
    assert np.array_equal(out, ref)
    assert np.count_nonzero(out) > 0.9 * out.size
    print("case: noalloc, T=3 L=20000 FP32 step 2 per channel", flush=True)
    small = _inputs("small", n_templates=3, n_samples_data=20_000)
    out, ref = _both(small, 2, _lib.KERNEL_FP32, False)
    assert np.array_equal(out, ref)
    print("case: noalloc, T=3 L=20000 bf16x3 step 2", flush=True)
    out, ref = _both(small, 2, _lib.KERNEL_BF16X3, True)
    assert np.array_equal(out, ref)


def test_noalloc_refuses_oversized_call():
    from fast_matched_filter_amd import _lib
    lib = _lib.require()
    limit = int(lib.fmf_noalloc_scratch_bytes())
    # 60 channels x 400 000 samples: 96 MB of window energies alone
    L, SC, N, T = 400_000, 60, 800, 1
    need = int(lib.fmf_workspace_bytes(N, L, T, SC, 1, 0, 0, 0))
    assert need > limit
    one = ct.c_void_p(16)  # never dereferenced: the call is refused first
    rc = lib.fmf_matched_filter_device_noalloc(one, one, one, one, one, 1, N, L, T, SC, 1,
                                               L - N + 1, one, 0, 1, 0, 0, ct.c_void_p(0))
    assert rc == 1
    assert b"workspace too small" in lib.fmf_last_error()
