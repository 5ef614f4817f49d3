"""fmf_copy_to_host, the device -> pageable host copy of the host-array paths
(faulting the destination's pages in with host threads, then streaming
through pinned staging buffers of 64 MiB): bit-exact copies of sizes below
the staged threshold, of exactly one staging buffer, and of several buffers
with a partial last one; into fresh, pre-filled and page-misaligned
destinations; ordered after work queued on the stream without a
synchronisation; and the host-array batched entry point whose output it
copies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIB = 1 << 20


def _src(nbytes, seed):
    import torch
    n = nbytes // 4
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda",
                         generator=g)


@pytest.mark.parametrize("nbytes", [4, 4096, 31 * MIB + 12, 64 * MIB, 150 * MIB + 4 * 3001])
def test_copy_bit_exact_fresh(nbytes):
    from fast_matched_filter_amd.device import copy_to_host
    print("case: copy of %d bytes into a fresh array" % nbytes, flush=True)
    src = _src(nbytes, nbytes % 997)
    got = copy_to_host(src)
    assert got.dtype == np.int32 and got.shape == (nbytes // 4,)
    assert np.array_equal(got, src.cpu().numpy())


def test_copy_into_prefilled_and_misaligned():
    from fast_matched_filter_amd.device import copy_to_host
    nbytes = 96 * MIB + 4 * 7
    print("case: copy of %d bytes into a pre-filled, misaligned view" % nbytes, flush=True)
    src = _src(nbytes, 3)
    host = np.full(nbytes // 4 + 3, 7, dtype=np.int32)
    view = host[1:-2]                       # 4 bytes past the array's start
    copy_to_host(src, out=view)
    assert np.array_equal(view, src.cpu().numpy())
    assert host[0] == 7 and host[-1] == 7 and host[-2] == 7


def test_copy_ordered_after_queued_work():
    import torch
    from fast_matched_filter_amd.device import copy_to_host
    print("case: copy queued behind a fill on the same stream", flush=True)
    n = 80 * MIB // 4
    src = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for v in (1.0, 2.0, 3.0):
        src.fill_(v)                        # queued, not synchronised
        src[::4097].add_(v)
        got = copy_to_host(src)
        expect = np.full(n, v, dtype=np.float32)
        expect[::4097] += v
        assert np.array_equal(got, expect), v


def test_copy_size_mismatch_refused():
    import torch
    from fast_matched_filter_amd.device import copy_to_host
    src = torch.zeros(16, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        copy_to_host(src, out=np.empty(15, dtype=np.float32))


@pytest.mark.parametrize("chunks", [None, 3, 12])
def test_batched_host_array_output(chunks):
    """matched_filter_batched returns its rows through the staged copy, in
    one launch or in several whose copies overlap the next launch: equal to
    the device-resident result of one launch bit for bit."""
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd.synthetic import make_workload
    w = make_workload("batched", n_templates=12, n_stations=6, n_samples_data=2_000_000,
                      stations=(2, 5))
    print("case: matched_filter_batched T=12, S=6, L=2e6, host output, chunks %s" % chunks,
          flush=True)
    args = (w["templates"], w["lengths"], w["moveouts"], w["weights"], w["data"], 1)
    host = fmf.matched_filter_batched(*args, host_chunks=chunks)
    dev = fmf.matched_filter_batched(*args, return_device=True).cpu().numpy()
    assert host.nbytes > 32 * MIB                      # the staged path
    assert np.array_equal(host, dev)
