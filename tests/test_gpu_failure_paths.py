"""Error paths of the host-pointer call on the GPU, forced from the host with
the fault-injection switch (fmf_debug_inject), so every kernel still exits
normally: the nth device allocation of a call fails as out-of-memory for
n = 1..6 (inputs, workspace, outputs) and the injected stream creation fails
as a device error.  Each failure raises the mapped Python exception, the
next call gives the same output as before, and the device memory the calls
held is returned (free memory back to its level before the failures)."""
import numpy as np
import pytest

from tests.golden import cases

pytestmark = pytest.mark.gpu


def test_injected_failures_release_everything():
    import torch
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd import _lib
    lib = _lib.require()
    inp = cases.small()
    args = (inp["templates"], inp["moveouts"], inp["weights"], inp["data"], inp["step"])
    print("case: reference call before injection", flush=True)
    ref = fmf.matched_filter(*args, arch="hip_mfma", check_zeros=False)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for nth in range(1, 7):
        print("case: allocation %d fails" % nth, flush=True)
        assert lib.fmf_debug_inject(_lib.INJECT_ALLOC, nth) == 0
        with pytest.raises(MemoryError, match="injected"):
            fmf.matched_filter(*args, arch="hip_mfma", check_zeros=False)
        assert lib.fmf_debug_inject(_lib.INJECT_ALLOC, 0) == 0
    print("case: stream creation fails", flush=True)
    assert lib.fmf_debug_inject(_lib.INJECT_STREAM, 1) == 0
    with pytest.raises(RuntimeError, match="injected"):
        fmf.matched_filter(*args, arch="hip", check_zeros=False)
    out = fmf.matched_filter(*args, arch="hip_mfma", check_zeros=False)
    assert np.array_equal(out, ref)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert abs(free1 - free0) <= 64 << 20, (free0, free1)
