"""Seeded sweep of small random configurations through the Python API against
the oracle's restatement of the reference (oracle/cpu.py, itself
bit-identical to the compiled reference on the golden cases):

  * arch 'hip' (FP32 kernel) and 'hip_mfma' (bf16x3) within 1e-4, network
    sum and per-channel layout;
  * exact_order=True bit for bit (short normalization);
  * 'full' normalization (archs 'precise' and 'hip_mfma') within 1e-4 of the
    reference's 'precise' arch;

over 1-4 templates, 1-5 stations, 1-3 components, template lengths from 1 to
1500 samples (past the MFMA kernel's reach, where its caller runs the FP32
kernel), steps 1-3, moveouts from -40 to +300, channels of weight 0 and
zeroed gaps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-4
SEEDS = list(range(48))


def _config(seed):
    rng = np.random.default_rng(1000 + seed)
    T = int(rng.integers(1, 5))
    S = int(rng.integers(1, 6))
    C = int(rng.integers(1, 4))
    N = int(rng.choice([1, 7, 37, 200, 333, 800, 1300, 1500]))
    step = int(rng.choice([1, 1, 2, 3]))
    mo = rng.integers(-40, 300, size=(T, S, C)).astype(np.int32)
    L = N + int(mo.max()) + int(rng.integers(300, 9000)) * step
    data = rng.standard_normal((S, C, L)).astype(np.float32)
    for _ in range(int(rng.integers(0, 3))):  # zero gaps on some channels
        s, c = int(rng.integers(S)), int(rng.integers(C))
        a = int(rng.integers(0, L))
        data[s, c, a:a + int(rng.integers(N, 3 * N + 100))] = 0.0
    templates = rng.standard_normal((T, S, C, N)).astype(np.float32)
    weights = rng.uniform(0.1, 1.0, size=(T, S, C)).astype(np.float32)
    if S * C > 1:
        weights[rng.random((T, S, C)) < 0.2] = 0.0
    return dict(templates=templates, moveouts=mo, weights=weights, data=data, step=step)


def _run(inp, **kw):
    import fast_matched_filter_amd as fmf
    assert fmf.HIP_LOADED, "libfmf_hip.so must load on an MI355X"
    return fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"], inp["data"],
                              inp["step"], check_zeros=False, **kw)


@pytest.mark.parametrize("seed", SEEDS)
def test_random_config_against_reference_restatement(seed, oracle_built):
    inp = _config(seed)
    T, S, C, N = inp["templates"].shape
    print("sweep seed %d: T=%d S=%d C=%d N=%d L=%d step=%d" %
          (seed, T, S, C, N, inp["data"].shape[-1], inp["step"]), flush=True)
    ref = oracle_built.run(inp["templates"], inp["moveouts"], inp["weights"], inp["data"],
                           inp["step"])
    for arch in ("hip", "hip_mfma"):
        got = _run(inp, arch=arch)
        assert got.shape == ref.shape
        err = float(np.abs(got - ref).max())
        assert err <= TOL, (arch, err)
    exact = _run(inp, arch="hip", exact_order=True)
    assert np.array_equal(exact.view(np.uint32), ref.view(np.uint32))
    if seed % 3 == 0:
        ref_full = oracle_built.run(inp["templates"], inp["moveouts"], inp["weights"],
                                    inp["data"], inp["step"], kind="precise", normalize="full")
        for arch in ("precise", "hip_mfma"):
            got = _run(inp, arch=arch, normalize="full")
            assert float(np.abs(got - ref_full).max()) <= TOL, arch
    if seed % 3 == 1:
        ref_ns = oracle_built.run(inp["templates"], inp["moveouts"], inp["weights"],
                                  inp["data"], inp["step"], network_sum=False)
        for arch in ("hip", "hip_mfma"):
            got = _run(inp, arch=arch, network_sum=False)
            assert got.shape == ref_ns.shape
            assert float(np.abs(got - ref_ns).max()) <= TOL, arch
