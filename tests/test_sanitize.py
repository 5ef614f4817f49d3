"""Sanitizer builds run on the CPU (no GPU needed):

* the oracle's C restatement driven on edge-case inputs under ASan + UBSan
  (oracle/sanitize_driver.c, `make -C oracle sanitize`);
* the library's host driver (fmf_api.hip, host code instrumented with ASan +
  UBSan) fed hostile arguments through every C entry point
  (tests/native/host_checks.cpp, `make -C tests/native`): each call is
  refused with its documented status before it touches a device;
* the fault-injection switch from Python: an injected stream failure comes
  back as RuntimeError before any device call, and the switch rejects
  unknown points.
"""
import os
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=timeout, env=env)
    assert r.returncode == 0, "%s failed (%d):\n%s\n%s" % (cmd, r.returncode, r.stdout[-3000:],
                                                          r.stderr[-3000:])
    return r.stdout


def test_sanitized_oracle():
    _run(["make", "-s", "-C", "oracle", "sanitize"])
    out = _run([os.path.join(ROOT, "oracle", "_build", "sanitize_oracle")])
    assert "all checks hold" in out


def test_sanitized_host_checks():
    _run(["make", "-s", "-j8", "-C", "tests/native"], timeout=900)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    out = _run([os.path.join(ROOT, "_build", "sanitize", "host_checks")], env=env)
    assert "all statuses as documented" in out


def test_injected_stream_failure_raises():
    import numpy as np
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd import _lib
    lib = _lib.require()
    tp = np.ones((1, 1, 1, 4), dtype=np.float32)
    d = np.arange(32, dtype=np.float32).reshape(1, 1, 32)
    m = np.zeros((1, 1), dtype=np.int32)
    w = np.ones((1, 1), dtype=np.float32)
    assert lib.fmf_debug_inject(_lib.INJECT_STREAM, 1) == 0
    with pytest.raises(RuntimeError, match="injected"):
        fmf.matched_filter(tp, m, w, d, 1, arch="hip", check_zeros=False)
    assert lib.fmf_debug_inject(7, 1) == 1
    assert lib.fmf_debug_inject(_lib.INJECT_ALLOC, 0) == 0
