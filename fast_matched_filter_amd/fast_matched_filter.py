"""Python API of the MI355X matched-filter engine.

``matched_filter`` keeps the reference's signature and behaviour
(fast_matched_filter/fast_matched_filter.py:66-362): the same keyword
arguments, input formats, output layouts, zero-tail convention and
``check_zeros`` messages.  Every ``arch`` runs on the GPU through
libfmf_hip.so:

=================  =====================================================
arch               kernel and semantics
=================  =====================================================
``'hip'``          FP32 vector-ALU kernel; threshold before the sqrt
``'hip_mfma'``     bf16x3 split-precision MFMA kernel, FP32 accumulation
``'gpu'``          same as ``'hip'`` (the reference's GPU arch)
``'cpu'``          same as ``'hip'``; 'full' is refused as in the reference
``'precise'``      FP32 kernel with the threshold after the sqrt
                   (matched_filter.c:650-655)
``'variable_precise'``  per-channel template lengths (``n_samples_template``),
                   Fisher-z network combination (matched_filter.c:321-412)
=================  =====================================================

``exact_order=True`` (FP32 kernel, short normalization) reproduces the
reference CPU build's float rounding sequence, giving bit-identical results.
"""
import ctypes as ct
import warnings

import numpy as np

from . import _lib

__all__ = ["matched_filter", "matched_filter_batched", "prepare_inputs", "test_matched_filter",
           "ARCHS"]

ARCHS = ("hip", "hip_mfma", "gpu", "cpu", "precise", "variable_precise")

_F = ct.POINTER(ct.c_float)
_I = ct.POINTER(ct.c_int)


def prepare_inputs(templates, moveouts, weights, data, step):
    """Input normalisation of fast_matched_filter.py:179-274.

    Returns None when the dimensions are incompatible (after printing the
    reference's message), else a dict of C-contiguous float32/int32 arrays
    and the sizes T, S, C, N, L, n_corr.
    """
    impossible = False
    if templates.ndim > data.ndim:
        n_templates = int(templates.shape[0])
        assert templates.shape[1] == data.shape[0]  # stations
        n_stations = int(templates.shape[1])
        if templates.ndim == 4:
            assert templates.shape[2] == data.shape[1]  # components
            n_components = int(templates.shape[2])
        elif templates.ndim == 3:
            n_components = 1
        else:
            impossible = True
    elif templates.ndim == data.ndim:
        n_templates = 1
        assert templates.shape[0] == data.shape[0]
        n_stations = int(templates.shape[0])
        if templates.ndim == 3:
            assert templates.shape[1] == data.shape[1]
            n_components = int(templates.shape[1])
        elif templates.ndim == 2:
            n_components = 1
        else:
            impossible = True
    else:
        impossible = True
    if impossible:
        print("Template and data dimensions are not compatible!")
        return None

    n_samples_template = int(templates.shape[-1])
    T, S, C, N = n_templates, n_stations, n_components, n_samples_template
    templates = templates.reshape(T, S, C, N)
    n_samples_data = int(data.shape[-1])
    data = data.reshape(S, C, n_samples_data)

    moveouts = np.asarray(moveouts)
    weights = np.asarray(weights)
    assert moveouts.shape == weights.shape
    full = T * S * C
    # (T, S) -> (T, S, C) by repeating each entry over components
    if moveouts.shape != (T, S, C):
        if full / moveouts.size == C:
            moveouts = np.repeat(moveouts, C).reshape(T, S, C)
        elif full / moveouts.size == 1.0:
            moveouts = moveouts.reshape(T, S, C)
    if weights.shape != (T, S, C):
        if full / weights.size == C:
            weights = np.repeat(weights, C).reshape(T, S, C)
        elif full / weights.size == 1.0:
            weights = weights.reshape(T, S, C)

    n_corr = int((n_samples_data - N) / step + 1)
    # sum of squares in the input dtype, then float32 (:256-261)
    sst = np.sum(templates ** 2, axis=-1)
    # Flattened as the reference flattens them (:258-268), without a copy
    # where the input already is C-contiguous in the right dtype (the library
    # only reads these arrays).
    return dict(
        templates=np.ascontiguousarray(templates.reshape(-1), dtype=np.float32),
        sst=np.ascontiguousarray(sst.reshape(-1), dtype=np.float32),
        moveouts=np.ascontiguousarray(moveouts.reshape(-1), dtype=np.int32),
        weights=np.ascontiguousarray(weights.reshape(-1), dtype=np.float32),
        data=np.ascontiguousarray(data.reshape(-1), dtype=np.float32),
        T=T, S=S, C=C, N=N, L=n_samples_data, n_corr=n_corr, step=int(step),
    )


def _check_zeros(cc, check_zeros, network_sum, n_corr, moveouts, weights, step,
                 n_templates):
    """Zero-count messages of fast_matched_filter.py:322-361, including its
    slicing of the per-channel layout."""
    if (check_zeros is not False) and (check_zeros != "first") and (check_zeros != "all"):
        print("check_zeros should be False, 'first', or 'all'. Set it to "
              "the default value: 'first'")
        check_zeros = "first"
    if not check_zeros:
        return
    stop = int(n_corr - moveouts.max() / step)
    if check_zeros == "first":
        if network_sum:
            zeros = np.sum(cc[0:1, :stop] == 0.0, axis=-1)
        else:
            zeros = np.sum(cc[0:1, :, :, :stop] == 0.0, axis=(1, 2, 3))
    else:
        zeros = np.sum(cc[..., :stop].reshape(n_templates, -1) == 0.0, axis=-1)
    # fast_matched_filter.py:348-351 indexes the flattened weights, so the
    # per-channel threshold is 10 times (first weight > 0), not 10 times the
    # active channels of template 0
    msg_threshold = 10
    if not network_sum:
        msg_threshold *= int(np.asarray(weights).reshape(-1)[0] > 0.0)
    for t in range(zeros.shape[0]):
        if zeros[t] > msg_threshold:
            print("{} correlation computations were skipped on the {:d}-th "
                  "template. Can be caused by zeros in data, or too low "
                  "amplitudes (try to increase the gain).".format(zeros[t], t))


def matched_filter(templates, moveouts, weights, data, step,
                   n_samples_template=None, arch="hip", check_zeros="first",
                   normalize="short", network_sum=True, exact_order=False,
                   strict_kernel=False):
    """Compute the correlation coefficients between `templates` and `data`.

    Same contract as the reference ``matched_filter``
    (fast_matched_filter.py:66-151): templates (T,S,C,N) / (T,traces,N) /
    single-template 3D-2D forms, moveouts and weights (T,S,C) or (T,S), data
    (S,C,L) or (traces,L), `step` in samples.  Returns float32
    (T, n_corr) when `network_sum` else (T, n_corr, S, C), with
    n_corr = (L - N) // step + 1 and zeros outside each template's valid lags.
    `n_samples_template` is not supported by these archs (the reference
    accepts it only with arch='variable_precise').

    With arch='hip_mfma', sizes the MFMA tiling does not take (step other
    than 1 or 2, L >= 2^29, N > 1344 - 31*step with the default 32x32 tile
    shapes) run on the FP32 kernel with a RuntimeWarning;
    ``strict_kernel=True`` raises RuntimeError instead.
    ``_lib.last_kernel()`` names the kernel of the last call.
    """
    assert normalize in ("short", "full"), "Only know short or full normalization methods"
    arch = arch.lower()
    if arch not in ARCHS:
        raise ValueError("arch must be one of {}, got {!r}".format(ARCHS, arch))
    if normalize == "full":
        norm = 1
        assert arch != "cpu", "Full normalization not supported with cpu arch - try arch=precise"
    else:
        norm = 0
    if arch == "variable_precise":
        return _variable_precise(templates, moveouts, weights, data, step,
                                 n_samples_template, check_zeros, norm,
                                 network_sum)
    if n_samples_template is not None:
        raise NotImplementedError(
            "per-channel template lengths need arch='variable_precise'")

    lib = _lib.require()
    p = prepare_inputs(templates, moveouts, weights, data, step)
    if p is None:
        return None
    T, S, C, n_corr = p["T"], p["S"], p["C"], p["n_corr"]
    if network_sum:
        cc = np.zeros(T * n_corr, dtype=np.float32)
    else:
        cc = np.zeros(T * S * C * n_corr, dtype=np.float32)

    kernel = _lib.KERNEL_BF16X3 if arch == "hip_mfma" else _lib.KERNEL_FP32
    flags = _lib.FLAG_THRESHOLD_AFTER_SQRT if arch == "precise" else 0
    if exact_order:
        flags |= _lib.FLAG_EXACT_ORDER
    if strict_kernel and kernel == _lib.KERNEL_BF16X3:
        flags |= _lib.FLAG_STRICT_KERNEL
    rc = lib.fmf_matched_filter_host(
        p["templates"].ctypes.data_as(_F), p["sst"].ctypes.data_as(_F),
        p["moveouts"].ctypes.data_as(_I), p["data"].ctypes.data_as(_F),
        p["weights"].ctypes.data_as(_F), p["step"], p["N"], p["L"], T, S, C,
        n_corr, cc.ctypes.data_as(_F), norm, int(bool(network_sum)), kernel,
        flags)
    _lib.check(rc, "matched_filter")
    if kernel == _lib.KERNEL_BF16X3 and _lib.last_kernel() != "bf16x3":
        warnings.warn("arch='hip_mfma': the MFMA tiling does not take step %d with "
                      "N=%d; the FP32 kernel ran instead" % (p["step"], p["N"]),
                      RuntimeWarning, stacklevel=2)

    if network_sum:
        cc = cc.reshape(T, n_corr)
    else:
        cc = cc.reshape(T, n_corr, S, C)
    _check_zeros(cc, check_zeros, network_sum, n_corr, p["moveouts"],
                 p["weights"], p["step"], T)
    return cc


def _variable_precise(templates, moveouts, weights, data, step,
                      n_samples_template, check_zeros, norm, network_sum):
    """arch='variable_precise' (fast_matched_filter.py:216-224, 292-313)."""
    if not network_sum:
        raise NotImplementedError("Variable precise without network sum is not yet implemented.")
    lib = _lib.require()
    p = prepare_inputs(templates, moveouts, weights, data, step)
    if p is None:
        return None
    T, S, C, n_corr = p["T"], p["S"], p["C"], p["n_corr"]
    if n_samples_template is None:
        lengths = np.full(T * S * C, p["N"], dtype=np.int32)
    else:
        n_samples_template = np.asarray(n_samples_template)
        assert n_samples_template.shape == np.asarray(moveouts).shape
        lengths = np.ascontiguousarray(n_samples_template.flatten(), dtype=np.int32)
        if lengths.size * C == T * S * C and lengths.size != T * S * C:
            lengths = np.repeat(lengths, C)
        if lengths.max() != p["N"]:
            raise ValueError("templates.shape[-1] must equal max(n_samples_template)")
    cc = np.zeros(T * n_corr, dtype=np.float32)
    rc = lib.matched_filter_variable_precise(
        p["templates"].ctypes.data_as(_F), p["sst"].ctypes.data_as(_F),
        p["moveouts"].ctypes.data_as(_I), p["data"].ctypes.data_as(_F),
        p["weights"].ctypes.data_as(_F), p["step"], lengths.ctypes.data_as(_I),
        p["L"], T, S, C, n_corr, cc.ctypes.data_as(_F), norm)
    _lib.check(rc, "matched_filter_variable_precise")
    cc = cc.reshape(T, n_corr)
    _check_zeros(cc, check_zeros, True, n_corr, p["moveouts"], p["weights"],
                 p["step"], T)
    return cc


def matched_filter_batched(templates, lengths, moveouts, weights, data, step,
                           normalize="short", persistent_blocks=0,
                           return_device=False, host_chunks=None):
    """Templates of different lengths in one launch (bf16x3 MFMA kernel).

    templates (T, S, C, Nmax) zero-padded beyond each template's length,
    lengths (T,) samples, moveouts/weights (T, S, C) or (T, S), data (S, C, L),
    step 1 or 2.  Row t equals ``matched_filter(templates[t:t+1, ..., :lengths[t]],
    ..., arch='hip_mfma')`` on its valid lags; the output is
    (T, (L - min(lengths)) // step + 1) with zeros outside each row's lags.

    With a host output (``return_device=False``) the templates run in
    ``host_chunks`` launches (default: one per 8 GB of output, at most 8) and
    the rows of each launch are copied to the host while the next computes;
    every row is the same as from one launch.
    """
    import torch
    from .device import copy_to_host, matched_filter_batched_device
    assert normalize in ("short", "full")
    p = prepare_inputs(templates, moveouts, weights, data, step)
    if p is None:
        return None
    T, S, C, N, L = p["T"], p["S"], p["C"], p["N"], p["L"]
    lengths = np.ascontiguousarray(np.asarray(lengths).reshape(-1), dtype=np.int32)
    if lengths.size != T:
        raise ValueError("lengths must have one entry per template")
    if lengths.min() < 1 or lengths.max() > N:
        raise ValueError("lengths must lie in [1, templates.shape[-1]]")
    distinct = np.unique(lengths)
    # Sum of squares over each template's own length, as a call on that
    # length group computes it (fast_matched_filter.py:256-261).
    tp4 = np.asarray(templates).reshape(T, S, C, N)
    sst = np.empty((T, S, C), dtype=np.float32)
    for n in distinct:
        rows = np.nonzero(lengths == n)[0]
        sst[rows] = np.sum(tp4[rows][..., :n] ** 2, axis=-1)
    p["sst"] = np.ascontiguousarray(sst.reshape(-1))
    n_corr = int((L - int(lengths.min())) // step + 1)
    dev = torch.device("cuda", torch.cuda.current_device())
    to = lambda a: torch.from_numpy(a).to(dev)
    out = torch.empty((T, n_corr), dtype=torch.float32, device=dev)
    d_tp, d_sst, d_mo, d_w = to(p["templates"]), to(p["sst"]), to(p["moveouts"]), to(p["weights"])
    d_data, d_len = to(p["data"]), to(lengths)
    norm = 1 if normalize == "full" else 0
    if return_device:
        matched_filter_batched_device(d_tp, d_sst, d_mo, d_w, d_data, d_len, distinct, step,
                                      N, L, T, S, C, n_corr, out, normalize=norm,
                                      persistent_blocks=persistent_blocks)
        return out
    if host_chunks is None:
        host_chunks = min(8, max(1, -(-T * n_corr * 4 // (8 << 30))))
    k = max(1, min(int(host_chunks), T))
    bounds = [T * i // k for i in range(k + 1)]
    host = np.empty((T, n_corr), dtype=np.float32)
    SC, NP = S * C, N
    compute = torch.cuda.current_stream(dev)
    copier = torch.cuda.Stream(device=dev)
    pending = None  # (rows, event) of the launch whose rows wait for their copy
    for i in range(k):
        a, b = bounds[i], bounds[i + 1]
        if b > a:
            matched_filter_batched_device(
                d_tp[a * SC * NP:b * SC * NP], d_sst[a * SC:b * SC], d_mo[a * SC:b * SC],
                d_w[a * SC:b * SC], d_data, d_len[a:b], np.unique(lengths[a:b]), step, N, L,
                b - a, S, C, n_corr, out[a:b], normalize=norm,
                persistent_blocks=persistent_blocks, stream=compute)
            ev = torch.cuda.Event()
            ev.record(compute)
        if pending is not None:  # copied while launch i computes
            (pa, pb), pev = pending
            copier.wait_event(pev)
            copy_to_host(out[pa:pb], out=host[pa:pb], stream=copier)
        pending = ((a, b), ev) if b > a else None
    if pending is not None:
        (pa, pb), pev = pending
        copier.wait_event(pev)
        copy_to_host(out[pa:pb], out=host[pa:pb], stream=copier)
    return host


def test_matched_filter(n_templates=1, n_stations=1, n_components=1, template_duration=10,
                        data_duration=86400, sampling_rate=100, step=1, arch="cpu",
                        check_zeros="first", normalize="short", network_sum=True, seed=None):
    """Self-detection harness with the reference's keyword arguments and
    return values (fast_matched_filter.py:365-569), plus ``seed``.

    Random data (uniform in [0, 1), mean removed per channel) of
    ``data_duration * sampling_rate`` samples; every template is cut from the
    data at a random time in the first half of the day with random moveouts in
    [0, 10 s) per (station, component), weights 1, so the network cc at the
    template's own lag equals n_stations * n_components (each channel's cc is
    1).  With arch='variable_precise' each station gets a random length in
    [N/2, N] and the templates are zero beyond it.  Prints the run time and
    returns (templates, moveouts, data, step, cc_sum, run_time), with
    n_samples_template before run_time for 'variable_precise'.
    """
    import time
    rng = np.random.default_rng(seed)
    n_data = data_duration * sampling_rate
    n_tp = template_duration * sampling_rate
    if float(int(n_data)) != float(n_data) or float(int(n_tp)) != float(n_tp):
        print("The durations times the sampling rate must give whole numbers of samples; "
              "adjust the parameters.")
        return None
    n_data, n_tp = int(n_data), int(n_tp)
    t_times = rng.random(n_templates) * (data_duration / 2)
    if step != 1:  # on the step grid, so that the self-detection lag is computed
        t_times = np.round(t_times / (step / sampling_rate)) * (step / sampling_rate)
    moveouts = np.int32(np.round(rng.random((n_templates, n_stations, n_components))
                                 * 10.0 * sampling_rate))
    data = rng.random((n_stations, n_components, n_data), dtype=np.float32)
    data -= data.mean(axis=-1, keepdims=True)
    templates = np.zeros((n_templates, n_stations, n_components, n_tp), dtype=np.float32)
    for t in range(n_templates):
        start_t = int(t_times[t] * sampling_rate)
        for s in range(n_stations):
            for c in range(n_components):
                a = start_t + int(moveouts[t, s, c])
                templates[t, s, c] = data[s, c, a:a + n_tp]
    weights = np.ones((n_templates, n_stations, n_components), dtype=np.float32)
    lengths = None
    kw = dict(arch=arch, check_zeros=check_zeros, normalize=normalize, network_sum=network_sum)
    if arch == "variable_precise":
        per_station = np.int32(np.round(rng.random((n_templates, n_stations)) * n_tp / 2
                                        + n_tp / 2))
        lengths = np.repeat(per_station[:, :, None], n_components, axis=2)
        templates = templates[..., :int(lengths.max())].copy()
        for t in range(n_templates):
            for s in range(n_stations):
                templates[t, s, :, lengths[t, s, 0]:] = 0.0
        kw["n_samples_template"] = lengths
    t0 = time.time()
    cc_sum = matched_filter(templates, moveouts, weights, data, step, **kw)
    run_time = time.time() - t0
    print("Matched filter ({}) for {} templates on {} stations/{} components over {} samples "
          "({} step) ran in {:.3f}s".format(arch, n_templates, n_stations, n_components,
                                             n_data, step, run_time))
    if arch == "variable_precise":
        return templates, moveouts, data, step, cc_sum, lengths, run_time
    return templates, moveouts, data, step, cc_sum, run_time
