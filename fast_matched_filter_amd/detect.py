"""Detection post-processing on device-resident cc rows.

The reference returns the (T, n_corr) cc rows and leaves detection to the
caller (fast_matched_filter.py:318-362).  Here the usual next steps run on
the GPU, so only the detections cross PCIe:

* ``median_mad``: per-template median and MAD (median absolute deviation)
  over the template's valid lags, exactly as ``numpy.median`` computes them
  in float32;
* ``detect``: threshold = median + n_mad * MAD, and a lag is a detection when
  its cc exceeds the threshold and is the first maximum of the cc within
  +/- ``min_separation`` lags (inside the valid range);
* ``matched_filter_detect``: the bf16x3 (or FP32) matched filter followed by
  the detection, with the cc rows kept in HBM.

Valid lags are those the reference computes (matched_filter.c:53-79); the
zero tail beyond them is excluded from the statistics.
"""
import numpy as np
import torch

from . import _lib
from .fast_matched_filter import prepare_inputs

__all__ = ["lag_ranges", "median_mad", "detect", "matched_filter_detect"]


def _dev(x, dtype):
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            x = x.cuda()
        return x.to(dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda").to(dtype).contiguous()


def _stream():
    return torch.cuda.current_stream().cuda_stream


def lag_ranges(moveouts, n_samples_template, n_samples_data, step, n_corr):
    """Valid lag range [lo, hi) per template (device int32 (T, 2)) for
    moveouts (T, S, C) or (T, S*C) in samples."""
    mo = _dev(moveouts, torch.int32)
    T = mo.shape[0]
    mo = mo.reshape(T, -1).contiguous()
    out = torch.empty((T, 2), dtype=torch.int32, device=mo.device)
    rc = _lib.require().fmf_lag_ranges_device(
        mo.data_ptr(), T, mo.shape[1], int(step), int(n_samples_template),
        int(n_samples_data), int(n_corr), out.data_ptr(), _stream())
    _lib.check(rc, "fmf_lag_ranges_device")
    return out


def _workspace(T, n_corr, device):
    need = int(_lib.require().fmf_detect_workspace_bytes(T, n_corr))
    return torch.empty(max(need, 1), dtype=torch.uint8, device=device)


def _rows(cc):
    cc = _dev(cc, torch.float32)
    if cc.dim() != 2:
        raise ValueError("cc must be (n_templates, n_corr)")
    return cc


def median_mad(cc, ranges=None, workspace=None):
    """Exact per-row median and MAD of cc (T, n_corr) over [lo, hi) of
    ``ranges`` (T, 2) (None: whole rows).  Returns two device float32 (T,)."""
    cc = _rows(cc)
    T, n_corr = cc.shape
    rg = None if ranges is None else _dev(ranges, torch.int32)
    ws = _workspace(T, n_corr, cc.device) if workspace is None else workspace
    med = torch.empty(T, dtype=torch.float32, device=cc.device)
    mad = torch.empty(T, dtype=torch.float32, device=cc.device)
    rc = _lib.require().fmf_median_mad_device(
        cc.data_ptr(), T, n_corr, 0 if rg is None else rg.data_ptr(), med.data_ptr(),
        mad.data_ptr(), ws.data_ptr(), ws.numel(), _stream())
    _lib.check(rc, "fmf_median_mad_device")
    return med, mad


def detect(cc, n_mad=8.0, min_separation=0, ranges=None, capacity=1 << 20,
           median=None, mad=None):
    """Detections in cc (T, n_corr): dict of numpy arrays ``template``,
    ``lag``, ``value`` in (template, lag) order, and per template ``median``,
    ``mad``, ``threshold``.  If more than ``capacity`` detections exist the
    call is repeated once with room for all of them.  ``median`` and ``mad``
    (T,) replace the rows' own statistics when given (both or neither), e.g.
    estimates from a sample when most of a row was not computed."""
    cc = _rows(cc)
    T, n_corr = cc.shape
    rg = None if ranges is None else _dev(ranges, torch.int32)
    ws = _workspace(T, n_corr, cc.device)
    if (median is None) != (mad is None):
        raise ValueError("give both median and mad, or neither")
    if median is None:
        med, mad = median_mad(cc, rg, ws)
    else:
        med, mad = _dev(median, torch.float32), _dev(mad, torch.float32)
        if med.numel() != T or mad.numel() != T:
            raise ValueError("median and mad need one value per template")
    lib = _lib.require()
    n_det = torch.zeros(1, dtype=torch.int64, device=cc.device)
    while True:
        cap = max(int(capacity), 1)
        det_t = torch.empty(cap, dtype=torch.int32, device=cc.device)
        det_l = torch.empty(cap, dtype=torch.int32, device=cc.device)
        det_v = torch.empty(cap, dtype=torch.float32, device=cc.device)
        rc = lib.fmf_detect_device(
            cc.data_ptr(), T, n_corr, 0 if rg is None else rg.data_ptr(), med.data_ptr(),
            mad.data_ptr(), float(n_mad), int(min_separation), det_t.data_ptr(),
            det_l.data_ptr(), det_v.data_ptr(), cap, n_det.data_ptr(), ws.data_ptr(),
            ws.numel(), _stream())
        _lib.check(rc, "fmf_detect_device")
        n = int(n_det.item())
        if n <= cap:
            break
        capacity = n
    med_h = med.cpu().numpy()
    mad_h = mad.cpu().numpy()
    thr = med_h + np.float32(n_mad) * mad_h
    return {"template": det_t[:n].cpu().numpy(), "lag": det_l[:n].cpu().numpy(),
            "value": det_v[:n].cpu().numpy(), "median": med_h, "mad": mad_h,
            "threshold": thr.astype(np.float32)}


def matched_filter_detect(templates, moveouts, weights, data, step, n_mad=8.0,
                          min_separation=None, arch="hip_mfma", normalize="short"):
    """Matched filter (network sum) and detection in one call; the cc rows
    stay on the GPU.  ``min_separation`` defaults to the template length.
    Returns the dict of ``detect`` plus ``n_corr``."""
    from .device import matched_filter_device
    p = prepare_inputs(templates, moveouts, weights, data, step)
    T, S, C, n_corr = p["T"], p["S"], p["C"], p["n_corr"]
    dev = torch.device("cuda")
    up = lambda a, dt: torch.from_numpy(a).to(dev).to(dt).contiguous()
    cc = torch.empty((T, n_corr), dtype=torch.float32, device=dev)
    kernel = _lib.KERNEL_BF16X3 if arch == "hip_mfma" else _lib.KERNEL_FP32
    flags = _lib.FLAG_THRESHOLD_AFTER_SQRT if arch == "precise" else 0
    mo = up(p["moveouts"], torch.int32)
    matched_filter_device(up(p["templates"], torch.float32), up(p["sst"], torch.float32),
                          mo, up(p["weights"], torch.float32), up(p["data"], torch.float32),
                          p["step"], p["N"], p["L"], T, S, C, n_corr, cc,
                          normalize=1 if normalize == "full" else 0, network_sum=True,
                          kernel=kernel, flags=flags)
    rg = lag_ranges(mo.reshape(T, S * C), p["N"], p["L"], p["step"], n_corr)
    sep = p["N"] if min_separation is None else int(min_separation)
    out = detect(cc, n_mad=n_mad, min_separation=sep, ranges=rg)
    out["n_corr"] = n_corr
    return out
