"""Device-resident entry point for callers that keep their arrays in HBM.

Thin wrapper over ``fmf_matched_filter_device`` of libfmf_hip.so.  Tensors are
torch tensors on the current HIP device (torch is used only to own device
memory and streams); layouts are the flattened reference layouts
(templates (T,S,C,N), sum of squares / moveouts / weights (T,S,C), data
(S,C,L)).  The call is stream-ordered and does not synchronise.
"""
import torch

from . import _lib

__all__ = ["workspace_bytes", "matched_filter_device", "noalloc_scratch_bytes",
           "matched_filter_device_noalloc", "copy_to_host"]


def workspace_bytes(n_samples_template, n_samples_data, n_templates,
                    n_stations, n_components, normalize=0,
                    kernel=_lib.KERNEL_FP32, flags=0):
    return int(_lib.require().fmf_workspace_bytes(
        n_samples_template, n_samples_data, n_templates, n_stations,
        n_components, normalize, kernel, flags))


def _ptr(t, dtype):
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor of {}".format(dtype))
    return t.data_ptr()


def matched_filter_device(templates, sum_square_templates, moveouts, weights,
                          data, step, n_samples_template, n_samples_data,
                          n_templates, n_stations, n_components, n_corr, out,
                          normalize=0, network_sum=True,
                          kernel=_lib.KERNEL_FP32, flags=0, workspace=None,
                          stream=None):
    lib = _lib.require()
    need = workspace_bytes(n_samples_template, n_samples_data, n_templates,
                           n_stations, n_components, normalize, kernel, flags)
    if workspace is None:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=data.device)
    if workspace.numel() < need:
        raise ValueError("workspace needs {} bytes".format(need))
    if stream is None:
        stream = torch.cuda.current_stream(data.device)
    rc = lib.fmf_matched_filter_device(
        _ptr(templates, torch.float32), _ptr(sum_square_templates, torch.float32),
        _ptr(moveouts, torch.int32), _ptr(data, torch.float32),
        _ptr(weights, torch.float32), int(step), int(n_samples_template),
        int(n_samples_data), int(n_templates), int(n_stations),
        int(n_components), int(n_corr), _ptr(out, torch.float32),
        int(normalize), int(bool(network_sum)), int(kernel), int(flags),
        workspace.data_ptr(), int(workspace.numel()), stream.cuda_stream)
    _lib.check(rc, "fmf_matched_filter_device")
    return out


def copy_to_host(src, out=None, stream=None):
    """Copy the contiguous device tensor `src` into the numpy array `out`
    (same number of bytes; a new array of src's shape and dtype if None) after
    the work queued on `stream`, through ``fmf_copy_to_host``: the
    destination's pages are faulted in by host threads while the GPU works,
    then the data streams through pinned staging buffers.  Returns `out`."""
    import numpy as np
    if not src.is_cuda or not src.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    if out is None:
        out = np.empty(tuple(src.shape), dtype=torch.empty(0, dtype=src.dtype).numpy().dtype)
    if not out.flags["C_CONTIGUOUS"] or not out.flags["WRITEABLE"]:
        raise ValueError("out must be a writeable C-contiguous array")
    nbytes = src.numel() * src.element_size()
    if out.nbytes != nbytes:
        raise ValueError("out holds {} bytes, src {}".format(out.nbytes, nbytes))
    if stream is None:
        stream = torch.cuda.current_stream(src.device)
    rc = _lib.require().fmf_copy_to_host(out.ctypes.data, src.data_ptr(), nbytes,
                                         stream.cuda_stream)
    _lib.check(rc, "fmf_copy_to_host")
    return out


def noalloc_scratch_bytes():
    """Size of the library's fixed per-GPU workspace used by
    ``matched_filter_device_noalloc``."""
    return int(_lib.require().fmf_noalloc_scratch_bytes())


def matched_filter_device_noalloc(templates, sum_square_templates, moveouts, weights,
                                  data, step, n_samples_template, n_samples_data,
                                  n_templates, n_stations, n_components, n_corr, out,
                                  normalize=0, network_sum=True,
                                  kernel=_lib.KERNEL_FP32, flags=0, stream=None):
    """``matched_filter_device`` without a workspace argument: the library's
    fixed per-GPU array serves as workspace and nothing is allocated (calls
    that need more than ``noalloc_scratch_bytes()`` raise ValueError).  Calls
    that overlap on one GPU race on that array."""
    lib = _lib.require()
    if stream is None:
        stream = torch.cuda.current_stream(data.device)
    rc = lib.fmf_matched_filter_device_noalloc(
        _ptr(templates, torch.float32), _ptr(sum_square_templates, torch.float32),
        _ptr(moveouts, torch.int32), _ptr(data, torch.float32),
        _ptr(weights, torch.float32), int(step), int(n_samples_template),
        int(n_samples_data), int(n_templates), int(n_stations),
        int(n_components), int(n_corr), _ptr(out, torch.float32),
        int(normalize), int(bool(network_sum)), int(kernel), int(flags),
        stream.cuda_stream)
    _lib.check(rc, "fmf_matched_filter_device_noalloc")
    return out


def batched_workspace_bytes(distinct_lengths, max_samples_template,
                            n_samples_data, n_templates, n_stations,
                            n_components):
    import ctypes as ct
    import numpy as np
    d = np.ascontiguousarray(distinct_lengths, dtype=np.int32)
    return int(_lib.require().fmf_batched_workspace_bytes(
        d.ctypes.data_as(ct.POINTER(ct.c_int)), int(d.size),
        int(max_samples_template), int(n_samples_data), int(n_templates),
        int(n_stations), int(n_components)))


def matched_filter_batched_device(templates, sum_square_templates, moveouts,
                                  weights, data, lengths, distinct_lengths,
                                  step, max_samples_template, n_samples_data,
                                  n_templates, n_stations, n_components,
                                  n_corr, out, normalize=0, flags=0,
                                  persistent_blocks=0, workspace=None,
                                  stream=None):
    """One launch over templates of different lengths (``lengths``: device
    int32 (T,); ``distinct_lengths``: the host list of its values)."""
    import ctypes as ct
    import numpy as np
    lib = _lib.require()
    d = np.ascontiguousarray(distinct_lengths, dtype=np.int32)
    need = batched_workspace_bytes(d, max_samples_template, n_samples_data,
                                   n_templates, n_stations, n_components)
    if need == 0:
        raise ValueError("invalid batched sizes or lengths")
    if workspace is None:
        workspace = torch.empty(need, dtype=torch.uint8, device=data.device)
    if workspace.numel() < need:
        raise ValueError("workspace needs {} bytes".format(need))
    if stream is None:
        stream = torch.cuda.current_stream(data.device)
    rc = lib.fmf_matched_filter_batched_device(
        _ptr(templates, torch.float32), _ptr(sum_square_templates, torch.float32),
        _ptr(moveouts, torch.int32), _ptr(data, torch.float32),
        _ptr(weights, torch.float32), _ptr(lengths, torch.int32),
        d.ctypes.data_as(ct.POINTER(ct.c_int)), int(d.size), int(step),
        int(max_samples_template), int(n_samples_data), int(n_templates),
        int(n_stations), int(n_components), int(n_corr), _ptr(out, torch.float32),
        int(normalize), int(flags), int(persistent_blocks), workspace.data_ptr(),
        int(workspace.numel()), stream.cuda_stream)
    _lib.check(rc, "fmf_matched_filter_batched_device")
    return out
