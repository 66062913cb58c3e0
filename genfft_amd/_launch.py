"""Calling a C entry point on a tensor's device and stream with little host work.

Every execute call of the C ABI takes the HIP stream as its last argument and
requires the plan's device to be current.  ``call_on`` resolves the caller's
stream as a raw handle and enters a device guard only when another device is
current, which keeps the per-call host cost of small transforms and of the
autograd backward (run on autograd's device thread) to a few microseconds.
"""
from __future__ import annotations

import torch

from . import _capi as C

# torch's raw current-stream query (an int handle, no Stream object); the
# public query is the fallback where this build lacks it
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def device_index(dev: torch.device) -> int:
    return dev.index if dev.index is not None else torch.cuda.current_device()


def stream_handle(idx: int, stream=None) -> int:
    """The HIP stream handle: ``stream`` when given, else device idx's current stream."""
    if stream is not None:
        return stream.cuda_stream
    if _raw_stream is not None:
        return _raw_stream(idx)
    return torch.cuda.current_stream(idx).cuda_stream


def call_on(idx: int, stream, fn, *args) -> None:
    """C.check(fn(*args, stream_handle)) with device idx current."""
    if torch.cuda.current_device() != idx:
        with torch.cuda.device(idx):
            return call_on(idx, stream, fn, *args)
    C.check(fn(*args, stream_handle(idx, stream)))
