"""Differentiable spectral filtering on real signals.

    y = IFFT(H_full * FFT(x)) / n        (x real [..., n], H complex [n/2 + 1])

H_full is the Hermitian extension of H; as in any c2r the imaginary parts of
bins 0 and n/2 are ignored, so y is real.  ``spectral_filter`` runs the fused
gfx950 kernels of the C ABI: the forward is one kernel (FFT -> filter ->
inverse FFT in LDS, one HBM round trip per row), the backward is one kernel
for gx = irfft(conj(H) * rfft(gy)) plus the per-row gH partials, and a small
reduction for gH = sum_rows w_k/n * rfft(gy)_k * conj(rfft(x)_k).

``spectral_filter_plain`` computes the same function and gradients through
the library's unfused entry points (rfft / irfft plans and elementwise
PyTorch ops, one launch per step): the reference the fused path is timed
against.  genFFT has no counterpart (it has no c2r, FFTReal.h:127-130).
"""
from __future__ import annotations

import ctypes
import threading

import torch

from . import _capi as C
from ._launch import call_on

_plans: dict = {}
_lock = threading.Lock()


class _FilterPlan:
    def __init__(self, prec: int, n: int, batch: int, device: torch.device):
        self._lib = C.lib()
        self._h = ctypes.c_void_p()
        self.index = device.index if device.index is not None else torch.cuda.current_device()
        with torch.cuda.device(self.index):
            C.check(self._lib.gfft_filter_plan_create(ctypes.byref(self._h), prec, n, batch))
        self.handle = self._h.value
        self.forward = self._lib.gfft_filter_forward
        self.backward = self._lib.gfft_filter_backward

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.gfft_filter_plan_destroy(self._h)


_COMPLEX_OF = {torch.float32: torch.complex64, torch.float64: torch.complex128}
_PREC_OF = {torch.float32: C.GFFT_F32, torch.float64: C.GFFT_F64}


def _check(x: torch.Tensor, H: torch.Tensor):
    if not (x.is_cuda and H.is_cuda):
        raise ValueError("spectral_filter runs on GPU tensors only")
    want = _COMPLEX_OF.get(x.dtype)
    if want is None or H.dtype != want:
        raise TypeError("x must be float32/float64 and H the matching complex dtype")
    if H.shape[-1] != x.shape[-1] // 2 + 1 or H.dim() != 1:
        raise ValueError("H must hold n/2 + 1 bins")
    if H.device != x.device:
        raise ValueError("x and H must be on the same device")


def _plan(x: torch.Tensor, H: torch.Tensor) -> _FilterPlan:
    """Validated plan for this (dtype, shape, device); arguments are checked once per key."""
    key = (x.dtype, x.shape, x.device, H.dtype, H.shape, H.device)
    p = _plans.get(key)
    if p is not None:
        return p
    _check(x, H)
    n = x.shape[-1]
    with _lock:
        p = _plans.get(key)
        if p is None:
            p = _plans[key] = _FilterPlan(_PREC_OF[x.dtype], n, x.numel() // n, x.device)
        return p


class SpectralFilter(torch.autograd.Function):
    """Fused forward and adjoint kernels over the C ABI."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, H: torch.Tensor) -> torch.Tensor:
        p = _plan(x, H)
        x = x.contiguous()
        H = H.contiguous()
        y = torch.empty_like(x)
        call_on(p.index, None, p.forward, p.handle, x.data_ptr(), H.data_ptr(), y.data_ptr(), 0)
        ctx.save_for_backward(x, H)
        ctx.plan = p
        return y

    @staticmethod
    def backward(ctx, gy: torch.Tensor):
        x, H = ctx.saved_tensors
        p = ctx.plan
        gy = gy.contiguous()
        gx = torch.empty_like(x)
        gH = torch.empty_like(H)
        call_on(p.index, None, p.backward, p.handle, x.data_ptr(), H.data_ptr(), gy.data_ptr(), gx.data_ptr(),
                gH.data_ptr())
        return gx, gH


def spectral_filter(x: torch.Tensor, H: torch.Tensor) -> torch.Tensor:
    """y = irfft(H * rfft(x)) over the last axis, fused, differentiable in x and H."""
    return SpectralFilter.apply(x, H)


_weights: dict = {}


def _edge_weights(bins: int, n: int, dtype, device) -> torch.Tensor:
    """w_k = 2/n, 1/n at bins 0 and n/2; built once per shape (no host copy per call)."""
    key = (bins, n, dtype, device)
    w = _weights.get(key)
    if w is None:
        w = torch.full((bins,), 2.0 / n, dtype=dtype, device=device)
        w[0] = 1.0 / n
        w[-1] = 1.0 / n
        _weights[key] = w
    return w


class _PlainFilter(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, H):
        from . import irfft, rfft
        n = x.shape[-1]
        X = rfft(x, half=True)
        y = irfft(X * H, n) / n
        ctx.save_for_backward(x, H)
        return y

    @staticmethod
    def backward(ctx, gy):
        from . import irfft, rfft
        x, H = ctx.saved_tensors
        n = x.shape[-1]
        G = rfft(gy.contiguous(), half=True)
        gx = irfft(G * H.conj(), n) / n
        X = rfft(x, half=True)
        w = _edge_weights(H.shape[-1], n, x.dtype, x.device)
        gH = (w * G * X.conj()).reshape(-1, H.shape[-1]).sum(0)
        return gx, gH


def spectral_filter_plain(x: torch.Tensor, H: torch.Tensor) -> torch.Tensor:
    """The same operator through unfused launches (rfft, multiply, irfft, scale)."""
    _check(x, H)
    return _PlainFilter.apply(x.contiguous(), H.contiguous())
