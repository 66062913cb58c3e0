"""genfft_amd -- MI355X (gfx950) FFT engine with genFFT's plan-object surface.

The compute path is the hand-written HIP library ``libgenfft_amd.so`` behind
the C ABI in ``include/genfft_amd.h``; this module is a thin ctypes layer that
takes PyTorch tensors (device memory and the current stream), mirroring the
reference classes for Python callers:

    genfft::FFT<T>         include/genFFT/fft.h:56-113      -> FFT
    genfft::FFTVert<T>     include/genFFT/fft.h:117-171     -> FFTVert
    genfft::DIT<T>         include/genFFT/fft.h:173-196     -> DIT
    genfft::FFT2D<T>       include/genFFT/fft.h:200-245     -> FFT2D
    genfft::RealFFT<T>     include/genFFT/FFTReal.h:188-221 -> RealFFT
    genfft::RealFFT2D<T>   include/genFFT/FFTReal.h:71-184  -> RealFFT2D
    separate_2x_real_FFT   include/genFFT/FFTReal.h:36-66   -> separate_2x_real_FFT

Same conventions as the reference: forward kernel exp(-2*pi*i*j*k/N),
unscaled inverse, natural order, power-of-two sizes.  Strides and row
pitches are in complex elements (RealFFT2D input pitch in scalars), exactly as
in the reference API.  Tensors must live on a GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import threading

import torch

from . import _capi as C
from ._capi import GfftError
from ._launch import call_on as _call_on, device_index as _dev_idx, stream_handle as _stream_handle

__all__ = [
    "Plan", "FFT", "FFTVert", "DIT", "FFT2D", "RealFFT", "RealFFT2D", "UnevenBatch",
    "separate_2x_real_FFT", "fft", "ifft", "rfft", "irfft", "fft2", "GfftError",
    "autotune", "save_tuning", "load_tuning", "clear_plan_cache",
]

_COMPLEX = {torch.complex64: C.GFFT_F32, torch.complex128: C.GFFT_F64}
_REAL = {torch.float32: C.GFFT_F32, torch.float64: C.GFFT_F64}
_REAL_OF = {torch.complex64: torch.float32, torch.complex128: torch.float64}
_COMPLEX_OF = {torch.float32: torch.complex64, torch.float64: torch.complex128}


def _prec(dtype: torch.dtype) -> int:
    if dtype in _COMPLEX:
        return _COMPLEX[dtype]
    if dtype in _REAL:
        return _REAL[dtype]
    raise TypeError(f"unsupported dtype {dtype}: use float32/float64 or complex64/complex128")


def _prec_complex(dtype: torch.dtype) -> int:
    if dtype not in _COMPLEX:
        raise TypeError(f"complex transforms take complex64 / complex128, not {dtype}")
    return _COMPLEX[dtype]


def _prec_real(dtype: torch.dtype) -> int:
    if dtype not in _REAL:
        raise TypeError(f"real transforms take float32 / float64, not {dtype}")
    return _REAL[dtype]


# tensor dtypes a plan's complex / real side accepts, by plan precision
# (16-bit storage: [..., 2] float16 / bfloat16 pairs, or complex32)
_COMPLEX_SIDE = {C.GFFT_F32: (torch.complex64,), C.GFFT_F64: (torch.complex128,),
                 C.GFFT_F16: (torch.float16, torch.complex32), C.GFFT_BF16: (torch.bfloat16,)}
_REAL_SIDE = {C.GFFT_F32: (torch.float32,), C.GFFT_F64: (torch.float64,)}


def _extent(t: torch.Tensor) -> int:
    """Bytes from t's first element to the end of its storage."""
    return t.untyped_storage().nbytes() - t.storage_offset() * t.element_size()


def _row_major(t: torch.Tensor) -> bool:
    """Row-major up to padding: ignoring dimensions of size 1, the innermost
    has stride 1 and each outer one a stride at least the span of the one
    inside it.  The library reads and writes from data_ptr() with the plan's
    own pitches, so a transposed or permuted view would be read as garbage."""
    span, first = 1, True
    for size, stride in zip(reversed(t.shape), reversed(t.stride())):
        if size == 1:
            continue
        if (first and stride != 1) or stride < span:
            return False
        first = False
        span = stride * size
    return True


def _check_gpu(*ts: torch.Tensor) -> torch.device:
    dev = None
    for t in ts:
        if not t.is_cuda:
            raise ValueError("genfft_amd works on GPU tensors only (got a CPU tensor)")
        if not _row_major(t):
            raise ValueError(f"tensor with strides {tuple(t.stride())} is not row-major (innermost stride 1, "
                             "outer strides covering the inner ones); call .contiguous()")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError("all tensors must be on the same device")
    return dev


class Plan:
    """One gfft_plan (include/genfft_amd.h); owns the plan's device workspace."""

    def __init__(self, kind: int, n: int, *, precision: int = C.GFFT_F32, n2: int = 0,
                 batch: int = 1, in_stride: int = 0, out_stride: int = 0, in_dist: int = 0,
                 out_dist: int = 0, half: bool = False, device=None):
        self.desc = C.GfftDesc(kind, precision, n, n2, batch, in_stride, out_stride, in_dist,
                               out_dist, C.GFFT_FLAG_HALF if half else 0, 0)
        self._h = ctypes.c_void_p()
        self._lib = C.lib()
        if device is not None:
            with torch.cuda.device(device):
                C.check(self._lib.gfft_plan_create(ctypes.byref(self._h), ctypes.byref(self.desc)))
        else:
            C.check(self._lib.gfft_plan_create(ctypes.byref(self._h), ctypes.byref(self.desc)))
        ib, ob = ctypes.c_size_t(), ctypes.c_size_t()
        C.check(self._lib.gfft_plan_extents(self._h, ctypes.byref(ib), ctypes.byref(ob)))
        self.in_bytes, self.out_bytes = ib.value, ob.value
        # a real side also takes a complex tensor of the same precision (its
        # interleaved scalars: real transforms in place over the spectrum)
        cplx, real = _COMPLEX_SIDE[precision], _REAL_SIDE.get(precision, ())
        if kind in (C.GFFT_R2C, C.GFFT_R2C_2D):
            self._in_dt, self._out_dt = real + cplx, cplx
        elif kind in (C.GFFT_C2R, C.GFFT_C2R_2D):
            self._in_dt, self._out_dt = cplx, real + cplx
        elif kind == C.GFFT_DIT:  # complex data, or its interleaved scalars (the T* overload, fft.h:180-188)
            self._in_dt = self._out_dt = cplx + real
        else:
            self._in_dt = self._out_dt = cplx

    def _check(self, inp: torch.Tensor, out: torch.Tensor) -> torch.device:
        """The buffers must be GPU tensors of the plan's types, each holding
        every byte the call reads or writes (no launch ever sees a short
        buffer)."""
        dev = _check_gpu(inp, out)
        for t, want, need, what in ((inp, self._in_dt, self.in_bytes, "input"),
                                    (out, self._out_dt, self.out_bytes, "output")):
            if t.dtype not in want:
                raise TypeError(f"{what} dtype {t.dtype} does not match the plan ({', '.join(map(str, want))})")
            if _extent(t) < need:
                raise ValueError(f"{what} holds {_extent(t)} bytes but the plan {'reads' if what == 'input' else 'writes'} "
                                 f"{need}")
        return dev

    def execute(self, inp: torch.Tensor, out: torch.Tensor, inverse: bool = False, stream=None) -> None:
        dev = self._check(inp, out)
        _call_on(_dev_idx(dev), stream, self._lib.gfft_execute, self._h, inp.data_ptr(), out.data_ptr(),
                 C.GFFT_INVERSE if inverse else C.GFFT_FORWARD)

    def execute_scaled(self, inp: torch.Tensor, out: torch.Tensor, scales: torch.Tensor, auto: bool,
                       inverse: bool = False, stream=None) -> None:
        """16-bit storage plans: out_b = scales[b] * FFT(in_b); with ``auto`` the
        library first writes power-of-two scales that keep FP16 finite."""
        dev = self._check(inp, out)
        _check_gpu(inp, scales)
        if scales.dtype != torch.float32 or not scales.is_contiguous() or scales.numel() < self.desc.batch:
            raise ValueError("scales must be a contiguous float32 tensor of batch elements")
        _call_on(_dev_idx(dev), stream, self._lib.gfft_execute_scaled, self._h, inp.data_ptr(), out.data_ptr(),
                 C.GFFT_INVERSE if inverse else C.GFFT_FORWARD, scales.data_ptr(),
                 C.GFFT_SCALE_AUTO if auto else C.GFFT_SCALE_GIVEN)

    def execute_scrambled(self, inp: torch.Tensor, out: torch.Tensor, inverse: bool = False, stream=None) -> None:
        """The plan on input whose elements (C2C) or rows (COLS) are in
        bit-reversed order (gfft_execute_scrambled); inp may be out."""
        dev = self._check(inp, out)
        _call_on(_dev_idx(dev), stream, self._lib.gfft_execute_scrambled, self._h, inp.data_ptr(), out.data_ptr(),
                 C.GFFT_INVERSE if inverse else C.GFFT_FORWARD)

    def execute_ptr(self, in_ptr: int, out_ptr: int, inverse: bool = False, stream_handle: int = 0) -> None:
        C.check(self._lib.gfft_execute(self._h, in_ptr, out_ptr,
                                       C.GFFT_INVERSE if inverse else C.GFFT_FORWARD, stream_handle))

    def profile(self, inp: torch.Tensor, out: torch.Tensor, inverse: bool = False, stream=None):
        """Runs once, returning the per-launch kernel durations in ms (HIP events
        recorded on the execution stream around each launch)."""
        dev = self._check(inp, out)
        ms = (ctypes.c_float * 4096)()
        cnt = ctypes.c_int()
        with torch.cuda.device(dev):
            C.check(self._lib.gfft_plan_profile(self._h, inp.data_ptr(), out.data_ptr(),
                                                C.GFFT_INVERSE if inverse else C.GFFT_FORWARD,
                                                _stream_handle(_dev_idx(dev), stream), ms, len(ms),
                                                ctypes.byref(cnt)))
        return [ms[i] for i in range(min(cnt.value, len(ms)))]

    def profile_launch(self, inp: torch.Tensor, out: torch.Tensor, launch: int, reps: int = 20,
                       inverse: bool = False, stream=None) -> float:
        """Mean duration in ms of launch `launch` run `reps` times back to back
        between one HIP event pair on the execution stream."""
        dev = self._check(inp, out)
        ms = ctypes.c_float()
        with torch.cuda.device(dev):
            C.check(self._lib.gfft_plan_profile_launch(self._h, inp.data_ptr(), out.data_ptr(),
                                                       C.GFFT_INVERSE if inverse else C.GFFT_FORWARD,
                                                       _stream_handle(_dev_idx(dev), stream), int(launch),
                                                       int(reps), ctypes.byref(ms)))
        return ms.value

    def workspace_bytes(self) -> int:
        v = ctypes.c_size_t()
        C.check(self._lib.gfft_plan_workspace(self._h, ctypes.byref(v)))
        return v.value

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(1 << 20)
        C.check(self._lib.gfft_plan_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.gfft_plan_destroy(h)
            self._h = None


class UnevenBatch:
    """Transforms of mixed power-of-two lengths in one plan (gfft_batch_plan).

    Job j reads ``in[in_offsets[j] : in_offsets[j] + sizes[j]]`` and writes
    ``out[out_offsets[j] : ...]`` (complex elements of flat buffers), forward or
    unscaled inverse per ``inverse[j]``.  ``mode``: "graph" (capture once,
    replay a HIP graph), "eager", or "per_job" (one plan launch per job).
    """

    MODES = {"graph": C.GFFT_BATCH_GRAPH, "eager": C.GFFT_BATCH_EAGER, "per_job": C.GFFT_BATCH_PER_JOB}

    def __init__(self, sizes, in_offsets, out_offsets, inverse, dtype: torch.dtype = torch.complex64, device=None):
        import numpy as np
        n = len(sizes)
        self._sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        self._in = np.ascontiguousarray(in_offsets, dtype=np.int64)
        self._out = np.ascontiguousarray(out_offsets, dtype=np.int64)
        self._dir = np.ascontiguousarray(np.asarray(inverse, dtype=bool).astype(np.int32))
        if not (len(self._in) == len(self._out) == len(self._dir) == n):
            raise ValueError("sizes, offsets and directions must have the same length")
        self._lib = C.lib()
        self._h = ctypes.c_void_p()
        p64 = ctypes.POINTER(ctypes.c_int64)
        args = (ctypes.byref(self._h), _prec(dtype), n, self._sizes.ctypes.data_as(p64),
                self._in.ctypes.data_as(p64), self._out.ctypes.data_as(p64),
                self._dir.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        if device is not None:
            with torch.cuda.device(device):
                C.check(self._lib.gfft_batch_plan_create(*args))
        else:
            C.check(self._lib.gfft_batch_plan_create(*args))

    def execute(self, inp: torch.Tensor, out: torch.Tensor, mode: str = "graph", stream=None) -> None:
        dev = _check_gpu(inp, out)
        _call_on(_dev_idx(dev), stream, self._lib.gfft_batch_execute, self._h, inp.data_ptr(), out.data_ptr(),
                 self.MODES[mode])

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(1 << 20)
        C.check(self._lib.gfft_batch_plan_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.gfft_batch_plan_destroy(h)
            self._h = None


def autotune(n: int, batch: int, dtype: torch.dtype = torch.complex64, reps: int = 10) -> dict:
    """Times the candidate kernel choices for a batched C2C of length n on the
    current device and keeps the fastest for later plans (gfft_autotune).
    Plans cached before the call keep their old choice; returns the report."""
    import json
    buf = ctypes.create_string_buffer(1 << 20)
    C.check(C.lib().gfft_autotune(_prec(dtype), n, batch, reps, buf, len(buf)))
    with _cache_lock:
        for k in [k for k in (_cache or {}) if k[0] == C.GFFT_C2C and k[1] == n]:
            del _cache[k]
    return json.loads(buf.value.decode())


def save_tuning(path: str) -> None:
    C.check(C.lib().gfft_tuning_save(path.encode()))


def load_tuning(path: str) -> None:
    C.check(C.lib().gfft_tuning_load(path.encode()))


# Plan cache: the analogue of the reference's per-(N, T) singleton
# (include/genFFT/FFTLevel.h:76-92), keyed by the full descriptor and device,
# and guarded by a lock (the reference's cache is not thread-safe).  It holds
# at most _CACHE_PLANS plans and _CACHE_BYTES of their device workspace,
# dropping the least recently used first (a plan still referenced elsewhere
# stays alive until that reference goes).
_cache: "collections.OrderedDict" = None
_cache_lock = threading.Lock()
_CACHE_PLANS = 64
_CACHE_BYTES = 4 << 30


def _plan(kind, n, precision, device, **kw) -> Plan:
    global _cache
    import collections
    key = (kind, n, precision, device.index if device.index is not None else torch.cuda.current_device(),
           tuple(sorted(kw.items())))
    with _cache_lock:
        if _cache is None:
            _cache = collections.OrderedDict()
        p = _cache.get(key)
        if p is not None:
            _cache.move_to_end(key)
            return p
        p = Plan(kind, n, precision=precision, device=device, **kw)
        _cache[key] = p
        held = sum(q.workspace_bytes() for q in _cache.values())
        while len(_cache) > 1 and (len(_cache) > _CACHE_PLANS or held > _CACHE_BYTES):
            _, old = _cache.popitem(last=False)
            held -= old.workspace_bytes()
        return p


def clear_plan_cache() -> None:
    """Drops every cached plan (their device workspace is freed once no
    caller holds them)."""
    with _cache_lock:
        if _cache is not None:
            _cache.clear()


def _batch_of(t: torch.Tensor, n: int) -> int:
    if t.numel() % n:
        raise ValueError(f"tensor of {t.numel()} elements is not a whole number of length-{n} rows")
    return max(1, t.numel() // n)


def _need_contig(*ts):
    for t in ts:
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


class FFT:
    """genfft::FFT<T>(n): batched 1D complex transforms over the last axis."""

    def __init__(self, n: int = 0, dtype: torch.dtype = torch.complex64):
        self.n = int(n)
        self.dtype = dtype
        self._prec = _prec_complex(dtype)

    def size(self) -> int:
        return self.n

    def __bool__(self) -> bool:
        return self.n > 0

    def transform(self, out: torch.Tensor, inp: torch.Tensor, inv: bool = False) -> None:
        """fft.h:80-85: natural-order transform, unscaled inverse."""
        dev = _check_gpu(out, inp)
        _need_contig(out, inp)
        b = _batch_of(inp, self.n)
        _plan(C.GFFT_C2C, self.n, self._prec, dev, batch=b).execute(inp, out, inverse=inv)

    def transform_no_scramble(self, inout: torch.Tensor, inv: bool = False) -> None:
        """fft.h:69-73: in place, input already in bit-reversed order."""
        dev = _check_gpu(inout)
        _need_contig(inout)
        b = _batch_of(inout, self.n)
        _plan(C.GFFT_C2C, self.n, self._prec, dev, batch=b).execute_scrambled(inout, inout, inverse=inv)

    def forward(self, out, inp):
        self.transform(out, inp, False)

    def inverse(self, out, inp):
        self.transform(out, inp, True)

    def transform_real(self, out: torch.Tensor, inp: torch.Tensor) -> None:
        """fft.h:90-94: real input, full n-bin spectrum."""
        dev = _check_gpu(out, inp)
        _need_contig(out, inp)
        b = _batch_of(inp, self.n)
        _plan(C.GFFT_R2C, self.n, self._prec, dev, batch=b).execute(inp, out)

    def transform_interleave(self, out: torch.Tensor, in1: torch.Tensor, in2: torch.Tensor) -> None:
        """fft.h:100-105: FFT of in1 + i*in2 (split later with separate_2x_real_FFT)."""
        dev = _check_gpu(out, in1, in2)
        _need_contig(out, in1, in2)
        b = _batch_of(in1, self.n)
        rdt = _REAL_OF[self.dtype]
        if in1.dtype != rdt or in2.dtype != rdt or out.dtype != self.dtype:
            raise TypeError(f"transform_interleave takes {rdt} inputs and a {self.dtype} output")
        if in2.numel() < b * self.n or out.numel() < b * self.n:
            raise ValueError("in2 and out need as many elements as in1")
        _call_on(_dev_idx(dev), None, C.lib().gfft_interleave_real, self._prec, out.data_ptr(), in1.data_ptr(),
                 in2.data_ptr(), self.n, b)
        _plan(C.GFFT_C2C, self.n, self._prec, dev, batch=b).execute(out, out)


class RealFFT:
    """genfft::RealFFT<T>(n).forward(out, in, half) (FFTReal.h:188-221)."""

    def __init__(self, n: int = 0, dtype: torch.dtype = torch.float32):
        self.n = int(n)
        self._prec = _prec_real(dtype)

    def size(self) -> int:
        return self.n

    def forward(self, out: torch.Tensor, inp: torch.Tensor, half: bool) -> None:
        dev = _check_gpu(out, inp)
        _need_contig(out, inp)
        b = _batch_of(inp, self.n)
        nout = self.n // 2 + 1 if half else self.n
        if out.numel() < b * nout:
            raise ValueError("output too small")
        od = out.numel() // b
        _plan(C.GFFT_R2C, self.n, self._prec, dev, batch=b, half=half, out_dist=od).execute(inp, out)

    def inverse(self, out: torch.Tensor, inp: torch.Tensor) -> None:
        """Half spectrum (n/2+1 bins) -> n reals, unscaled (the reference's
        RealFFT2D::inverse is a TODO, FFTReal.h:127-130; this is the 1D c2r)."""
        dev = _check_gpu(out, inp)
        _need_contig(out, inp)
        b = _batch_of(out, self.n)
        _plan(C.GFFT_C2R, self.n, self._prec, dev, batch=b).execute(inp, out, inverse=True)


class DIT:
    """genfft::DIT<T>(n).apply(out, in, half): the real-FFT unpack step."""

    def __init__(self, n: int = 0, dtype: torch.dtype = torch.complex64):
        self.n = int(n)
        self._prec = _prec_complex(dtype)

    def size(self) -> int:
        return self.n

    def apply(self, out: torch.Tensor, inp: torch.Tensor, half: bool) -> None:
        dev = _check_gpu(out, inp)
        _need_contig(out, inp)
        nin = max(1, self.n // 2)
        b = _batch_of(inp, nin)
        od = out.numel() // b
        _plan(C.GFFT_DIT, self.n, self._prec, dev, batch=b, half=half, out_dist=od).execute(inp, out)


class FFTVert:
    """genfft::FFTVert<T>(n): n-point FFTs down `cols` columns (fft.h:117-171)."""

    def __init__(self, n: int = 0, dtype: torch.dtype = torch.complex64):
        self.n = int(n)
        self._prec = _prec_complex(dtype)

    def size(self) -> int:
        return self.n

    def transform(self, out, out_stride, inp, in_stride=None, cols=None, inv: bool = False) -> None:
        if in_stride is None:  # transform(out, in, cols) overload, fft.h:158-163
            out, inp, cols = out, out_stride, inp
            out_stride = in_stride = cols
        dev = _check_gpu(out, inp)
        _plan(C.GFFT_COLS, self.n, self._prec, dev, n2=int(cols), in_stride=int(in_stride),
              out_stride=int(out_stride)).execute(inp, out, inverse=inv)

    def transform_no_scramble(self, data, stride, cols, inv: bool = False) -> None:
        """fft.h:132-136: in place, rows already in bit-reversed order."""
        dev = _check_gpu(data)
        _plan(C.GFFT_COLS, self.n, self._prec, dev, n2=int(cols), in_stride=int(stride),
              out_stride=int(stride)).execute_scrambled(data, data, inverse=inv)


class FFT2D:
    """genfft::FFT2D<T>(width, height) (fft.h:200-245)."""

    def __init__(self, width: int = 0, height: int = 0, dtype: torch.dtype = torch.complex64):
        self.width, self.height = int(width), int(height)
        self._prec = _prec_complex(dtype)

    def cols(self) -> int:
        return self.width

    def rows(self) -> int:
        return self.height

    def __bool__(self) -> bool:
        return self.width > 0 and self.height > 0

    def transform(self, out, out_stride, inp, in_stride, inv: bool = False, batch: int = 1) -> None:
        dev = _check_gpu(out, inp)
        _plan(C.GFFT_C2C_2D, self.width, self._prec, dev, n2=self.height, in_stride=int(in_stride),
              out_stride=int(out_stride), batch=batch).execute(inp, out, inverse=inv)


class RealFFT2D:
    """genfft::RealFFT2D<T>(width, height): forward, and the inverse the reference leaves as a TODO
    (FFTReal.h:71-184)."""

    def __init__(self, width: int = 0, height: int = 0, dtype: torch.dtype = torch.float32):
        self.width, self.height = int(width), int(height)
        self.dtype = dtype
        self._prec = _prec_real(dtype)

    def cols(self) -> int:
        return self.width

    def rows(self) -> int:
        return self.height

    def forward(self, out, out_stride, inp, in_stride) -> None:
        """FFTReal.h:83-104: ``inp`` real, pitch ``in_stride`` scalars (any,
        odd included); ``out`` the full complex spectrum, pitch ``out_stride``."""
        dev = _check_gpu(out, inp)
        _plan(C.GFFT_R2C_2D, self.width, self._prec, dev, n2=self.height, in_stride=int(in_stride),
              out_stride=int(out_stride)).execute(inp, out)

    def forward_2x(self, out, out_stride, in1, in_stride1, in2, in_stride2) -> None:
        """FFTReal.h:106-118: the 2D FFT of in1 + i*in2 (two real images, pitches
        in scalars, each its own), split afterwards per element as
        separate_2x_real_FFT does.  One launch packs every row, then the 2D
        transform runs in place on ``out``."""
        dev = _check_gpu(out, in1, in2)
        cdt = _COMPLEX_OF[self.dtype]
        if in1.dtype != self.dtype or in2.dtype != self.dtype or out.dtype != cdt:
            raise TypeError(f"forward_2x takes {self.dtype} inputs and a {cdt} output")
        w, h, p1, p2, po = self.width, self.height, int(in_stride1), int(in_stride2), int(out_stride)
        for t, pitch, what in ((in1, p1, "in1"), (in2, p2, "in2"), (out, po, "out")):
            if pitch < w:
                raise ValueError(f"{what} pitch {pitch} is smaller than the width {w}")
            if _extent(t) < ((h - 1) * pitch + w) * t.element_size():
                raise ValueError(f"{what} is too small for {h} rows at pitch {pitch}")
        _call_on(_dev_idx(dev), None, C.lib().gfft_interleave_real_2d, self._prec, out.data_ptr(), po,
                 in1.data_ptr(), p1, in2.data_ptr(), p2, w, h)
        _plan(C.GFFT_C2C_2D, w, self._prec, dev, n2=h, in_stride=po, out_stride=po).execute(out, out)

    def inverse(self, out, out_stride, inp, in_stride) -> None:
        """FFTReal.h:120-130 (a TODO in the reference): unscaled inverse of
        forward.  ``inp``: complex spectrum, columns 0..width/2 read, pitch
        ``in_stride`` complex; ``out``: real, pitch ``out_stride`` scalars."""
        dev = _check_gpu(out, inp)
        _plan(C.GFFT_C2R_2D, self.width, self._prec, dev, n2=self.height, in_stride=int(in_stride),
              out_stride=int(out_stride)).execute(inp, out, inverse=True)


def separate_2x_real_FFT(out1: torch.Tensor, out2: torch.Tensor, inp: torch.Tensor, n: int) -> None:
    """FFTReal.h:36-66 on device data (batched over rows of n)."""
    dev = _check_gpu(out1, out2, inp)
    _prec_complex(inp.dtype)
    if out1.dtype != inp.dtype or out2.dtype != inp.dtype:
        raise TypeError("separate_2x_real_FFT takes three tensors of one complex dtype")
    b = _batch_of(inp, n)
    if _extent(out1) < inp.numel() * inp.element_size() or _extent(out2) < inp.numel() * inp.element_size():
        raise ValueError("out1 and out2 need as many elements as the input")
    _call_on(_dev_idx(dev), None, C.lib().gfft_separate_2x_real, _prec(inp.dtype), out1.data_ptr(),
             out2.data_ptr(), inp.data_ptr(), n, b)


# ---------------------------------------------------------------------------
# functional helpers over the last axis (allocate the output)
# ---------------------------------------------------------------------------
def fft(x: torch.Tensor, inverse: bool = False) -> torch.Tensor:
    x = x.contiguous()
    out = torch.empty_like(x)
    FFT(x.shape[-1], x.dtype).transform(out, x, inverse)
    return out


def ifft(x: torch.Tensor) -> torch.Tensor:
    """Unscaled inverse (genFFT convention)."""
    return fft(x, inverse=True)


def rfft(x: torch.Tensor, half: bool = True) -> torch.Tensor:
    _prec_real(x.dtype)
    x = x.contiguous()
    n = x.shape[-1]
    out = torch.empty(*x.shape[:-1], n // 2 + 1 if half else n, dtype=_COMPLEX_OF[x.dtype], device=x.device)
    RealFFT(n, x.dtype).forward(out, x, half)
    return out


def irfft(X: torch.Tensor, n: int) -> torch.Tensor:
    """Unscaled c2r: n reals from n/2+1 Hermitian bins."""
    if X.shape[-1] != n // 2 + 1:
        raise ValueError(f"irfft of length {n} takes {n // 2 + 1} bins per row, got {X.shape[-1]}")
    X = X.contiguous()
    out = torch.empty(*X.shape[:-1], n, dtype=_REAL_OF[X.dtype], device=X.device)
    RealFFT(n, _REAL_OF[X.dtype]).inverse(out, X)
    return out


def fft2(x: torch.Tensor, inverse: bool = False) -> torch.Tensor:
    """2D transform over the last two axes (rows = x.shape[-2], cols = x.shape[-1])."""
    x = x.contiguous()
    H, W = x.shape[-2], x.shape[-1]
    b = max(1, x.numel() // (H * W))
    out = torch.empty_like(x)
    FFT2D(W, H, x.dtype).transform(out, W, x, W, inverse, batch=b)
    return out


from .autograd import SpectralFilter, spectral_filter, spectral_filter_plain  # noqa: E402

__all__ += ["SpectralFilter", "spectral_filter", "spectral_filter_plain"]


# ---------------------------------------------------------------------------
# 16-bit storage (FP16 / BF16 pairs, FP32 arithmetic); no reference counterpart
# ---------------------------------------------------------------------------
_HALF = {torch.float16: C.GFFT_F16, torch.bfloat16: C.GFFT_BF16}


def _as_pairs(x: torch.Tensor) -> torch.Tensor:
    if x.dtype == torch.complex32:
        return torch.view_as_real(x)
    if x.dtype not in _HALF or x.shape[-1] != 2:
        raise TypeError("16-bit transforms take complex32 or float16/bfloat16 [..., n, 2] tensors")
    return x


def _run16(kind, x, n, n2, batch, inverse, scale):
    pairs = _as_pairs(x).contiguous()
    dev = _check_gpu(pairs)
    prec = _HALF[pairs.dtype]
    plan = _plan(kind, n, prec, dev, n2=n2, batch=batch)
    out = torch.empty_like(pairs)
    if scale is None:
        plan.execute(pairs, out, inverse)
        scales = None
    else:
        auto = isinstance(scale, str)
        if auto and scale != "auto":
            raise ValueError("scale must be None, 'auto' or a float32 tensor of batch factors")
        scales = torch.empty(batch, dtype=torch.float32, device=dev) if auto else \
            scale.to(device=dev, dtype=torch.float32).contiguous()
        plan.execute_scaled(pairs, out, scales, auto, inverse)
    if x.dtype == torch.complex32:
        out = out.view(torch.complex32).squeeze(-1)
    return out, scales


def fft16(x: torch.Tensor, inverse: bool = False, scale=None):
    """Batched 1D FFT over the last complex axis with 16-bit storage.

    ``x``: complex32 [..., n] or float16 / bfloat16 [..., n, 2].  ``scale``:
    None (no scaling), "auto" (per-transform power of two keeping FP16
    finite) or a float32 tensor with one factor per transform.  Returns
    (y, scales) with y = scales[b] * FFT(x_b) in x's layout."""
    n = x.shape[-1] if x.dtype == torch.complex32 else x.shape[-2]
    batch = max(1, (x.numel() if x.dtype == torch.complex32 else x.numel() // 2) // n)
    return _run16(C.GFFT_C2C, x, n, 0, batch, inverse, scale)


def fft2_16(x: torch.Tensor, inverse: bool = False, scale=None):
    """2D FFT over the last two complex axes (rows, cols) with 16-bit storage;
    one scale per image.  Same layouts and scale modes as fft16."""
    cplx = x.dtype == torch.complex32
    H, W = (x.shape[-2], x.shape[-1]) if cplx else (x.shape[-3], x.shape[-2])
    batch = max(1, (x.numel() if cplx else x.numel() // 2) // (H * W))
    return _run16(C.GFFT_C2C_2D, x, W, H, batch, inverse, scale)


__all__ += ["fft16", "fft2_16"]
