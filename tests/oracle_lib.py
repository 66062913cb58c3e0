"""numpy front end of the CPU oracle (oracle/liboracle.so) for the tests.

Test infrastructure only: the library under test never imports this.
Complex arrays are numpy complex64/complex128; each wrapper calls the C
restatement in oracle/oracle.c (which cites the reference file:line it follows).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libgenfft_ref.so")

_o = None


def lib() -> ctypes.CDLL:
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle`")
        _o = ctypes.CDLL(ORACLE_SO)
    return _o


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


I64 = ctypes.c_int64


def _sfx(dtype) -> str:
    return "f64" if np.dtype(dtype) in (np.float64, np.complex128) else "f32"


def _cdt(dtype):
    return np.complex128 if _sfx(dtype) == "f64" else np.complex64


def _rdt(dtype):
    return np.float64 if _sfx(dtype) == "f64" else np.float32


# ---------------------------------------------------------------- exact refs
def twiddle(m: int, n: int) -> complex:
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().orc_twiddle(I64(m), I64(n), ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def dft(x: np.ndarray, inv: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, np.complex128)
    y = np.empty_like(x)
    lib().orc_dft(_p(x), _p(y), I64(x.size), int(inv))
    return y


def dft_bins(x: np.ndarray, bins, inv: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, np.complex128)
    b = np.ascontiguousarray(bins, np.int64)
    y = np.empty(b.size, np.complex128)
    lib().orc_dft_bins(_p(x), I64(x.size), int(inv), _p(b), I64(b.size), _p(y))
    return y


def fft_f64(x: np.ndarray, inv: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, np.complex128)
    y = np.empty_like(x)
    lib().orc_fft_f64(_p(x), _p(y), I64(x.size), int(inv))
    return y


def fft_ld(x: np.ndarray, inv: bool = False) -> np.ndarray:
    """Long-double FFT over the last axis, returned as complex128."""
    x = np.ascontiguousarray(x, np.complex128)
    n = x.shape[-1]
    y = np.empty_like(x)
    lib().orc_fft_ld(_p(x), _p(y), I64(n), I64(x.size // n), int(inv))
    return y


# ------------------------------------------------ genFFT restatement (T, T)
def c2c(x: np.ndarray, inv: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x)
    y = np.empty_like(x)
    getattr(lib(), "orc_c2c_" + _sfx(x.dtype))(_p(y), _p(x), I64(x.size), int(inv))
    return y


def transform_real(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x)
    y = np.empty(x.size, _cdt(x.dtype))
    getattr(lib(), "orc_c2c_real_" + _sfx(x.dtype))(_p(y), _p(x), I64(x.size))
    return y


def transform_interleave(x1: np.ndarray, x2: np.ndarray) -> np.ndarray:
    x1 = np.ascontiguousarray(x1)
    x2 = np.ascontiguousarray(x2, x1.dtype)
    y = np.empty(x1.size, _cdt(x1.dtype))
    getattr(lib(), "orc_c2c_interleave_" + _sfx(x1.dtype))(_p(y), _p(x1), _p(x2), I64(x1.size))
    return y


def unpack(z: np.ndarray, n: int, half: bool) -> np.ndarray:
    """DIT::apply: z holds max(1, n/2) complex values."""
    z = np.ascontiguousarray(z)
    f = np.zeros(max(n, 2), z.dtype)
    getattr(lib(), "orc_unpack_" + _sfx(z.dtype))(_p(f), _p(z), I64(n), int(half))
    return f[: (n // 2 + 1 if half else n)] if n > 1 else f[:1]


def real_forward(x: np.ndarray, half: bool) -> np.ndarray:
    x = np.ascontiguousarray(x)
    n = x.size
    f = np.zeros(max(n, 2), _cdt(x.dtype))
    getattr(lib(), "orc_real_forward_" + _sfx(x.dtype))(_p(f), _p(x), I64(n), int(half))
    return f[: (n // 2 + 1 if half else n)] if n > 1 else f[:1]


def separate_2x(z: np.ndarray):
    z = np.ascontiguousarray(z)
    a = np.empty_like(z)
    b = np.empty_like(z)
    getattr(lib(), "orc_separate_2x_" + _sfx(z.dtype))(_p(a), _p(b), _p(z), I64(z.size))
    return a, b


def vert(x: np.ndarray, n: int, cols: int, in_stride: int, out_stride: int, inv: bool = False,
         out: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x)
    y = np.zeros(n * out_stride, x.dtype) if out is None else out
    getattr(lib(), "orc_vert_" + _sfx(x.dtype))(_p(y), I64(out_stride), _p(x), I64(in_stride), I64(n),
                                                I64(cols), int(inv))
    return y


def fft2d(x: np.ndarray, width: int, height: int, in_stride: int, out_stride: int, inv: bool = False,
          out: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x)
    y = np.zeros(height * out_stride, x.dtype) if out is None else out
    getattr(lib(), "orc_fft2d_" + _sfx(x.dtype))(_p(y), I64(out_stride), _p(x), I64(in_stride), I64(width),
                                                 I64(height), int(inv))
    return y


def real2d_forward(x: np.ndarray, width: int, height: int, in_stride: int, out_stride: int,
                   out: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x)
    y = np.zeros(height * out_stride, _cdt(x.dtype)) if out is None else out
    getattr(lib(), "orc_real2d_forward_" + _sfx(x.dtype))(_p(y), I64(out_stride), _p(x), I64(in_stride),
                                                          I64(width), I64(height))
    return y


def c2c_batch_mt(x: np.ndarray, n: int, nthreads: int, inv: bool = False) -> np.ndarray:
    """The restatement over a batch, strided across POSIX threads."""
    x = np.ascontiguousarray(x)
    y = np.empty_like(x)
    lib().orc_c2c_batch_mt(_p(y), _p(x), I64(n), I64(x.size // n), int(inv),
                           int(x.dtype == np.complex128), int(nthreads))
    return y


# ---------------------------------------------------------------- metrics
def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def rel_l2_rows(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-row relative L2 over the last axis."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = np.linalg.norm(b, axis=-1)
    den[den == 0] = 1.0
    return np.linalg.norm(a - b, axis=-1) / den


def tol_c64(n: int) -> float:
    """Relative-L2 bound for complex64: 2e-7 * log2(N) (at least one level)."""
    return 2e-7 * max(1, int(np.log2(max(n, 2))))


def tol_c128(n: int) -> float:
    """Relative-L2 bound for complex128: 1e-15 * log2(N)."""
    return 1e-15 * max(1, int(np.log2(max(n, 2))))
