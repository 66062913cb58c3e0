"""Host-pointer execution (gfft_execute_host), the memory model of genFFT's
callers (include/genFFT/fft.h:80-85 on host arrays): the call returns with the
result in `out`.  Two transfer paths, chosen by the call's size a side:
  * up to 1 MiB: the launches read and write the plan's pinned buffers in place;
  * above: the runtime's pageable copies through device staging buffers.
Each path against the long-double oracle, forward and unscaled inverse, with
pitch gaps in the output left untouched, and calls of one plan repeated.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _plan(n, batch, prec=0, kind=None, out_dist=0, half=False):
    import genfft_amd as G
    from genfft_amd import _capi as C
    return G.Plan(C.GFFT_C2C if kind is None else kind, n, precision=prec, batch=batch, out_dist=out_dist, half=half)


def _exec_host(plan, x, y, inv=False):
    from genfft_amd import _capi as C
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    C.check(C.lib().gfft_execute_host(plan._h, P(x), P(y), C.GFFT_INVERSE if inv else C.GFFT_FORWARD))


# bytes a side: 64 B, 8 KiB, 64 KiB, 128 KiB, 1 MiB (in place); 2 MiB (pageable)
@pytest.mark.parametrize("n,batch", [(8, 1), (1024, 1), (4096, 2), (16384, 1), (1 << 17, 1), (1 << 18, 1)])
@pytest.mark.parametrize("prec", [0, 1])
def test_host_pointers_vs_oracle(gpu, n, batch, prec):
    dt = np.complex128 if prec else np.complex64
    tol = O.tol_c128(n) if prec else O.tol_c64(n)
    rng = np.random.default_rng(n + prec)
    x = (rng.uniform(-1, 1, (batch, n)) + 1j * rng.uniform(-1, 1, (batch, n))).astype(dt)
    p = _plan(n, batch, prec)
    for inv in (False, True):
        y = np.empty_like(x)
        _exec_host(p, x, y, inv)
        ref = O.fft_ld(x.astype(np.complex128), inv)
        assert O.rel_l2_rows(y, ref).max() <= tol, (n, batch, prec, inv)


@pytest.mark.parametrize("n", [512, 8192, 32768])
def test_host_output_gaps_untouched(gpu, n):
    """out_dist > n: the bytes between the rows keep the caller's values."""
    batch, od = 3, n + 5
    rng = np.random.default_rng(7 * n)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    p = _plan(n, batch, out_dist=od)
    y = np.full(batch * od, np.complex64(3 - 4j))
    _exec_host(p, x, y)
    got = y[: (batch - 1) * od + n].copy()
    rows = np.stack([got[b * od: b * od + n] for b in range(batch)])
    assert O.rel_l2_rows(rows, O.fft_ld(x.astype(np.complex128))).max() <= O.tol_c64(n)
    for b in range(batch - 1):
        assert np.all(y[b * od + n: (b + 1) * od] == np.complex64(3 - 4j))


def test_host_calls_repeat_on_one_plan(gpu):
    """Consecutive calls of one plan reuse its pinned buffers: each result is
    the transform of its own input."""
    n = 2048
    p = _plan(n, 1)
    rng = np.random.default_rng(3)
    for i in range(6):
        x = (rng.uniform(-1, 1, (1, n)) + 1j * rng.uniform(-1, 1, (1, n))).astype(np.complex64)
        y = np.empty_like(x)
        _exec_host(p, x, y, inv=bool(i & 1))
        ref = O.fft_ld(x.astype(np.complex128), bool(i & 1))
        assert O.rel_l2_rows(y, ref).max() <= O.tol_c64(n), i


def test_host_real_half_spectrum(gpu):
    """RealFFT::forward(out, in, half) on host arrays (include/genFFT/FFTReal.h:188-221)."""
    from genfft_amd import _capi as C
    n = 4096
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, n)).astype(np.float32)
    p = _plan(n, 2, kind=C.GFFT_R2C, half=True)
    y = np.empty((2, n // 2 + 1), np.complex64)
    _exec_host(p, x, y)
    ref = O.fft_ld(x.astype(np.complex128))[:, : n // 2 + 1]
    assert O.rel_l2_rows(y, ref).max() <= O.tol_c64(n)
