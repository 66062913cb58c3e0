"""gfft_execute_noalloc: the execute variant that allocates nothing (no
reference counterpart; for callers that cannot allocate at call time).

It runs a plan with the library's static per-GPU workspace instead of the
plan's own.  Checked through the C ABI against gfft_execute on a plan of the
same descriptor, bit for bit, on two calls in a row (the second with a
smaller input), plus the long-double oracle on sample rows, and the
documented refusals.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _create(C, kind, n, batch, flags=0, prec=0):
    h = ctypes.c_void_p()
    d = C.GfftDesc(kind, prec, n, 0, batch, 0, 0, 0, 0, flags, 0)
    C.check(C.lib().gfft_plan_create(ctypes.byref(h), ctypes.byref(d)))
    return h


def _stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_execute_twice(gpu, prec):
    from genfft_amd import _capi as C
    lib = C.lib()
    cdt = torch.complex64 if prec == 0 else torch.complex128
    # first call: a four-step length whose chunks fill the static workspace;
    # second: a smaller input through another plan on the same workspace
    calls = [(1 << 20, 8 if prec == 0 else 4), (1 << 16, 3)]
    for n, batch in calls:
        g = torch.Generator(device=gpu).manual_seed(n + batch)
        x = torch.complex(torch.randn(batch, n, device=gpu, generator=g, dtype=cdt.to_real()),
                          torch.randn(batch, n, device=gpu, generator=g, dtype=cdt.to_real()))
        a = _create(C, C.GFFT_C2C, n, batch, C.GFFT_FLAG_NO_WORKSPACE, prec)
        b = _create(C, C.GFFT_C2C, n, batch, 0, prec)
        try:
            ws = ctypes.c_size_t()
            C.check(lib.gfft_plan_workspace(a, ctypes.byref(ws)))
            assert 0 < ws.value <= C.GFFT_NOALLOC_SCRATCH_BYTES
            # the workspace-less plan refuses the allocating entry point
            y_bad = torch.empty_like(x)
            assert lib.gfft_execute(a, x.data_ptr(), y_bad.data_ptr(), 0, _stream()) == C.GFFT_ERR_INVALID_ARG
            for direction in (C.GFFT_FORWARD, C.GFFT_INVERSE):
                ya = torch.full_like(x, 3 + 4j)
                yb = torch.full_like(x, -1 - 2j)
                C.check(lib.gfft_execute_noalloc(a, x.data_ptr(), ya.data_ptr(), direction, _stream()))
                C.check(lib.gfft_execute(b, x.data_ptr(), yb.data_ptr(), direction, _stream()))
                torch.cuda.synchronize()
                assert torch.equal(ya, yb), (n, batch, direction)
                # and against the oracle on sample rows
                rows = [0, batch - 1]
                ref = O.fft_ld(x[rows].cpu().numpy().astype(np.complex128), inv=bool(direction))
                tol = O.tol_c64(n) if prec == 0 else O.tol_c128(n)
                assert O.rel_l2_rows(ya[rows].cpu().numpy(), ref).max() <= tol
        finally:
            lib.gfft_plan_destroy(a)
            lib.gfft_plan_destroy(b)


def test_noalloc_runs_a_plan_with_its_own_workspace_and_plans_without_any(gpu):
    """A plan created the ordinary way (with a workspace that fits) runs too;
    a single-pass plan needs no workspace at all."""
    from genfft_amd import _capi as C
    lib = C.lib()
    for n, batch in ((1 << 18, 2), (1024, 64)):
        x = torch.randn(batch, n, dtype=torch.complex64, device=gpu)
        p = _create(C, C.GFFT_C2C, n, batch)
        try:
            ya, yb = torch.empty_like(x), torch.empty_like(x)
            C.check(lib.gfft_execute_noalloc(p, x.data_ptr(), ya.data_ptr(), 0, _stream()))
            C.check(lib.gfft_execute(p, x.data_ptr(), yb.data_ptr(), 0, _stream()))
            torch.cuda.synchronize()
            assert torch.equal(ya, yb)
        finally:
            lib.gfft_plan_destroy(p)


def test_noalloc_refusals(gpu):
    """Beyond the static workspace, or with buffers that would need a staging
    copy, the call fails with a clear error and launches nothing."""
    from genfft_amd import _capi as C
    lib = C.lib()
    # 2^22 x 8 complex64 through the ordinary plan: a 256 MiB workspace
    n, batch = 1 << 22, 8
    x = torch.zeros(batch * n, dtype=torch.complex64, device=gpu)
    p = _create(C, C.GFFT_C2C, n, batch)
    try:
        assert lib.gfft_execute_noalloc(p, x.data_ptr(), x.data_ptr(), 0, _stream()) == C.GFFT_ERR_UNSUPPORTED
        assert b"static workspace" in lib.gfft_last_error()
    finally:
        lib.gfft_plan_destroy(p)
    # the same descriptor created without a workspace chunks to fit
    p = _create(C, C.GFFT_C2C, n, batch, C.GFFT_FLAG_NO_WORKSPACE)
    lib.gfft_plan_destroy(p)
    # overlapping buffers at unequal distances would need the staging copy
    m, b = 1024, 16
    buf = torch.zeros(b * (m + 2), dtype=torch.complex64, device=gpu)
    h = ctypes.c_void_p()
    d = C.GfftDesc(C.GFFT_C2C, 0, m, 0, b, 0, 0, m, m + 2, 0, 0)
    C.check(lib.gfft_plan_create(ctypes.byref(h), ctypes.byref(d)))
    try:
        assert lib.gfft_execute_noalloc(h, buf.data_ptr(), buf.data_ptr(), 0, _stream()) == C.GFFT_ERR_UNSUPPORTED
        assert b"staging" in lib.gfft_last_error()
    finally:
        lib.gfft_plan_destroy(h)
