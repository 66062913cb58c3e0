"""Seeded random plan shapes against the long-double oracle: batched 1D
complex transforms (genfft::FFT<T>::transform, include/genFFT/fft.h:80-85)
with padded batch distances and both directions, and column transforms
(FFTVert<T>::transform, fft.h:146-163) with random widths and row pitches.

The shapes reach the paths fixed sizes leave out: four-step column passes whose
tile count is not a multiple of the persistent grid (uneven runs of tiles per
workgroup), odd distances that disable the 16-byte vector paths, and partial
column tiles.  Every case also checks that nothing outside the transformed
elements is written.  Tolerances as test_gpu_c2c: relative L2 per transform,
complex64 2e-7*log2(N), complex128 1e-15*log2(N).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

SENTINEL = 7 + 5j


def _tol(prec, n):
    return O.tol_c64(max(n, 2)) if prec == 0 else O.tol_c128(max(n, 2))


def _c2c_cases():
    rng = np.random.default_rng(20261203)
    cases = []
    for i in range(36):
        prec = int(rng.integers(0, 2))
        # half the cases in four-step territory (complex64 rows > 16384, complex128 > 8192)
        k = int(rng.integers(15, 22)) if i % 2 else int(rng.integers(0, 15))
        n = 1 << k
        batch = int(rng.integers(1, max(2, min(8, (12 << 20) // n))))
        pad_in = int(rng.choice([0, 0, 1, 2, 6, 64]))
        pad_out = int(rng.choice([0, 0, 1, 3, 8]))
        inv = bool(rng.integers(0, 2))
        cases.append((prec, n, batch, n + pad_in, n + pad_out, inv))
    # tile counts just above the persistent grid (256 CUs x 1): uneven runs per workgroup
    cases += [(0, 1 << 20, 5, 1 << 20, 1 << 20, False), (0, 1 << 20, 7, (1 << 20) + 2, 1 << 20, True),
              (1, 1 << 20, 5, 1 << 20, 1 << 20, False), (0, 1 << 21, 3, 1 << 21, (1 << 21) + 4, False)]
    return cases


@pytest.mark.parametrize("prec,n,batch,din,dout,inv", _c2c_cases())
def test_random_c2c_shapes(gpu, prec, n, batch, din, dout, inv):
    import genfft_amd as G
    from genfft_amd import _capi as C
    dt = np.complex64 if prec == 0 else np.complex128
    rng = np.random.default_rng(n * 131 + batch * 7 + din + 3 * dout + inv)
    x_np = (rng.uniform(-1, 1, (batch, din)) + 1j * rng.uniform(-1, 1, (batch, din))).astype(dt)
    x = torch.from_numpy(x_np.reshape(-1)).to(gpu)
    y = torch.full((batch * dout,), SENTINEL, dtype=x.dtype, device=gpu)
    p = G.Plan(C.GFFT_C2C, n, precision=prec, batch=batch, in_dist=din, out_dist=dout)
    p.execute(x, y, inverse=inv)
    yy = y.cpu().numpy().reshape(batch, dout)
    rows = sorted({0, batch // 2, batch - 1})
    ref = O.fft_ld(x_np[rows, :n].astype(np.complex128), inv)
    assert O.rel_l2_rows(yy[rows, :n], ref).max() <= _tol(prec, n), (prec, n, batch, din, dout, inv)
    assert np.all(yy[:, n:] == SENTINEL)
    assert np.array_equal(x.cpu().numpy(), x_np.reshape(-1))  # out-of-place: input untouched


def _vert_cases():
    rng = np.random.default_rng(1203)
    cases = []
    for _ in range(16):
        prec = int(rng.integers(0, 2))
        n = 1 << int(rng.integers(1, 13))
        cols = int(rng.integers(1, 200))
        ins = cols + int(rng.choice([0, 1, 5, 32]))
        outs = cols + int(rng.choice([0, 2, 3]))
        inv = bool(rng.integers(0, 2))
        cases.append((prec, n, cols, ins, outs, inv))
    return cases


@pytest.mark.parametrize("prec,n,cols,ins,outs,inv", _vert_cases())
def test_random_vert_shapes(gpu, prec, n, cols, ins, outs, inv):
    import genfft_amd as G
    dt, tdt = (np.complex64, torch.complex64) if prec == 0 else (np.complex128, torch.complex128)
    rng = np.random.default_rng(n * 17 + cols * 3 + ins + outs + inv)
    x_np = (rng.uniform(-1, 1, (n, ins)) + 1j * rng.uniform(-1, 1, (n, ins))).astype(dt)
    out = torch.full((n * outs,), SENTINEL, dtype=tdt, device=gpu)
    G.FFTVert(n, tdt).transform(out, outs, torch.from_numpy(x_np.reshape(-1)).to(gpu), ins, cols, inv)
    y = out.cpu().numpy().reshape(n, outs)
    ref = O.fft_ld(np.ascontiguousarray(x_np[:, :cols].T).astype(np.complex128), inv).T
    assert O.rel_l2_rows(y[:, :cols].T, ref.T).max() <= _tol(prec, n), (prec, n, cols, ins, outs, inv)
    assert np.all(y[:, cols:] == SENTINEL)
