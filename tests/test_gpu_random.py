"""Seeded random plan shapes against the long-double oracle: batched 1D
complex transforms (genfft::FFT<T>::transform, include/genFFT/fft.h:80-85)
with padded batch distances and both directions, column transforms
(FFTVert<T>::transform, fft.h:146-163) with random widths and row pitches,
real-input transforms (RealFFT<T>::forward, FFTReal.h:188-221, half and
Hermitian-filled full output) and their c2r inverse with padded distances,
and 2D transforms (FFT2D<T>::transform, fft.h:214-218) with row pitches.

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


def _real_cases():
    """(kind, prec, n, batch, in pad, out pad, half): r2c (RealFFT::forward,
    FFTReal.h:188-221; half or Hermitian-filled full output) and c2r."""
    rng = np.random.default_rng(31)
    cases = []
    for i in range(30):
        kind = "r2c" if i % 3 else "c2r"
        prec = int(rng.integers(0, 2))
        k = int(rng.integers(15, 22)) if i % 2 else int(rng.integers(1 if kind == "c2r" else 0, 15))
        n = 1 << k
        batch = int(rng.integers(1, max(2, min(7, (12 << 20) // n))))
        # odd real-side pads run through a dense copy in scratch
        pin, pout = int(rng.choice([0, 0, 1, 2, 5])), int(rng.choice([0, 0, 1, 4]))
        half = bool(rng.integers(0, 2)) if kind == "r2c" else True
        cases.append((kind, prec, n, batch, pin, pout, half))
    return cases


def _hermitian_full(F: np.ndarray, n: int) -> np.ndarray:
    """n bins of a real signal's spectrum from its n/2 + 1 (imaginary parts of
    bins 0 and n/2 dropped, as a c2r does)."""
    F = F.copy()
    F[:, 0] = F[:, 0].real
    F[:, n // 2] = F[:, n // 2].real
    return np.concatenate([F, np.conj(F[:, 1:n // 2][:, ::-1])], axis=1)


@pytest.mark.parametrize("kind,prec,n,batch,pin,pout,half", _real_cases())
def test_random_real_shapes(gpu, kind, prec, n, batch, pin, pout, half):
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt = (np.float32, np.complex64) if prec == 0 else (np.float64, np.complex128)
    trdt, tcdt = (torch.float32, torch.complex64) if prec == 0 else (torch.float64, torch.complex128)
    rng = np.random.default_rng(n * 13 + batch + 5 * pin + 7 * pout + half)
    rows = sorted({0, batch // 2, batch - 1})
    nin = n // 2 + 1
    if kind == "r2c":
        bins = nin if half else n
        din, dout = n + pin, bins + pout
        x_np = rng.uniform(-1, 1, (batch, din)).astype(rdt)
        y = torch.full((batch * dout,), SENTINEL, dtype=tcdt, device=gpu)
        p = G.Plan(C.GFFT_R2C, n, precision=prec, batch=batch, half=half, in_dist=din, out_dist=dout)
        p.execute(torch.from_numpy(x_np.reshape(-1)).to(gpu), y)
        yy = y.cpu().numpy().reshape(batch, dout)
        ref = O.fft_ld(x_np[rows, :n].astype(np.complex128))[:, :bins]
        assert O.rel_l2_rows(yy[rows, :bins], ref).max() <= _tol(prec, n), (kind, prec, n, batch, pin, pout, half)
        assert np.all(yy[:, bins:] == SENTINEL)
    else:
        din, dout = nin + pin, n + pout
        F = (rng.standard_normal((batch, din)) + 1j * rng.standard_normal((batch, din))).astype(cdt)
        out = torch.full((batch * dout,), 7.0, dtype=trdt, device=gpu)
        p = G.Plan(C.GFFT_C2R, n, precision=prec, batch=batch, in_dist=din, out_dist=dout)
        p.execute(torch.from_numpy(F.reshape(-1)).to(gpu), out, inverse=True)
        yy = out.cpu().numpy().reshape(batch, dout)
        ref = O.fft_ld(_hermitian_full(F[rows, :nin].astype(np.complex128), n), inv=True).real
        assert O.rel_l2_rows(yy[rows, :n], ref).max() <= _tol(prec, n), (kind, prec, n, batch, pin, pout)
        assert np.all(yy[:, n:] == 7.0)


def _2d_cases():
    rng = np.random.default_rng(77)
    cases = []
    for _ in range(14):
        prec = int(rng.integers(0, 2))
        w, h = 1 << int(rng.integers(1, 12)), 1 << int(rng.integers(0, 12))
        if w * h > (1 << 21):
            h = max(1, (1 << 21) // w)
        ins, outs = w + int(rng.choice([0, 0, 2, 7])), w + int(rng.choice([0, 0, 1, 4]))
        inv = bool(rng.integers(0, 2))
        cases.append((prec, w, h, ins, outs, inv))
    return cases


@pytest.mark.parametrize("prec,w,h,ins,outs,inv", _2d_cases())
def test_random_2d_shapes(gpu, prec, w, h, ins, outs, inv):
    """FFT2D<T>::transform (fft.h:214-218): rows of width w, h of them, row
    pitches ins / outs; separable reference (rows, then columns)."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    cdt, tcdt = (np.complex64, torch.complex64) if prec == 0 else (np.complex128, torch.complex128)
    rng = np.random.default_rng(w * 31 + h + ins + outs + inv)
    x_np = (rng.uniform(-1, 1, (h, ins)) + 1j * rng.uniform(-1, 1, (h, ins))).astype(cdt)
    y = torch.full((h * outs,), SENTINEL, dtype=tcdt, device=gpu)
    p = G.Plan(C.GFFT_C2C_2D, w, precision=prec, n2=h, in_stride=ins, out_stride=outs)
    p.execute(torch.from_numpy(x_np.reshape(-1)).to(gpu), y, inverse=inv)
    yy = y.cpu().numpy().reshape(h, outs)
    ref = O.fft_ld(x_np[:, :w].astype(np.complex128), inv)
    ref = O.fft_ld(np.ascontiguousarray(ref.T), inv).T
    assert O.rel_l2(yy[:, :w], ref) <= _tol(prec, w * h), (prec, w, h, ins, outs, inv)
    assert np.all(yy[:, w:] == SENTINEL)
