"""GPU parity of the column and 2D paths:

  FFTVert<T>::transform    include/genFFT/fft.h:146-163   (GFFT_COLS)
  FFT2D<T>::transform      include/genFFT/fft.h:214-218   (GFFT_C2C_2D)
  RealFFT2D<T>::forward    include/genFFT/FFTReal.h:83-104 (GFFT_R2C_2D)

The column table is the reference's own (test/test_dispatch.cpp:96-120),
checked on every column (the reference checks only the first n outputs,
test/fft_test_impl.h:121).  2D results are compared with the separable
long-double oracle; padded pitches must leave the gap bytes untouched.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

VERT_TABLE = [(1, 1), (1, 40), (2, 1), (2, 7), (2, 999), (4, 1), (4, 7), (4, 999), (8, 1), (8, 3), (8, 7),
              (8, 33), (8, 47), (8, 1023), (16, 1), (16, 47), (16, 999), (256, 1), (256, 1023), (1024, 1),
              (1024, 767), (65536, 1), (65536, 31)]


def _tol(prec, n):
    return O.tol_c64(n) if prec == "f32" else O.tol_c128(n)


def _cdt(prec):
    return (np.complex64, torch.complex64) if prec == "f32" else (np.complex128, torch.complex128)


def _rand(shape, dt, rng):
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(dt)


def ref2d(x2d, inv=False):
    """Separable long-double 2D DFT: rows, then columns."""
    r = O.fft_ld(x2d.astype(np.complex128), inv)
    return O.fft_ld(np.ascontiguousarray(r.T), inv).T


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n,cols", VERT_TABLE)
def test_vert_table(gpu, prec, n, cols):
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    rng = np.random.default_rng(n * 1000 + cols)
    x = _rand((n, cols), ndt, rng)
    tx = torch.from_numpy(x.reshape(-1)).to(gpu)
    f = G.FFTVert(n, tdt)
    out = torch.empty_like(tx)
    inv = torch.empty_like(tx)
    f.transform(out, cols, tx, cols, cols, False)
    f.transform(inv, cols, out, cols, cols, True)
    y = out.cpu().numpy().reshape(n, cols)
    ref = O.fft_ld(np.ascontiguousarray(x.T).astype(np.complex128)).T
    assert O.rel_l2_rows(y.T, ref.T).max() <= _tol(prec, n)
    assert O.rel_l2(inv.cpu().numpy().reshape(n, cols) / n, x) <= 2 * _tol(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n,cols,ins,outs", [(8, 7, 9, 8), (256, 33, 40, 35), (4096, 5, 7, 6), (1 << 14, 3, 3, 4)])
def test_vert_padded_pitch(gpu, golden, prec, n, cols, ins, outs):
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    rng = np.random.default_rng(3)
    x = _rand((n, ins), ndt, rng)
    out = torch.full((n * outs,), 7 + 5j, dtype=tdt, device=gpu)
    G.FFTVert(n, tdt).transform(out, outs, torch.from_numpy(x.reshape(-1)).to(gpu), ins, cols, False)
    y = out.cpu().numpy().reshape(n, outs)
    ref = O.fft_ld(np.ascontiguousarray(x[:, :cols].T).astype(np.complex128)).T
    assert O.rel_l2_rows(y[:, :cols].T, ref.T).max() <= _tol(prec, n)
    assert np.all(y[:, cols:] == 7 + 5j)


@pytest.mark.parametrize("big_in,big_out", [(True, False), (False, True)])
def test_vert_pitch_beyond_31_bit_tile_offsets(gpu, big_in, big_out):
    """complex64 column tiles address 16-byte vectors with 32-bit in-tile
    offsets; a row pitch that puts a tile's last row 2^31 bytes or more from
    its first takes the element path instead (4.3 GB buffers)."""
    import genfft_amd as G
    n, cols, big = 1024, 64, (1 << 19) + 2
    assert (n - 1) * big * 8 >= 1 << 31
    ins, outs = (big if big_in else cols), (big if big_out else cols)
    rng = np.random.default_rng(11)
    x = _rand((n, cols), np.complex64, rng)
    tin = torch.empty((n - 1) * ins + cols, dtype=torch.complex64, device=gpu)
    tin.as_strided((n, cols), (ins, 1)).copy_(torch.from_numpy(x).to(gpu))
    out = torch.empty((n - 1) * outs + cols, dtype=torch.complex64, device=gpu)
    f = G.FFTVert(n, torch.complex64)
    f.transform(out, outs, tin, ins, cols, False)
    y = out.as_strided((n, cols), (outs, 1)).cpu().numpy()
    ref = O.fft_ld(np.ascontiguousarray(x.T).astype(np.complex128)).T
    assert O.rel_l2_rows(y.T, ref.T).max() <= _tol("f32", n)
    del tin, out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_vert_matches_genfft_golden(gpu, golden, prec):
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    for (n, cols, ins, outs) in [(8, 7, 7, 7), (16, 47, 47, 47), (256, 33, 40, 35), (2, 999, 999, 999),
                                 (4, 7, 9, 8), (1024, 5, 5, 5)]:
        key = f"vert_{prec}_n{n}_c{cols}_i{ins}_o{outs}"
        eps = 1e-8 + n * 1e-12 if prec == "f64" else 1e-5 + n * 1e-8
        for inv in (0, 1):
            out = torch.full((n * outs,), 7 + 5j, dtype=tdt, device=gpu)
            G.FFTVert(n, tdt).transform(out, outs, torch.from_numpy(golden[key + "_in"]).to(gpu), ins, cols,
                                        bool(inv))
            assert np.max(np.abs(out.cpu().numpy() - golden[key + ("_inv" if inv else "_fwd")])) <= eps


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h,ins,outs", [(32, 16, 33, 35), (8, 64, 8, 8), (64, 4, 64, 70), (16, 1, 16, 16)])
def test_fft2d_matches_genfft_golden(gpu, golden, prec, w, h, ins, outs):
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    key = f"fft2d_{prec}_w{w}_h{h}_i{ins}_o{outs}"
    eps = 1e-8 + w * h * 1e-12 if prec == "f64" else 1e-5 + w * h * 1e-8
    for inv in (0, 1):
        out = torch.full((h * outs,), 7 + 5j, dtype=tdt, device=gpu)
        G.FFT2D(w, h, tdt).transform(out, outs, torch.from_numpy(golden[key + "_in"]).to(gpu), ins, bool(inv))
        assert np.max(np.abs(out.cpu().numpy() - golden[key + ("_inv" if inv else "_fwd")])) <= eps


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (1, 8), (4, 2), (32, 16), (256, 128), (1024, 4096), (4096, 4096)])
def test_fft2d_vs_oracle(gpu, prec, w, h):
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    rng = np.random.default_rng(w * 7 + h)
    x = _rand((h, w), ndt, rng)
    tx = torch.from_numpy(x).to(gpu)
    for inv in (False, True):
        y = G.fft2(tx, inverse=inv).cpu().numpy()
        if w * h <= (1 << 20):
            ref = ref2d(x, inv)
            assert O.rel_l2(y, ref) <= _tol(prec, w * h)
        else:  # large: row/column spot checks through the separable definition
            r = O.fft_ld(x.astype(np.complex128), inv)       # all row transforms
            for c in (0, 1, w // 3, w - 1):
                col = O.fft_ld(np.ascontiguousarray(r[:, c]), inv)
                assert O.rel_l2(y[:, c], col) <= _tol(prec, w * h)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_fft2d_batched_and_in_place(gpu, prec):
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    rng = np.random.default_rng(11)
    x = _rand((3, 64, 128), ndt, rng)
    tx = torch.from_numpy(x).to(gpu)
    y = G.fft2(tx).cpu().numpy()
    for b in range(3):
        assert O.rel_l2(y[b], ref2d(x[b])) <= _tol(prec, 64 * 128)
    t1 = torch.from_numpy(x[0].copy()).to(gpu)
    G.FFT2D(128, 64, tdt).transform(t1, 128, t1, 128, False)
    assert O.rel_l2(t1.cpu().numpy(), ref2d(x[0])) <= _tol(prec, 64 * 128)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h,ins,outs", [(32, 16, 34, 33), (16, 8, 16, 16), (4, 32, 4, 4), (64, 2, 64, 64)])
def test_real2d_matches_genfft_golden(gpu, golden, prec, w, h, ins, outs):
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    key = f"real2d_{prec}_w{w}_h{h}_i{ins}_o{outs}"
    eps = 1e-8 + w * h * 1e-12 if prec == "f64" else 1e-5 + w * h * 1e-8
    out = torch.full((h * outs,), 7 + 5j, dtype=tdt, device=gpu)
    G.RealFFT2D(w, h, torch.float32 if prec == "f32" else torch.float64).forward(
        out, outs, torch.from_numpy(golden[key + "_in"]).to(gpu), ins)
    assert np.max(np.abs(out.cpu().numpy() - golden[key + "_fwd"])) <= eps


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h", [(2, 2), (8, 4), (256, 64), (4096, 4096)])
def test_real2d_vs_oracle(gpu, prec, w, h):
    import genfft_amd as G
    rdt = np.float32 if prec == "f32" else np.float64
    ndt, tdt = _cdt(prec)
    rng = np.random.default_rng(w + h)
    x = rng.standard_normal((h, w)).astype(rdt)
    out = torch.empty(h * w, dtype=tdt, device=gpu)
    G.RealFFT2D(w, h, torch.float32 if prec == "f32" else torch.float64).forward(
        out, w, torch.from_numpy(x).to(gpu), w)
    y = out.cpu().numpy().reshape(h, w)
    if w * h <= (1 << 16):
        assert O.rel_l2(y, ref2d(x.astype(np.complex128))) <= _tol(prec, w * h)
    else:
        r = O.fft_ld(x.astype(np.complex128))
        for c in (0, 1, w // 2, w // 2 + 1, w - 1):
            col = O.fft_ld(np.ascontiguousarray(r[:, c]))
            assert O.rel_l2(y[:, c], col) <= _tol(prec, w * h)
    # 2D Hermitian symmetry X[k][j] = conj X[-k][-j]
    idx_r = (-np.arange(h)) % h
    idx_c = (-np.arange(w)) % w
    # (columns 0 and w/2 are computed, not filled, so they agree to rounding)
    assert O.rel_l2(y, np.conj(y[idx_r][:, idx_c])) <= 2 * _tol(prec, w * h)


def test_fft2d_8192_round_trip(gpu):
    """complex64 8192 x 8192 (the largest 2D config): inverse(forward(x)) = N*x
    and Parseval, as size-independent properties."""
    import genfft_amd as G
    h = w = 8192
    g = torch.Generator(device=gpu).manual_seed(2)
    x = torch.complex(torch.randn(h, w, device=gpu, generator=g), torch.randn(h, w, device=gpu, generator=g))
    X = G.fft2(x)
    ex = torch.linalg.vector_norm(x.to(torch.complex128)).item() ** 2
    eX = torch.linalg.vector_norm(X.to(torch.complex128)).item() ** 2
    assert abs(eX / (h * w) - ex) / ex < 1e-5
    z = G.fft2(X, inverse=True)
    err = (torch.linalg.vector_norm((z / (h * w) - x).to(torch.complex128)) /
           torch.linalg.vector_norm(x.to(torch.complex128))).item()
    assert err <= 2 * O.tol_c64(h * w)


# ---------------------------------------------------------------------------
# RealFFT2D<T>::inverse (FFTReal.h:120-130; the reference body is
# assert(!"TODO")): the unscaled inverse of forward, GFFT_C2R_2D.
# ---------------------------------------------------------------------------
def _ifft2_ld_real(X: np.ndarray) -> np.ndarray:
    """Re(unscaled inverse 2D DFT) of a full spectrum [..., h, w], long double."""
    rows = O.fft_ld(X, inv=True)
    cols = np.swapaxes(O.fft_ld(np.ascontiguousarray(np.swapaxes(rows, -1, -2)), inv=True), -1, -2)
    return cols.real


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h", [(2, 1), (4, 1), (2, 8), (8, 4), (32, 16), (256, 64), (64, 4096), (1024, 256)])
def test_real2d_inverse_vs_oracle(gpu, prec, w, h):
    """inverse of a Hermitian spectrum (the forward's output) against the
    oracle's inverse 2D DFT (real part), and inverse(forward(x)) = w*h*x."""
    import genfft_amd as G
    rdt = np.float32 if prec == "f32" else np.float64
    ndt, tdt = _cdt(prec)
    rt = torch.float32 if prec == "f32" else torch.float64
    rng = np.random.default_rng(3 * w + h)
    x = rng.standard_normal((h, w)).astype(rdt)
    f = G.RealFFT2D(w, h, rt)
    spec = torch.empty(h * w, dtype=tdt, device=gpu)
    f.forward(spec, w, torch.from_numpy(x).to(gpu), w)
    back = torch.full((h * w,), 7.0, dtype=rt, device=gpu)
    f.inverse(back, w, spec, w)
    y = back.cpu().numpy().reshape(h, w)
    ref = _ifft2_ld_real(spec.cpu().numpy().reshape(h, w).astype(np.complex128))
    assert O.rel_l2(y, ref) <= _tol(prec, w * h)
    assert O.rel_l2(y / (w * h), x) <= 2 * _tol(prec, w * h)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_real2d_inverse_padded_and_batched(gpu, prec):
    """Padded pitches (input pitch w/2+1 and w+3 complex, output pitch w+2
    scalars), pitch gaps untouched, and three planes in one plan."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt = np.float32 if prec == "f32" else np.float64
    ndt, tdt = _cdt(prec)
    rt = torch.float32 if prec == "f32" else torch.float64
    w, h = 64, 32
    rng = np.random.default_rng(5)
    x = rng.standard_normal((h, w)).astype(rdt)
    spec = torch.empty(h * w, dtype=tdt, device=gpu)
    G.RealFFT2D(w, h, rt).forward(spec, w, torch.from_numpy(x).to(gpu), w)
    S = spec.view(h, w)
    for ins in (w // 2 + 1, w + 3):
        inp = torch.full((h, ins), 99 + 99j, dtype=tdt, device=gpu)
        inp[:, :w // 2 + 1] = S[:, :w // 2 + 1]     # only columns 0..w/2 are read
        outs = w + 2
        out = torch.full((h, outs), -5.0, dtype=rt, device=gpu)
        G.RealFFT2D(w, h, rt).inverse(out, outs, inp, ins)
        o = out.cpu().numpy()
        assert O.rel_l2(o[:, :w] / (w * h), x) <= 2 * _tol(prec, w * h)
        assert np.all(o[:, w:] == -5.0), "write into the output pitch gap"
    # three planes, dense, one plan
    xs = rng.standard_normal((3, h, w)).astype(rdt)
    specs = torch.stack([torch.empty(h * w, dtype=tdt, device=gpu) for _ in range(3)])
    for i in range(3):
        G.RealFFT2D(w, h, rt).forward(specs[i], w, torch.from_numpy(xs[i]).to(gpu), w)
    outb = torch.empty(3, h, w, dtype=rt, device=gpu)
    p = G.Plan(C.GFFT_C2R_2D, w, precision=0 if prec == "f32" else 1, n2=h, batch=3)
    p.execute(specs, outb, inverse=True)
    assert O.rel_l2(outb.cpu().numpy() / (w * h), xs) <= 2 * _tol(prec, w * h)


@pytest.mark.parametrize("prec,n", [("f32", 4096), ("f32", 8192), ("f64", 4096)])
def test_real2d_inverse_large_round_trip(gpu, prec, n):
    """n x n real planes (the 2D configs): inverse(forward(x)) = n^2 x."""
    import genfft_amd as G
    ndt, tdt = _cdt(prec)
    rt = torch.float32 if prec == "f32" else torch.float64
    g = torch.Generator(device=gpu).manual_seed(n)
    x = torch.randn(n, n, device=gpu, generator=g, dtype=rt)
    spec = torch.empty(n * n, dtype=tdt, device=gpu)
    f = G.RealFFT2D(n, n, rt)
    f.forward(spec, n, x, n)
    back = torch.empty(n, n, device=gpu, dtype=rt)
    f.inverse(back, n, spec, n)
    err = (torch.linalg.vector_norm((back / (n * n) - x).double()) / torch.linalg.vector_norm(x.double())).item()
    assert err <= 2 * _tol(prec, n * n)
