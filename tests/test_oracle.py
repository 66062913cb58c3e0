"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

1. The genFFT restatement must reproduce genFFT itself: bit-exactly against the
   generic scalar backend (same operations in the same order, FMA contraction
   off) and within the reference's own FFT_Eps tolerances (test/test_util.h:59-72)
   against its AVX/FMA backend, on the golden vectors in tests/golden/.
2. The exact references (long-double DFT with integer argument reduction,
   long-double and FP64 FFTs) must satisfy closed forms that a dropped term,
   a wrong sign, a wrong index or a transposed operand would break.
"""
import numpy as np
import pytest

import oracle_lib as O

SIZES_C2C = [1, 2, 4, 8, 16, 32, 64, 512, 2048]
SIZES_REAL = [2, 4, 8, 16, 64, 1024]


def fft_eps(sfx, n):
    # test/test_util.h:59-72
    return 1e-8 + n * 1e-12 if sfx == "f64" else 1e-5 + n * 1e-8


# ---------------------------------------------------------------- 1. vs genFFT
@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("n", SIZES_C2C)
@pytest.mark.parametrize("inv", [0, 1])
def test_c2c_matches_genfft(golden, sfx, n, inv):
    x = golden[f"c2c_{sfx}_n{n}_in"]
    y = O.c2c(x, bool(inv))
    d = "inv" if inv else "fwd"
    assert np.array_equal(y, golden[f"c2c_{sfx}_n{n}_{d}_generic"])
    nat = golden[f"c2c_{sfx}_n{n}_{d}_native"]
    assert np.max(np.abs(y - nat)) <= fft_eps(sfx, n)


@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("n", SIZES_REAL)
@pytest.mark.parametrize("half", [0, 1])
def test_real_forward_matches_genfft(golden, sfx, n, half):
    x = golden[f"real_{sfx}_n{n}_in"]
    f = O.real_forward(x, bool(half))
    key = f"real_{sfx}_n{n}_{'half' if half else 'full'}"
    g = golden[key + "_generic"]
    lim = n // 2 + 1 if half else n
    assert np.array_equal(f, g[:lim])
    assert np.max(np.abs(f - golden[key + "_native"][:lim])) <= fft_eps(sfx, n)
    # the reference's sentinel check (test/fft_test_impl.h:102-105)
    assert np.all(g[lim:] == 43 + 21j)


@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("n", SIZES_REAL + [12, 24, 40, 100, 1000])
@pytest.mark.parametrize("half", [0, 1])
def test_unpack_matches_genfft(golden, sfx, n, half):
    """adjust_DIT_impl, bit-exact against genFFT's generic DIT; the sizes that
    are not powers of two run genFFT's dynamic FFTDIT<-1>."""
    z = golden[f"dit_{sfx}_n{n}_in"]
    f = O.unpack(z, n, bool(half))
    g = golden[f"dit_{sfx}_n{n}_{'half' if half else 'full'}"]
    assert np.array_equal(f, g[: f.size])


@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("n", SIZES_REAL)
def test_transform_real_and_interleave(golden, sfx, n):
    x = golden[f"real_{sfx}_n{n}_in"]
    x2 = golden[f"treal_{sfx}_n{n}_in2"]
    assert np.max(np.abs(O.transform_real(x) - golden[f"treal_{sfx}_n{n}"])) <= fft_eps(sfx, n)
    assert np.max(np.abs(O.transform_interleave(x, x2) - golden[f"tinter_{sfx}_n{n}"])) <= fft_eps(sfx, n)


@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("n", [8, 64])
def test_separate_matches_genfft(golden, sfx, n):
    a, b = O.separate_2x(golden[f"sep_{sfx}_n{n}_in"])
    assert np.array_equal(a, golden[f"sep_{sfx}_n{n}_x"])
    assert np.array_equal(b, golden[f"sep_{sfx}_n{n}_y"])


VERT = [(8, 7, 7, 7), (16, 47, 47, 47), (256, 33, 40, 35), (2, 999, 999, 999), (4, 7, 9, 8), (1024, 5, 5, 5)]


@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("n,cols,ins,outs", VERT)
@pytest.mark.parametrize("inv", [0, 1])
def test_vert_matches_genfft(golden, sfx, n, cols, ins, outs, inv):
    key = f"vert_{sfx}_n{n}_c{cols}_i{ins}_o{outs}"
    x = golden[key + "_in"]
    g = golden[key + ("_inv" if inv else "_fwd")]
    y = np.full(n * outs, 7 + 5j, x.dtype)
    O.vert(x, n, cols, ins, outs, bool(inv), out=y)
    assert np.max(np.abs(y - g)) <= fft_eps(sfx, n)
    # pitch gaps untouched
    assert np.all(y.reshape(n, outs)[:, cols:] == 7 + 5j)


FFT2D = [(32, 16, 33, 35), (8, 64, 8, 8), (64, 4, 64, 70), (16, 1, 16, 16)]


@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("w,h,ins,outs", FFT2D)
@pytest.mark.parametrize("inv", [0, 1])
def test_fft2d_matches_genfft(golden, sfx, w, h, ins, outs, inv):
    key = f"fft2d_{sfx}_w{w}_h{h}_i{ins}_o{outs}"
    x = golden[key + "_in"]
    y = np.full(h * outs, 7 + 5j, x.dtype)
    O.fft2d(x, w, h, ins, outs, bool(inv), out=y)
    assert np.max(np.abs(y - golden[key + ("_inv" if inv else "_fwd")])) <= fft_eps(sfx, w * h)


REAL2D = [(32, 16, 34, 33), (16, 8, 16, 16), (4, 32, 4, 4), (64, 2, 64, 64)]


@pytest.mark.parametrize("sfx", ["f32", "f64"])
@pytest.mark.parametrize("w,h,ins,outs", REAL2D)
def test_real2d_matches_genfft(golden, sfx, w, h, ins, outs):
    key = f"real2d_{sfx}_w{w}_h{h}_i{ins}_o{outs}"
    x = golden[key + "_in"]
    cdt = np.complex64 if sfx == "f32" else np.complex128
    y = np.full(h * outs, 7 + 5j, cdt)
    O.real2d_forward(x, w, h, ins, outs, out=y)
    assert np.max(np.abs(y - golden[key + "_fwd"])) <= fft_eps(sfx, w * h)


# ----------------------------------------------------- 2. exact references
def test_twiddle_special_values():
    for n in [4, 8, 64, 1 << 20]:
        assert O.twiddle(0, n) == 1 + 0j
        assert O.twiddle(n // 4, n) == -1j          # W_N^(N/4) = -i exactly
        assert O.twiddle(n // 2, n) == -1 + 0j
        assert O.twiddle(3 * n // 4, n) == 1j
    r = np.float64(np.sqrt(0.5))
    for n in [8, 1024]:
        w = O.twiddle(n // 8, n)
        assert w.real == r and w.imag == -r          # FP64-rounded sqrt(1/2)
    # exact integer reduction: m and m + k*n give the same value
    assert O.twiddle(5, 96) == O.twiddle(5 + 7 * 96, 96)


def test_dft_closed_forms():
    rng = np.random.default_rng(1)
    for n in [1, 2, 8, 64, 256]:
        d = np.zeros(n, complex)
        d[0] = 1
        assert np.allclose(O.dft(d), np.ones(n), atol=1e-15)            # delta -> ones
        assert np.allclose(O.dft(np.ones(n)), n * d, atol=1e-12)        # constant -> N delta
        for m in [1, n // 3, n - 1]:
            tone = np.exp(2j * np.pi * m * np.arange(n) / n)             # e^{+2 pi i m j/N} -> N delta_m
            want = np.zeros(n, complex)
            want[m % n] = n
            assert np.allclose(O.dft(tone), want, atol=1e-11)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        X = O.dft(x)
        assert np.isclose(np.vdot(X, X).real, n * np.vdot(x, x).real, rtol=1e-13)  # Parseval
        s = 3 % n
        assert np.allclose(O.dft(np.roll(x, s)), X * np.exp(-2j * np.pi * s * np.arange(n) / n),
                           atol=1e-11)                                  # shift theorem
        assert np.allclose(O.dft(X, inv=True), n * x, atol=1e-10)       # unscaled inverse


def test_dft_vs_numpy_and_bins():
    rng = np.random.default_rng(2)
    for n in [16, 1024, 4096]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        X = O.dft(x)
        assert O.rel_l2(X, np.fft.fft(x)) < 1e-15
        bins = [0, 1, n // 2, n - 1, 37 % n]
        assert np.allclose(O.dft_bins(x, bins), X[np.array(bins) % n], rtol=0, atol=1e-12)


def test_fft_ld_and_f64_vs_exact_dft():
    rng = np.random.default_rng(3)
    for n in [2, 8, 128, 2048, 8192]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        X = O.dft(x)
        # both are rounded once to double at the end: at most ~half an ulp
        # apart, independent of n, while the FP64 FFT grows with log2(n)
        assert O.rel_l2(O.fft_ld(x), X) < 1.2e-16
        assert O.rel_l2(O.fft_f64(x), X) < 2e-16 * max(1, np.log2(n))
        assert O.rel_l2(O.fft_ld(X, inv=True), n * x) < 2.5e-16


def test_fft_ld_large_spot_bins():
    rng = np.random.default_rng(4)
    n = 1 << 18
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = O.fft_ld(x)
    bins = rng.integers(0, n, 16)
    exact = O.dft_bins(x, bins)
    assert O.rel_l2(X[bins], exact) < 1.2e-16
    assert O.rel_l2(X, np.fft.fft(x)) < 1e-15


def test_batched_fft_ld_rows_are_independent():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, 64)) + 1j * rng.standard_normal((3, 64))
    Y = O.fft_ld(x)
    for r in range(3):
        assert O.rel_l2(Y[r], O.dft(x[r])) < 1e-16


def test_real_paths_closed_forms():
    rng = np.random.default_rng(6)
    for n in [2, 4, 16, 256]:
        x = rng.standard_normal(n).astype(np.float64)
        F = O.real_forward(x, half=False)
        assert np.isclose(F[0].real, x.sum()) and F[0].imag == 0                  # F_0 = sum x
        assert np.isclose(F[n // 2].real, (x * (-1.0) ** np.arange(n)).sum())     # F_{N/2}
        assert np.allclose(F[1:], np.conj(F[1:][::-1]), atol=1e-12)             # Hermitian
        assert O.rel_l2(F, O.dft(x.astype(complex))) < 1e-15


def test_multithreaded_batch_equals_serial():
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(16 * 256) + 1j * rng.standard_normal(16 * 256)).astype(np.complex64)
    y = O.c2c_batch_mt(x, 256, 4)
    for b in range(16):
        assert np.array_equal(y[b * 256:(b + 1) * 256], O.c2c(x[b * 256:(b + 1) * 256]))


def test_planted_mistakes_are_caught():
    """A sign flip or index slip in an FFT fails the exact-DFT check used above."""
    rng = np.random.default_rng(8)
    n = 64
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = O.dft(x)
    wrong_sign = O.fft_ld(x, inv=True)            # conjugated kernel
    wrong_index = np.roll(O.fft_ld(x), 1)          # off-by-one output order
    conj_input = O.fft_ld(np.conj(x))              # transposed re/im handling
    for bad in (wrong_sign, wrong_index, conj_input):
        assert O.rel_l2(bad, X) > 1e-3
