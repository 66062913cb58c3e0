"""GPU parity of the real-input paths:

  RealFFT<T>::forward(out, in, half)   include/genFFT/FFTReal.h:204-213
  DIT<T>::apply(out, in, half)         include/genFFT/fft.h:180-188
  FFT<T>::transform_real / _interleave include/genFFT/fft.h:90-105
  separate_2x_real_FFT                 include/genFFT/FFTReal.h:36-66
  c2r (half spectrum -> reals)          no reference counterpart (FFTReal.h:127-130
                                       is a TODO); pinned by the round trip and
                                       the oracle's inverse DFT

Sizes follow test/test_real_fft.cpp:58-61 (2 .. 2^22, half and full) and
test/test_dit.cpp:57-61.  The half-spectrum case keeps the reference's
sentinel check: nothing past bin n/2 may be written (test/fft_test_impl.h:88-105).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZES = [1 << k for k in range(1, 23)]


def _tol(prec, n):
    return O.tol_c64(n) if prec == "f32" else O.tol_c128(n)


def _dts(prec):
    return (np.float32, np.complex64, torch.float32, torch.complex64) if prec == "f32" else \
        (np.float64, np.complex128, torch.float64, torch.complex128)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("half", [False, True])
def test_real_forward_vs_oracle(gpu, prec, n, half):
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    b = max(2, min(64, (1 << 16) // n))
    rng = np.random.default_rng(n + half)
    x = rng.uniform(-1, 1, (b, n)).astype(rdt)
    nout = n // 2 + 1 if half else n
    # each row of `out` holds n complex slots, like the reference test's buffer
    out = torch.full((b, n), 43 + 21j, dtype=tcdt, device=gpu)
    from genfft_amd import _capi as C
    p = G.Plan(C.GFFT_R2C, n, precision=0 if prec == "f32" else 1, batch=b, half=half, out_dist=n)
    p.execute(torch.from_numpy(x).to(gpu), out)
    y = out.cpu().numpy()
    rows = [0, b - 1]
    ref = O.fft_ld(x[rows].astype(np.complex128))[:, :nout]
    assert O.rel_l2_rows(y[rows, :nout], ref).max() <= _tol(prec, n)
    assert np.all(y[:, nout:] == 43 + 21j), "write past the half spectrum"


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 1024])
@pytest.mark.parametrize("half", [False, True])
def test_real_forward_matches_genfft(gpu, golden, prec, n, half):
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    x = golden[f"real_{prec}_n{n}_in"]
    nout = n // 2 + 1 if half else n
    y = G.rfft(torch.from_numpy(x).to(gpu)[None], half=half).cpu().numpy()[0]
    g = golden[f"real_{prec}_n{n}_{'half' if half else 'full'}_native"][:nout]
    eps = 1e-8 + n * 1e-12 if prec == "f64" else 1e-5 + n * 1e-8
    assert np.max(np.abs(y - g)) <= eps


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64, 1024, 1 << 14, 1 << 20])
@pytest.mark.parametrize("half", [False, True])
def test_dit_apply(gpu, prec, n, half):
    """DIT(n).apply on the n/2-point FFT of the packed signal equals the
    full FFT of the real signal (test/fft_test_impl.h:60-82), in place too."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    rng = np.random.default_rng(5 + n)
    x = rng.uniform(-1, 1, max(n, 2)).astype(rdt)[:n]
    if n == 1:
        z = np.array([x[0] + 0j], cdt)
    else:
        z = O.fft_ld((x[0::2] + 1j * x[1::2]).astype(np.complex128)).astype(cdt)
    nout = n // 2 + 1 if half else n
    want = O.unpack(z, n, half)
    tz = torch.from_numpy(z).to(gpu)
    out = torch.full((max(n, 2),), 43 + 21j, dtype=tcdt, device=gpu)
    p = G.Plan(C.GFFT_DIT, n, precision=0 if prec == "f32" else 1, half=half)
    p.execute(tz, out)
    y = out.cpu().numpy()
    assert O.rel_l2(y[:nout], want) <= _tol(prec, n)
    if n > 1:
        full = O.fft_ld(x.astype(np.complex128))[:nout]
        assert O.rel_l2(y[:nout], full) <= 2 * _tol(prec, n)
    assert np.all(y[nout:] == 43 + 21j)
    # in place: F overwrites Z in a buffer of n slots
    if n >= 2:
        buf = torch.full((n,), 43 + 21j, dtype=tcdt, device=gpu)
        buf[: n // 2] = tz
        p.execute(buf, buf)
        assert O.rel_l2(buf.cpu().numpy()[:nout], want) <= _tol(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [2, 4, 16, 1024, 1 << 15, 1 << 20])
def test_c2r_round_trip(gpu, prec, n):
    """irfft(rfft(x)) = n * x (unscaled), and c2r equals the oracle's inverse DFT
    of the Hermitian-extended spectrum."""
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    rng = np.random.default_rng(77 + n)
    x = rng.standard_normal((3, n)).astype(rdt)
    tx = torch.from_numpy(x).to(gpu)
    F = G.rfft(tx, half=True)
    back = G.irfft(F, n).cpu().numpy()
    assert O.rel_l2(back / n, x) <= 2 * _tol(prec, n)
    Fn = F.cpu().numpy()
    full = np.concatenate([Fn, np.conj(Fn[:, 1:n // 2][:, ::-1])], axis=1) if n > 2 else Fn
    ref = O.fft_ld(full.astype(np.complex128), inv=True).real
    assert O.rel_l2(back, ref) <= _tol(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 1024])
def test_transform_real_and_interleave(gpu, golden, prec, n):
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    x = golden[f"real_{prec}_n{n}_in"]
    x2 = golden[f"treal_{prec}_n{n}_in2"]
    eps = 1e-8 + n * 1e-12 if prec == "f64" else 1e-5 + n * 1e-8
    f = G.FFT(n, tcdt)
    out = torch.empty(n, dtype=tcdt, device=gpu)
    f.transform_real(out, torch.from_numpy(x).to(gpu))
    assert np.max(np.abs(out.cpu().numpy() - golden[f"treal_{prec}_n{n}"])) <= eps
    f.transform_interleave(out, torch.from_numpy(x).to(gpu), torch.from_numpy(x2).to(gpu))
    z = out.cpu().numpy()
    assert np.max(np.abs(z - golden[f"tinter_{prec}_n{n}"])) <= eps
    # separate_2x_real_FFT recovers both spectra, aliasing out1 with the input
    o1 = out.clone()
    o2 = torch.empty_like(out)
    G.separate_2x_real_FFT(o1, o2, o1, n)
    a, b = O.separate_2x(z)
    sep_tol = 1e-6 if prec == "f32" else 1e-14
    assert np.max(np.abs(o1.cpu().numpy() - a)) <= sep_tol
    assert np.max(np.abs(o2.cpu().numpy() - b)) <= sep_tol
    assert np.max(np.abs(o1.cpu().numpy() - golden[f"treal_{prec}_n{n}"])) <= 2 * eps


def test_real_full_size_properties(gpu):
    """R2C at the benchmark shape (float32, N = 2^20, batch 256): Hermitian
    consistency with the full spectrum, F_0 = sum x, spot rows vs the oracle."""
    import genfft_amd as G
    n, b = 1 << 20, 256
    g = torch.Generator(device=gpu).manual_seed(1)
    x = torch.randn(b, n, device=gpu, generator=g)
    h = G.rfft(x, half=True)
    s = x.to(torch.float64).sum(dim=1)
    assert ((h[:, 0].real.to(torch.float64) - s).abs() / s.abs().clamp_min(1)).max().item() < 1e-4
    for r in (0, 255):
        ref = O.fft_ld(x[r].cpu().numpy().astype(np.complex128))[: n // 2 + 1]
        assert O.rel_l2(h[r].cpu().numpy(), ref) <= O.tol_c64(n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("lg", range(15, 22))
@pytest.mark.parametrize("pad", [0, 3])
def test_real_half_fused_unpack_vs_oracle(gpu, prec, lg, pad):
    """Half spectra of four-step lengths, where the unpack runs inside the
    last pass (row k_c and its mirror n1 - k_c in one workgroup): every bin
    of every transform against the oracle, dense (odd distance n/2 + 1) and
    padded output; the pad between spectra stays untouched."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    n, b = 1 << lg, 3
    nout = n // 2 + 1
    od = nout + pad
    rng = np.random.default_rng(lg * 10 + pad)
    x = rng.standard_normal((b, n)).astype(rdt)
    out = torch.full((b * od,), 43 + 21j, dtype=tcdt, device=gpu)
    p = G.Plan(C.GFFT_R2C, n, precision=0 if prec == "f32" else 1, batch=b, half=True, out_dist=od)
    fused = n // 2 > (16384 if prec == "f32" else 8192)
    assert ("r2c_rows_unpack" in p.describe()) == fused
    p.execute(torch.from_numpy(x).to(gpu), out)
    y = out.cpu().numpy().reshape(b, od)
    ref = O.fft_ld(x.astype(np.complex128))[:, :nout]
    assert O.rel_l2_rows(y[:, :nout], ref).max() <= _tol(prec, n)
    assert np.all(y[:, nout:] == 43 + 21j), "write between spectra"
    assert np.all(y[:, 0].imag == 0) and np.all(y[:, nout - 1].imag == 0)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_real_half_fused_matches_unfused_and_chunks(gpu, prec, monkeypatch):
    """N = 2^20: the fused plan and the separate-unpack plan (whole batch, and
    chunked through a small scratch with an even output distance) agree to
    rounding; spot rows against the oracle."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    n, b = 1 << 20, 5
    nout = n // 2 + 1
    g = torch.Generator(device=gpu).manual_seed(7)
    x = torch.randn(b, n, device=gpu, generator=g, dtype=trdt)
    small = str(2 * (n // 2) * (8 if prec == "f32" else 16))
    outs = {}
    for name, env, od, want in (("fused", {}, nout, "r2c_rows_unpack"),
                                ("unfused", {"GFFT_R2C_FUSE": "0"}, nout, "unpack n="),
                                ("chunked", {"GFFT_CHUNK_BYTES": small}, nout + 1, "unpack n=")):
        with monkeypatch.context() as m:
            for k, v in env.items():
                m.setenv(k, v)
            p = G.Plan(C.GFFT_R2C, n, precision=0 if prec == "f32" else 1, batch=b, half=True, out_dist=od)
        assert want in p.describe(), name
        y = torch.empty(b, od, dtype=tcdt, device=gpu)
        p.execute(x, y)
        outs[name] = y[:, :nout].cpu().numpy()
    assert outs["chunked"].shape == outs["fused"].shape
    tol = _tol(prec, n)
    for name in ("chunked", "unfused"):
        assert O.rel_l2_rows(outs[name], outs["fused"]).max() <= tol
    ref = O.fft_ld(x[[0, b - 1]].cpu().numpy().astype(np.complex128))[:, :nout]
    assert O.rel_l2_rows(outs["fused"][[0, b - 1]], ref).max() <= tol


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_real_half_in_place_dense(gpu, prec, monkeypatch):
    """in == out with the dense layout (n reals in, n/2 + 1 complex out per
    transform): each transform's bins overlap the next transform's input, so
    every input must be read before a bin is written, also when the batch
    exceeds the chunk size."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    n, b = 1 << 16, 6
    nout = n // 2 + 1
    rng = np.random.default_rng(3)
    x = rng.standard_normal((b, n)).astype(rdt)
    buf = torch.zeros(b * nout, dtype=tcdt, device=gpu)
    xin = torch.view_as_real(buf).reshape(-1)[: b * n].view(b, n)
    xin.copy_(torch.from_numpy(x).to(gpu))
    monkeypatch.setenv("GFFT_CHUNK_BYTES", str(2 * (n // 2) * (8 if prec == "f32" else 16)))
    p = G.Plan(C.GFFT_R2C, n, precision=0 if prec == "f32" else 1, batch=b, half=True)
    assert "r2c_rows_unpack" in p.describe()
    p.execute(xin, buf.view(b, nout))
    y = buf.view(b, nout).cpu().numpy()
    ref = O.fft_ld(x.astype(np.complex128))[:, :nout]
    assert O.rel_l2_rows(y, ref).max() <= _tol(prec, n)


def _hermitian_full(F: np.ndarray, n: int) -> np.ndarray:
    """The n-bin spectrum of a real signal from its n/2 + 1 bins (imaginary
    parts of bins 0 and n/2 dropped, as any c2r does)."""
    F = F.copy()
    F[:, 0] = F[:, 0].real
    F[:, n // 2] = F[:, n // 2].real
    return np.concatenate([F, np.conj(F[:, 1:n // 2][:, ::-1])], axis=1)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("lg", range(15, 23))
@pytest.mark.parametrize("pad", [0, 5])
def test_c2r_fused_pack_vs_oracle(gpu, prec, lg, pad):
    """c2r of four-step lengths, where the pack runs inside the first pass
    (column c and its mirror n2 - c in one workgroup): every sample of every
    transform against the oracle's inverse DFT of the Hermitian-extended
    spectrum (non-zero imaginary parts at bins 0 and n/2 included, which a
    c2r ignores), dense and padded distances."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    n, b = 1 << lg, 3
    nin = n // 2 + 1
    ind, outd = nin + pad, n + 2 * pad
    rng = np.random.default_rng(lg * 3 + pad)
    F = (rng.standard_normal((b, nin)) + 1j * rng.standard_normal((b, nin))).astype(cdt)
    src = torch.zeros(b, ind, dtype=tcdt, device=gpu)
    src[:, :nin] = torch.from_numpy(F).to(gpu)
    out = torch.full((b, outd), 7.0, dtype=trdt, device=gpu)
    p = G.Plan(C.GFFT_C2R, n, precision=0 if prec == "f32" else 1, batch=b, in_dist=ind, out_dist=outd)
    p.execute(src, out, inverse=True)
    y = out.cpu().numpy()
    ref = O.fft_ld(_hermitian_full(F.astype(np.complex128), n), inv=True).real
    assert O.rel_l2_rows(y[:, :n], ref).max() <= _tol(prec, n)
    assert np.all(y[:, n:] == 7.0), "write between outputs"


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c2r_fused_in_place_and_matches_unfused(gpu, prec, monkeypatch):
    """N = 2^20: the fused c2r in place (reals over their own spectrum, dense
    n/2 + 1 -> n layout) and the separate-pack plan agree to rounding."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    n, b = 1 << 20, 4
    nin = n // 2 + 1
    g = torch.Generator(device=gpu).manual_seed(11)
    F = G.rfft(torch.randn(b, n, device=gpu, generator=g, dtype=trdt), half=True)
    p = G.Plan(C.GFFT_C2R, n, precision=0 if prec == "f32" else 1, batch=b, out_dist=2 * nin)
    assert "c2r_pack_cols" in p.describe()
    buf = F.clone()
    p.execute(buf, buf, inverse=True)  # out_dist 2 * nin reals: transform b's reals over its own bins
    fused = torch.view_as_real(buf).reshape(b, 2 * nin)[:, :n].cpu().numpy()
    with monkeypatch.context() as m:
        m.setenv("GFFT_R2C_FUSE", "0")
        q = G.Plan(C.GFFT_C2R, n, precision=0 if prec == "f32" else 1, batch=b)
    assert "c2r_pack_cols" not in q.describe()
    y = torch.empty(b, n, dtype=trdt, device=gpu)
    q.execute(F, y, inverse=True)
    assert O.rel_l2_rows(fused, y.cpu().numpy()).max() <= _tol(prec, n)


def test_overlapping_buffers_are_staged(gpu, monkeypatch):
    """Plans whose launches write output before reading all input (a batch
    run in chunks; real rows in one tile pass) copy an overlapping input
    aside first: in == out with other distances gives the right transform,
    and the in-place aligned layout needs no copy."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rng = np.random.default_rng(21)
    # chunked complex64 batch, in_dist n, out_dist n + 2, same buffer
    n, b = 1 << 16, 4
    monkeypatch.setenv("GFFT_CHUNK_BYTES", str(n * 8))
    x = (rng.standard_normal((b, n)) + 1j * rng.standard_normal((b, n))).astype(np.complex64)
    buf = torch.zeros(b * (n + 2), dtype=torch.complex64, device=gpu)
    buf[: b * n] = torch.from_numpy(x.reshape(-1)).to(gpu)
    p = G.Plan(C.GFFT_C2C, n, batch=b, in_dist=n, out_dist=n + 2)
    p.execute(buf, buf)
    y = buf.cpu().numpy().reshape(b, n + 2)[:, :n]
    assert O.rel_l2_rows(y, O.fft_ld(x.astype(np.complex128))).max() <= O.tol_c64(n)
    monkeypatch.delenv("GFFT_CHUNK_BYTES")
    for prec in ("f32", "f64"):
        rdt, cdt, trdt, tcdt = _dts(prec)
        pr = 0 if prec == "f32" else 1
        # dense in-place r2c of rows that fit one tile: n reals -> n/2 + 1 bins per row
        n, b = 4096, 37
        nin = n // 2 + 1
        xr = rng.standard_normal((b, n)).astype(rdt)
        buf = torch.zeros(b * nin, dtype=tcdt, device=gpu)
        torch.view_as_real(buf).reshape(-1)[: b * n].copy_(torch.from_numpy(xr.reshape(-1)).to(gpu))
        p = G.Plan(C.GFFT_R2C, n, precision=pr, batch=b, half=True)
        assert "real_rows" in p.describe()
        p.execute(torch.view_as_real(buf).reshape(-1)[: b * n], buf)
        assert O.rel_l2_rows(buf.view(b, nin).cpu().numpy(),
                             O.fft_ld(xr.astype(np.complex128))[:, :nin]).max() <= _tol(prec, n)
        # and back: dense in-place c2r (n/2 + 1 bins -> n reals per row)
        F = buf.view(b, nin).cpu().numpy()
        q = G.Plan(C.GFFT_C2R, n, precision=pr, batch=b)
        assert "real_rows" in q.describe()
        q.execute(buf, torch.view_as_real(buf).reshape(-1)[: b * n], inverse=True)
        back = torch.view_as_real(buf).reshape(-1)[: b * n].view(b, n).cpu().numpy()
        ref = O.fft_ld(_hermitian_full(F.astype(np.complex128), n), inv=True).real
        assert O.rel_l2_rows(back, ref).max() <= _tol(prec, n)
        assert O.rel_l2_rows(back / n, xr).max() <= 2 * _tol(prec, n)
    # 2D real forward in place with the dense pitches (M reals in, M complex out)
    w, h = 256, 64
    img = rng.standard_normal((h, w)).astype(np.float32)
    buf = torch.zeros(h * w, dtype=torch.complex64, device=gpu)
    torch.view_as_real(buf).reshape(-1)[: h * w].copy_(torch.from_numpy(img.reshape(-1)).to(gpu))
    p = G.Plan(C.GFFT_R2C_2D, w, n2=h)
    p.execute(torch.view_as_real(buf).reshape(-1)[: h * w], buf)
    ref = O.fft_ld(O.fft_ld(img.astype(np.complex128)).T.copy()).T
    assert O.rel_l2(buf.view(h, w).cpu().numpy(), ref) <= O.tol_c64(w * h)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [4, 64, 4096])
def test_c2r_ignores_imag_of_dc_and_nyquist(gpu, prec, n):
    """A c2r reads bins 0 and n/2 as real (the Hermitian convention): random
    imaginary parts there leave the output unchanged, and the output equals
    the inverse DFT of the Hermitian extension built from the real parts."""
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    rng = np.random.default_rng(n)
    F = (rng.standard_normal((2, n // 2 + 1)) + 1j * rng.standard_normal((2, n // 2 + 1))).astype(cdt)
    Fr = F.copy()
    Fr[:, 0] = F[:, 0].real
    Fr[:, n // 2] = F[:, n // 2].real
    a = G.irfft(torch.from_numpy(F).to(gpu), n).cpu().numpy()
    b = G.irfft(torch.from_numpy(Fr).to(gpu), n).cpu().numpy()
    assert np.array_equal(a, b)
    full = np.concatenate([Fr, np.conj(Fr[:, 1:n // 2][:, ::-1])], axis=1).astype(np.complex128)
    ref = O.fft_ld(full, inv=True).real
    assert O.rel_l2(a, ref) <= _tol(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [12, 24, 40, 100, 1000])
@pytest.mark.parametrize("half", [False, True])
def test_dit_non_pow2_matches_genfft(gpu, golden, prec, n, half):
    """DIT<T>(n) for n divisible by 4 but not a power of two against genFFT's
    dynamic FFTDIT<-1> outputs, with the reference's sentinel past the bins."""
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    z = golden[f"dit_{prec}_n{n}_in"]
    g = golden[f"dit_{prec}_n{n}_{'half' if half else 'full'}"]
    nout = n // 2 + 1 if half else n
    out = torch.full((n,), 43 + 21j, dtype=tcdt, device=gpu)
    G.DIT(n, tcdt).apply(out, torch.from_numpy(z).to(gpu), half)
    y = out.cpu().numpy()
    eps = 1e-12 if prec == "f64" else 1e-5
    assert np.max(np.abs(y[:nout] - g[:nout])) <= eps * max(1.0, np.max(np.abs(g[:nout])))
    assert np.all(y[nout:] == 43 + 21j)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [4100, 3 << 18])
def test_dit_non_pow2_is_the_real_fft(gpu, prec, n):
    """unpack(FFT_{n/2}(x_even + i x_odd)) = FFT_n(x) for n a multiple of 4,
    the n/2-point input made by numpy's FP64 FFT, checked against numpy's FP64
    FFT of the real signal and against the oracle's unpack."""
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (2, n))
    z = np.fft.fft(x[:, 0::2] + 1j * x[:, 1::2], axis=-1).astype(cdt)
    out = torch.empty((2, n), dtype=tcdt, device=gpu)
    G.DIT(n, tcdt).apply(out, torch.from_numpy(z).to(gpu), False)
    y = out.cpu().numpy()
    assert O.rel_l2(y, np.fft.fft(x, axis=-1)) <= 2 * _tol(prec, n)
    for r in range(2):
        assert O.rel_l2(y[r], O.unpack(z[r], n, False)) <= _tol(prec, n)
