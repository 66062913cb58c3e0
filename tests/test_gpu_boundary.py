"""GPU parity at the edges of the drop-in boundary.

  RealFFT2D<T>::forward      include/genFFT/FFTReal.h:83-104, any scalar
                             in_stride (FFTReal.h:83; rows read at :144-165)
  RealFFT2D<T>::forward_2x   include/genFFT/FFTReal.h:106-118 (rows at :167-180)
  real-side pitches and distances that are odd (R2C input, C2R output,
  C2R_2D output pitch): the transform runs on a dense copy in scratch
  overlapping input / output buffers: staged unless exactly in place

Expected values: genFFT's own outputs (tests/golden, within genFFT's
FFT_Eps-style absolute bound used by the other golden tests) and the
long-double oracle (rel-L2 2e-7*log2(N) complex64, 1e-15*log2(N) complex128).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _tol(prec, n):
    return O.tol_c64(n) if prec == "f32" else O.tol_c128(n)


def _dts(prec):
    if prec == "f32":
        return np.float32, np.complex64, torch.float32, torch.complex64
    return np.float64, np.complex128, torch.float64, torch.complex128


def _eps(prec, w, h):
    return 1e-8 + w * h * 1e-12 if prec == "f64" else 1e-5 + w * h * 1e-8


def ref2d(x2d, inv=False):
    """Separable long-double 2D DFT: rows, then columns."""
    r = O.fft_ld(x2d.astype(np.complex128), inv)
    return O.fft_ld(np.ascontiguousarray(r.T), inv).T


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h,ins,outs", [(32, 16, 33, 35), (8, 8, 9, 8), (64, 4, 65, 64), (16, 1, 17, 16)])
def test_real2d_odd_input_pitch_matches_genfft(gpu, golden, prec, w, h, ins, outs):
    """RealFFT2D::forward with an odd scalar input pitch, against genFFT's
    own output on the same input (the pitch gaps of the output untouched)."""
    import genfft_amd as G
    _, _, trdt, tcdt = _dts(prec)
    key = f"real2d_{prec}_w{w}_h{h}_i{ins}_o{outs}"
    out = torch.full((h * outs,), 7 + 5j, dtype=tcdt, device=gpu)
    G.RealFFT2D(w, h, trdt).forward(out, outs, torch.from_numpy(golden[key + "_in"]).to(gpu), ins)
    y = out.cpu().numpy()
    assert np.max(np.abs(y - golden[key + "_fwd"])) <= _eps(prec, w, h)
    # and against the separable oracle on the w columns of every row
    x = golden[key + "_in"].reshape(h, ins)[:, :w]
    assert O.rel_l2(y.reshape(h, outs)[:, :w], ref2d(x)) <= _tol(prec, w * h)
    assert np.all(y.reshape(h, outs)[:, w:] == 7 + 5j)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h,ins,outs", [(32, 16, 33, 35), (16, 8, 16, 16), (8, 32, 9, 8), (64, 2, 64, 66)])
def test_real2d_forward_2x_matches_genfft(gpu, golden, prec, w, h, ins, outs):
    """RealFFT2D::forward_2x (the 2D FFT of in1 + i*in2) against genFFT on the
    same inputs, both at one pitch: genFFT offsets in2's second half by
    in_stride1 (FFTReal.h:178), which agrees with in_stride2 only then."""
    import genfft_amd as G
    _, _, trdt, tcdt = _dts(prec)
    key = f"real2d2x_{prec}_w{w}_h{h}_i{ins}_o{outs}"
    out = torch.full((h * outs,), 7 + 5j, dtype=tcdt, device=gpu)
    G.RealFFT2D(w, h, trdt).forward_2x(out, outs, torch.from_numpy(golden[key + "_in1"]).to(gpu), ins,
                                       torch.from_numpy(golden[key + "_in2"]).to(gpu), ins)
    y = out.cpu().numpy()
    assert np.max(np.abs(y - golden[key + "_fwd"])) <= _eps(prec, w, h)
    x1 = golden[key + "_in1"].reshape(h, ins)[:, :w].astype(np.float64)
    x2 = golden[key + "_in2"].reshape(h, ins)[:, :w].astype(np.float64)
    assert O.rel_l2(y.reshape(h, outs)[:, :w], ref2d(x1 + 1j * x2)) <= _tol(prec, w * h)
    assert np.all(y.reshape(h, outs)[:, w:] == 7 + 5j)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_real2d_forward_2x_unequal_pitches(gpu, prec):
    """Each input keeps its own pitch (odd and even), checked against the
    separable oracle of in1 + i*in2, and separating the spectrum per element
    (X = (Z_k + conj Z_-k)/2, Y = -i (Z_k - conj Z_-k)/2) recovers both real
    images' 2D spectra."""
    import genfft_amd as G
    rdt, _, trdt, tcdt = _dts(prec)
    w, h, p1, p2, po = 64, 32, 67, 64, 70
    rng = np.random.default_rng(77)
    a = rng.standard_normal((h, p1)).astype(rdt)
    b = rng.standard_normal((h, p2)).astype(rdt)
    out = torch.full((h * po,), 7 + 5j, dtype=tcdt, device=gpu)
    G.RealFFT2D(w, h, trdt).forward_2x(out, po, torch.from_numpy(a.reshape(-1)).to(gpu), p1,
                                       torch.from_numpy(b.reshape(-1)).to(gpu), p2)
    Z = out.cpu().numpy().reshape(h, po)
    assert np.all(Z[:, w:] == 7 + 5j)
    Z = Z[:, :w].astype(np.complex128)
    A, B = a[:, :w].astype(np.float64), b[:, :w].astype(np.float64)
    assert O.rel_l2(Z, ref2d(A + 1j * B)) <= _tol(prec, w * h)
    Zm = np.conj(Z[(-np.arange(h)) % h][:, (-np.arange(w)) % w])
    assert O.rel_l2((Z + Zm) / 2, ref2d(A)) <= 2 * _tol(prec, w * h)
    assert O.rel_l2(-0.5j * (Z - Zm), ref2d(B)) <= 2 * _tol(prec, w * h)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n,batch,half", [(1024, 5, True), (1024, 3, False), (64, 7, True), (1 << 15, 3, True),
                                          (1 << 20, 2, True), (1 << 18, 3, False)])
def test_r2c_odd_input_distance(gpu, prec, n, batch, half):
    """R2C rows n + 1 reals apart (odd distance): every transform against the
    long-double oracle, and the planner shows the dense copy."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    rng = np.random.default_rng(n + batch)
    din = n + 1
    nout = n // 2 + 1 if half else n
    x = rng.standard_normal((batch, din)).astype(rdt)
    p = G.Plan(C.GFFT_R2C, n, precision=0 if prec == "f32" else 1, batch=batch, half=half, in_dist=din)
    assert p.describe().splitlines()[1].startswith("repack")
    y = torch.full((batch, nout), 7 + 5j, dtype=tcdt, device=gpu)
    p.execute(torch.from_numpy(x.reshape(-1)).to(gpu), y)
    ref = O.fft_ld(x[:, :n].astype(np.complex128))[:, :nout]
    assert O.rel_l2_rows(y.cpu().numpy(), ref).max() <= _tol(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n,batch", [(1024, 5), (64, 3), (1 << 15, 3), (1 << 20, 2)])
def test_c2r_odd_output_distance(gpu, prec, n, batch):
    """C2R writing rows n + 3 reals apart: the gaps stay untouched and each
    row is the unscaled inverse of its half spectrum."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rdt, cdt, trdt, tcdt = _dts(prec)
    rng = np.random.default_rng(3 * n + batch)
    x = rng.standard_normal((batch, n))
    F = O.fft_ld(x.astype(np.complex128))[:, : n // 2 + 1].astype(cdt)
    dout = n + 3
    out = torch.full((batch * dout,), 7.0, dtype=trdt, device=gpu)
    p = G.Plan(C.GFFT_C2R, n, precision=0 if prec == "f32" else 1, batch=batch, out_dist=dout)
    assert p.describe().splitlines()[-1].startswith("repack")
    p.execute(torch.from_numpy(F.reshape(-1)).to(gpu), out, inverse=True)
    y = out.cpu().numpy().reshape(batch, dout)
    assert O.rel_l2_rows(y[:, :n] / n, x).max() <= 2 * _tol(prec, n)
    assert np.all(y[:, n:] == 7.0)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("w,h,outs", [(32, 16, 33), (256, 64, 257), (4096, 8, 4097)])
def test_c2r2d_odd_output_pitch(gpu, prec, w, h, outs):
    """RealFFT2D::inverse into an odd scalar pitch: inverse(forward(x)) = w*h*x."""
    import genfft_amd as G
    rdt, cdt, trdt, tcdt = _dts(prec)
    rng = np.random.default_rng(w + h)
    x = rng.standard_normal((h, w)).astype(rdt)
    spec = torch.empty(h * w, dtype=tcdt, device=gpu)
    f = G.RealFFT2D(w, h, trdt)
    f.forward(spec, w, torch.from_numpy(x.reshape(-1)).to(gpu), w)
    out = torch.full((h * outs,), 7.0, dtype=trdt, device=gpu)
    f.inverse(out, outs, spec, w)
    y = out.cpu().numpy().reshape(h, outs)
    assert O.rel_l2(y[:, :w] / (w * h), x) <= 2 * _tol(prec, w * h)
    assert np.all(y[:, w:] == 7.0)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_single_pass_in_place_with_other_distance(gpu, prec):
    """One tile launch reads and writes the same buffer at unequal distances
    (in_dist n, out_dist n + 2): workgroups store while others have yet to
    load, so the input is staged first; the result matches the oracle."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    _, cdt, _, tcdt = _dts(prec)
    n, b = 1024, 4096 if prec == "f32" else 1024
    rng = np.random.default_rng(31)
    x = (rng.standard_normal((b, n)) + 1j * rng.standard_normal((b, n))).astype(cdt)
    buf = torch.zeros(b * (n + 2), dtype=tcdt, device=gpu)
    buf[: b * n] = torch.from_numpy(x.reshape(-1)).to(gpu)
    p = G.Plan(C.GFFT_C2C, n, precision=0 if prec == "f32" else 1, batch=b, in_dist=n, out_dist=n + 2)
    assert p.describe().count("\ntile") == 1  # a single pass
    p.execute(buf, buf)
    y = buf.cpu().numpy().reshape(b, n + 2)[:, :n]
    rows = np.arange(0, b, 97)
    assert O.rel_l2_rows(y[rows], O.fft_ld(x[rows].astype(np.complex128))).max() <= _tol(prec, n)


def test_2d_in_place_with_other_image_distance(gpu):
    """Batched 2D, equal row pitch, unequal image distance, one buffer: staged,
    every image against the separable oracle."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    w, h, b = 256, 128, 6
    rng = np.random.default_rng(41)
    x = (rng.standard_normal((b, h, w)) + 1j * rng.standard_normal((b, h, w))).astype(np.complex64)
    dout = h * w + 2 * w
    buf = torch.zeros(b * dout, dtype=torch.complex64, device=gpu)
    buf[: b * h * w] = torch.from_numpy(x.reshape(-1)).to(gpu)
    p = G.Plan(C.GFFT_C2C_2D, w, n2=h, batch=b, in_dist=h * w, out_dist=dout)
    p.execute(buf, buf)
    y = buf.cpu().numpy()
    for i in range(b):
        assert O.rel_l2(y[i * dout: i * dout + h * w].reshape(h, w), ref2d(x[i])) <= O.tol_c64(w * h)


def test_exactly_in_place(gpu):
    """in == out with equal distances runs in place (no staging copy) and
    gives the out-of-place result bit for bit."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    n, b = 1024, 256
    x = torch.randn(b, n, dtype=torch.complex64, device=gpu)
    want = G.fft(x)
    p = G.Plan(C.GFFT_C2C, n, batch=b)
    p.execute(x, x)
    assert torch.equal(x, want)


def test_python_wrappers_reject_wrong_types_and_short_buffers(gpu):
    """The tensor layer checks dtypes and byte extents against the plan
    before any launch."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    n = 1024
    p = G.Plan(C.GFFT_C2C, n, batch=4)
    ok = torch.zeros(4 * n, dtype=torch.complex64, device=gpu)
    with pytest.raises(TypeError):
        p.execute(torch.zeros(8 * n, dtype=torch.float32, device=gpu), ok)
    with pytest.raises(TypeError):
        p.execute(ok, torch.zeros(4 * n, dtype=torch.complex128, device=gpu))
    with pytest.raises(ValueError):
        p.execute(torch.zeros(3 * n, dtype=torch.complex64, device=gpu), ok)
    with pytest.raises(ValueError):
        p.execute(ok, torch.zeros(4 * n - 1, dtype=torch.complex64, device=gpu))
    r = G.Plan(C.GFFT_R2C, n, batch=2, half=True)
    with pytest.raises(TypeError):  # a real side takes its reals, or its complex dtype (in place)
        r.execute(torch.zeros(2 * n, dtype=torch.float64, device=gpu),
                  torch.zeros(2 * (n // 2 + 1), dtype=torch.complex64, device=gpu))
    with pytest.raises(TypeError):  # the complex side must be complex
        r.execute(torch.zeros(2 * n, dtype=torch.float32, device=gpu),
                  torch.zeros(4 * (n // 2 + 1), dtype=torch.float32, device=gpu))
    with pytest.raises(ValueError):
        G.irfft(torch.zeros(2, n // 2, dtype=torch.complex64, device=gpu), n)
    with pytest.raises(TypeError):
        G.fft(torch.zeros(4, n, dtype=torch.float32, device=gpu))
    # a view into a larger buffer is accepted: the bytes past it belong to the storage
    big = torch.zeros(8 * n, dtype=torch.complex64, device=gpu)
    p.execute(big[n:], big[n:])


def test_plans_reject_non_row_major_tensors(gpu):
    """A plan reads and writes from the tensor's first element with its own
    pitches, so the Python layer accepts row-major tensors (outer strides may
    pad) and rejects transposed or permuted views instead of reading them as
    garbage.  The half spectrum of a 2D real image, a padded view and its
    dense copy give the same image."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    w, h = 64, 32
    g = torch.Generator(device=gpu).manual_seed(9)
    img = torch.randn(h, w, device=gpu, generator=g)
    spec = torch.fft.rfft2(img)                                     # h x (w/2 + 1); ROCm's torch may return it
    if not spec.is_contiguous():                                    # column-major, which a plan rejects
        with pytest.raises(ValueError, match="row-major"):
            G.Plan(C.GFFT_C2R_2D, w, precision=0, n2=h, in_stride=w // 2 + 1, out_stride=w).execute(
                spec, torch.empty(h, w, device=gpu), inverse=True)
    spec = spec.contiguous()
    wide = torch.zeros(h, w, dtype=torch.complex64, device=gpu)
    wide[:, : w // 2 + 1] = spec
    view = wide[:, : w // 2 + 1]                                    # pitch w, row-major with padding
    dense_plan = G.Plan(C.GFFT_C2R_2D, w, precision=0, n2=h, in_stride=w // 2 + 1, out_stride=w)
    wide_plan = G.Plan(C.GFFT_C2R_2D, w, precision=0, n2=h, in_stride=w, out_stride=w)
    a, b = torch.empty(h, w, device=gpu), torch.empty(h, w, device=gpu)
    dense_plan.execute(spec, a, inverse=True)
    wide_plan.execute(view, b, inverse=True)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert O.rel_l2(a.cpu().numpy() / (w * h), img.cpu().numpy()) <= 2 * O.tol_c64(w * h)
    x = torch.randn(64, 32, dtype=torch.complex64, device=gpu, generator=g)
    p = G.Plan(C.GFFT_C2C, 32, precision=0, batch=64)
    y = torch.empty(64, 32, dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError, match="row-major"):
        p.execute(x.t().contiguous().t(), y)                         # column-major storage
    with pytest.raises(ValueError, match="row-major"):
        p.execute(x, torch.empty(32, 64, dtype=torch.complex64, device=gpu).t())
