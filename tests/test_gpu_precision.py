"""16-bit storage modes (GFFT_F16 / GFFT_BF16): FP16 or BF16 pairs in HBM,
FP32 arithmetic in LDS, optional per-transform scale.  No reference
counterpart (genFFT computes in float/double only); the transforms keep
genFFT's conventions (forward exp(-2 pi i jk/N), unscaled inverse, natural
order; include/genFFT/fft.h:56-113).

Every result is compared with two references on the dequantized input (so
the input's own rounding is not counted):
  * the oracle's long-double FFT (``fft_ld``) -- the FP64-or-better oracle;
  * the GPU complex128 plan of the same transform.
Tolerance: each pass through HBM rounds once to 16 bits (relative error at
most u = 2^-11 for FP16, 2^-8 for BF16, per element) and the FP32 arithmetic
adds ~1e-7 log2 N, so rel-L2 <= passes * u with passes = the tile launches
the plan runs (1 single-pass, 2 four-step, 2-3 for 2D).  The measured
figures are printed (pytest -s) for the record.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

U = {"f16": 2.0 ** -11, "bf16": 2.0 ** -8}
TDT = {"f16": torch.float16, "bf16": torch.bfloat16}


def _pairs(z: np.ndarray, fmt: str) -> torch.Tensor:
    """complex -> [..., n, 2] 16-bit pairs (round to nearest even), on the CPU."""
    return torch.from_numpy(np.stack([z.real, z.imag], -1).astype(np.float32)).to(TDT[fmt])


def _deq(t: torch.Tensor) -> np.ndarray:
    f = t.float().cpu().numpy().astype(np.float64)
    return f[..., 0] + 1j * f[..., 1]


def _passes(plan_text: str) -> int:
    return sum(1 for ln in plan_text.splitlines() if ln.startswith("tile"))


def _max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def _report(tag, y, ref_oracle, ref_c128):
    r1, m1 = O.rel_l2(y, ref_oracle), _max_abs(y, ref_oracle)
    r2, m2 = O.rel_l2(y, ref_c128), _max_abs(y, ref_c128)
    print(f"{tag}: vs FP64 oracle rel-L2 {r1:.3e} max-abs {m1:.3e}; vs GPU c128 rel-L2 {r2:.3e} max-abs {m2:.3e}")
    return r1, r2


def _plan_passes(fmt, kind, n, n2=0, batch=1):
    import genfft_amd as G
    from genfft_amd import _capi as C
    p = G.Plan(kind, n, precision=C.GFFT_F16 if fmt == "f16" else C.GFFT_BF16, n2=n2, batch=batch)
    return _passes(p.describe())


SIZES = [2, 4, 16, 64, 256, 1024, 4096, 16384, 1 << 16, 1 << 20]


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("inverse", [False, True])
def test_c2c_16bit_vs_oracle_and_c128(gpu, fmt, n, inverse):
    import genfft_amd as G
    from genfft_amd import _capi as C
    b = max(2, min(64, (1 << 18) // n))
    rng = np.random.default_rng(n * 2 + inverse)
    z = rng.standard_normal((b, n)) + 1j * rng.standard_normal((b, n))
    x = _pairs(z, fmt)
    xd = _deq(x)
    scale = "auto" if n >= (1 << 16) else None
    y, s = G.fft16(x.to(gpu), inverse=inverse, scale=scale)
    yd = _deq(y)
    if s is not None:
        yd = yd / s.cpu().numpy()[:, None].astype(np.float64)
    rows = [0, b - 1]
    ref_o = O.fft_ld(xd[rows], inv=inverse)
    ref_g = G.fft(torch.from_numpy(xd).to(gpu), inverse=inverse).cpu().numpy()
    r1, r2 = _report(f"{fmt} c2c n={n} b={b} inv={int(inverse)}", yd[rows], ref_o, ref_g[rows])
    tol = _plan_passes(fmt, C.GFFT_C2C, n, batch=b) * U[fmt]
    assert r1 <= tol and O.rel_l2(yd, ref_g) <= tol


def test_fp16_overflows_without_scale_and_auto_scale_fixes_it(gpu):
    """|x| ~ 1e3 with a coherent component, N = 2^20: |X_0| ~ 1e3 * 2^20
    overflows FP16 unscaled; the automatic per-transform scale keeps every
    stored value finite and the descaled result accurate."""
    import genfft_amd as G
    n, b = 1 << 20, 3
    rng = np.random.default_rng(1)
    z = 1e3 * (0.5 + rng.uniform(-0.5, 0.5, (b, n)) + 1j * rng.uniform(-1, 1, (b, n)))
    z[1] *= 1e-3   # a small transform in the same batch gets its own, larger scale
    x = _pairs(z, "f16")
    assert torch.isfinite(x.float()).all()
    y0, _ = G.fft16(x.to(gpu))
    assert not torch.isfinite(y0.float()).all(), "expected FP16 overflow without a scale"
    y, s = G.fft16(x.to(gpu), scale="auto")
    assert torch.isfinite(y.float()).all()
    # the scale is the documented power of two: largest 2^-k <= 1 with n * max(|re|+|im|) * 2^-k <= 2^15
    xd = _deq(x)
    m = np.max(np.abs(xd.real) + np.abs(xd.imag), axis=1)
    q = (m * n / 2.0 ** 15).astype(np.float32).astype(np.float64)
    want = np.where(q > 1, 2.0 ** -np.ceil(np.log2(q)), 1.0)
    assert np.array_equal(s.cpu().numpy().astype(np.float64), want)
    assert s[1] > s[0]
    yd = _deq(y) / s.cpu().numpy()[:, None].astype(np.float64)
    ref = O.fft_ld(xd)
    _report("f16 N=2^20 |x|~1e3 auto scale", yd, ref, ref)
    for r in range(b):
        assert O.rel_l2(yd[r], ref[r]) <= 2 * U["f16"]


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_given_scales(gpu, fmt):
    import genfft_amd as G
    n, b = 4096, 4
    rng = np.random.default_rng(2)
    z = rng.standard_normal((b, n)) + 1j * rng.standard_normal((b, n))
    x = _pairs(z, fmt)
    scales = torch.tensor([1.0, 0.5, 0.125, 2.0 ** -10], dtype=torch.float32)
    y, s = G.fft16(x.to(gpu), scale=scales)
    yd = _deq(y) / scales.numpy()[:, None].astype(np.float64)
    ref = O.fft_ld(_deq(x))
    for r in range(b):
        assert O.rel_l2(yd[r], ref[r]) <= U[fmt]


def _fft2_ld(z, inverse=False):
    rows = O.fft_ld(z, inv=inverse)
    return np.swapaxes(O.fft_ld(np.ascontiguousarray(np.swapaxes(rows, -1, -2)), inv=inverse), -1, -2)


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("w,h,b", [(64, 32, 3), (512, 256, 2), (16, 4096, 1)])
def test_2d_16bit_vs_oracle(gpu, fmt, w, h, b):
    import genfft_amd as G
    from genfft_amd import _capi as C
    rng = np.random.default_rng(w + h)
    z = rng.standard_normal((b, h, w)) + 1j * rng.standard_normal((b, h, w))
    x = _pairs(z, fmt)
    y, s = G.fft2_16(x.to(gpu), scale="auto")
    yd = _deq(y) / s.cpu().numpy()[:, None, None].astype(np.float64)
    xd = _deq(x)
    ref_o = _fft2_ld(xd)
    ref_g = G.fft2(torch.from_numpy(xd).to(gpu)).cpu().numpy()
    r1, r2 = _report(f"{fmt} 2D {w}x{h} b={b}", yd, ref_o, ref_g)
    tol = _plan_passes(fmt, C.GFFT_C2C_2D, w, h, b) * U[fmt]
    assert r1 <= tol and r2 <= tol


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_2d_8192_squared_vs_c128(gpu, fmt):
    """The 8192^2 image: against the GPU complex128 2D plan (64 Mi points;
    the long-double oracle pins the smaller 2D shapes)."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    g = torch.Generator(device=gpu).manual_seed(3)
    n = 8192
    xr = torch.randn(1, n, n, 2, device=gpu, generator=g).to(TDT[fmt])
    y, s = G.fft2_16(xr, scale="auto")
    xd = torch.view_as_complex(xr.double().contiguous())
    ref = G.fft2(xd)
    yd = torch.view_as_complex(y.double().contiguous()) / s.double()[0]
    rel = (torch.linalg.vector_norm(yd - ref) / torch.linalg.vector_norm(ref)).item()
    mx = (yd - ref).abs().max().item()
    print(f"{fmt} 2D 8192x8192: vs GPU c128 rel-L2 {rel:.3e} max-abs {mx:.3e} scale {s[0].item():g}")
    assert rel <= _plan_passes(fmt, C.GFFT_C2C_2D, n, n, 1) * U[fmt]


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("rows,cols", [(64, 7), (256, 999), (1024, 33)])
def test_cols_16bit_odd_widths(gpu, fmt, rows, cols):
    """FFTVert shapes (odd column counts, test/test_dispatch.cpp:96-120) with
    16-bit storage and padded pitches: scalar access path."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rng = np.random.default_rng(rows + cols)
    ip, op = cols + 3, cols + 1
    z = rng.standard_normal((rows, ip)) + 1j * rng.standard_normal((rows, ip))
    x = _pairs(z, fmt).to(gpu)
    out = torch.full((rows, op, 2), 7.0, dtype=TDT[fmt], device=gpu)
    p = G.Plan(C.GFFT_COLS, rows, precision=C.GFFT_F16 if fmt == "f16" else C.GFFT_BF16, n2=cols,
               in_stride=ip, out_stride=op)
    p.execute(x, out)
    yd = _deq(out)
    ref = np.swapaxes(O.fft_ld(np.ascontiguousarray(np.swapaxes(_deq(x)[:, :cols], 0, 1))), 0, 1)
    assert O.rel_l2(yd[:, :cols], ref) <= _passes(p.describe()) * U[fmt]
    assert torch.all(out[:, cols:].float() == 7.0), "write into the pitch gap"


# ---------------------------------------------------------------------------
# The precision workload: N = 2^20 batch 64 and 2D 8192^2, inputs with
# log-uniform magnitudes over 1e-3 .. 1e3 and uniform phase, in three modes
# (FP16 + per-transform scale, BF16, full FP32).
# ---------------------------------------------------------------------------
def _log_uniform(shape, gpu, seed):
    g = torch.Generator(device=gpu).manual_seed(seed)
    mag = torch.pow(10.0, torch.rand(shape, device=gpu, generator=g, dtype=torch.float64) * 6 - 3)
    ph = torch.rand(shape, device=gpu, generator=g, dtype=torch.float64) * (2 * np.pi)
    return torch.polar(mag, ph)


def _mode_input(z, mode):
    """(what the transform reads, its exact complex128 value)"""
    if mode == "f32":
        x = z.to(torch.complex64)
        return x, x.to(torch.complex128)
    x = torch.view_as_real(z).to(TDT[mode]).contiguous()
    return x, torch.view_as_complex(x.double().contiguous())


def _mode_run(mode, x, two_d):
    import genfft_amd as G
    if mode == "f32":
        y = G.fft2(x) if two_d else G.fft(x)
        return y.to(torch.complex128), None
    y, s = (G.fft2_16 if two_d else G.fft16)(x, scale="auto")
    yd = torch.view_as_complex(y.double().contiguous())
    shape = (-1,) + (1,) * (yd.dim() - 1)
    return yd / s.double().view(shape), s


def _gpu_err(a, b):
    return ((torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item(), (a - b).abs().max().item())


MODE_TOL = {"f16": 2 * U["f16"], "bf16": 2 * U["bf16"], "f32": O.tol_c64(1 << 20)}


@pytest.mark.parametrize("mode", ["f16", "bf16", "f32"])
def test_precision_workload_n2e20_b64(gpu, mode):
    import genfft_amd as G
    n, b = 1 << 20, 64
    z = _log_uniform((b, n), gpu, 20)
    x, xd = _mode_input(z, mode)
    y, s = _mode_run(mode, x, False)
    ref_g = G.fft(xd)
    rg, mg = _gpu_err(y, ref_g)
    rows = [0, b - 1]
    ref_o = O.fft_ld(xd[rows].cpu().numpy())
    yo = y[rows].cpu().numpy()
    ro, mo = O.rel_l2(yo, ref_o), _max_abs(yo, ref_o)
    print(f"precision N=2^20 b=64 {mode}: vs FP64 oracle (rows 0, 63) rel-L2 {ro:.3e} max-abs {mo:.3e}; "
          f"vs GPU c128 rel-L2 {rg:.3e} max-abs {mg:.3e}"
          + (f"; scales {s.min().item():g}..{s.max().item():g}" if s is not None else ""))
    assert ro <= MODE_TOL[mode] and rg <= MODE_TOL[mode]
    if mode == "f16":
        # these inputs overflow FP16 without a scale (|X| reaches ~1e5 and more)
        assert ref_g.abs().max().item() > 65504


def _fft_ld_rows_mt(x: np.ndarray) -> np.ndarray:
    from concurrent.futures import ThreadPoolExecutor
    out = np.empty_like(x)
    chunks = np.array_split(np.arange(x.shape[0]), 16)

    def work(idx):
        out[idx[0]:idx[-1] + 1] = O.fft_ld(x[idx[0]:idx[-1] + 1])

    with ThreadPoolExecutor(16) as ex:
        list(ex.map(work, chunks))
    return out


@pytest.mark.parametrize("mode", ["f16", "bf16", "f32"])
def test_precision_workload_2d_8192(gpu, mode):
    """8192 x 8192: against the GPU complex128 2D plan everywhere and the
    long-double oracle on four whole output columns (rows transformed by the
    oracle, then the four columns)."""
    import genfft_amd as G
    n = 8192
    z = _log_uniform((1, n, n), gpu, 21)
    x, xd = _mode_input(z, mode)
    y, s = _mode_run(mode, x, True)
    ref_g = G.fft2(xd)
    rg, mg = _gpu_err(y, ref_g)
    cols = [0, 1, 4097, 8191]
    rows_spec = _fft_ld_rows_mt(xd[0].cpu().numpy())
    ref_o = np.ascontiguousarray(O.fft_ld(np.ascontiguousarray(rows_spec[:, cols].T)).T)
    yo = y[0][:, cols].cpu().numpy()
    ro, mo = O.rel_l2(yo, ref_o), _max_abs(yo, ref_o)
    print(f"precision 2D 8192^2 {mode}: vs FP64 oracle (4 columns) rel-L2 {ro:.3e} max-abs {mo:.3e}; "
          f"vs GPU c128 rel-L2 {rg:.3e} max-abs {mg:.3e}"
          + (f"; scale {s[0].item():g}" if s is not None else ""))
    tol = {"f16": 3 * U["f16"], "bf16": 3 * U["bf16"], "f32": 2 * O.tol_c64(n * n)}[mode]
    assert ro <= tol and rg <= tol
