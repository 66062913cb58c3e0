"""GPU parity of the fused spectral filter and its adjoint.

    y  = irfft(H * rfft(x))                       (forward, one fused kernel)
    gx = irfft(conj(H) * rfft(gy))                (adjoint in x)
    gH = sum_rows w_k/n * rfft(gy)_k * conj(rfft(x)_k),  w = 1 at k = 0, n/2, else 2

genFFT has no c2r and no autograd (include/genFFT/FFTReal.h:127-130); the
operator is built on its forward real transform (FFTReal.h:204-213) and the
library's c2r.  Checks:
  * forward and both gradients against FP64 references computed with the
    oracle's long-double FFT (``fft_ld``), Hermitian extension written out;
  * the dot-product identity <A x, g> = <x, A^T g>, which pins the adjoint
    independently of the formula above;
  * torch.autograd.gradcheck (FP64 central differences, eps = 1e-6) on a
    [4, 256] slice through the complex128 kernels, for x and complex H;
  * the fused path against the unfused one (library rfft / irfft launches
    plus elementwise ops) on the same inputs, and the fused path faster.
"""
import time

import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _h_full(H: np.ndarray, n: int) -> np.ndarray:
    """Hermitian extension of n/2+1 bins; bins 0 and n/2 keep their real parts only."""
    L = n // 2
    Hf = np.zeros(n, np.complex128)
    Hf[: L + 1] = H
    Hf[0] = H[0].real
    Hf[L] = H[L].real
    Hf[L + 1:] = np.conj(H[1:L][::-1])
    return Hf


def _filter_ref(x: np.ndarray, H: np.ndarray, conj_h=False) -> np.ndarray:
    n = x.shape[-1]
    Hf = _h_full(np.conj(H) if conj_h else H, n)
    X = O.fft_ld(x.astype(np.complex128))
    return O.fft_ld(X * Hf, inv=True).real / n


def _grad_h_ref(x: np.ndarray, gy: np.ndarray) -> np.ndarray:
    n = x.shape[-1]
    L = n // 2
    X = O.fft_ld(x.astype(np.complex128))[:, : L + 1]
    G = O.fft_ld(gy.astype(np.complex128))[:, : L + 1]
    w = np.full(L + 1, 2.0 / n)
    w[0] = w[L] = 1.0 / n
    gH = (w * G * np.conj(X)).sum(0)
    gH[0] = gH[0].real
    gH[L] = gH[L].real
    return gH


def _dts(prec):
    return (np.float64, np.complex128, torch.float64, torch.complex128) if prec == "f64" else \
        (np.float32, np.complex64, torch.float32, torch.complex64)


def _tol(prec, n):
    return O.tol_c128(n) if prec == "f64" else O.tol_c64(n)


def _inputs(prec, n, b, seed):
    rdt, cdt, _, _ = _dts(prec)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((b, n)).astype(rdt)
    H = (1.0 + 0.3 * rng.standard_normal(n // 2 + 1) + 0.3j * rng.standard_normal(n // 2 + 1)).astype(cdt)
    gy = rng.standard_normal((b, n)).astype(rdt)
    return x, H, gy


CASES = [("f32", n) for n in (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)] + \
        [("f64", n) for n in (64, 128, 256, 512, 1024, 2048, 4096, 8192)]


@pytest.mark.parametrize("prec,n", CASES)
def test_filter_forward_backward_vs_fp64(gpu, prec, n):
    import genfft_amd as G
    b = 8
    x, H, gy = _inputs(prec, n, b, 1000 + n)
    tx = torch.from_numpy(x).to(gpu).requires_grad_(True)
    tH = torch.from_numpy(H).to(gpu).requires_grad_(True)
    y = G.spectral_filter(tx, tH)
    y.backward(torch.from_numpy(gy).to(gpu))
    tol = _tol(prec, n)
    assert O.rel_l2(y.detach().cpu().numpy(), _filter_ref(x, H)) <= 2 * tol
    assert O.rel_l2(tx.grad.cpu().numpy(), _filter_ref(gy, H, conj_h=True)) <= 2 * tol
    assert O.rel_l2(tH.grad.cpu().numpy(), _grad_h_ref(x, gy)) <= 4 * tol
    # gradients of the ignored imaginary parts of bins 0 and n/2 are exactly zero
    g = tH.grad.cpu().numpy()
    assert g[0].imag == 0 and g[n // 2].imag == 0


@pytest.mark.parametrize("prec,n,b", [("f32", 64, 1000), ("f64", 64, 1000), ("f32", 16384, 256),
                                      ("f32", 1024, 129), ("f64", 256, 65)])
def test_filter_grad_h_many_rows(gpu, prec, n, b):
    """gH summed over many rows: two reduction passes with a ragged last segment."""
    import genfft_amd as G
    x, H, gy = _inputs(prec, n, b, 31 * b + n)
    tx = torch.from_numpy(x).to(gpu)
    tH = torch.from_numpy(H).to(gpu).requires_grad_(True)
    G.spectral_filter(tx, tH).backward(torch.from_numpy(gy).to(gpu))
    assert O.rel_l2(tH.grad.cpu().numpy(), _grad_h_ref(x, gy)) <= 4 * _tol(prec, n)
    # deterministic: the same call gives the same bits
    g1 = tH.grad.clone()
    tH.grad = None
    G.spectral_filter(tx, tH).backward(torch.from_numpy(gy).to(gpu))
    assert torch.equal(g1, tH.grad)


@pytest.mark.parametrize("prec,n", [("f32", 1024), ("f64", 256), ("f64", 4096)])
def test_filter_adjoint_identity(gpu, prec, n):
    """<A x, g> = <x, A^T g> for the operator and the x-adjoint the backward kernel applies."""
    import genfft_amd as G
    x, H, gy = _inputs(prec, n, 4, 7 * n)
    tx = torch.from_numpy(x).to(gpu).requires_grad_(True)
    tH = torch.from_numpy(H).to(gpu)
    y = G.spectral_filter(tx, tH)
    y.backward(torch.from_numpy(gy).to(gpu))
    lhs = float(np.sum(y.detach().cpu().numpy().astype(np.float64) * gy))
    rhs = float(np.sum(x.astype(np.float64) * tx.grad.cpu().numpy()))
    scale = np.linalg.norm(x) * np.linalg.norm(gy)
    assert abs(lhs - rhs) <= 4 * _tol(prec, n) * scale


def test_filter_gradcheck_fp64(gpu):
    """FP64 central differences (eps 1e-6) on a [4, 256] slice, x and complex H."""
    import genfft_amd as G
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.standard_normal((4, 256))).to(gpu).requires_grad_(True)
    H = torch.from_numpy(1 + 0.01 * rng.standard_normal(129) + 0.01j * rng.standard_normal(129)).to(gpu)
    H.requires_grad_(True)
    assert torch.autograd.gradcheck(G.spectral_filter, (x, H), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_filter_loss_central_differences(gpu):
    """d/dh of the loss mean|y - t|^2 along random directions: central
    differences (h = 1e-6) against the fused gradients, FP64."""
    import genfft_amd as G
    rng = np.random.default_rng(11)
    n, b = 256, 4
    x = torch.from_numpy(rng.standard_normal((b, n))).to(gpu).requires_grad_(True)
    H = torch.from_numpy(1 + 0.01 * rng.standard_normal(n // 2 + 1) + 0j).to(gpu).requires_grad_(True)
    t = torch.roll(x.detach(), 3, -1) + 0.1 * torch.from_numpy(rng.standard_normal((b, n))).to(gpu)

    def loss(xv, Hv):
        return ((G.spectral_filter(xv, Hv) - t) ** 2).mean()

    loss(x, H).backward()
    h = 1e-6
    for trial in range(4):
        dx = torch.from_numpy(rng.standard_normal((b, n))).to(gpu)
        dH = torch.from_numpy(rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)).to(gpu)
        with torch.no_grad():
            num = (loss(x + h * dx, H + h * dH) - loss(x - h * dx, H - h * dH)).item() / (2 * h)
        ana = float((x.grad * dx).sum() + (H.grad.real * dH.real + H.grad.imag * dH.imag).sum())
        assert abs(num - ana) <= 1e-7 * max(1.0, abs(ana)), (trial, num, ana)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_fused_matches_unfused(gpu, prec):
    import genfft_amd as G
    _, _, trdt, tcdt = _dts(prec)
    n, b = 2048, 16
    x, H, gy = _inputs(prec, n, b, 3)
    outs = []
    for fn in (G.spectral_filter, G.spectral_filter_plain):
        tx = torch.from_numpy(x).to(gpu).requires_grad_(True)
        tH = torch.from_numpy(H).to(gpu).requires_grad_(True)
        y = fn(tx, tH)
        y.backward(torch.from_numpy(gy).to(gpu))
        outs.append([y.detach().cpu().numpy(), tx.grad.cpu().numpy(), tH.grad.cpu().numpy()])
    tol = _tol(prec, n)
    for a, r in zip(outs[0], outs[1]):
        assert O.rel_l2(a, r) <= 4 * tol


def test_filter_backward_on_several_streams(gpu):
    """The backward's per-row gH partials live in the plan, shared by every
    call: calls on three streams, enqueued back to back with no host
    synchronisation, each give the gradients of a call run alone."""
    import genfft_amd as G
    g = torch.Generator(device=gpu).manual_seed(5)
    n, b = 16384, 256
    xs = [torch.randn(b, n, device=gpu, generator=g) for _ in range(3)]
    gys = [torch.randn(b, n, device=gpu, generator=g) for _ in range(3)]
    H = (1 + 0.01 * torch.randn(n // 2 + 1, device=gpu, generator=g)).to(torch.complex64)

    def grads(x, gy):
        tx, tH = x.clone().requires_grad_(True), H.clone().requires_grad_(True)
        G.spectral_filter(tx, tH).backward(gy)
        return tx.grad, tH.grad

    ref = []
    for x, gy in zip(xs, gys):
        ref.append(grads(x, gy))
        torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=gpu) for _ in range(3)]
    for rep in range(2):
        got = []
        for x, gy, s in zip(xs, gys, streams):
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                got.append(grads(x, gy))
        torch.cuda.synchronize()
        for (gx, gH), (rx, rH) in zip(got, ref):
            assert torch.equal(gx, rx) and torch.equal(gH, rH), rep


def test_filter_in_place_and_batch_shapes(gpu):
    """Leading axes are flattened into rows; a plan is reused across calls."""
    import genfft_amd as G
    rng = np.random.default_rng(9)
    x = rng.standard_normal((2, 3, 512)).astype(np.float32)
    H = (1 + 0.1 * rng.standard_normal(257)).astype(np.complex64)
    y = G.spectral_filter(torch.from_numpy(x).to(gpu), torch.from_numpy(H).to(gpu)).cpu().numpy()
    ref = _filter_ref(x.reshape(6, 512), H).reshape(2, 3, 512)
    assert O.rel_l2(y, ref) <= 2 * O.tol_c64(512)


def test_filter_strided_inputs_and_gradients(gpu):
    """Non-contiguous x, H and upstream gradient give the contiguous results."""
    import genfft_amd as G
    g = torch.Generator(device=gpu).manual_seed(4)
    xt = torch.randn(512, 8, device=gpu, generator=g)
    Hs = torch.randn(2 * 257, dtype=torch.complex64, device=gpu, generator=g)[::2]
    gyt = torch.randn(512, 8, device=gpu, generator=g)
    outs = []
    for x, H, gy in ((xt.t(), Hs, gyt.t()), (xt.t().contiguous(), Hs.contiguous(), gyt.t().contiguous())):
        x = x.detach().requires_grad_(True)
        H = H.detach().requires_grad_(True)
        y = G.spectral_filter(x, H)
        y.backward(gy)
        outs.append((y.detach(), x.grad, H.grad))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_filter_rejects_bad_arguments(gpu):
    import genfft_amd as G
    from genfft_amd import GfftError
    x = torch.zeros(4, 100, device=gpu)
    with pytest.raises(GfftError):
        G.spectral_filter(x, torch.zeros(51, dtype=torch.complex64, device=gpu))
    with pytest.raises(ValueError):
        G.spectral_filter(torch.zeros(4, 64, device=gpu), torch.zeros(30, dtype=torch.complex64, device=gpu))
    with pytest.raises(TypeError):
        G.spectral_filter(torch.zeros(4, 64, device=gpu), torch.zeros(33, dtype=torch.complex128, device=gpu))


def test_fused_faster_than_unfused(gpu):
    """x [256, 16384] f32, H = 1 + 0.01 N(0,1) over 8193 bins, loss mean|y - t|^2
    with t = shift(x, 3) + 0.1 N(0,1): one forward + backward, fused vs unfused."""
    import genfft_amd as G
    g = torch.Generator(device=gpu).manual_seed(0)
    x = torch.randn(256, 16384, device=gpu, generator=g).requires_grad_(True)
    H = (1 + 0.01 * torch.randn(8193, device=gpu, generator=g)).to(torch.complex64).requires_grad_(True)
    t = torch.roll(x.detach(), 3, -1) + 0.1 * torch.randn(256, 16384, device=gpu, generator=g)

    def step(fn):
        x.grad = None
        H.grad = None
        ((fn(x, H) - t) ** 2).mean().backward()

    # A GPU that sat idle runs its first ~20 ms of work at lower clocks: warm
    # each form for 50 ms of calls, then keep the best of three timed rounds,
    # the two forms taking turns.
    times = {}
    for fn in (G.spectral_filter, G.spectral_filter_plain):
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 0.05:
            step(fn)
            torch.cuda.synchronize()
    for _ in range(3):
        for fn in (G.spectral_filter, G.spectral_filter_plain):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(10):
                step(fn)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 10
            times[fn.__name__] = min(ms, times.get(fn.__name__, ms))
    print("fused vs unfused ms:", times)
    assert times["spectral_filter"] < times["spectral_filter_plain"]
