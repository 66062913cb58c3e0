"""GPU parity of the batched 1D complex transform (genfft::FFT<T>::transform,
include/genFFT/fft.h:80-85) against the long-double oracle and genFFT's own
outputs.  Tolerances (relative L2 per transform): complex64 2e-7*log2(N),
complex128 1e-15*log2(N).  Sizes follow test/test_fft.cpp:59-60 and
test/test_dispatch.cpp:58-59 (2 .. 2^22), plus N = 1.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZES = [1 << k for k in range(0, 23)]


def _batch(n, prec):
    # ~2-4 MiB per case, at least 2 transforms; long-double reference stays fast
    return max(2, min(256, (1 << 17) // n))


def _rand(shape, dt, rng):
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(dt)


def _tol(prec, n):
    return O.tol_c64(n) if prec == "f32" else O.tol_c128(n)


def _run(x_np, inv, dev):
    import genfft_amd as G
    x = torch.from_numpy(x_np).to(dev)
    y = torch.empty_like(x)
    G.FFT(x_np.shape[-1], x.dtype).transform(y, x, inv)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", SIZES)
def test_c2c_forward_inverse_vs_oracle(gpu, prec, n):
    rng = np.random.default_rng(n)
    dt = np.complex64 if prec == "f32" else np.complex128
    b = _batch(n, prec)
    x = _rand((b, n), dt, rng)
    rows = [0, b - 1] if n >= (1 << 16) else list(range(b))
    for inv in (False, True):
        print(f"c2c {prec} n={n} batch={b} inv={inv}", flush=True)
        y = _run(x, inv, gpu)
        ref = O.fft_ld(x[rows].astype(np.complex128), inv)
        err = O.rel_l2_rows(y[rows], ref)
        assert err.max() <= _tol(prec, n), (n, inv, err.max())


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [2, 16, 1024, 1 << 14, 1 << 15, 1 << 20, 1 << 22])
def test_inverse_identity(gpu, prec, n):
    """test/test_fft.cpp:41-57: inverse(forward(x)) / n == x."""
    import genfft_amd as G
    rng = np.random.default_rng(100 + n)
    dt = np.complex64 if prec == "f32" else np.complex128
    x = torch.from_numpy(_rand((2, n), dt, rng)).to(gpu)
    f = G.FFT(n, x.dtype)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    f.forward(y, x)
    f.inverse(z, y)
    err = O.rel_l2(z.cpu().numpy() / n, x.cpu().numpy())
    assert err <= 2 * _tol(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 512, 2048])
def test_matches_genfft_golden(gpu, golden, prec, n):
    """Same inputs as genFFT, results within genFFT's FFT_Eps (test/test_util.h:59-72)."""
    x = golden[f"c2c_{prec}_n{n}_in"]
    eps = 1e-8 + n * 1e-12 if prec == "f64" else 1e-5 + n * 1e-8
    for inv in (False, True):
        y = _run(x[None, :], inv, gpu)[0]
        g = golden[f"c2c_{prec}_n{n}_{'inv' if inv else 'fwd'}_native"]
        assert np.max(np.abs(y - g)) <= eps


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [64, 4096, 1 << 16])
def test_in_place(gpu, prec, n):
    import genfft_amd as G
    rng = np.random.default_rng(7)
    dt = np.complex64 if prec == "f32" else np.complex128
    x_np = _rand((3, n), dt, rng)
    x = torch.from_numpy(x_np).to(gpu)
    G.FFT(n, x.dtype).transform(x, x, False)
    err = O.rel_l2_rows(x.cpu().numpy(), O.fft_ld(x_np.astype(np.complex128)))
    assert err.max() <= _tol(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [8, 256, 1 << 15])
def test_strided_batch_leaves_gaps(gpu, prec, n):
    """in_dist / out_dist larger than n: only the n elements of each row are written."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rng = np.random.default_rng(9)
    dt = np.complex64 if prec == "f32" else np.complex128
    b, di, do = 5, n + 3, n + 5
    x_np = _rand((b, di), dt, rng)
    x = torch.from_numpy(x_np.reshape(-1)).to(gpu)
    y = torch.full((b * do,), 7 + 5j, dtype=x.dtype, device=gpu)
    p = G.Plan(C.GFFT_C2C, n, precision=0 if prec == "f32" else 1, batch=b, in_dist=di, out_dist=do)
    p.execute(x, y)
    yy = y.cpu().numpy().reshape(b, do)
    ref = O.fft_ld(x_np[:, :n].astype(np.complex128))
    assert O.rel_l2_rows(yy[:, :n], ref).max() <= _tol(prec, n)
    assert np.all(yy[:, n:] == 7 + 5j)


def test_closed_forms_on_gpu(gpu):
    """delta -> ones, constant -> N*delta, tone -> N*delta_m at N = 2^20."""
    import genfft_amd as G
    n = 1 << 20
    x = torch.zeros(3, n, dtype=torch.complex128, device=gpu)
    x[0, 0] = 1
    x[1, :] = 1
    m = 12345
    # exact integer reduction of the tone's phase: (m*j mod n) keeps the
    # argument below 2*pi, so the input itself is accurate to ~1e-16
    j = torch.arange(n, device=gpu, dtype=torch.int64)
    ph = ((m * j) % n).to(torch.float64) * (2 * np.pi / n)
    x[2] = torch.polar(torch.ones_like(ph), ph)
    y = G.fft(x).cpu().numpy()
    assert np.max(np.abs(y[0] - 1)) < 1e-15
    want = np.zeros(n)
    want[0] = n
    assert np.max(np.abs(y[1] - want)) < 1e-9
    want = np.zeros(n)
    want[m] = n
    assert O.rel_l2(y[2], want) < O.tol_c128(n)


def test_linearity_full_size(gpu):
    """Size-independent property at the benchmark shape (complex64, N = 2^20, batch 128)."""
    import genfft_amd as G
    n, b = 1 << 20, 128
    g = torch.Generator(device=gpu).manual_seed(0)
    x = torch.randn(b, n, dtype=torch.complex64, device=gpu, generator=g)
    y = torch.randn(b, n, dtype=torch.complex64, device=gpu, generator=g)
    fx, fy, fs = G.fft(x), G.fft(y), G.fft(x + 2 * y)
    lhs = fs.to(torch.complex128)
    rhs = fx.to(torch.complex128) + 2 * fy.to(torch.complex128)
    rel = (torch.linalg.vector_norm(lhs - rhs, dim=1) / torch.linalg.vector_norm(rhs, dim=1)).max().item()
    assert rel < 2 * O.tol_c64(n)
    # Parseval per transform
    ex = torch.linalg.vector_norm(x.to(torch.complex128), dim=1) ** 2 * n
    eX = torch.linalg.vector_norm(fx.to(torch.complex128), dim=1) ** 2
    assert ((eX - ex).abs() / ex).max().item() < 1e-5
    # spot transforms against the oracle
    for r in (0, 77, 127):
        ref = O.fft_ld(x[r].cpu().numpy().astype(np.complex128))
        assert O.rel_l2(fx[r].cpu().numpy(), ref) <= O.tol_c64(n)


@pytest.mark.parametrize("kind", ["four_step", "c2r"])
def test_shared_plan_across_streams_and_threads(gpu, kind):
    """One plan with a workspace driven from several streams (and host
    threads) at once gives the same bits as running it serially: calls on a
    new stream wait for the plan's previous work (genFFT's plans are reentrant
    across threads for distinct buffers, FFTLevel.h:76-92)."""
    import threading

    import genfft_amd as G
    from genfft_amd import _capi as C
    g = torch.Generator(device=gpu).manual_seed(4)
    if kind == "four_step":
        n, b = 1 << 20, 8
        p = G.Plan(C.GFFT_C2C, n, batch=b)
        xs = [torch.randn(b, n, dtype=torch.complex64, device=gpu, generator=g) for _ in range(4)]
        ys = [torch.empty_like(x) for x in xs]
        inv = False
    else:
        n, b = 1 << 18, 8
        p = G.Plan(C.GFFT_C2R, n, batch=b)
        xs = [G.rfft(torch.randn(b, n, device=gpu, generator=g)) for _ in range(4)]
        ys = [torch.empty(b, n, device=gpu) for _ in range(4)]
        inv = True
    assert p.workspace_bytes() > 0
    ref = []
    for x in xs:
        y = torch.empty_like(ys[0])
        p.execute(x, y, inverse=inv)
        torch.cuda.synchronize()
        ref.append(y.clone())
    streams = [torch.cuda.Stream(device=gpu) for _ in range(4)]
    for rep in range(3):
        for x, y, s in zip(xs, ys, streams):     # back-to-back on four streams, no host sync
            p.execute(x, y, inverse=inv, stream=s)
        torch.cuda.synchronize()
        for y, r in zip(ys, ref):
            assert torch.equal(y, r)
        for y in ys:
            y.zero_()
        torch.cuda.synchronize()

    def worker(i):
        for _ in range(5):
            p.execute(xs[i], ys[i], inverse=inv, stream=streams[i])
        streams[i].synchronize()

    th = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    for y, r in zip(ys, ref):
        assert torch.equal(y, r)


def _bitrev_perm(n):
    lg = n.bit_length() - 1
    return np.array([int(format(i, f"0{lg}b")[::-1], 2) if lg else 0 for i in range(n)])


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [1, 2, 8, 1024, 1 << 14, 1 << 20])
def test_transform_no_scramble_on_device(gpu, prec, n):
    """FFT<T>::transform_no_scramble (fft.h:69-73) on device data: input in
    bit-reversed order, in place, natural-order output -- the same bits as the
    natural-order transform of the unscrambled data."""
    import genfft_amd as G
    cdt = torch.complex64 if prec == "f32" else torch.complex128
    b = 3
    g = torch.Generator(device=gpu).manual_seed(n)
    x = torch.randn(b, n, dtype=cdt, device=gpu, generator=g)
    perm = torch.from_numpy(_bitrev_perm(n)).to(gpu)
    scr = torch.empty_like(x)
    scr[:, perm] = x
    for inv in (False, True):
        want = G.fft(x, inverse=inv)
        got = scr.clone()
        G.FFT(n, cdt).transform_no_scramble(got, inv)
        assert torch.equal(got, want)


@pytest.mark.parametrize("n,cols,stride", [(8, 7, 9), (1024, 33, 40), (4096, 16, 16)])
def test_vert_no_scramble_on_device(gpu, n, cols, stride):
    """FFTVert<T>::transform_no_scramble (fft.h:132-136): rows in bit-reversed order."""
    import genfft_amd as G
    g = torch.Generator(device=gpu).manual_seed(cols)
    x = torch.randn(n, stride, dtype=torch.complex64, device=gpu, generator=g)
    want = torch.zeros_like(x)
    G.FFTVert(n, torch.complex64).transform(want, stride, x, stride, cols)
    perm = torch.from_numpy(_bitrev_perm(n)).to(gpu)
    scr = x.clone()
    scr[perm] = x
    G.FFTVert(n, torch.complex64).transform_no_scramble(scr, stride, cols)
    assert torch.equal(scr[:, :cols], want[:, :cols])


@pytest.mark.parametrize("prec,batch", [("f32", 128), ("f32", 13), ("f32", 9), ("f64", 11), ("f64", 64)])
def test_overlapped_four_step_matches_sequential_passes(gpu, prec, batch, monkeypatch):
    """The overlapped four-step (the batch in chunks through a two-slot ring,
    each launch pairing the column pass of chunk k with the row pass of chunk
    k - 1) computes every transform exactly as the two sequential passes do:
    bit-identical results for ragged last chunks, both directions and in
    place, and spot transforms within tolerance of the oracle."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    n = 1 << 20
    pr = 0 if prec == "f32" else 1
    cdt = torch.complex64 if prec == "f32" else torch.complex128
    g = torch.Generator(device=gpu).manual_seed(batch)
    x = torch.randn(batch, n, dtype=cdt, device=gpu, generator=g)
    monkeypatch.setenv("GFFT_DUAL", "0")
    plain = G.Plan(C.GFFT_C2C, n, precision=pr, batch=batch)
    monkeypatch.setenv("GFFT_DUAL", "1")  # an autotuner candidate, not the default
    dual = G.Plan(C.GFFT_C2C, n, precision=pr, batch=batch)
    monkeypatch.delenv("GFFT_DUAL")
    assert "dual" not in plain.describe() and "dual" in dual.describe()
    for inv in (False, True):
        print(f"overlapped four-step {prec} batch={batch} inv={inv}", flush=True)
        ya, yb = torch.empty_like(x), torch.empty_like(x)
        dual.execute(x, ya, inverse=inv)
        plain.execute(x, yb, inverse=inv)
        torch.cuda.synchronize()
        assert torch.equal(ya, yb)
        for r in (0, batch // 2, batch - 1):
            ref = O.fft_ld(x[r].cpu().numpy().astype(np.complex128), inv)
            assert O.rel_l2(ya[r].cpu().numpy(), ref) <= _tol(prec, n)
    # in place, exactly: the chunks' outputs land on their own inputs
    z = x.clone()
    dual.execute(z, z)
    want = torch.empty_like(x)
    plain.execute(x, want)
    torch.cuda.synchronize()
    assert torch.equal(z, want)


def test_profile_launch_times_one_launch(gpu):
    """gfft_plan_profile_launch: each launch of a four-step plan repeated
    between one event pair gives a positive time, leaves an out-of-place
    result unchanged (the passes read the input and the scratch only), and
    rejects a launch index outside the plan and non-positive reps."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    n, batch = 1 << 20, 4
    g = torch.Generator(device=gpu).manual_seed(5)
    x = torch.randn(batch, n, dtype=torch.complex64, device=gpu, generator=g)
    p = G.Plan(C.GFFT_C2C, n, precision=0, batch=batch)
    y = torch.empty_like(x)
    p.execute(x, y)
    torch.cuda.synchronize()
    want = y.clone()
    nl = len([ln for ln in p.describe().splitlines() if ln.strip() and not ln.startswith("plan ")])
    assert nl == 2
    for i in range(nl):
        ms = p.profile_launch(x, y, i, reps=5)
        assert 0.0 < ms < 1000.0
    torch.cuda.synchronize()
    assert torch.equal(y, want)
    for bad_launch, bad_reps in ((nl, 1), (-1, 1), (0, 0)):
        with pytest.raises(G.GfftError):
            p.profile_launch(x, y, bad_launch, reps=bad_reps)
