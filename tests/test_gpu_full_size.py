"""GPU parity at the full configuration sizes: batched 1D complex64 and
complex128 forward and inverse (genfft::FFT<T>::transform,
include/genFFT/fft.h:80-85) for N = 2^10, 2^14, 2^18, 2^20, 2^22 at the
batch that moves about 1 GiB per run, i.e. the chunking, tile counts and
grids the benchmark runs.  test_gpu_c2c.py checks the same sizes against the
oracle at reduced batch.

Every transform of the batch is checked through the round trip
inverse(forward(x)) = N x.  The first and last transforms, in each direction,
are checked against the long-double oracle.  Tolerances (relative L2 per
transform), as in test_gpu_c2c.py: complex64 2e-7*log2(N), complex128
1e-15*log2(N); the round trip is allowed twice that.  Entries are i.i.d.
N(0,1), as the configuration states.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

CASES = [("f32", 10, 131072), ("f32", 14, 8192), ("f32", 18, 512), ("f32", 20, 128), ("f32", 22, 32),
         ("f64", 10, 65536), ("f64", 14, 4096), ("f64", 18, 256), ("f64", 20, 64), ("f64", 22, 16)]


@pytest.mark.parametrize("prec,lg,batch", CASES)
def test_full_size_forward_inverse(gpu, prec, lg, batch):
    import genfft_amd as G
    n = 1 << lg
    cdt, rdt = (torch.complex64, torch.float32) if prec == "f32" else (torch.complex128, torch.float64)
    tol = O.tol_c64(n) if prec == "f32" else O.tol_c128(n)
    print(f"full size {prec} n=2^{lg} batch={batch}", flush=True)
    g = torch.Generator(device=gpu).manual_seed(1000 * lg + batch)
    x = torch.complex(torch.randn(batch, n, generator=g, device=gpu, dtype=rdt),
                      torch.randn(batch, n, generator=g, device=gpu, dtype=rdt))
    f = G.FFT(n, cdt)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    f.transform(y, x, False)
    f.transform(z, y, True)
    torch.cuda.synchronize()
    # every transform of the batch: inverse(forward(x)) = N x
    err = torch.linalg.vector_norm(z / n - x, dim=1) / torch.linalg.vector_norm(x, dim=1)
    worst = int(torch.argmax(err))
    assert float(err[worst]) <= 2 * tol, f"round trip of transform {worst}: {float(err[worst]):.3g} > {2 * tol:.3g}"
    for r in (0, batch - 1):
        fwd = O.fft_ld(x[r].cpu().numpy().astype(np.complex128), False)
        assert O.rel_l2(y[r].cpu().numpy(), fwd) <= tol, f"forward transform {r}"
        inv = O.fft_ld(y[r].cpu().numpy().astype(np.complex128), True)
        assert O.rel_l2(z[r].cpu().numpy(), inv) <= tol, f"inverse transform {r}"


def test_plain_small_configuration(gpu):
    """genfft::FFT<float>(1024) forward, then the unscaled inverse, on 4096
    independent complex64 transforms, re/im i.i.d. uniform[-1, 1) from seed
    0: inverse(forward(x)) == 1024 x within 2e-7*log2(N) relative L2 per
    transform, and every forward transform against the FP64 naive DFT (the
    DFT matrix with exactly reduced angles 2*pi*((j*k) mod N)/N)."""
    import genfft_amd as G
    n, batch = 1024, 4096
    rng = np.random.default_rng(0)
    xs = (rng.uniform(-1, 1, (batch, n)) + 1j * rng.uniform(-1, 1, (batch, n))).astype(np.complex64)
    x = torch.from_numpy(xs).to(gpu)
    f = G.FFT(n, torch.complex64)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    f.transform(y, x, False)
    f.transform(z, y, True)
    torch.cuda.synchronize()
    tol = O.tol_c64(n)
    rt = torch.linalg.vector_norm(z / n - x, dim=1) / torch.linalg.vector_norm(x, dim=1)
    assert float(rt.max()) <= tol, f"round trip {float(rt.max()):.3g} > {tol:.3g}"
    j = np.arange(n)
    W = np.exp(-2j * np.pi * ((np.outer(j, j) % n) / n))
    ref = xs.astype(np.complex128) @ W.T
    errs = O.rel_l2_rows(y.cpu().numpy(), ref)
    assert float(errs.max()) <= tol, f"forward vs naive DFT {float(errs.max()):.3g} > {tol:.3g}"
