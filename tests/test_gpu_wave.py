"""The one-wavefront 1024-point complex64 kernel (genfft_amd/csrc/gfft_wave.h).

A whole row sits in the 64 lanes x 16 registers of one wave and its three
radix stages exchange data between lanes with DPP row shifts (bank-masked),
v_permlane16_swap / v_permlane32_swap and DPP quad permutations: no LDS and
no barrier.  The autotuner may pick it as a single-pass row kernel; forced
through a tuning file here, it must match the oracle (the long-double FFT)
within the complex64 bound 2e-7 * log2 N, forward and unscaled inverse, on:
  * ragged batches, where the last workgroup's waves have no row;
  * padded row pitches (in_dist / out_dist > n);
  * a 4099-row random batch against numpy's float64 FFT as a second reference.
"""
import json

import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu

N = 1024


def _force(path, variant):
    import genfft_amd as G
    from genfft_amd import _capi as C
    # the tuning file format: one entry per line
    entry = {"precision": 0, "n": N, "n1_log2": 0, "col_m": 0, "row_m": 0, "pipe_mask": 3, "row_variant": variant}
    with open(path, "w") as f:
        f.write('{"version": 1, "entries": [\n ' + json.dumps(entry) + '\n]}\n')
    C.check(C.lib().gfft_tuning_clear())
    G.load_tuning(path)
    G.clear_plan_cache()


@pytest.fixture(scope="module")
def wave_variant(gpu, tmp_path_factory):
    """Registry index of the wave kernel: the tuner candidate whose plan has no LDS."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    path = str(tmp_path_factory.mktemp("wave") / "one.json")
    rep = G.autotune(N, 64, torch.complex64, reps=1)
    found = None
    for c in rep["candidates"]:
        _force(path, c["row_variant"])
        if " lds=0 " in G.Plan(C.GFFT_C2C, N, batch=64).describe():
            found = c["row_variant"]
    assert found is not None, "no wave-kernel candidate among the tuner's single-pass kernels"
    _force(path, found)
    yield found
    C.check(C.lib().gfft_tuning_clear())
    G.clear_plan_cache()


@pytest.mark.parametrize("batch", [1, 3, 4, 5, 257])
def test_wave_kernel_vs_oracle(gpu, wave_variant, batch):
    import genfft_amd as G
    from genfft_amd import _capi as C
    rng = np.random.default_rng(batch)
    x = (rng.uniform(-1, 1, (batch, N)) + 1j * rng.uniform(-1, 1, (batch, N))).astype(np.complex64)
    p = G.Plan(C.GFFT_C2C, N, batch=batch)
    assert " lds=0 " in p.describe()
    xt = torch.from_numpy(x).to(gpu)
    for inv in (False, True):
        y = torch.empty_like(xt)
        p.execute(xt, y, inverse=inv)
        ref = O.fft_ld(x.astype(np.complex128), inv)
        assert O.rel_l2_rows(y.cpu().numpy(), ref).max() <= O.tol_c64(N)


def test_wave_kernel_padded_pitch(gpu, wave_variant):
    import genfft_amd as G
    from genfft_amd import _capi as C
    batch, di, do = 7, N + 40, N + 8
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((batch, di)) + 1j * rng.standard_normal((batch, di))).astype(np.complex64)
    p = G.Plan(C.GFFT_C2C, N, batch=batch, in_dist=di, out_dist=do)
    assert " lds=0 " in p.describe()
    xt = torch.from_numpy(x).to(gpu)
    y = torch.full((batch, do), complex(7, 7), dtype=torch.complex64, device=gpu)
    p.execute(xt.view(-1), y.view(-1))
    got = y.cpu().numpy()
    assert O.rel_l2_rows(got[:, :N], O.fft_ld(x[:, :N].astype(np.complex128))).max() <= O.tol_c64(N)
    assert np.all(got[:, N:] == np.complex64(7 + 7j))  # the padding is never written


def test_wave_kernel_vs_numpy(gpu, wave_variant):
    import genfft_amd as G
    rng = np.random.default_rng(4099)
    x = (rng.standard_normal((4099, N)) + 1j * rng.standard_normal((4099, N))).astype(np.complex64)
    y = G.fft(torch.from_numpy(x).to(gpu)).cpu().numpy()
    ref = np.fft.fft(x.astype(np.complex128), axis=-1)
    err = np.linalg.norm(y - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
    assert err.max() <= O.tol_c64(N)
