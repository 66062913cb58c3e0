"""The autotuner (gfft_autotune): candidate radix orders / tile widths /
threads for single-pass lengths and four-step splits / pipelined kernels for
long lengths are timed on the device; the tuned plan must keep parity with the
oracle and the table must survive a JSON save / load round trip."""
import json
import os

import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _parity(gpu, n, batch=3, prec=0):
    import genfft_amd as G
    rng = np.random.default_rng(n)
    x = (rng.uniform(-1, 1, (batch, n)) + 1j * rng.uniform(-1, 1, (batch, n))).astype(
        np.complex128 if prec else np.complex64)
    tol = O.tol_c128(n) if prec else O.tol_c64(n)
    for inv in (False, True):
        y = G.fft(torch.from_numpy(x).to(gpu), inverse=inv).cpu().numpy()
        assert O.rel_l2_rows(y, O.fft_ld(x.astype(np.complex128), inv)).max() <= tol


@pytest.mark.parametrize("n,batch", [(1024, 4096), (1 << 16, 64)])
def test_autotune_picks_fastest_and_keeps_parity(gpu, tmp_path, n, batch):
    import genfft_amd as G
    from genfft_amd import _capi as C
    rep = G.autotune(n, batch, torch.complex64, reps=5)
    cands = rep["candidates"]
    assert len(cands) >= 3
    best = rep["best"]
    assert best["ms"] == min(c["ms"] for c in cands)
    _parity(gpu, n)
    path = str(tmp_path / "tune.json")
    G.save_tuning(path)
    saved = json.load(open(path))
    assert any(e["n"] == n for e in saved["entries"])
    C.check(C.lib().gfft_tuning_clear())
    G.load_tuning(path)
    p = G.Plan(C.GFFT_C2C, n, batch=batch)
    text = p.describe()
    if n <= 16384:
        assert "tile L=%d" % n in text
    else:
        assert f"L={1 << best['n1_log2']} " in text
    _parity(gpu, n)
    C.check(C.lib().gfft_tuning_clear())


@pytest.mark.parametrize("prec,n,batch", [(0, 256, 4096), (0, 512, 2048), (0, 1024, 1024), (0, 2048, 512),
                                          (0, 4096, 256), (0, 8192, 128), (0, 16384, 64), (0, 1 << 20, 4),
                                          (1, 256, 2048), (1, 512, 1024), (1, 1024, 512), (1, 2048, 256),
                                          (1, 4096, 128), (1, 8192, 64), (1, 1 << 16, 8)])
def test_every_tuner_candidate_keeps_parity(gpu, tmp_path, prec, n, batch):
    """Each candidate the tuner can pick (registered radix orders, tile
    widths, threads per tile, four-step splits, pipelined forms), forced
    through a tuning file one at a time, matches the oracle."""
    import genfft_amd as G
    from genfft_amd import _capi as C
    rep = G.autotune(n, batch, torch.complex128 if prec else torch.complex64, reps=1)
    path = str(tmp_path / "one.json")
    try:
        for c in rep["candidates"]:
            with open(path, "w") as f:
                f.write('{"version": 1, "entries": [\n')
                f.write(f' {{"precision": {prec}, "n": {n}, "n1_log2": {c["n1_log2"]}, "col_m": {c["col_m"]}, '
                        f'"row_m": {c["row_m"]}, "pipe_mask": {c["pipe_mask"]}, "row_variant": {c["row_variant"]}}}\n')
                f.write(']}\n')
            C.check(C.lib().gfft_tuning_clear())
            G.load_tuning(path)
            with G._cache_lock:
                G._cache.clear()
            _parity(gpu, n, prec=prec)
    finally:
        C.check(C.lib().gfft_tuning_clear())
        with G._cache_lock:
            G._cache.clear()
