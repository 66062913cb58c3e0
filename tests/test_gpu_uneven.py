"""Uneven-size batches (gfft_batch_plan): thousands of transforms of mixed
power-of-two lengths, mixed directions, arbitrary offsets, in one plan.

No reference counterpart (genFFT transforms one sequence per call,
include/genFFT/fft.h:80-85); every job is checked against the long-double
oracle with the per-transform tolerance, and the three execution modes
(persistent launch + graph replay, eager, one plan launch per job) must agree.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu


def make_jobs(count, kmin, kmax, seed, gap=0):
    rng = np.random.default_rng(seed)
    ks = rng.integers(kmin, kmax + 1, count)
    sizes = (1 << ks).astype(np.int64)
    inv = rng.integers(0, 2, count).astype(bool)
    offs = np.zeros(count, np.int64)
    pos = 0
    for j in range(count):
        offs[j] = pos
        pos += int(sizes[j]) + gap
    return sizes, offs, inv, pos


def check_jobs(x, y, sizes, offs_in, offs_out, inv, prec, rows):
    tol = O.tol_c64 if prec == "f32" else O.tol_c128
    for j in rows:
        n = int(sizes[j])
        src = x[offs_in[j]:offs_in[j] + n].astype(np.complex128)
        ref = O.fft_ld(src, bool(inv[j]))
        err = O.rel_l2(y[offs_out[j]:offs_out[j] + n], ref)
        assert err <= tol(n), (j, n, bool(inv[j]), err)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_uneven_all_lengths_all_jobs(gpu, prec):
    """lengths 1 .. 2^16 (every class, incl. the four-step ones), odd offsets."""
    import genfft_amd as G
    dt = torch.complex64 if prec == "f32" else torch.complex128
    sizes, offs, inv, total = make_jobs(700, 0, 16, seed=1, gap=1)
    rng = np.random.default_rng(2)
    x = (rng.uniform(-1, 1, total) + 1j * rng.uniform(-1, 1, total)).astype(
        np.complex64 if prec == "f32" else np.complex128)
    out_offs = offs + 3
    b = G.UnevenBatch(sizes, offs, out_offs, inv, dt)
    tx = torch.from_numpy(x).to(gpu)
    ty = torch.full((total + 3,), 7 + 5j, dtype=dt, device=gpu)
    b.execute(tx, ty, mode="eager")
    y = ty.cpu().numpy()
    check_jobs(x, y, sizes, offs, out_offs, inv, prec, range(len(sizes)))
    # gaps between outputs untouched
    covered = np.zeros(total + 3, bool)
    for j in range(len(sizes)):
        covered[out_offs[j]:out_offs[j] + sizes[j]] = True
    assert np.all(y[~covered] == 7 + 5j)


def test_uneven_graph_replay_matches_eager_and_per_job(gpu):
    """The uneven-batch configuration's shape (lengths 2^6..2^16, 50/50 directions) at
    reduced count; graph replayed several times with eager work in between."""
    import genfft_amd as G
    sizes, offs, inv, total = make_jobs(2000, 6, 16, seed=3)
    g = torch.Generator(device=gpu).manual_seed(4)
    x = (torch.rand(total, device=gpu, generator=g) * 2 - 1) + 1j * (torch.rand(total, device=gpu, generator=g) * 2 - 1)
    x = x.to(torch.complex64)
    b = G.UnevenBatch(sizes, offs, offs, inv, torch.complex64)
    ref = torch.empty_like(x)
    b.execute(x, ref, mode="eager")
    per = torch.empty_like(x)
    b.execute(x, per, mode="per_job")
    torch.cuda.synchronize()
    assert torch.equal(ref, per) or (ref - per).abs().max().item() < 1e-3
    for rep in range(4):
        y = torch.zeros_like(x)
        b.execute(x, y, mode="graph")
        torch.cuda.synchronize()
        scratch = torch.randn(1 << 20, device=gpu)      # unrelated eager work between replays
        (scratch * 2).sum().item()
        assert torch.equal(y, ref), f"graph replay {rep} differs from the eager result"
    xs = x.cpu().numpy()
    rows = np.random.default_rng(5).choice(len(sizes), 200, replace=False)
    check_jobs(xs, ref.cpu().numpy(), sizes, offs, offs, inv, "f32", rows)


def test_uneven_in_place_and_large_classes(gpu):
    import genfft_amd as G
    sizes = np.array([1 << 13, 1 << 14, 1 << 15, 1 << 17, 1 << 20, 64, 1 << 13], np.int64)
    inv = np.array([0, 1, 0, 1, 0, 1, 1], bool)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    total = int(sizes.sum())
    rng = np.random.default_rng(6)
    x = (rng.standard_normal(total) + 1j * rng.standard_normal(total)).astype(np.complex64)
    b = G.UnevenBatch(sizes, offs, offs, inv, torch.complex64)
    t = torch.from_numpy(x).to(gpu)
    b.execute(t, t, mode="graph")   # in place
    check_jobs(x, t.cpu().numpy(), sizes, offs, offs, inv, "f32", range(len(sizes)))


def test_uneven_full_configuration_sample(gpu):
    """The benchmark configuration itself: 20,000 complex64 transforms of
    2^k, k uniform in 6..16 (seed 0), entries U[-1, 1), 50/50 directions, run
    as the captured graph (replayed twice); 500 random jobs against the
    long-double oracle, every class represented."""
    import genfft_amd as G
    rng = np.random.default_rng(0)
    count = 20000
    sizes = (1 << rng.integers(6, 17, count)).astype(np.int64)
    inv = rng.integers(0, 2, count).astype(bool)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    total = int(sizes.sum())
    g = torch.Generator(device=gpu).manual_seed(0)
    x = torch.complex(torch.rand(total, device=gpu, generator=g) * 2 - 1,
                      torch.rand(total, device=gpu, generator=g) * 2 - 1)
    y = torch.empty_like(x)
    b = G.UnevenBatch(sizes, offs, offs, inv, torch.complex64)
    b.execute(x, y, mode="graph")
    b.execute(x, y, mode="graph")
    torch.cuda.synchronize()
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    pick = np.random.default_rng(1).choice(count, 500, replace=False)
    assert len(set(int(sizes[j]) for j in pick)) == 11
    check_jobs(xs, ys, sizes, offs, offs, inv, "f32", pick)
