"""CPU-side checks of the C ABI (no GPU needed).

* the library loads and exports every function include/genfft_amd.h declares;
* descriptor validation returns the documented status codes;
* the planner (dry run, gfft_plan_explain) proves every launch in bounds for
  the sizes the reference tests use and the benchmark shapes, and picks the
  single-pass / four-step / sub-pass structure it documents.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from genfft_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "genfft_amd.h")).read()
    return sorted(set(re.findall(r"GFFT_API\s+(?:const\s+)?\w+\s*\*?\s*(gfft_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    declared = header_functions()
    assert len(declared) >= 12
    assert sorted(C.EXPORTS) == declared
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.gfft_version() == 1


def test_only_the_c_abi_is_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIB_PATH], capture_output=True, text=True).stdout
    funcs = [ln.split()[-1] for ln in out.splitlines() if " T " in ln]
    assert sorted(f for f in funcs if f.startswith("gfft_")) == header_functions()
    assert not [f for f in funcs if not f.startswith("gfft_") and not f.startswith("_")]


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64, C.GFFT_F16, C.GFFT_BF16])
def test_column_passes_take_contiguous_tile_runs(prec):
    """Persistent column passes of 8- and 16-byte elements give each workgroup
    one contiguous run of tiles; transposing row passes and 16-bit storage
    interleave tiles across workgroups."""
    store = prec in (C.GFFT_F16, C.GFFT_BF16)
    st, text = explain(desc(prec=prec, n=1 << 20, batch=64))
    assert st == C.GFFT_OK, C.lib().gfft_last_error()
    t = tiles(text)
    assert [field(x, "inp") for x in t] == ["1", "0"]
    for x in t:
        want = "-1" if field(x, "inp") == "1" and not store and field(x, "pipe") == "1" else "0"
        assert field(x, "runs") == want, x


def _launch_kinds(text):
    return [ln.split()[0] for ln in text.splitlines()[1:]]


@pytest.mark.parametrize("n", [2, 64, 1024, 1 << 15, 1 << 20])
def test_odd_real_side_layouts_run_on_a_dense_copy(n):
    """Real sides move as complex pairs, so an odd real pitch or distance is
    planned on a dense copy in scratch: R2C / R2C_2D input repacked first,
    C2R / C2R_2D output repacked last; the even layouts plan without it, and
    the bounds proof holds for both (genFFT's RealFFT2D::forward takes any
    scalar in_stride, include/genFFT/FFTReal.h:83)."""
    cases = [
        (desc(kind=C.GFFT_R2C, n=n, batch=3, ind=n + 1), "first"),
        (desc(kind=C.GFFT_R2C, n=n, batch=3, ind=n + 1, flags=C.GFFT_FLAG_HALF), "first"),
        (desc(kind=C.GFFT_C2R, n=n, batch=3, outd=n + 1), "last"),
    ]
    if n <= 4096:
        cases += [
            (desc(kind=C.GFFT_R2C_2D, n=n, n2=16, ins=n + 1), "first"),
            (desc(kind=C.GFFT_R2C_2D, n=n, n2=16, batch=2, ind=16 * n + 1), "first"),
            (desc(kind=C.GFFT_C2R_2D, n=n, n2=16, outs=n + 3), "last"),
        ]
    for d, where in cases:
        st, text = explain(d)
        assert st == C.GFFT_OK, C.lib().gfft_last_error()
        kinds = _launch_kinds(text)
        assert kinds.count("repack") == 1 and kinds[0 if where == "first" else -1] == "repack", text
        # the same shape with even layouts needs no copy
        d.in_stride -= d.in_stride % 2
        d.out_stride -= d.out_stride % 2
        d.in_dist -= d.in_dist % 2
        d.out_dist -= d.out_dist % 2
        st, text = explain(d)
        assert st == C.GFFT_OK and "repack" not in text, text


def test_repack_keeps_buffers_apart():
    """The repacked plans read the caller's input only in their first launch
    (input side) or write the caller's output only in their last (output
    side): every other launch works in scratch."""
    for d, side in ((desc(kind=C.GFFT_R2C, n=1 << 20, batch=2, ind=(1 << 20) + 1), "in"),
                    (desc(kind=C.GFFT_C2R, n=1 << 20, batch=2, outd=(1 << 20) + 1), "out")):
        st, text = explain(d)
        assert st == C.GFFT_OK
        lines = text.splitlines()[1:]
        inner = lines[1:] if side == "in" else lines[:-1]
        assert all(f"{'in' if side == 'in' else 'out'}=" + ("in+" if side == "in" else "out+") not in ln
                   for ln in inner), text


@pytest.mark.parametrize("prec,n,batch,col,row", [
    (C.GFFT_F32, 1 << 20, 128, "16,4,16,0", "16,16,4,0"),   # measured fastest orders (MI355X)
    (C.GFFT_F32, 1 << 22, 32, "16,8,16,0", "16,16,8,0"),
    (C.GFFT_F64, 1 << 22, 16, "16,8,16,0", "16,16,8,0"),
    (C.GFFT_F32, 1 << 14, 8192, None, "16,4,16,16"),
])
def test_tile_radix_orders(prec, n, batch, col, row):
    """The planner's kernels for the benchmark shapes carry the radix orders
    their measurements chose (smallest radix in the middle of a column pass)."""
    st, text = explain(desc(prec=prec, n=n, batch=batch))
    assert st == C.GFFT_OK, C.lib().gfft_last_error()
    t = tiles(text)
    if col is None:
        assert len(t) == 1 and field(t[0], "rad") == row
    else:
        assert [field(x, "rad") for x in t] == [col, row]


def test_single_1024_row_takes_the_wave_kernel():
    """One complex64 1024-point row runs on the one-wavefront kernel (no LDS,
    radix stages 16, 16, 4); two rows, complex128 and 16-bit storage keep the
    LDS tile kernels."""
    def first_tile(**kw):
        st, text = explain(desc(n=1024, **kw))
        assert st == C.GFFT_OK, C.lib().gfft_last_error()
        return tiles(text)[0]
    t = first_tile(batch=1)
    assert field(t, "lds") == "0" and field(t, "rad") == "16,16,4,0"
    assert field(first_tile(batch=1, ind=1100, outd=1030), "lds") == "0"
    for kw in (dict(batch=2), dict(batch=1, prec=C.GFFT_F64), dict(batch=1, prec=C.GFFT_F16)):
        assert field(first_tile(**kw), "lds") != "0"


@pytest.mark.parametrize("batch,m,nt", [(16, "2", "128"), (255, "2", "128"), (1024, "4", "256"), (8192, "4", "256")])
def test_small_row_batches_take_the_narrow_tile(batch, m, nt):
    """complex64 1024-point rows: while the default 4-row tiles would number
    fewer than the 256 CUs, the 2-row tile of 128 threads runs instead."""
    st, text = explain(desc(n=1024, batch=batch))
    assert st == C.GFFT_OK, C.lib().gfft_last_error()
    t = tiles(text)[0]
    assert (field(t, "M"), field(t, "NT")) == (m, nt)


@pytest.mark.parametrize("n,m", [(256, 16), (512, 8), (1024, 4), (2048, 2)])
def test_batches_past_the_infinity_cache_take_the_wide_row_tile(n, m):
    """complex64 single-pass batches of >= 512 MiB run on the double-width row
    tile (512 threads, 64 KB of LDS), smaller ones and complex128 on the
    default tile."""
    def first_tile(**kw):
        st, text = explain(desc(n=n, **kw))
        assert st == C.GFFT_OK, C.lib().gfft_last_error()
        return tiles(text)[0]
    big, small = (1 << 30) // (8 * n), (256 << 20) // (8 * n)
    t = first_tile(batch=big)
    assert (field(t, "M"), field(t, "NT")) == (str(2 * m), "512")
    t = first_tile(batch=small)
    assert (field(t, "M"), field(t, "NT")) == (str(m), "256")
    shape = lambda t: (field(t, "M"), field(t, "NT"), field(t, "rad"))
    assert shape(first_tile(batch=big // 2, prec=C.GFFT_F64)) == shape(first_tile(batch=2, prec=C.GFFT_F64))


@pytest.mark.parametrize("n,batch,col_m,row_m", [
    (1 << 18, 1, "16", "8"),      # 16 row tiles of 16: narrow to 8
    (1 << 20, 1, "8", "8"),       # 64 column and row tiles: both narrow
    (1 << 20, 4, "16", "16"),     # 256 tiles: the default width
    (1 << 20, 128, "16", "16"),
])
def test_small_four_step_batches_take_narrow_tiles(n, batch, col_m, row_m):
    """A four-step pass whose default tiles would number fewer than the 256
    CUs runs on the narrowest registered tile of its length."""
    st, text = explain(desc(n=n, batch=batch))
    assert st == C.GFFT_OK, C.lib().gfft_last_error()
    assert [field(t, "M") for t in tiles(text)] == [col_m, row_m]


def _random_descs(count, seed):
    """Random plan descriptors over every kind, precision, size class, batch
    and padded stride/distance the ABI accepts (and some it rejects)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        kind = int(rng.integers(0, 8))
        prec = int(rng.integers(0, 4)) if kind in (C.GFFT_C2C, C.GFFT_COLS, C.GFFT_C2C_2D) else int(rng.integers(0, 2))
        batch = int(rng.integers(1, 6))
        pad = lambda: int(rng.choice([0, 0, 1, 2, 3, 8, 33]))
        ins = outs = ind = outd = n2 = flags = 0
        if kind in (C.GFFT_C2C_2D, C.GFFT_R2C_2D, C.GFFT_C2R_2D):
            n, n2 = 1 << int(rng.integers(1, 14)), 1 << int(rng.integers(0, 14))
            if rng.integers(0, 2):
                ins, outs = n + pad(), n + pad()
                if kind == C.GFFT_R2C_2D:
                    ins += ins % 2
                if kind == C.GFFT_C2R_2D:
                    ins, outs = n // 2 + 1 + pad(), n + 2 * pad()
            batch = 1
        elif kind == C.GFFT_COLS:
            n, n2 = 1 << int(rng.integers(0, 17)), int(rng.integers(1, 400))
            ins, outs = n2 + pad(), n2 + pad()
        else:
            n = 1 << int(rng.integers(0, 24))
            if kind == C.GFFT_DIT and rng.integers(0, 3) == 0:
                n = 4 * int(rng.integers(1, 3000))  # genFFT's dynamic DIT: any n divisible by 4
            flags = int(rng.integers(0, 2)) if kind in (C.GFFT_R2C, C.GFFT_DIT) else 0
            if rng.integers(0, 2):
                ind, outd = n + pad(), n + pad()
                if kind == C.GFFT_C2R:
                    ind, outd = n // 2 + 1 + pad(), n + 2 * pad()
                elif kind == C.GFFT_R2C:
                    ind = n + 2 * pad()
            if n > (1 << 20):
                batch = 1
        out.append(desc(kind=kind, prec=prec, n=n, n2=n2, batch=batch, ins=ins, outs=outs, ind=ind, outd=outd,
                        flags=flags))
    return out


def test_random_descriptors_plan_in_bounds_or_fail_cleanly():
    """Every random descriptor either plans -- and the planner's bounds proof
    holds for every launch -- or is rejected with a documented status; no
    descriptor reaches an internal error."""
    planned = 0
    for d in _random_descs(600, 424242):
        st, text = explain(d)
        msg = C.lib().gfft_last_error() or b""
        assert st in (C.GFFT_OK, C.GFFT_ERR_INVALID_ARG, C.GFFT_ERR_UNSUPPORTED), (st, msg)
        assert b"internal" not in msg or st == C.GFFT_OK, (d.kind, d.precision, d.n, d.n2, d.batch, msg)
        if st == C.GFFT_OK:
            planned += 1
            assert text.startswith("plan ")
    assert planned > 400


def test_pattern_copy_probe_loads_and_rejects_bad_shapes():
    """The benchmark's tile-pattern copy probe is a library of its own (not
    part of the C ABI); shapes it does not cover fail before any device call."""
    import os
    path = os.path.join(os.path.dirname(C.LIB_PATH), "libgfft_pattern_copy.so")
    lib = ctypes.CDLL(path)
    fn = lib.gfft_pattern_copy
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    for n1, n2, m, mode, order in ((1024, 1024, 16, 3, 0), (1000, 1024, 16, 0, 0), (1024, 1024, 8, 0, 0),
                                   (1024, 1024, 32, 1, 0), (1024, 1024, 1, 0, 0), (1024, 1024, 16, 0, 2)):
        assert fn(None, None, 128, n1, n2, m, mode, order, None) == -1
    assert fn(16, 32, 128, 1024, 1024, 16, 0, 1, None) == -1  # pointers not 16-byte aligned


def desc(kind=C.GFFT_C2C, prec=C.GFFT_F32, n=1024, n2=0, batch=1, ins=0, outs=0, ind=0, outd=0, flags=0):
    return C.GfftDesc(kind, prec, n, n2, batch, ins, outs, ind, outd, flags, 0)


def explain(d):
    buf = ctypes.create_string_buffer(1 << 20)
    st = C.lib().gfft_plan_explain(ctypes.byref(d), buf, len(buf))
    return st, buf.value.decode()


def test_status_codes_and_messages():
    lib = C.lib()
    for code in range(6):
        assert lib.gfft_error_string(code)
    h = ctypes.c_void_p()
    assert lib.gfft_plan_create(ctypes.byref(h), None) == C.GFFT_ERR_INVALID_ARG
    bad = [
        (desc(n=1000), C.GFFT_ERR_UNSUPPORTED),               # not a power of two
        (desc(n=1 << 24), C.GFFT_ERR_UNSUPPORTED),            # beyond 2^23
        (desc(prec=7), C.GFFT_ERR_INVALID_ARG),
        (desc(kind=42), C.GFFT_ERR_INVALID_ARG),
        (desc(batch=0), C.GFFT_ERR_INVALID_ARG),
        (desc(ind=-1), C.GFFT_ERR_INVALID_ARG),
        (desc(kind=C.GFFT_C2C_2D, n=64, n2=48), C.GFFT_ERR_UNSUPPORTED),
        (desc(kind=C.GFFT_COLS, n=64, n2=0), C.GFFT_ERR_INVALID_ARG),
        (desc(kind=C.GFFT_C2C_2D, n=64, n2=64, ins=32), C.GFFT_ERR_INVALID_ARG),  # pitch < width
    ]
    for d, code in bad:
        assert lib.gfft_plan_create(ctypes.byref(h), ctypes.byref(d)) == code
        assert lib.gfft_last_error()
        assert explain(d)[0] == code


def test_plan_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    d = desc()
    assert C.lib().gfft_plan_create(ctypes.byref(h), ctypes.byref(d)) == C.GFFT_ERR_NO_DEVICE


def tiles(text):
    return [ln for ln in text.splitlines() if ln.startswith("tile")]


def field(line, name):
    return re.search(rf"\b{name}=([^ ]+)", line).group(1)


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
@pytest.mark.parametrize("k", range(0, 24))
def test_c2c_plans_in_bounds(prec, k):
    n = 1 << k
    batch = max(1, (1 << 27) // (n * (8 << prec)))  # the ~1 GiB benchmark batches
    for b in sorted({1, 3, batch}):
        st, text = explain(desc(prec=prec, n=n, batch=b))
        assert st == C.GFFT_OK, text or C.lib().gfft_last_error()
        t = tiles(text)
        maxrow = 16384 if prec == C.GFFT_F32 else 8192
        if n == 1:
            assert "copy" in text and not t
        elif n <= maxrow:
            assert len(t) == 1 and field(t[0], "L") == str(n)
        else:
            assert len(t) % 2 == 0 and len(t) >= 2       # column pass + transposing row pass per chunk
            L1, L2 = int(field(t[0], "L")), int(field(t[1], "L"))
            assert L1 * L2 == n and field(t[0], "tw") == "1"
            assert field(t[0], "inp") == "1" and field(t[1], "outp") == "1"


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
def test_real_and_dit_plans_in_bounds(prec):
    for k in range(0, 24):
        n = 1 << k
        for half in (0, 1):
            for kind in (C.GFFT_R2C, C.GFFT_DIT, C.GFFT_C2R):
                for b in (1, 5):
                    st, text = explain(desc(kind=kind, prec=prec, n=n, batch=b, flags=half))
                    assert st == C.GFFT_OK, (kind, n, half, C.lib().gfft_last_error())


VERT_TABLE = [(1, 1), (1, 40), (2, 1), (2, 7), (2, 999), (4, 1), (4, 7), (4, 999), (8, 1), (8, 3), (8, 7),
              (8, 33), (8, 47), (8, 1023), (16, 1), (16, 47), (16, 999), (256, 1), (256, 1023), (1024, 1),
              (1024, 767), (65536, 1), (65536, 31)]  # test/test_dispatch.cpp:96-120


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
@pytest.mark.parametrize("n,cols", VERT_TABLE)
def test_cols_plans_in_bounds(prec, n, cols):
    for ins, outs in ((0, 0), (cols + 3, cols + 1)):
        st, text = explain(desc(kind=C.GFFT_COLS, prec=prec, n=n, n2=cols, ins=ins, outs=outs))
        assert st == C.GFFT_OK, C.lib().gfft_last_error()


def test_vector_paths_need_31_bit_tile_offsets():
    """complex64 tiles move 16-byte vectors addressed as a uniform base plus a
    32-bit per-thread offset: a column tile whose last element lies 2^31
    bytes or more from its first falls back to element accesses."""
    for stride, vec in ((1 << 16, "1"), ((1 << 19) + 2, "0")):
        st, text = explain(desc(kind=C.GFFT_COLS, n=1024, n2=64, ins=stride, outs=64))
        assert st == C.GFFT_OK, C.lib().gfft_last_error()
        t = tiles(text)
        assert len(t) == 1 and field(t[0], "inp") == "1"
        assert (int(field(t[0], "L")) - 1) * stride * 8 >= (1 << 31) or vec == "1"
        assert field(t[0], "vin") == vec and field(t[0], "vout") == "1"


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (1, 8), (32, 16), (8, 64), (4096, 4096), (8192, 8192),
                                 (16, 4096), (4096, 2)])
def test_2d_plans_in_bounds(prec, w, h):
    for kind in (C.GFFT_C2C_2D, C.GFFT_R2C_2D):
        if kind == C.GFFT_R2C_2D and w < 2:
            continue
        for ins, outs, b in ((0, 0, 1), (w + 2, w + 6, 1), (0, 0, 3)):
            st, text = explain(desc(kind=kind, prec=prec, n=w, n2=h, ins=ins, outs=outs, batch=b))
            if prec == C.GFFT_F64 and w > 8192:
                assert st == C.GFFT_ERR_UNSUPPORTED
                continue
            assert st == C.GFFT_OK, (kind, w, h, C.lib().gfft_last_error())


def test_bench_shape_structure():
    """complex64 N=2^20 x 128: 1024-point column pass (with twiddle) + 1024-point transposing row pass."""
    st, text = explain(desc(n=1 << 20, batch=128))
    assert st == 0
    t = tiles(text)
    assert field(t[0], "L") == "1024" and field(t[1], "L") == "1024"
    assert "scratch_bytes=" in text


def test_filter_plan_argument_checks():
    """Argument and size checks of the fused filter come before any device use."""
    import torch
    lib = C.lib()
    h = ctypes.c_void_p()
    assert lib.gfft_filter_plan_create(None, C.GFFT_F32, 1024, 1) == C.GFFT_ERR_INVALID_ARG
    assert lib.gfft_filter_plan_create(ctypes.byref(h), 7, 1024, 1) == C.GFFT_ERR_INVALID_ARG
    assert lib.gfft_filter_plan_create(ctypes.byref(h), C.GFFT_F32, 1024, 0) == C.GFFT_ERR_INVALID_ARG
    for prec, n in [(C.GFFT_F32, 100), (C.GFFT_F32, 32), (C.GFFT_F32, 32768), (C.GFFT_F64, 16384)]:
        assert lib.gfft_filter_plan_create(ctypes.byref(h), prec, n, 4) == C.GFFT_ERR_UNSUPPORTED
        assert b"fused filter" in lib.gfft_last_error()
    assert lib.gfft_filter_forward(None, None, None, None, 0, None) == C.GFFT_ERR_INVALID_ARG
    assert lib.gfft_filter_backward(None, None, None, None, None, None, None) == C.GFFT_ERR_INVALID_ARG
    if not torch.cuda.is_available():
        assert lib.gfft_filter_plan_create(ctypes.byref(h), C.GFFT_F32, 1024, 4) == C.GFFT_ERR_NO_DEVICE


def test_spectral_filter_rejects_host_tensors():
    """The autograd wrapper validates before it plans, and never caches a failed key."""
    import torch

    import genfft_amd as G
    from genfft_amd import autograd as A
    x, H = torch.zeros(4, 64), torch.zeros(33, dtype=torch.complex64)
    for fn in (G.spectral_filter, G.spectral_filter_plain):
        for _ in range(2):
            with pytest.raises(ValueError, match="GPU tensors only"):
                fn(x, H)
    assert not any(k[2].type == "cpu" for k in A._plans)


@pytest.mark.parametrize("prec", [C.GFFT_F16, C.GFFT_BF16])
def test_16bit_storage_plans_in_bounds(prec):
    """FP16 / BF16 storage: C2C 2..2^23, 2D (incl. 8192^2) and the COLS table
    plan in bounds with four-byte elements, on the 16-bit kernel table."""
    for k in range(1, 24):
        n = 1 << k
        for b in (1, 3, max(1, (1 << 28) // (n * 4))):
            st, text = explain(desc(prec=prec, n=n, batch=b))
            assert st == C.GFFT_OK, (n, b, C.lib().gfft_last_error())
    st, text = explain(desc(prec=prec, n=1 << 20, batch=64))
    assert "scratch_bytes=268435456" in text   # 2^20 x 64 four-byte elements
    for w, h in ((8192, 8192), (16, 4096), (4096, 2), (2, 1)):
        st, text = explain(desc(kind=C.GFFT_C2C_2D, prec=prec, n=w, n2=h))
        assert st == C.GFFT_OK, (w, h, C.lib().gfft_last_error())
    for n, cols in VERT_TABLE:
        if n < 2:
            continue
        st, text = explain(desc(kind=C.GFFT_COLS, prec=prec, n=n, n2=cols, ins=cols + 3, outs=cols + 1))
        assert st == C.GFFT_OK, (n, cols, C.lib().gfft_last_error())


def test_16bit_storage_rejections():
    lib = C.lib()
    for kind in (C.GFFT_R2C, C.GFFT_C2R, C.GFFT_DIT, C.GFFT_R2C_2D):
        assert explain(desc(kind=kind, prec=C.GFFT_F16, n=64, n2=8))[0] == C.GFFT_ERR_UNSUPPORTED
    assert explain(desc(prec=C.GFFT_BF16, n=1))[0] == C.GFFT_ERR_UNSUPPORTED
    assert explain(desc(prec=4, n=64))[0] == C.GFFT_ERR_INVALID_ARG
    assert lib.gfft_execute_scaled(None, None, None, 0, None, 0, None) == C.GFFT_ERR_INVALID_ARG


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
@pytest.mark.parametrize("w,h", [(2, 1), (4, 1), (2, 8), (32, 16), (8, 64), (1024, 512), (4096, 4096),
                                 (8192, 8192), (16, 4096), (4096, 2), (32768, 4)])
def test_c2r_2d_plans_in_bounds(prec, w, h):
    """RealFFT2D::inverse (FFTReal.h:120-130): column inverse into scratch
    region B, then row c2r into the real output; dense, padded and batched."""
    for ins, outs, b in ((0, 0, 1), (w // 2 + 1, w + 2, 1), (w + 3, w + 6, 1), (0, 0, 3)):
        st, text = explain(desc(kind=C.GFFT_C2R_2D, prec=prec, n=w, n2=h, ins=ins, outs=outs, batch=b))
        if (prec == C.GFFT_F64 and w // 2 > 4096 * 2) or w // 2 > (1 << 22):
            continue
        assert st == C.GFFT_OK, (w, h, ins, outs, b, C.lib().gfft_last_error())
        # rows of 4 .. 32768 (complex64) / 16384 (complex128) reals: pack fused into the row pass
        fused = 4 <= w <= (32768 if prec == C.GFFT_F32 else 16384)
        assert ("real_rows mode=c2r" in text) == fused and ("pack n=" in text) != fused, (w, text)


def test_c2r_2d_rejections():
    assert explain(desc(kind=C.GFFT_C2R_2D, n=1, n2=4))[0] == C.GFFT_ERR_UNSUPPORTED
    assert explain(desc(kind=C.GFFT_C2R_2D, n=16, n2=4, outs=15))[0] == C.GFFT_ERR_INVALID_ARG
    assert explain(desc(kind=C.GFFT_C2R_2D, n=16, n2=4, ins=8))[0] == C.GFFT_ERR_INVALID_ARG
    assert explain(desc(kind=8, n=16, n2=4))[0] == C.GFFT_ERR_INVALID_ARG


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
def test_r2c_half_fuses_the_unpack_into_the_row_pass(prec, monkeypatch):
    """Half-spectrum R2C of a four-step length n/2 = n1 * 512 runs two launches
    (column pass, fused row-and-unpack pass) per chunk; full spectra, single
    pass lengths and n1 beyond the column table keep the separate unpack.  The
    bounds proof holds for dense (odd), padded and chunked outputs."""
    maxrow = 16384 if prec == C.GFFT_F32 else 8192
    for k in range(1, 24):
        n = 1 << k
        L = n // 2
        rows = 2 <= L <= maxrow                      # one row-tile pass, unpack as its epilogue
        fused = L > maxrow and L // 512 <= 2048     # four-step, unpack in the last pass
        for od in (0, n // 2 + 2, n // 2 + 7):
            for b in (1, 4):
                st, text = explain(desc(kind=C.GFFT_R2C, prec=prec, n=n, batch=b, outd=od, flags=C.GFFT_FLAG_HALF))
                assert st == C.GFFT_OK, (n, od, b, C.lib().gfft_last_error())
                assert ("r2c_rows_unpack" in text) == fused, (n, text)
                assert ("real_rows mode=r2c" in text) == rows, (n, text)
                assert ("unpack n=" in text) == (not fused and not rows and n > 1), (n, text)
        st, text = explain(desc(kind=C.GFFT_R2C, prec=prec, n=n, batch=2))
        assert st == C.GFFT_OK and "r2c_rows_unpack" not in text
        assert ("real_rows mode=r2c" in text) == rows
    # the fused plan never chunks (every input is read before a bin is
    # written); past the chunk size an even output distance keeps the
    # chunked separate-unpack path, an odd one (full scratch either way) fuses
    monkeypatch.setenv("GFFT_CHUNK_BYTES", str(2 * (1 << 19) * (8 if prec == C.GFFT_F32 else 16)))
    st, text = explain(desc(kind=C.GFFT_R2C, prec=prec, n=1 << 20, batch=5, flags=C.GFFT_FLAG_HALF))
    assert st == C.GFFT_OK and text.count("r2c_rows_unpack") == 1 and text.count("tile ") == 1
    st, text = explain(desc(kind=C.GFFT_R2C, prec=prec, n=1 << 20, batch=5, outd=(1 << 19) + 2,
                            flags=C.GFFT_FLAG_HALF))
    assert st == C.GFFT_OK and "r2c_rows_unpack" not in text and text.count("tile ") == 6
    st, text = explain(desc(kind=C.GFFT_R2C, prec=prec, n=1 << 20, batch=2, outd=(1 << 19) + 2,
                            flags=C.GFFT_FLAG_HALF))
    assert st == C.GFFT_OK and "r2c_rows_unpack" in text
    monkeypatch.setenv("GFFT_R2C_FUSE", "0")
    st, text = explain(desc(kind=C.GFFT_R2C, prec=prec, n=1 << 20, batch=5, flags=C.GFFT_FLAG_HALF))
    assert st == C.GFFT_OK and "r2c_rows_unpack" not in text and "unpack n=" in text


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
def test_c2r_fuses_the_pack_into_the_column_pass(prec, monkeypatch):
    """C2R of a four-step length n/2 = n1 * n2 with a fused shape for n1:
    the pack-and-column launch plus the transposing row pass; other lengths
    keep the separate pack.  Bounds proofs hold for padded distances."""
    maxrow = 16384 if prec == C.GFFT_F32 else 8192
    shapes = {16, 32, 64, 128, 256, 512} | ({1024} if prec == C.GFFT_F32 else set())
    for k in range(1, 24):
        n = 1 << k
        L = n // 2
        for ind, outd, b in ((0, 0, 1), (L + 4, n + 6, 3)):
            st, text = explain(desc(kind=C.GFFT_C2R, prec=prec, n=n, batch=b, ind=ind, outd=outd))
            assert st == C.GFFT_OK, (n, C.lib().gfft_last_error())
            fused = "c2r_pack_cols" in text
            if fused:
                assert L > maxrow and len(tiles(text)) == 1 and "pack n=" not in text
                n1 = int(re.search(r"c2r_pack_cols L=(\d+)", text).group(1))
                assert n1 in shapes and L % n1 == 0
            else:
                assert L <= maxrow or L // 1024 not in shapes or n == 1 << 23, (n, text)
    assert "c2r_pack_cols" in explain(desc(kind=C.GFFT_C2R, prec=prec, n=1 << 20, batch=2))[1]
    monkeypatch.setenv("GFFT_R2C_FUSE", "0")
    assert "c2r_pack_cols" not in explain(desc(kind=C.GFFT_C2R, prec=prec, n=1 << 20, batch=2))[1]


def test_4096_column_tile_serves_columns_not_four_step():
    """complex64: 4096-row column transforms (2D 4096^2, FFTVert) run one
    4-column tile pass; the four-step of 2^22 keeps 2048 x 2048 and the fused
    real FFT of 2^22 keeps the separate unpack (n1 = 4096 would need it)."""
    st, text = explain(desc(kind=C.GFFT_C2C_2D, n=4096, n2=4096))
    t = tiles(text)
    assert st == C.GFFT_OK and len(t) == 2
    assert field(t[1], "L") == "4096" and field(t[1], "M") == "4" and field(t[1], "inp") == "1"
    st, text = explain(desc(kind=C.GFFT_COLS, n=4096, n2=999, ins=1001, outs=1000))
    t = tiles(text)
    assert st == C.GFFT_OK and len(t) == 1 and field(t[0], "L") == "4096"
    st, text = explain(desc(n=1 << 22, batch=32))
    t = tiles(text)
    assert st == C.GFFT_OK and [field(x, "L") for x in t] == ["2048", "2048"]
    st, text = explain(desc(kind=C.GFFT_R2C, n=1 << 22, batch=2, flags=C.GFFT_FLAG_HALF))
    assert st == C.GFFT_OK and "r2c_rows_unpack" not in text


@pytest.mark.parametrize("prec", [C.GFFT_F32, C.GFFT_F64])
def test_dit_plans_for_multiples_of_four(prec):
    """DIT<T>(n) takes any n divisible by 4, as genFFT's dynamic FFTDIT<-1>
    (generic/fft_dit_impl_generic.inl:117-153); other kinds stay powers of two."""
    for n in (12, 20, 100, 1000, 4100, 3 << 18, (1 << 23) - 4):
        for half in (0, 1):
            st, text = explain(desc(kind=C.GFFT_DIT, prec=prec, n=n, batch=3, flags=half))
            assert st == C.GFFT_OK, (n, C.lib().gfft_last_error())
            assert "unpack" in text
    for n in (6, 10, 7):
        assert explain(desc(kind=C.GFFT_DIT, prec=prec, n=n))[0] == C.GFFT_ERR_UNSUPPORTED
    assert explain(desc(kind=C.GFFT_C2C, prec=prec, n=12))[0] == C.GFFT_ERR_UNSUPPORTED
    assert explain(desc(kind=C.GFFT_R2C, prec=prec, n=12))[0] == C.GFFT_ERR_UNSUPPORTED


def test_workspace_less_plans_fit_the_static_workspace():
    """GFFT_FLAG_NO_WORKSPACE plans (run through gfft_execute_noalloc) cut a
    four-step batch into chunks that fit the static workspace; a plan that
    cannot fit it is refused with a clear error."""
    st, text = explain(desc(n=1 << 20, batch=128, flags=C.GFFT_FLAG_NO_WORKSPACE))
    assert st == C.GFFT_OK, C.lib().gfft_last_error()
    scratch = int(text.split("scratch_bytes=")[1].split()[0])
    assert 0 < scratch <= C.GFFT_NOALLOC_SCRATCH_BYTES
    assert text.count("\ntile") == 2 * 128 // (C.GFFT_NOALLOC_SCRATCH_BYTES // ((1 << 20) * 8))
    # the same descriptor with its own workspace runs the batch in one chunk
    st, text = explain(desc(n=1 << 20, batch=128))
    assert st == C.GFFT_OK and text.count("\ntile") == 2
    # an 8192 x 8192 2D plane needs its column sub-pass intermediate whole
    st, _ = explain(desc(kind=C.GFFT_C2C_2D, n=8192, n2=8192, flags=C.GFFT_FLAG_NO_WORKSPACE))
    assert st == C.GFFT_ERR_UNSUPPORTED and b"static workspace" in C.lib().gfft_last_error()
    # single-pass plans need no workspace at all
    st, text = explain(desc(n=1024, batch=4096, flags=C.GFFT_FLAG_NO_WORKSPACE))
    assert st == C.GFFT_OK and "scratch_bytes=0" in text
