"""Writes tests/golden/genfft_golden.npz from genFFT itself.

Runs only where the reference tree is present: it loads
oracle/_ref/libgenfft_ref.so (compiled from the reference's own headers by
oracle/Makefile) and records inputs and genFFT's outputs for the entry points
on the hot path.  The fixture holds data only: seeded inputs and the outputs of
  - FFT<T>::transform<inv>        (include/genFFT/fft.h:80-85), both backends
  - RealFFT<T>::forward           (include/genFFT/FFTReal.h:204-213)
  - DIT<T>::apply                 (include/genFFT/fft.h:180-188)
  - separate_2x_real_FFT          (include/genFFT/FFTReal.h:36-66)
  - FFT<T>::transform_real / transform_interleave (fft.h:90-105)
  - FFTVert<T>::transform         (fft.h:146-150), incl. odd column counts
  - FFT2D<T>::transform           (fft.h:214-218), padded pitches
  - RealFFT2D<T>::forward         (FFTReal.h:83-104), incl. odd scalar input pitches
  - RealFFT2D<T>::forward_2x      (FFTReal.h:106-118), equal input pitches only:
                                  genFFT offsets in2's second half by in_stride1
                                  (FFTReal.h:178), right only when the pitches agree
Usage:  python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libgenfft_ref.so")


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/libgenfft_ref.so missing: run `make -C oracle` with the reference tree present")
    r = ctypes.CDLL(REF)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rng = np.random.default_rng(20240601)
    out = {}

    def cplx(n, dt):
        return (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(dt)

    for dt, sfx in [(np.complex64, "f32"), (np.complex128, "f64")]:
        rdt = np.float32 if sfx == "f32" else np.float64
        for n in [1, 2, 4, 8, 16, 32, 64, 512, 2048]:
            x = cplx(n, dt)
            out[f"c2c_{sfx}_n{n}_in"] = x
            for inv in (0, 1):
                for backend, bname in ((0, "native"), (1, "generic")):
                    y = np.empty_like(x)
                    getattr(r, "ref_c2c_" + sfx)(P(y), P(x), n, inv, backend)
                    out[f"c2c_{sfx}_n{n}_{'inv' if inv else 'fwd'}_{bname}"] = y
        for n in [2, 4, 8, 16, 64, 1024]:
            x = rng.uniform(-1, 1, n).astype(rdt)
            out[f"real_{sfx}_n{n}_in"] = x
            for half in (0, 1):
                for backend, bname in ((0, "native"), (1, "generic")):
                    f = np.full(n, 43 + 21j, dt)  # sentinel past the written bins
                    getattr(r, "ref_real_forward_" + sfx)(P(f), P(x), n, half, backend)
                    out[f"real_{sfx}_n{n}_{'half' if half else 'full'}_{bname}"] = f
            z = cplx(max(1, n // 2), dt)
            out[f"dit_{sfx}_n{n}_in"] = z
            for half in (0, 1):
                f = np.full(n, 43 + 21j, dt)
                getattr(r, "ref_dit_apply_" + sfx)(P(f), P(z), n, half, 1)
                out[f"dit_{sfx}_n{n}_{'half' if half else 'full'}"] = f
            x2 = rng.uniform(-1, 1, n).astype(rdt)
            out[f"treal_{sfx}_n{n}_in2"] = x2
            y = np.empty(n, dt)
            getattr(r, "ref_transform_real_" + sfx)(P(y), P(x), n)
            out[f"treal_{sfx}_n{n}"] = y
            y = np.empty(n, dt)
            getattr(r, "ref_transform_interleave_" + sfx)(P(y), P(x), P(x2), n)
            out[f"tinter_{sfx}_n{n}"] = y
        # DIT<T>(n) for n divisible by 4 but not a power of two: genFFT's
        # dynamic FFTDIT<-1> (generic/fft_dit_impl_generic.inl:117-153)
        # (own generator: the inputs of every other entry stay as they were)
        rng_dit = np.random.default_rng(1212 if sfx == "f32" else 1213)
        for n in [12, 24, 40, 100, 1000]:
            z = (rng_dit.uniform(-1, 1, n // 2) + 1j * rng_dit.uniform(-1, 1, n // 2)).astype(dt)
            out[f"dit_{sfx}_n{n}_in"] = z
            for half in (0, 1):
                f = np.full(n, 43 + 21j, dt)
                getattr(r, "ref_dit_apply_" + sfx)(P(f), P(z), n, half, 1)
                out[f"dit_{sfx}_n{n}_{'half' if half else 'full'}"] = f
        for n in [8, 64]:
            z = cplx(n, dt)
            out[f"sep_{sfx}_n{n}_in"] = z
            a = np.empty_like(z)
            b = np.empty_like(z)
            getattr(r, "ref_separate_2x_" + sfx)(P(a), P(b), P(z), n)
            out[f"sep_{sfx}_n{n}_x"] = a
            out[f"sep_{sfx}_n{n}_y"] = b
        # FFTVert: (n, cols, in_stride, out_stride) incl. odd column counts
        for (n, cols, ins, outs) in [(8, 7, 7, 7), (16, 47, 47, 47), (256, 33, 40, 35), (2, 999, 999, 999),
                                     (4, 7, 9, 8), (1024, 5, 5, 5)]:
            x = cplx(n * ins, dt)
            key = f"vert_{sfx}_n{n}_c{cols}_i{ins}_o{outs}"
            out[key + "_in"] = x
            for inv in (0, 1):
                y = np.full(n * outs, 7 + 5j, dt)
                getattr(r, "ref_vert_" + sfx)(P(y), outs, P(x), ins, n, cols, inv)
                out[key + ("_inv" if inv else "_fwd")] = y
        for (w, h, ins, outs) in [(32, 16, 33, 35), (8, 64, 8, 8), (64, 4, 64, 70), (16, 1, 16, 16)]:
            x = cplx(h * ins, dt)
            key = f"fft2d_{sfx}_w{w}_h{h}_i{ins}_o{outs}"
            out[key + "_in"] = x
            for inv in (0, 1):
                y = np.full(h * outs, 7 + 5j, dt)
                getattr(r, "ref_fft2d_" + sfx)(P(y), outs, P(x), ins, w, h, inv)
                out[key + ("_inv" if inv else "_fwd")] = y
        for (w, h, ins, outs) in [(32, 16, 34, 33), (16, 8, 16, 16), (4, 32, 4, 4), (64, 2, 64, 64)]:
            x = rng.uniform(-1, 1, h * ins).astype(rdt)
            key = f"real2d_{sfx}_w{w}_h{h}_i{ins}_o{outs}"
            out[key + "_in"] = x
            y = np.full(h * outs, 7 + 5j, dt)
            getattr(r, "ref_real2d_forward_" + sfx)(P(y), outs, P(x), ins, w, h)
            out[key + "_fwd"] = y
        # own generator, so the inputs of every earlier entry stay as they were
        rng2 = np.random.default_rng(4242 if sfx == "f32" else 4243)
        # RealFFT2D::forward with odd scalar input pitches (FFTReal.h:83)
        for (w, h, ins, outs) in [(32, 16, 33, 35), (8, 8, 9, 8), (64, 4, 65, 64), (16, 1, 17, 16)]:
            x = rng2.uniform(-1, 1, h * ins).astype(rdt)
            key = f"real2d_{sfx}_w{w}_h{h}_i{ins}_o{outs}"
            out[key + "_in"] = x
            y = np.full(h * outs, 7 + 5j, dt)
            getattr(r, "ref_real2d_forward_" + sfx)(P(y), outs, P(x), ins, w, h)
            out[key + "_fwd"] = y
        # RealFFT2D::forward_2x, both inputs at one pitch (odd and even)
        for (w, h, ins, outs) in [(32, 16, 33, 35), (16, 8, 16, 16), (8, 32, 9, 8), (64, 2, 64, 66)]:
            x1 = rng2.uniform(-1, 1, h * ins).astype(rdt)
            x2 = rng2.uniform(-1, 1, h * ins).astype(rdt)
            key = f"real2d2x_{sfx}_w{w}_h{h}_i{ins}_o{outs}"
            out[key + "_in1"] = x1
            out[key + "_in2"] = x2
            y = np.full(h * outs, 7 + 5j, dt)
            getattr(r, "ref_real2d_forward_2x_" + sfx)(P(y), outs, P(x1), ins, P(x2), ins, w, h)
            out[key + "_fwd"] = y
    path = os.path.join(HERE, "genfft_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {len(out)} arrays to {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main()
