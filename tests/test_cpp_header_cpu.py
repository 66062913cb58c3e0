"""CPU checks of the C++ drop-in surface: the header compiles standalone and
keeps the reference's class/method names (include/genFFT/fft.h, FFTReal.h)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_compiles_and_exposes_reference_names(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <genFFT/fft.h>
#include <genFFT/FFTReal.h>
template <class T> void use() {
    genfft::FFT<T> f(1024); genfft::FFTVert<T> v(64); genfft::DIT<T> d(64);
    genfft::FFT2D<T> f2(32, 16); genfft::RealFFT<T> r(64); genfft::RealFFT2D<T> r2(32, 16);
    std::complex<T> *o = nullptr; const std::complex<T> *i = nullptr; T *ro = nullptr; const T *ri = nullptr;
    if (false) {
        f.template transform<false>(o, i); f.template transform<true>(o, i);
        f.forward(o, i); f.inverse(o, i); f.transform_real(o, ri); f.transform_interleave(o, ri, ri);
        f.template transform_no_scramble<false>(o);
        v.template transform<false>(o, 8, i, 8, 8); v.template transform<true>(o, i, 8);
        d.apply(o, i, true); d.apply(ro, ri, false);
        f2.template transform<false>(o, 32, i, 32);
        r.forward(o, ri, true); r.inverse(ro, i);
        r2.forward(o, 32, ri, 32); r2.forward_2x(o, 32, ri, 32, ri, 32);
        genfft::separate_2x_real_FFT(o, o, i, 64);
    }
    (void)f.size(); (void)bool(f); (void)v.size(); (void)d.size(); (void)f2.cols(); (void)f2.rows();
}
int main() { use<float>(); use<double>(); return 0; }
''')
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", f"-I{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_header_accepts_the_reference_factory_spellings(tmp_path):
    """genFFT's plug-in seam: the classes take a factory template parameter
    (fft.h:47-56, :117, :173; FFTReal.h:188) and callers name the backends
    (FFTBackend.h:40-70), e.g. the reference's test_x86.cpp forcing a tier.
    Every spelling compiles and names this engine's factory."""
    src = tmp_path / "f.cpp"
    src.write_text(r'''
#include <genFFT/fft.h>
#include <genFFT/FFTReal.h>
#include <type_traits>
template <class T> void use() {
    genfft::FFT<T, genfft::impl_generic::GetImpl> a(1024);
    genfft::FFT<T, genfft::impl_native::GetImpl> b(1024);
    genfft::FFT<T, genfft::impl_x86_dispatch::GetImpl> c(1024);
    genfft::FFT<T, genfft::backend::GetImpl> d(1024);
    genfft::FFTVert<T, genfft::impl_generic::GetVertImpl> v(64);
    genfft::DIT<T, genfft::impl_x86_dispatch::GetDITImpl> t(64);
    genfft::RealFFT<T, genfft::impl_native::GetImpl, genfft::impl_native::GetDITImpl> r(64);
    genfft::FFTImplFactory<T> *f = genfft::impl_generic::GetImpl;
    genfft::FFTVertImplFactory<T> *fv = genfft::backend::GetVertImpl;
    genfft::FFTDITImplFactory<T> *fd = genfft::impl_native::GetDITImpl;
    static_assert(std::is_same<decltype(f(8, T())), genfft::FFTImplPtr<T>>::value, "factory returns the impl pointer");
    (void)a.size(); (void)b.size(); (void)c.size(); (void)d.size(); (void)v.size(); (void)t.size(); (void)r;
    (void)f; (void)fv; (void)fd;
}
int main() { use<float>(); use<double>(); return 0; }
''')
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", f"-I{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
