// ref_genfft.cpp -- extern "C" shim over genFFT itself, compiled from the
// reference's own headers where they lie (-I <reference>/include) into
// oracle/_ref/libgenfft_ref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: used to write the golden vectors under
// tests/golden/, to validate the C restatement in oracle.c, and as the
// "reference" CPU baseline timed by bench.py.  The shipped library never
// links or loads it.
//
// Two backends are exposed, selected by `backend`:
//   0 = genFFT's native x86 backend (impl_native, include/genFFT/FFTBackend.h:51-60),
//       built here with -mavx2 -mfma, i.e. the AVX_FMA kernels;
//   1 = genFFT's generic scalar backend (impl_generic, passed explicitly as
//       the plan factory template argument, include/genFFT/fft.h:56).
// Pointers are interleaved complex arrays; strides are in complex elements
// except where the reference API uses scalars (RealFFT2D input stride).
#include <genFFT/fft.h>

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdint>
#include <thread>
#include <vector>

using genfft::FFT;

namespace {

template <class T>
using GenericFFT = genfft::FFT<T, genfft::impl_generic::GetImpl>;

template <class T>
void c2c(T *out, const T *in, int n, int inv, int backend)
{
    auto *o = reinterpret_cast<std::complex<T> *>(out);
    auto *i = reinterpret_cast<const std::complex<T> *>(in);
    if (backend == 1) {
        GenericFFT<T> f(n);
        if (inv) f.template transform<true>(o, i); else f.template transform<false>(o, i);
    } else {
        FFT<T> f(n);
        if (inv) f.template transform<true>(o, i); else f.template transform<false>(o, i);
    }
}

template <class T>
void c2c_batch_mt(T *out, const T *in, int n, int64_t batch, int inv, int nthreads)
{
    // One shared plan, batch strided across threads (plan creation first:
    // genFFT's plan cache is not thread-safe, include/genFFT/FFTLevel.h:76-92).
    FFT<T> f(n);
    auto work = [&](int tid) {
        for (int64_t b = tid; b < batch; b += nthreads) {
            auto *o = reinterpret_cast<std::complex<T> *>(out) + b * n;
            auto *i = reinterpret_cast<const std::complex<T> *>(in) + b * n;
            if (inv) f.template transform<true>(o, i); else f.template transform<false>(o, i);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work, t);
    for (auto &t : th) t.join();
}

template <class T>
void real_forward_batch_mt(T *out, int64_t out_dist, const T *in, int n, int64_t batch, int half,
                           int nthreads)
{
    genfft::RealFFT<T> f(n);
    auto work = [&](int tid) {
        for (int64_t b = tid; b < batch; b += nthreads)
            f.forward(reinterpret_cast<std::complex<T> *>(out) + b * out_dist, in + b * n, half != 0);
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work, t);
    for (auto &t : th) t.join();
}

}  // namespace

#define REF_API extern "C" __attribute__((visibility("default")))

REF_API void ref_c2c_f32(float *out, const float *in, int n, int inv, int backend) { c2c(out, in, n, inv, backend); }
REF_API void ref_c2c_f64(double *out, const double *in, int n, int inv, int backend) { c2c(out, in, n, inv, backend); }

REF_API void ref_c2c_batch_mt(void *out, const void *in, int n, int64_t batch, int inv, int dbl, int nthreads)
{
    if (dbl) c2c_batch_mt((double *)out, (const double *)in, n, batch, inv, nthreads);
    else c2c_batch_mt((float *)out, (const float *)in, n, batch, inv, nthreads);
}

// Seconds per forward + unscaled inverse pair of one complex64 transform of
// length n on this thread, as genFFT's own benchmark times it
// (test/fft_bench.cpp:41-100: one plan, the same buffers, calls in a loop):
// best of `trials` runs of `reps` pairs, after one untimed pair.
REF_API double ref_c2c_pair_seconds(int n, int reps, int trials)
{
    std::vector<std::complex<float>> a((size_t)n), b((size_t)n);
    for (int i = 0; i < n; i++) a[(size_t)i] = {(float)(i % 7) - 3.0f, (float)(i % 5) - 2.0f};
    FFT<float> f(n);
    f.transform<false>(b.data(), a.data());
    f.transform<true>(a.data(), b.data());
    double best = 1e30;
    for (int t = 0; t < trials; t++) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; r++) {
            f.transform<false>(b.data(), a.data());
            f.transform<true>(a.data(), b.data());
        }
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        best = std::min(best, s / reps);
    }
    return best;
}

REF_API void ref_real_forward_batch_mt(void *out, int64_t out_dist, const void *in, int n, int64_t batch,
                                       int half, int dbl, int nthreads)
{
    if (dbl) real_forward_batch_mt((double *)out, out_dist, (const double *)in, n, batch, half, nthreads);
    else real_forward_batch_mt((float *)out, out_dist, (const float *)in, n, batch, half, nthreads);
}

#define REF_TYPED(T, SFX)                                                                          \
    REF_API void ref_real_forward_##SFX(T *out, const T *in, int n, int half, int backend)         \
    {                                                                                              \
        auto *o = reinterpret_cast<std::complex<T> *>(out);                                       \
        if (backend == 1) {                                                                        \
            genfft::RealFFT<T, genfft::impl_generic::GetImpl, genfft::impl_generic::GetDITImpl> f(n); \
            f.forward(o, in, half != 0);                                                           \
        } else {                                                                                   \
            genfft::RealFFT<T> f(n);                                                               \
            f.forward(o, in, half != 0);                                                           \
        }                                                                                          \
    }                                                                                              \
    REF_API void ref_dit_apply_##SFX(T *out, const T *in, int n, int half, int backend)            \
    {                                                                                              \
        if (backend == 1) {                                                                        \
            genfft::DIT<T, genfft::impl_generic::GetDITImpl> d(n);                                 \
            d.apply(out, in, half != 0);                                                           \
        } else {                                                                                   \
            genfft::DIT<T> d(n);                                                                   \
            d.apply(out, in, half != 0);                                                           \
        }                                                                                          \
    }                                                                                              \
    REF_API void ref_separate_2x_##SFX(T *out1, T *out2, const T *in, int n)                       \
    {                                                                                              \
        genfft::separate_2x_real_FFT(reinterpret_cast<std::complex<T> *>(out1),                  \
                                     reinterpret_cast<std::complex<T> *>(out2),                   \
                                     reinterpret_cast<const std::complex<T> *>(in), n);           \
    }                                                                                              \
    REF_API void ref_transform_real_##SFX(T *out, const T *in, int n)                              \
    {                                                                                              \
        FFT<T> f(n);                                                                               \
        f.transform_real(reinterpret_cast<std::complex<T> *>(out), in);                           \
    }                                                                                              \
    REF_API void ref_transform_interleave_##SFX(T *out, const T *in1, const T *in2, int n)         \
    {                                                                                              \
        FFT<T> f(n);                                                                               \
        f.transform_interleave(reinterpret_cast<std::complex<T> *>(out), in1, in2);               \
    }                                                                                              \
    REF_API void ref_vert_##SFX(T *out, int os, const T *in, int is, int n, int cols, int inv)     \
    {                                                                                              \
        genfft::FFTVert<T> f(n);                                                                   \
        auto *o = reinterpret_cast<std::complex<T> *>(out);                                       \
        auto *i = reinterpret_cast<const std::complex<T> *>(in);                                  \
        if (inv) f.template transform<true>(o, os, i, is, cols);                                   \
        else f.template transform<false>(o, os, i, is, cols);                                      \
    }                                                                                              \
    REF_API void ref_fft2d_##SFX(T *out, int os, const T *in, int is, int width, int height,       \
                                 int inv)                                                          \
    {                                                                                              \
        genfft::FFT2D<T> f(width, height);                                                         \
        auto *o = reinterpret_cast<std::complex<T> *>(out);                                       \
        auto *i = reinterpret_cast<const std::complex<T> *>(in);                                  \
        if (inv) f.template transform<true>(o, os, i, is);                                         \
        else f.template transform<false>(o, os, i, is);                                           \
    }                                                                                              \
    REF_API void ref_real2d_forward_##SFX(T *out, int os, const T *in, int is, int width,          \
                                          int height)                                              \
    {                                                                                              \
        genfft::RealFFT2D<T> f(width, height);                                                     \
        f.forward(reinterpret_cast<std::complex<T> *>(out), os, in, is);                          \
    }                                                                                              \
    REF_API void ref_real2d_forward_2x_##SFX(T *out, int os, const T *in1, int is1, const T *in2,  \
                                             int is2, int width, int height)                       \
    {                                                                                              \
        genfft::RealFFT2D<T> f(width, height);                                                     \
        f.forward_2x(reinterpret_cast<std::complex<T> *>(out), os, in1, is1, in2, is2);           \
    }

REF_TYPED(float, f32)
REF_TYPED(double, f64)
