// genfft_amd/fft.h -- genFFT's plan-object surface on the gfx950 engine.
//
// Drop-in for <genFFT/fft.h> + <genFFT/FFTReal.h>: the same class templates,
// constructors and methods, implemented over the C ABI in <genfft_amd.h>.
//
//   genfft::FFT<T>(n)              include/genFFT/fft.h:56-113
//   genfft::FFTVert<T>(n)          include/genFFT/fft.h:117-171
//   genfft::DIT<T>(n)              include/genFFT/fft.h:173-196
//   genfft::FFT2D<T>(w, h)         include/genFFT/fft.h:200-245
//   genfft::RealFFT2D<T>(w, h)     include/genFFT/FFTReal.h:71-184
//   genfft::RealFFT<T>(n)          include/genFFT/FFTReal.h:188-221
//   genfft::separate_2x_real_FFT   include/genFFT/FFTReal.h:36-66
//
// Pointers may be host memory (the reference's model: the call stages the
// data through the device and returns with the result in place) or device
// memory (the call is enqueued on the stream set with genfft::set_stream,
// default the null stream, and returns without waiting).
//
// Differences from the reference, all additive or stricter:
//   * failures throw genfft::error (the reference asserts, fft_impl_generic.h:276-278);
//   * plans are cached per (kind, size, strides, precision) behind a mutex --
//     the reference's weak_ptr singleton cache is not thread-safe
//     (FFTLevel.h:76-92);
//   * FFT<T>::forward / inverse(out, in) exist, as the reference README shows
//     (README.txt:28-31); transform_many<inv>(out, in, batch) runs a batch;
//   * RealFFT<T>::inverse(out, in) is the unscaled half-spectrum -> real
//     transform the reference lacks (FFTReal.h:127-130, README.txt:51-52).
//
// The backend seam keeps the reference's spelling: FFT, FFTVert, DIT and
// RealFFT take the factory template parameters of fft.h:56, :117, :173 and
// FFTReal.h:188, and the factory names of every reference backend
// (backend::, impl_generic::, impl_native::, impl_x86_dispatch::, all
// FFTBackend.h:40-70) name the gfx950 engine's factories, so code that picks
// a CPU tier compiles unchanged and runs on the GPU.
#ifndef GENFFT_AMD_FFT_H
#define GENFFT_AMD_FFT_H

#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../genfft_amd.h"

namespace genfft {

using stride_t = ptrdiff_t;  // include/genFFT/FFTTypes.h:35
using index_t = int;         // include/genFFT/FFTTypes.h:36

struct error : std::runtime_error {
    int status;
    error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

namespace detail {

inline void check(int st)
{
    if (st != GFFT_OK) {
        const char *m = gfft_last_error();
        throw error(st, std::string("genfft_amd: ") + gfft_error_string(st) + ": " + (m ? m : ""));
    }
}

inline void *&stream_ref()
{
    static void *s = nullptr;
    return s;
}

template <class T>
constexpr int precision()
{
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                  "genfft supports float and double");
    return std::is_same<T, double>::value ? GFFT_F64 : GFFT_F32;
}

struct PlanHandle {
    gfft_plan h = nullptr;
    explicit PlanHandle(const gfft_desc &d) { check(gfft_plan_create(&h, &d)); }
    ~PlanHandle() { gfft_plan_destroy(h); }
    PlanHandle(const PlanHandle &) = delete;
    PlanHandle &operator=(const PlanHandle &) = delete;
};

using PlanPtr = std::shared_ptr<PlanHandle>;

// Process-wide plan cache (the factory seam of include/genFFT/fft.h:47-48),
// guarded by a mutex; entries live while any user object holds them.
inline PlanPtr get_plan(int kind, int prec, int64_t n, int64_t n2, int64_t batch, int64_t in_stride,
                        int64_t out_stride, int64_t in_dist, int64_t out_dist, int flags)
{
    using Key = std::tuple<int, int, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int>;
    static std::mutex mu;
    static std::map<Key, std::weak_ptr<PlanHandle>> cache;
    Key key{kind, prec, n, n2, batch, in_stride, out_stride, in_dist, out_dist, flags};
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end())
        if (auto p = it->second.lock()) return p;
    gfft_desc d{};
    d.kind = kind;
    d.precision = prec;
    d.n = n;
    d.n2 = n2;
    d.batch = batch;
    d.in_stride = in_stride;
    d.out_stride = out_stride;
    d.in_dist = in_dist;
    d.out_dist = out_dist;
    d.flags = flags;
    auto p = std::make_shared<PlanHandle>(d);
    cache[key] = p;
    return p;
}

inline bool on_device(const void *a, const void *b)
{
    const bool da = gfft_is_device_pointer(a) != 0, db = gfft_is_device_pointer(b) != 0;
    if (da != db) throw error(GFFT_ERR_INVALID_ARG, "genfft_amd: input and output must both be host or both device memory");
    return da;
}

inline void run(const PlanPtr &p, const void *in, void *out, bool inv)
{
    const int dir = inv ? GFFT_INVERSE : GFFT_FORWARD;
    if (on_device(in, out)) check(gfft_execute(p->h, in, out, dir, stream_ref()));
    else check(gfft_execute_host(p->h, in, out, dir));
}

inline int64_t bitrev(int64_t i, int64_t n)
{
    int64_t r = 0;
    for (int64_t m = n >> 1; m; m >>= 1, i >>= 1) r = (r << 1) | (i & 1);
    return r;
}

}  // namespace detail

/// Stream used for device-pointer calls (a hipStream_t; nullptr = null stream).
inline void set_stream(void *hip_stream) { detail::stream_ref() = hip_stream; }

// ---- the backend seam (include/genFFT/fft.h:41-52, FFTBackend.h:40-70) ----
namespace impl {
/// The reference's backend objects hold a CPU tier's butterflies
/// (FFTLevel.h:43-117, FFTDIT.h:40-50).  Every class here runs on the gfx950
/// engine, so a factory's result records only the size it accepted; the
/// plans themselves come from the engine's plan cache when a class runs.
template <class T>
struct FFTBase {
    int n = 0;
};
template <class T>
struct FFTVertBase {
    int n = 0;
};
template <class T>
struct FFTDITBase {
    int n = 0;
};
}  // namespace impl

template <typename T>
using FFTImplPtr = std::shared_ptr<impl::FFTBase<T>>;
template <typename T>
using FFTVertImplPtr = std::shared_ptr<impl::FFTVertBase<T>>;
template <typename T>
using FFTDITImplPtr = std::shared_ptr<impl::FFTDITBase<T>>;
template <typename T>
using FFTImplFactory = FFTImplPtr<T>(int n, T);
template <typename T>
using FFTVertImplFactory = FFTVertImplPtr<T>(int n, T);
template <typename T>
using FFTDITImplFactory = FFTDITImplPtr<T>(int n, T);

namespace detail {
inline bool pow2_size(int n) { return n >= 1 && n <= (1 << 23) && (n & (n - 1)) == 0; }

template <class B>
std::shared_ptr<B> make_impl(int n, bool ok, const char *what)
{
    if (!ok) throw error(GFFT_ERR_UNSUPPORTED, std::string("genfft_amd: unsupported ") + what + " size " + std::to_string(n));
    auto p = std::make_shared<B>();
    p->n = n;
    return p;
}
}  // namespace detail

/// The gfx950 engine's factories, one overload per scalar type as the
/// reference's (generic/fft_impl_generic.h:245-283, fft_dit_impl_generic.inl:117-153).
namespace impl_gfx950 {
inline FFTImplPtr<float> GetImpl(int n, float)
{
    return detail::make_impl<impl::FFTBase<float>>(n, detail::pow2_size(n), "FFT");
}
inline FFTImplPtr<double> GetImpl(int n, double)
{
    return detail::make_impl<impl::FFTBase<double>>(n, detail::pow2_size(n), "FFT");
}
inline FFTVertImplPtr<float> GetVertImpl(int n, float)
{
    return detail::make_impl<impl::FFTVertBase<float>>(n, detail::pow2_size(n), "FFTVert");
}
inline FFTVertImplPtr<double> GetVertImpl(int n, double)
{
    return detail::make_impl<impl::FFTVertBase<double>>(n, detail::pow2_size(n), "FFTVert");
}
// DIT: powers of two and, as genFFT's dynamic FFTDIT<-1>, any n divisible by 4
inline FFTDITImplPtr<float> GetDITImpl(int n, float)
{
    return detail::make_impl<impl::FFTDITBase<float>>(n, detail::pow2_size(n) || (n > 0 && n % 4 == 0), "DIT");
}
inline FFTDITImplPtr<double> GetDITImpl(int n, double)
{
    return detail::make_impl<impl::FFTDITBase<double>>(n, detail::pow2_size(n) || (n > 0 && n % 4 == 0), "DIT");
}
}  // namespace impl_gfx950

namespace backend {
using impl_gfx950::GetDITImpl;
using impl_gfx950::GetImpl;
using impl_gfx950::GetVertImpl;
}  // namespace backend
namespace impl_generic {
using impl_gfx950::GetDITImpl;
using impl_gfx950::GetImpl;
using impl_gfx950::GetVertImpl;
}  // namespace impl_generic
namespace impl_native {
using impl_gfx950::GetDITImpl;
using impl_gfx950::GetImpl;
using impl_gfx950::GetVertImpl;
}  // namespace impl_native
namespace impl_x86_dispatch {
using impl_gfx950::GetDITImpl;
using impl_gfx950::GetImpl;
using impl_gfx950::GetVertImpl;
}  // namespace impl_x86_dispatch

/// A 1D FFT for densely packed data (include/genFFT/fft.h:54-113).
template <class T, FFTImplFactory<T> *factory = backend::GetImpl>
struct FFT {
    FFT() = default;
    explicit FFT(int n)
        : n(n), impl(factory(n, T())),
          plan(detail::get_plan(GFFT_C2C, detail::precision<T>(), n, 0, 1, 0, 0, 0, 0, 0))
    {
    }

    /// fft.h:80-85: natural-order transform; the inverse is unscaled.
    template <bool inv>
    void transform(std::complex<T> *out, const std::complex<T> *in)
    {
        detail::run(plan, in, out, inv);
    }

    void forward(std::complex<T> *out, const std::complex<T> *in) { transform<false>(out, in); }
    void inverse(std::complex<T> *out, const std::complex<T> *in) { transform<true>(out, in); }

    /// `batch` transforms at distance n (not in the reference).
    template <bool inv>
    void transform_many(std::complex<T> *out, const std::complex<T> *in, int batch)
    {
        auto p = detail::get_plan(GFFT_C2C, detail::precision<T>(), n, 0, batch, 0, 0, 0, 0, 0);
        detail::run(p, in, out, inv);
    }

    /// fft.h:69-73: in-place transform of data already in bit-reversed order
    /// (device data: gathered into natural order on the device first).
    template <bool inv>
    void transform_no_scramble(std::complex<T> *inout)
    {
        if (gfft_is_device_pointer(inout)) {
            detail::check(gfft_execute_scrambled(plan->h, inout, inout, inv ? GFFT_INVERSE : GFFT_FORWARD,
                                                 detail::stream_ref()));
            return;
        }
        std::vector<std::complex<T>> nat(n);
        for (int i = 0; i < n; i++) nat[detail::bitrev(i, n)] = inout[i];
        detail::check(gfft_execute_host(plan->h, nat.data(), inout, inv ? GFFT_INVERSE : GFFT_FORWARD));
    }

    /// fft.h:90-94: real input, full n-bin spectrum.
    void transform_real(std::complex<T> *out, const T *in)
    {
        auto p = detail::get_plan(GFFT_R2C, detail::precision<T>(), n, 0, 1, 0, 0, 0, 0, 0);
        detail::run(p, in, out, false);
    }

    /// fft.h:100-105: FFT of in1 + i*in2; split with separate_2x_real_FFT.
    void transform_interleave(std::complex<T> *out, const T *in1, const T *in2)
    {
        if (detail::on_device(in1, out)) {
            detail::check(gfft_interleave_real(detail::precision<T>(), out, in1, in2, n, 1, detail::stream_ref()));
            detail::check(gfft_execute(plan->h, out, out, GFFT_FORWARD, detail::stream_ref()));
        } else {
            for (int i = 0; i < n; i++) out[i] = std::complex<T>(in1[i], in2[i]);
            detail::check(gfft_execute_host(plan->h, out, out, GFFT_FORWARD));
        }
    }

    int size() const noexcept { return n; }
    explicit operator bool() const noexcept { return (bool)impl; }

private:
    int n = 0;
    FFTImplPtr<T> impl;
    detail::PlanPtr plan;
};

/// Column-wise 1D FFT for multiple columns (include/genFFT/fft.h:115-171).
/// Strides are in complex elements.
template <class T, FFTVertImplFactory<T> *factory = backend::GetVertImpl>
struct FFTVert {
    FFTVert() = default;
    explicit FFTVert(int n) : n(n), impl(factory(n, T())) {}

    template <bool inv>
    void transform(std::complex<T> *out, stride_t out_stride, const std::complex<T> *in, stride_t in_stride,
                   index_t cols)
    {
        auto p = detail::get_plan(GFFT_COLS, detail::precision<T>(), n, cols, 1, in_stride, out_stride, 0, 0, 0);
        detail::run(p, in, out, inv);
    }

    template <bool inv>
    void transform(std::complex<T> *out, const std::complex<T> *in, index_t cols)
    {
        transform<inv>(out, cols, in, cols, cols);
    }

    /// fft.h:132-136: rows already in bit-reversed order (host or device data).
    template <bool inv>
    void transform_no_scramble(std::complex<T> *data, stride_t stride, index_t cols)
    {
        if (gfft_is_device_pointer(data)) {
            auto p = detail::get_plan(GFFT_COLS, detail::precision<T>(), n, cols, 1, stride, stride, 0, 0, 0);
            detail::check(gfft_execute_scrambled(p->h, data, data, inv ? GFFT_INVERSE : GFFT_FORWARD,
                                                 detail::stream_ref()));
            return;
        }
        std::vector<std::complex<T>> nat((size_t)n * stride);
        for (int r = 0; r < n; r++)
            std::memcpy(&nat[(size_t)detail::bitrev(r, n) * stride], data + (size_t)r * stride,
                        sizeof(std::complex<T>) * cols);
        transform<inv>(data, stride, nat.data(), stride, cols);
    }

    int size() const noexcept { return n; }
    explicit operator bool() const noexcept { return (bool)impl; }

private:
    int n = 0;
    FFTVertImplPtr<T> impl;
};

/// Real-FFT unpack step (include/genFFT/fft.h:173-196).
template <class T, FFTDITImplFactory<T> *factory = backend::GetDITImpl>
class DIT {
public:
    DIT() = default;
    explicit DIT(int n) : n(n), impl(factory(n, T())) {}

    void apply(T *out, const T *in, bool half)
    {
        auto p = detail::get_plan(GFFT_DIT, detail::precision<T>(), n, 0, 1, 0, 0, 0, 0, half ? GFFT_FLAG_HALF : 0);
        detail::run(p, in, out, false);
    }
    void apply(std::complex<T> *out, const std::complex<T> *in, bool half) { apply((T *)out, (const T *)in, half); }

    int size() const noexcept { return n; }
    explicit operator bool() const noexcept { return (bool)impl; }

private:
    int n = 0;
    FFTDITImplPtr<T> impl;
};

/// 2D FFT (include/genFFT/fft.h:198-245); strides in complex elements.
template <class T>
class FFT2D {
public:
    FFT2D() = default;
    FFT2D(int width, int height) : w(width), h(height) {}

    template <bool inv>
    void transform(std::complex<T> *out, stride_t out_stride, const std::complex<T> *in, stride_t in_stride)
    {
        auto p = detail::get_plan(GFFT_C2C_2D, detail::precision<T>(), w, h, 1, in_stride, out_stride, 0, 0, 0);
        detail::run(p, in, out, inv);
    }

    int cols() const noexcept { return w; }
    int rows() const noexcept { return h; }
    explicit operator bool() const noexcept { return w > 0 && h > 0; }

private:
    int w = 0, h = 0;
};

/// separate_2x_real_FFT (include/genFFT/FFTReal.h:34-66): in may alias out1.
template <class T>
void separate_2x_real_FFT(std::complex<T> *out1, std::complex<T> *out2, const std::complex<T> *in, int N)
{
    if (gfft_is_device_pointer(in)) {
        detail::check(gfft_separate_2x_real(detail::precision<T>(), out1, out2, in, N, 1, detail::stream_ref()));
        return;
    }
    const std::complex<T> z0 = in[0];
    out1[0] = std::complex<T>(z0.real(), 0);
    out2[0] = std::complex<T>(z0.imag(), 0);
    for (int i = 1; i <= N / 2; i++) {
        const int k = N - i;
        const std::complex<T> a = in[i], b = in[k];
        const T xr = (a.real() + b.real()) * T(0.5), xi = (a.imag() - b.imag()) * T(0.5);
        const T yr = (b.imag() + a.imag()) * T(0.5), yi = (b.real() - a.real()) * T(0.5);
        out1[i] = {xr, xi};
        out2[i] = {yr, yi};
        out1[k] = {xr, -xi};
        out2[k] = {yr, -yi};
    }
}

/// 2D real-input forward FFT (include/genFFT/FFTReal.h:69-184): out_stride in
/// complex elements, in_stride in scalars; the full spectrum is written.
template <class T>
class RealFFT2D {
public:
    RealFFT2D() = default;
    RealFFT2D(int width, int height) : w(width), h(height) {}

    void forward(std::complex<T> *out, int out_stride, const T *in, int in_stride)
    {
        auto p = detail::get_plan(GFFT_R2C_2D, detail::precision<T>(), w, h, 1, in_stride, out_stride, 0, 0, 0);
        detail::run(p, in, out, false);
    }

    /// FFTReal.h:106-118: the 2D FFT of in1 + i*in2 (split per element with
    /// X = (Z_k + conj Z_-k)/2, Y = -i (Z_k - conj Z_-k)/2).
    void forward_2x(std::complex<T> *out, int out_stride, const T *in1, int in_stride1, const T *in2, int in_stride2)
    {
        const bool dev = detail::on_device(in1, out);
        if (dev) {
            // one launch packs every row, then the 2D transform runs in place
            detail::check(gfft_interleave_real_2d(detail::precision<T>(), out, out_stride, in1, in_stride1, in2,
                                                  in_stride2, w, h, detail::stream_ref()));
        } else {
            for (int r = 0; r < h; r++)
                for (int c = 0; c < w; c++)
                    out[(size_t)r * out_stride + c] =
                        std::complex<T>(in1[(size_t)r * in_stride1 + c], in2[(size_t)r * in_stride2 + c]);
        }
        FFT2D<T>(w, h).template transform<false>(out, out_stride, out, out_stride);
    }

    /// FFTReal.h:120-130 (the reference's body is assert(!"TODO")): the
    /// unscaled inverse of forward -- reads columns 0..width/2 of the complex
    /// spectrum `in` (pitch in complex elements), writes width x height reals
    /// (pitch in scalars); inverse(forward(x)) = width * height * x.
    void inverse(T *out, int out_stride, const std::complex<T> *in, int in_stride)
    {
        auto p = detail::get_plan(GFFT_C2R_2D, detail::precision<T>(), w, h, 1, in_stride, out_stride, 0, 0, 0);
        detail::run(p, in, out, true);
    }

    int cols() const { return w; }
    int rows() const { return h; }

private:
    int w = 0, h = 0;
};

/// Real-input 1D FFT (include/genFFT/FFTReal.h:186-221).  The factories are
/// consulted as the reference does (the n/2-point transform and the n-point
/// unpack); the transform itself is one engine plan of the whole real FFT.
template <class T, FFTImplFactory<T> *factory = backend::GetImpl,
          FFTDITImplFactory<T> *dit_factory = backend::GetDITImpl>
struct RealFFT {
    RealFFT() = default;
    RealFFT(int n) : n(n)
    {
        if (n > 1) {
            impl = factory(n / 2, T());
            dit = dit_factory(n, T());
        }
    }

    /// half: write n/2+1 bins; otherwise all n (upper half by conjugate symmetry).
    void forward(std::complex<T> *out, const T *in, bool half)
    {
        auto p = detail::get_plan(GFFT_R2C, detail::precision<T>(), n, 0, 1, 0, 0, 0, 0, half ? GFFT_FLAG_HALF : 0);
        detail::run(p, in, out, false);
    }

    /// n/2+1 Hermitian bins -> n reals, unscaled: inverse(forward(x)) = n*x.
    void inverse(T *out, const std::complex<T> *in)
    {
        auto p = detail::get_plan(GFFT_C2R, detail::precision<T>(), n, 0, 1, 0, 0, 0, 0, 0);
        detail::run(p, in, out, true);
    }

    int size() const noexcept { return n; }

private:
    FFTImplPtr<T> impl;
    FFTDITImplPtr<T> dit;
    int n = 0;
};

}  // namespace genfft

#endif  // GENFFT_AMD_FFT_H
