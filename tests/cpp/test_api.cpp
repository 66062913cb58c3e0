// test_api.cpp -- the reference's own test cases, re-expressed against the
// drop-in header (#include <genFFT/fft.h> resolves to the gfx950 engine).
//
// Mirrors, with the same inputs and tolerances:
//   test/test_fft.cpp:41-66        FFT_float_test.InverseIdentity, 2 .. 2^22
//   test/test_dispatch.cpp:45-65   FFT_test_pow2 float/double,     2 .. 2^20
//   test/test_dispatch.cpp:84-126  FFT_vert_test_pow2 table (every column checked)
//   test/test_real_fft.cpp:42-67   FFT_real_test half/full + sentinel, 2 .. 2^22
//   test/test_dit.cpp:45-67        FFT_DIT_test_pow2 in/out of place, 2 .. 2^20
// plus FFT2D / RealFFT2D (forward with odd input pitches, forward_2x) /
// transform_real / transform_interleave / separate_2x_real_FFT, which the
// reference leaves untested, the factory template parameters spelled with the
// reference's backend names, and the device-pointer path.
//
// Inputs: DummyData exactly as test/test_util.h:36-57 (std::mt19937_64 with
// the default seed, std::uniform_real_distribution<T>(-1, 1)).  Expected
// values: the long-double oracle (oracle/liboracle.so, orc_fft_ld), with
// genFFT's absolute FFT_Eps (test/test_util.h:59-72) and the relative-L2
// bound 2e-7*log2(N) (c64) / 1e-15*log2(N) (c128).
#include <genFFT/fft.h>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

extern "C" void orc_fft_ld(const double *in, double *out, int64_t n, int64_t batch, int inv);

static int g_fail = 0, g_pass = 0;

static int genfft_amd_bitrev(int i, int n)
{
    int r = 0;
    for (int m = n >> 1; m; m >>= 1, i >>= 1) r = (r << 1) | (i & 1);
    return r;
}
#define EXPECT(cond, ...)                                               \
    do {                                                                \
        if (!(cond)) {                                                  \
            std::printf("  FAIL %s:%d: ", __FILE__, __LINE__);          \
            std::printf(__VA_ARGS__);                                   \
            std::printf("\n");                                          \
            g_fail++;                                                   \
            return;                                                     \
        }                                                               \
    } while (0)

template <typename T>
void DummyData(std::vector<std::complex<T>> &vec, bool real)
{
    std::mt19937_64 rng;
    std::uniform_real_distribution<T> dist(-1, 1);
    for (auto &c : vec) {
        c.real(dist(rng));
        c.imag(real ? 0 : dist(rng));
    }
}

template <typename T>
void DummyData(std::vector<T> &vec)
{
    std::mt19937_64 rng;
    std::uniform_real_distribution<T> dist(-1, 1);
    for (auto &x : vec) x = dist(rng);
}

template <typename T>
double FFT_Eps(int n) { return sizeof(T) == 8 ? 1e-8 + n * 1e-12 : 1e-5 + n * 1e-8; }

template <typename T>
double RelTol(int n)
{
    const int lg = std::max(1, (int)std::lround(std::log2((double)std::max(n, 2))));
    return (sizeof(T) == 8 ? 1e-15 : 2e-7) * lg;
}

// long-double oracle of a complex sequence (or of a real one, imag = 0)
template <typename T>
std::vector<std::complex<double>> oracle(const std::complex<T> *x, int n, bool inv, int stride = 1)
{
    std::vector<std::complex<double>> in(n), out(n);
    for (int i = 0; i < n; i++) in[i] = std::complex<double>(x[(size_t)i * stride]);
    orc_fft_ld((const double *)in.data(), (double *)out.data(), n, 1, inv);
    return out;
}

template <typename T, typename U>
double rel_l2(const std::complex<T> *a, const std::complex<U> *b, int n, int sa = 1, int sb = 1)
{
    double num = 0, den = 0;
    for (int i = 0; i < n; i++) {
        std::complex<double> x(a[(size_t)i * sa]), y(b[(size_t)i * sb]);
        num += std::norm(x - y);
        den += std::norm(y);
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

template <typename T, typename U>
double max_abs(const std::complex<T> *a, const std::complex<U> *b, int n)
{
    double m = 0;
    for (int i = 0; i < n; i++) m = std::max(m, std::abs(std::complex<double>(a[i]) - std::complex<double>(b[i])));
    return m;
}

// --------------------------------------------------------------------------
void InverseIdentity(int n)  // test/test_fft.cpp:41-57
{
    genfft::FFT<float> fft(n);
    std::vector<std::complex<float>> in(n), out(n), invout(n);
    DummyData(in, false);
    fft.transform<false>(out.data(), in.data());
    fft.transform<true>(invout.data(), out.data());
    const float eps = 1e-4f, norm = 1.0f / n;
    for (int i = 0; i < n; i++) {
        std::complex<float> x = invout[i] * norm;
        EXPECT(std::abs(in[i].real() - x.real()) <= eps && std::abs(in[i].imag() - x.imag()) <= eps,
               "InverseIdentity n=%d i=%d", n, i);
    }
    g_pass++;
}

template <typename T>
void TestFFT_Pow2(int n)  // test/fft_test_impl.h:35-58
{
    genfft::FFT<T> fft(n);
    std::vector<std::complex<T>> in(n), out(n), invout(n);
    DummyData(in, false);
    fft.template transform<false>(out.data(), in.data());
    fft.template transform<true>(invout.data(), out.data());
    auto ref = oracle(in.data(), n, false);
    auto ref_inv = oracle(ref.data(), n, true);
    EXPECT(max_abs(out.data(), ref.data(), n) <= FFT_Eps<T>(n), "FFT_Pow2 n=%d fwd abs", n);
    EXPECT(max_abs(invout.data(), ref_inv.data(), n) / n <= FFT_Eps<T>(n), "FFT_Pow2 n=%d inv abs", n);
    EXPECT(rel_l2(out.data(), ref.data(), n) <= RelTol<T>(n), "FFT_Pow2 n=%d relL2 %g", n,
           rel_l2(out.data(), ref.data(), n));
    EXPECT(rel_l2(invout.data(), ref_inv.data(), n) <= RelTol<T>(n), "FFT_Pow2 n=%d inv relL2", n);
    // README usage: forward / inverse aliases
    std::vector<std::complex<T>> o2(n);
    fft.forward(o2.data(), in.data());
    EXPECT(max_abs(o2.data(), out.data(), n) == 0, "forward alias n=%d", n);
    g_pass++;
}

template <typename T>
void TestRealFFT_Pow2(int n, bool half)  // test/fft_test_impl.h:84-106
{
    genfft::RealFFT<T> fft(n);
    std::complex<T> fill = {43, 21};
    std::vector<std::complex<T>> out(n, fill);
    std::vector<T> in(n);
    DummyData(in);
    fft.forward(out.data(), in.data(), half);
    std::vector<std::complex<T>> cin(n);
    for (int i = 0; i < n; i++) cin[i] = in[i];
    auto ref = oracle(cin.data(), n, false);
    const int limit = half ? n / 2 + 1 : n;
    for (int i = 0; i < limit; i++)
        EXPECT(std::abs(out[i].real() - ref[i].real()) <= FFT_Eps<T>(n) &&
                   std::abs(out[i].imag() - ref[i].imag()) <= FFT_Eps<T>(n),
               "RealFFT n=%d half=%d i=%d", n, (int)half, i);
    EXPECT(rel_l2(out.data(), ref.data(), limit) <= RelTol<T>(n), "RealFFT n=%d relL2", n);
    for (int i = limit; i < n; i++) EXPECT(out[i] == fill, "Corruption detected @ index %d (n=%d)", i, n);
    g_pass++;
}

template <typename T>
void TestDIT_Pow2(int n, bool in_place)  // test/fft_test_impl.h:60-82
{
    std::vector<T> real_input(n);
    std::vector<std::complex<T>> complex_input(n), out1(n), out2(n), out3(n);
    DummyData(real_input);
    for (int i = 0; i < n; i++) complex_input[i] = real_input[i];
    genfft::FFT<T> half_fft(n == 1 ? 1 : n / 2);
    genfft::FFT<T> full_fft(n);
    half_fft.template transform<false>(out1.data(), (std::complex<T> *)real_input.data());
    full_fft.template transform<false>(out2.data(), complex_input.data());
    genfft::DIT<T> dit(n);
    std::complex<T> *out_ptr = in_place ? out1.data() : out3.data();
    dit.apply(out_ptr, out1.data(), false);
    for (int i = 0; i < n; i++)
        EXPECT(std::abs(out_ptr[i].real() - out2[i].real()) <= FFT_Eps<T>(n) &&
                   std::abs(out_ptr[i].imag() - out2[i].imag()) <= FFT_Eps<T>(n),
               "DIT n=%d in_place=%d @ %d", n, (int)in_place, i);
    g_pass++;
}

template <typename T>
void TestFFTVert_Pow2(int n, int cols)  // test/fft_test_impl.h:108-131, all columns
{
    genfft::FFTVert<T> fft(n);
    std::vector<std::complex<T>> in((size_t)n * cols), out((size_t)n * cols), invout((size_t)n * cols);
    DummyData(in, false);
    fft.template transform<false>(out.data(), cols, in.data(), cols, cols);
    fft.template transform<true>(invout.data(), cols, out.data(), cols, cols);
    for (int c = 0; c < cols; c++) {
        auto ref = oracle(in.data() + c, n, false, cols);
        EXPECT(rel_l2(out.data() + c, ref.data(), n, cols, 1) <= RelTol<T>(n), "Vert n=%d cols=%d col=%d", n, cols, c);
        double err = 0;
        for (int r = 0; r < n; r++)
            err = std::max(err, std::abs(std::complex<double>(invout[(size_t)r * cols + c]) / (double)n -
                                         std::complex<double>(in[(size_t)r * cols + c])));
        EXPECT(err <= FFT_Eps<T>(n), "Vert inverse n=%d cols=%d col=%d", n, cols, c);
    }
    g_pass++;
}

template <typename T>
void TestFFT2D(int w, int h, int in_stride, int out_stride)
{
    genfft::FFT2D<T> fft(w, h);
    std::vector<std::complex<T>> in((size_t)h * in_stride);
    DummyData(in, false);
    std::complex<T> fill = {7, 5};
    std::vector<std::complex<T>> out((size_t)h * out_stride, fill);
    fft.template transform<false>(out.data(), out_stride, in.data(), in_stride);
    // separable oracle: rows then columns
    std::vector<std::complex<double>> rows((size_t)h * w);
    for (int r = 0; r < h; r++) {
        auto rr = oracle(in.data() + (size_t)r * in_stride, w, false);
        for (int c = 0; c < w; c++) rows[(size_t)r * w + c] = rr[c];
    }
    for (int c = 0; c < w; c++) {
        auto col = oracle(rows.data() + c, h, false, w);
        EXPECT(rel_l2(out.data() + c, col.data(), h, out_stride, 1) <= RelTol<T>(w * h), "FFT2D %dx%d col %d", w, h, c);
    }
    for (int r = 0; r < h; r++)
        for (int c = w; c < out_stride; c++) EXPECT(out[(size_t)r * out_stride + c] == fill, "FFT2D pitch gap written");
    g_pass++;
}

template <typename T>
void TestRealFFT2D(int w, int h)
{
    genfft::RealFFT2D<T> fft(w, h);
    std::vector<T> in((size_t)h * w);
    DummyData(in);
    std::vector<std::complex<T>> out((size_t)h * w), cin((size_t)h * w);
    fft.forward(out.data(), w, in.data(), w);
    for (size_t i = 0; i < in.size(); i++) cin[i] = in[i];
    genfft::FFT2D<T> cfft(w, h);
    std::vector<std::complex<T>> ref((size_t)h * w);
    cfft.template transform<false>(ref.data(), w, cin.data(), w);
    EXPECT(rel_l2(out.data(), ref.data(), w * h) <= 2 * RelTol<T>(w * h), "RealFFT2D %dx%d", w, h);
    // inverse (a TODO in the reference, FFTReal.h:120-130): unscaled, output pitch w + 2 scalars
    std::vector<T> back((size_t)h * (w + 2), T(-3));
    fft.inverse(back.data(), w + 2, out.data(), w);
    double num = 0, den = 0;
    bool gap_ok = true;
    for (int r = 0; r < h; r++) {
        for (int c = 0; c < w; c++) {
            const double d = (double)back[(size_t)r * (w + 2) + c] / ((double)w * h) - (double)in[(size_t)r * w + c];
            num += d * d;
            den += (double)in[(size_t)r * w + c] * in[(size_t)r * w + c];
        }
        gap_ok = gap_ok && back[(size_t)r * (w + 2) + w] == T(-3) && back[(size_t)r * (w + 2) + w + 1] == T(-3);
    }
    EXPECT(std::sqrt(num / den) <= 2 * RelTol<T>(w * h) && gap_ok, "RealFFT2D::inverse %dx%d", w, h);
    g_pass++;
}

// RealFFT2D::forward with an odd scalar input pitch (FFTReal.h:83 takes any
// in_stride) and forward_2x (FFTReal.h:106-118) on host and device data, each
// input at its own pitch, against the 2D transform of in1 + i*in2.
template <typename T>
void TestRealFFT2D_Pitches(int w, int h, int p1, int p2, int po)
{
    genfft::RealFFT2D<T> fft(w, h);
    std::vector<T> a((size_t)h * p1), b((size_t)h * p2);
    DummyData(a);
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<T> dist(-1, 1);
    for (auto &v : b) v = dist(rng);
    std::vector<std::complex<T>> ca((size_t)h * w), cz((size_t)h * w);
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            ca[(size_t)r * w + c] = a[(size_t)r * p1 + c];
            cz[(size_t)r * w + c] = std::complex<T>(a[(size_t)r * p1 + c], b[(size_t)r * p2 + c]);
        }
    genfft::FFT2D<T> cfft(w, h);
    std::vector<std::complex<T>> ra((size_t)h * w), rz((size_t)h * w);
    cfft.template transform<false>(ra.data(), w, ca.data(), w);
    cfft.template transform<false>(rz.data(), w, cz.data(), w);
    const std::complex<T> fill(7, 5);
    std::vector<std::complex<T>> out((size_t)h * po, fill), got((size_t)h * w);
    auto dense = [&](const std::vector<std::complex<T>> &v) {
        for (int r = 0; r < h; r++)
            for (int c = 0; c < w; c++) got[(size_t)r * w + c] = v[(size_t)r * po + c];
        bool gaps = true;
        for (int r = 0; r < h; r++)
            for (int c = w; c < po; c++) gaps = gaps && v[(size_t)r * po + c] == fill;
        return gaps;
    };
    fft.forward(out.data(), po, a.data(), p1);
    EXPECT(dense(out) && rel_l2(got.data(), ra.data(), w * h) <= 2 * RelTol<T>(w * h),
           "RealFFT2D::forward %dx%d in_stride %d", w, h, p1);
    std::fill(out.begin(), out.end(), fill);
    fft.forward_2x(out.data(), po, a.data(), p1, b.data(), p2);
    EXPECT(dense(out) && rel_l2(got.data(), rz.data(), w * h) <= 2 * RelTol<T>(w * h),
           "RealFFT2D::forward_2x host %dx%d pitches %d %d", w, h, p1, p2);
    // device data: one packing launch, then the 2D transform in place
    void *da = nullptr, *db = nullptr, *dz = nullptr;
    EXPECT(hipMalloc(&da, a.size() * sizeof(T)) == hipSuccess && hipMalloc(&db, b.size() * sizeof(T)) == hipSuccess &&
               hipMalloc(&dz, out.size() * sizeof(std::complex<T>)) == hipSuccess, "hipMalloc");
    std::fill(out.begin(), out.end(), fill);
    EXPECT(hipMemcpy(da, a.data(), a.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(db, b.data(), b.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess &&
               hipMemcpy(dz, out.data(), out.size() * sizeof(std::complex<T>), hipMemcpyHostToDevice) == hipSuccess,
           "H2D");
    fft.forward_2x((std::complex<T> *)dz, po, (const T *)da, p1, (const T *)db, p2);
    EXPECT(hipMemcpy(out.data(), dz, out.size() * sizeof(std::complex<T>), hipMemcpyDeviceToHost) == hipSuccess, "D2H");
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dz);
    EXPECT(dense(out) && rel_l2(got.data(), rz.data(), w * h) <= 2 * RelTol<T>(w * h),
           "RealFFT2D::forward_2x device %dx%d pitches %d %d", w, h, p1, p2);
    g_pass++;
}

// The reference's factory template parameters (fft.h:56, :117, :173,
// FFTReal.h:188): code that names a CPU backend compiles and runs on the GPU.
template <typename T>
void TestFactorySpellings(int n)
{
    std::vector<std::complex<T>> in(n), out(n), ref(n);
    DummyData(in, false);
    genfft::FFT<T, genfft::impl_generic::GetImpl> f(n);
    EXPECT((bool)f && f.size() == n, "FFT<T, impl_generic::GetImpl> plan");
    f.template transform<false>(out.data(), in.data());
    genfft::FFT<T> g(n);
    g.template transform<false>(ref.data(), in.data());
    bool same = true;
    for (int i = 0; i < n; i++) same = same && out[i] == ref[i];
    EXPECT(same, "FFT<T, impl_generic::GetImpl> and FFT<T> agree bit for bit, n=%d", n);
    genfft::RealFFT<T, genfft::impl_native::GetImpl, genfft::impl_native::GetDITImpl> rf(n);
    std::vector<T> x(n);
    DummyData(x);
    std::vector<std::complex<T>> h(n / 2 + 1);
    rf.forward(h.data(), x.data(), true);
    std::vector<std::complex<T>> cx(n);
    for (int i = 0; i < n; i++) cx[i] = x[i];
    auto rx = oracle(cx.data(), n, false);
    EXPECT(rel_l2(h.data(), rx.data(), n / 2 + 1) <= RelTol<T>(n), "RealFFT<T, impl_native...> n=%d", n);
    genfft::FFTVert<T, genfft::impl_x86_dispatch::GetVertImpl> v(n);
    genfft::DIT<T, genfft::backend::GetDITImpl> d(n);
    EXPECT((bool)v && (bool)d, "FFTVert / DIT with named factories");
    g_pass++;
}

template <typename T>
void TestRealHelpers(int n)
{
    std::vector<T> x(n), y(n);
    DummyData(x);
    std::mt19937_64 rng(5);
    std::uniform_real_distribution<T> dist(-1, 1);
    for (auto &v : y) v = dist(rng);
    genfft::FFT<T> fft(n);
    std::vector<std::complex<T>> fx(n), z(n), o1(n), o2(n), cx(n), cy(n);
    fft.transform_real(fx.data(), x.data());
    for (int i = 0; i < n; i++) cx[i] = x[i], cy[i] = y[i];
    auto rx = oracle(cx.data(), n, false), ry = oracle(cy.data(), n, false);
    EXPECT(rel_l2(fx.data(), rx.data(), n) <= RelTol<T>(n), "transform_real n=%d", n);
    fft.transform_interleave(z.data(), x.data(), y.data());
    genfft::separate_2x_real_FFT(o1.data(), o2.data(), z.data(), n);
    EXPECT(rel_l2(o1.data(), rx.data(), n) <= 2 * RelTol<T>(n), "separate x n=%d", n);
    EXPECT(rel_l2(o2.data(), ry.data(), n) <= 2 * RelTol<T>(n), "separate y n=%d", n);
    // c2r (not in the reference): inverse(forward(x)) = n*x
    genfft::RealFFT<T> rf(n);
    std::vector<std::complex<T>> h(n / 2 + 1);
    std::vector<T> back(n);
    rf.forward(h.data(), x.data(), true);
    rf.inverse(back.data(), h.data());
    double e = 0, d = 0;
    for (int i = 0; i < n; i++) e += std::pow(back[i] / n - x[i], 2), d += std::pow(x[i], 2);
    EXPECT(std::sqrt(e / d) <= 2 * RelTol<T>(n), "c2r round trip n=%d", n);
    g_pass++;
}

template <typename T>
void TestDevicePointers(int n, int batch)
{
    std::vector<std::complex<T>> in((size_t)n * batch), out((size_t)n * batch);
    DummyData(in, false);
    void *din = nullptr, *dout = nullptr;
    const size_t bytes = in.size() * sizeof(std::complex<T>);
    EXPECT(hipMalloc(&din, bytes) == hipSuccess && hipMalloc(&dout, bytes) == hipSuccess, "hipMalloc");
    EXPECT(hipMemcpy(din, in.data(), bytes, hipMemcpyHostToDevice) == hipSuccess, "H2D");
    genfft::FFT<T> fft(n);
    fft.template transform_many<false>((std::complex<T> *)dout, (const std::complex<T> *)din, batch);
    EXPECT(hipMemcpy(out.data(), dout, bytes, hipMemcpyDeviceToHost) == hipSuccess, "D2H");
    (void)hipFree(din);
    (void)hipFree(dout);
    for (int b : {0, batch - 1}) {
        auto ref = oracle(in.data() + (size_t)b * n, n, false);
        EXPECT(rel_l2(out.data() + (size_t)b * n, ref.data(), n) <= RelTol<T>(n), "device ptr n=%d b=%d", n, b);
    }
    g_pass++;
}

// FFT<T>::transform_no_scramble (fft.h:69-73) on device memory: the input is
// stored in bit-reversed order, the output comes back in natural order.
template <typename T>
void TestNoScrambleDevice(int n)
{
    std::vector<std::complex<T>> in(n), scr(n), out(n);
    DummyData(in, false);
    for (int i = 0; i < n; i++) scr[genfft_amd_bitrev(i, n)] = in[i];
    void *d = nullptr;
    const size_t bytes = (size_t)n * sizeof(std::complex<T>);
    EXPECT(hipMalloc(&d, bytes) == hipSuccess, "hipMalloc");
    EXPECT(hipMemcpy(d, scr.data(), bytes, hipMemcpyHostToDevice) == hipSuccess, "H2D");
    genfft::FFT<T> fft(n);
    fft.template transform_no_scramble<false>((std::complex<T> *)d);
    EXPECT(hipMemcpy(out.data(), d, bytes, hipMemcpyDeviceToHost) == hipSuccess, "D2H");
    (void)hipFree(d);
    auto ref = oracle(in.data(), n, false);
    EXPECT(rel_l2(out.data(), ref.data(), n) <= RelTol<T>(n), "transform_no_scramble device n=%d", n);
    g_pass++;
}

void TestErrors()
{
    bool threw = false;
    try {
        genfft::FFT<float> bad(1000);  // not a power of two
    } catch (const genfft::error &e) {
        threw = e.status == GFFT_ERR_UNSUPPORTED;
    }
    EXPECT(threw, "non-power-of-two size must throw genfft::error(GFFT_ERR_UNSUPPORTED)");
    genfft::FFT<float> empty;
    EXPECT(!empty && empty.size() == 0, "default-constructed plan is empty (fft.h:59, :108)");
    g_pass++;
}

int main(int argc, char **argv)
{
    const bool quick = argc > 1 && std::string(argv[1]) == "quick";
    const int maxk = quick ? 16 : 22;
    for (int k = 1; k <= maxk; k++) InverseIdentity(1 << k);
    for (int k = 1; k <= std::min(maxk, 20); k++) {
        TestFFT_Pow2<float>(1 << k);
        TestFFT_Pow2<double>(1 << k);
    }
    const int vert[][2] = {{1, 1}, {1, 40}, {2, 1}, {2, 7}, {2, 999}, {4, 1}, {4, 7}, {4, 999}, {8, 1}, {8, 3},
                           {8, 7}, {8, 33}, {8, 47}, {8, 1023}, {16, 1}, {16, 47}, {16, 999}, {256, 1}, {256, 1023},
                           {1024, 1}, {1024, 767}, {65536, 1}, {65536, 31}};
    for (auto &v : vert) {
        TestFFTVert_Pow2<float>(v[0], v[1]);
        TestFFTVert_Pow2<double>(v[0], v[1]);
    }
    for (int k = 1; k <= maxk; k++)
        for (bool half : {false, true}) {
            TestRealFFT_Pow2<float>(1 << k, half);
            TestRealFFT_Pow2<double>(1 << k, half);
        }
    for (int k = 1; k <= std::min(maxk, 20); k++)
        for (bool ip : {false, true}) {
            TestDIT_Pow2<float>(1 << k, ip);
            TestDIT_Pow2<double>(1 << k, ip);
        }
    TestFFT2D<float>(32, 16, 33, 35);
    TestFFT2D<double>(32, 16, 33, 35);
    TestFFT2D<float>(8, 64, 8, 8);
    TestFFT2D<float>(256, 4096, 256, 260);
    TestRealFFT2D<float>(32, 16);
    TestRealFFT2D<double>(64, 128);
    TestRealFFT2D_Pitches<float>(32, 16, 33, 35, 34);
    TestRealFFT2D_Pitches<double>(64, 8, 64, 67, 64);
    TestRealFFT2D_Pitches<float>(256, 64, 257, 256, 260);
    TestFactorySpellings<float>(1024);
    TestFactorySpellings<double>(1 << 16);
    for (int n : {2, 16, 1024, 1 << 15}) {
        TestRealHelpers<float>(n);
        TestRealHelpers<double>(n);
    }
    TestDevicePointers<float>(1 << 20, 4);
    TestDevicePointers<double>(4096, 16);
    for (int n : {1, 2, 16, 1024, 1 << 18}) {
        TestNoScrambleDevice<float>(n);
        TestNoScrambleDevice<double>(n);
    }
    TestErrors();
    std::printf("genfft C++ API tests: %d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
}
