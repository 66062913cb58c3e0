/*
 * oracle.c -- CPU oracle for the genfft_amd GPU FFT engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, the smoke check and bench.py's CPU-baseline
 * leg use it, and only as the checker.
 *
 * Two families of functions live here:
 *
 *  1. Exact references (long double, integer argument reduction):
 *       orc_twiddle       -- cos/-sin of 2*pi*m/n with m reduced mod n first
 *       orc_dft           -- naive O(n^2) DFT, exact (j*k mod n) reduction
 *       orc_dft_bins      -- the same DFT evaluated at selected output bins
 *       orc_fft_f64       -- FP64 radix-2 FFT whose twiddles come from
 *                            orc_twiddle (no angle recurrences)
 *       orc_fft_ld        -- the same FFT computed in long double, batched:
 *                            the full-array reference for large N
 *     The reference's own oracle (test/fft_ref_impl.h:37-76, :120-135)
 *     accumulates the angle with `a += da`; these keep its structure
 *     (scramble + radix-2 in place, naive DFT) but compute every twiddle
 *     directly, which is what an FP64/complex128 tolerance of
 *     1e-15*log2(N) needs.
 *
 *  2. A restatement of genFFT's own CPU algorithms in T = float / double
 *     (the "repo's own CPU transform" the GPU results are judged against):
 *       orc_scramble_*        include/genFFT/FFTUtil.h:51-98
 *       orc_c2c_*             include/genFFT/fft.h:80-85 +
 *                             generic/fft_impl_generic.h:36-120
 *       orc_unpack_*          generic/fft_dit_impl_generic.inl:27-61
 *                             (+ DITTwiddle, FFTTwiddle.h:96-128)
 *       orc_real_forward_*    include/genFFT/FFTReal.h:204-213
 *       orc_separate_2x_*     include/genFFT/FFTReal.h:36-66
 *       orc_vert_*            include/genFFT/fft.h:146-150 +
 *                             generic/fft_impl_generic.h:124-243
 *       orc_fft2d_*           include/genFFT/fft.h:214-241
 *       orc_real2d_forward_*  include/genFFT/FFTReal.h:83-104, :144-165
 *     Data are interleaved complex (re, im) arrays; strides are in complex
 *     elements unless stated otherwise, exactly as in the reference API.
 *
 * Every function is pinned by tests/test_oracle.py against golden vectors
 * produced by genFFT itself (tests/golden/) and against closed forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* 1. exact references                                                       */
/* ------------------------------------------------------------------------ */

static const long double ORC_PI =
    3.141592653589793238462643383279502884197169399375105820974944592307816L;

/* exp(-2*pi*i*m/n) in long double, m reduced to [0, n) and folded into the
 * first octant so the argument handed to cosl/sinl is at most pi/4. */
static void twiddle_ld(int64_t m, int64_t n, long double *c, long double *s)
{
    m %= n;
    if (m < 0)
        m += n;
    /* angle = 2*pi*m/n; work with a = 8*m (in units of 2*pi/(8n)) */
    int64_t a = 8 * m;          /* in [0, 8n) */
    int64_t q = a / n;          /* octant 0..7 */
    int64_t r = a - q * n;      /* remainder in [0, n) */
    /* fold: odd octants measure from the upper edge */
    long double x;
    int64_t rr = (q & 1) ? (n - r) : r; /* in [0, n] */
    x = (2.0L * ORC_PI * (long double)rr) / (8.0L * (long double)n);
    long double cx = cosl(x), sx = sinl(x);
    long double C, S; /* cos, sin of the true angle */
    switch (q) {
    case 0: C = cx;  S = sx;  break;
    case 1: C = sx;  S = cx;  break;
    case 2: C = -sx; S = cx;  break;
    case 3: C = -cx; S = sx;  break;
    case 4: C = -cx; S = -sx; break;
    case 5: C = -sx; S = -cx; break;
    case 6: C = sx;  S = -cx; break;
    default: C = cx; S = -sx; break;
    }
    *c = C;
    *s = -S;
}

/* FP64-rounded twiddle {cos(2*pi*m/n), -sin(2*pi*m/n)} (FFTTwiddle.h:48-49
 * convention: forward kernel exp(-2*pi*i*m/n)). */
ORC_API void orc_twiddle(int64_t m, int64_t n, double *re, double *im)
{
    long double c, s;
    twiddle_ld(m, n, &c, &s);
    *re = (double)c;
    *im = (double)s;
}

/* Naive DFT with exact (j*k mod n) reduction and long-double accumulation.
 * Structure follows reference_impl::DFT (test/fft_ref_impl.h:120-135), with
 * the angle recurrence replaced by a directly computed table. */
ORC_API void orc_dft(const double *in, double *out, int64_t n, int inv)
{
    long double *wr = (long double *)malloc(sizeof(long double) * n);
    long double *wi = (long double *)malloc(sizeof(long double) * n);
    for (int64_t m = 0; m < n; m++) {
        twiddle_ld(m, n, &wr[m], &wi[m]);
        if (inv)
            wi[m] = -wi[m];
    }
    for (int64_t k = 0; k < n; k++) {
        long double ar = 0, ai = 0;
        int64_t idx = 0;
        for (int64_t j = 0; j < n; j++) {
            long double xr = in[2 * j], xi = in[2 * j + 1];
            ar += xr * wr[idx] - xi * wi[idx];
            ai += xr * wi[idx] + xi * wr[idx];
            idx += k;
            if (idx >= n)
                idx -= n;
        }
        out[2 * k] = (double)ar;
        out[2 * k + 1] = (double)ai;
    }
    free(wr);
    free(wi);
}

/* The same exact DFT at selected output bins only: O(n) per bin. */
ORC_API void orc_dft_bins(const double *in, int64_t n, int inv,
                          const int64_t *bins, int64_t nbins, double *out)
{
    for (int64_t b = 0; b < nbins; b++) {
        int64_t k = bins[b] % n;
        long double ar = 0, ai = 0;
        int64_t idx = 0;
        for (int64_t j = 0; j < n; j++) {
            long double c, s;
            twiddle_ld(idx, n, &c, &s);
            if (inv)
                s = -s;
            long double xr = in[2 * j], xi = in[2 * j + 1];
            ar += xr * c - xi * s;
            ai += xr * s + xi * c;
            idx += k;
            if (idx >= n)
                idx -= n;
        }
        out[2 * b] = (double)ar;
        out[2 * b + 1] = (double)ai;
    }
}

/* bit reversal of i within log2n bits */
static int64_t bitrev(int64_t i, int log2n)
{
    int64_t r = 0;
    for (int b = 0; b < log2n; b++) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

static int ilog2(int64_t n)
{
    int l = 0;
    while (((int64_t)1 << l) < n)
        l++;
    return l;
}

/* FP64 radix-2 FFT (scramble + in-place DIT, test/fft_ref_impl.h:37-96) with
 * every twiddle taken from the long-double table, so its error stays at a
 * few ulp * log2(n).  `in` and `out` are n interleaved complex doubles. */
ORC_API void orc_fft_f64(const double *in, double *out, int64_t n, int inv)
{
    int lg = ilog2(n);
    double *x = (double *)malloc(sizeof(double) * 2 * n);
    for (int64_t i = 0; i < n; i++) {
        int64_t r = bitrev(i, lg);
        x[2 * r] = in[2 * i];
        x[2 * r + 1] = in[2 * i + 1];
    }
    double *twr = (double *)malloc(sizeof(double) * (n / 2 + 1));
    double *twi = (double *)malloc(sizeof(double) * (n / 2 + 1));
    for (int64_t m = 0; m < n / 2; m++) {
        orc_twiddle(m, n, &twr[m], &twi[m]);
        if (inv)
            twi[m] = -twi[m];
    }
    for (int64_t len = 2; len <= n; len <<= 1) {
        int64_t half = len >> 1, step = n / len;
        for (int64_t base = 0; base < n; base += len) {
            for (int64_t k = 0; k < half; k++) {
                double wr = twr[k * step], wi = twi[k * step];
                double *e = x + 2 * (base + k), *o = x + 2 * (base + k + half);
                double tr = o[0] * wr - o[1] * wi;
                double ti = o[0] * wi + o[1] * wr;
                o[0] = e[0] - tr;
                o[1] = e[1] - ti;
                e[0] += tr;
                e[1] += ti;
            }
        }
    }
    memcpy(out, x, sizeof(double) * 2 * n);
    free(x);
    free(twr);
    free(twi);
}

/* The same radix-2 structure computed entirely in long double (x87, 64-bit
 * mantissa) with long-double twiddles: a full-array reference whose own
 * error (~1e-19 relative) is far below the complex128 tolerance.  Batched:
 * `batch` transforms at distance n. */
ORC_API void orc_fft_ld(const double *in, double *out, int64_t n, int64_t batch, int inv)
{
    int lg = ilog2(n);
    long double *x = (long double *)malloc(sizeof(long double) * 2 * n);
    long double *twr = (long double *)malloc(sizeof(long double) * (n / 2 + 1));
    long double *twi = (long double *)malloc(sizeof(long double) * (n / 2 + 1));
    for (int64_t m = 0; m < n / 2; m++) {
        twiddle_ld(m, n, &twr[m], &twi[m]);
        if (inv)
            twi[m] = -twi[m];
    }
    for (int64_t b = 0; b < batch; b++) {
        const double *src = in + 2 * b * n;
        for (int64_t i = 0; i < n; i++) {
            int64_t r = bitrev(i, lg);
            x[2 * r] = src[2 * i];
            x[2 * r + 1] = src[2 * i + 1];
        }
        for (int64_t len = 2; len <= n; len <<= 1) {
            int64_t half = len >> 1, step = n / len;
            for (int64_t base = 0; base < n; base += len) {
                for (int64_t k = 0; k < half; k++) {
                    long double wr = twr[k * step], wi = twi[k * step];
                    long double *e = x + 2 * (base + k), *o = x + 2 * (base + k + half);
                    long double tr = o[0] * wr - o[1] * wi;
                    long double ti = o[0] * wi + o[1] * wr;
                    o[0] = e[0] - tr;
                    o[1] = e[1] - ti;
                    e[0] += tr;
                    e[1] += ti;
                }
            }
        }
        double *dst = out + 2 * b * n;
        for (int64_t i = 0; i < 2 * n; i++)
            dst[i] = (double)x[i];
    }
    free(x);
    free(twr);
    free(twi);
}

/* ------------------------------------------------------------------------ */
/* 2. restatement of genFFT's CPU algorithms, T = float and T = double        */
/* ------------------------------------------------------------------------ */

#define ORC_DEFINE_GENFFT(T, SFX)                                              \
                                                                               \
/* FFTUtil.h:51-83 -- bit-reversal copy; `stride` is in units of U.         \
 * in_elem / out_elem: scalars per element of the source / destination    \
 * (2 for complex, 1 for real).  Real sources fill the imaginary part    \
 * with zero, as the reference's complex<T> = T conversion does.          */ \
static void scramble_##SFX(T *out, int64_t out_stride, int out_cplx,          \
                           const T *in, int in_cplx, int64_t n)               \
{                                                                              \
    int lg = ilog2(n);                                                         \
    for (int64_t i = 0; i < n; i++) {                                          \
        int64_t r = bitrev(i, lg);                                             \
        T *o = out + (out_cplx ? 2 : 1) * r * out_stride;                      \
        if (in_cplx) {                                                         \
            o[0] = in[2 * i];                                                  \
            if (out_cplx) o[1] = in[2 * i + 1];                                \
        } else {                                                               \
            o[0] = in[i];                                                      \
            if (out_cplx) o[1] = 0;                                            \
        }                                                                      \
    }                                                                          \
}                                                                              \
                                                                               \
/* Twiddle<N,T> (FFTTwiddle.h:38-52): N/2 complex {cos, -sin}(2*pi*k/N),   \
 * computed in double as std::cos(M_PI*i/N) and rounded once to T.         */ \
static void stage_twiddle_##SFX(T *t, int64_t N)                              \
{                                                                              \
    for (int64_t i = 0; i < N; i += 2) {                                       \
        t[i] = (T)cos(M_PI * (double)i / (double)N);                           \
        t[i + 1] = (T)-sin(M_PI * (double)i / (double)N);                      \
    }                                                                          \
}                                                                              \
                                                                               \
/* FFTGeneric<4,T> leaf (fft_impl_generic.h:66-98) on 4 complex values.    */ \
static void leaf4_##SFX(T *d, int inv)                                         \
{                                                                              \
    T tr = d[2], ti = d[3];                                                    \
    d[2] = d[0] - tr; d[3] = d[1] - ti; d[0] += tr; d[1] += ti;                \
    tr = d[6]; ti = d[7];                                                      \
    d[6] = inv ? ti - d[5] : d[5] - ti;                                        \
    d[7] = inv ? d[4] - tr : tr - d[4];                                        \
    d[4] += tr; d[5] += ti;                                                    \
    tr = d[4]; ti = d[5];                                                      \
    d[4] = d[0] - tr; d[5] = d[1] - ti; d[0] += tr; d[1] += ti;                \
    tr = d[6]; ti = d[7];                                                      \
    d[6] = d[2] - tr; d[7] = d[3] - ti; d[2] += tr; d[3] += ti;                \
}                                                                              \
                                                                               \
/* FFTGeneric<N,T>::transform_impl (fft_impl_generic.h:36-63), in place on  \
 * n bit-reversed complex values; recursion unrolled into levels.          */ \
static void dit_inplace_##SFX(T *data, int64_t n, int inv)                     \
{                                                                              \
    if (n <= 1) return;                                                        \
    if (n == 2) {                                                              \
        T tr = data[2], ti = data[3];                                          \
        data[2] = data[0] - tr; data[3] = data[1] - ti;                        \
        data[0] += tr; data[1] += ti;                                          \
        return;                                                                \
    }                                                                          \
    for (int64_t b = 0; b < n; b += 4) leaf4_##SFX(data + 2 * b, inv);         \
    T *tw = (T *)malloc(sizeof(T) * n);                                        \
    for (int64_t N = 8; N <= n; N <<= 1) {                                     \
        /* FFTGeneric<N/2 ... > in the reference's scalar counting: a level  \
         * handling N complex values uses Twiddle<N,T>.                    */ \
        stage_twiddle_##SFX(tw, N);                                            \
        for (int64_t base = 0; base < n; base += N) {                          \
            T *d = data + 2 * base;                                            \
            for (int64_t i = 0; i < N; i += 2) {                               \
                T wr = tw[i], wi = tw[i + 1];                                  \
                T *o = d + N;                                                  \
                T tempr = inv ? o[i] * wr + o[i + 1] * wi                      \
                              : o[i] * wr - o[i + 1] * wi;                     \
                T tempi = inv ? o[i + 1] * wr - o[i] * wi                      \
                              : o[i] * wi + o[i + 1] * wr;                     \
                o[i] = d[i] - tempr;                                           \
                o[i + 1] = d[i + 1] - tempi;                                   \
                d[i] += tempr;                                                 \
                d[i + 1] += tempi;                                             \
            }                                                                  \
        }                                                                      \
    }                                                                          \
    free(tw);                                                                  \
}                                                                              \
                                                                               \
/* FFT<T>::transform<inv>(out, in) (fft.h:80-85): natural-order, unscaled.  */ \
ORC_API void orc_c2c_##SFX(T *out, const T *in, int64_t n, int inv)           \
{                                                                              \
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }                    \
    if (out == in) {                                                           \
        T *tmp = (T *)malloc(sizeof(T) * 2 * n);                               \
        scramble_##SFX(tmp, 1, 1, in, 1, n);                                   \
        dit_inplace_##SFX(tmp, n, inv);                                        \
        memcpy(out, tmp, sizeof(T) * 2 * n);                                   \
        free(tmp);                                                             \
        return;                                                                \
    }                                                                          \
    scramble_##SFX(out, 1, 1, in, 1, n);                                       \
    dit_inplace_##SFX(out, n, inv);                                            \
}                                                                              \
                                                                               \
/* Batched form (batch transforms at distance n, used for baselines).      */ \
ORC_API void orc_c2c_batch_##SFX(T *out, const T *in, int64_t n,             \
                                 int64_t batch, int inv)                       \
{                                                                              \
    for (int64_t b = 0; b < batch; b++)                                        \
        orc_c2c_##SFX(out + 2 * b * n, in + 2 * b * n, n, inv);                \
}                                                                              \
                                                                               \
/* FFT<T>::transform_real (fft.h:90-94): real input, full n-bin output.     */ \
ORC_API void orc_c2c_real_##SFX(T *out, const T *in, int64_t n)              \
{                                                                              \
    scramble_##SFX(out, 1, 1, in, 0, n);                                       \
    dit_inplace_##SFX(out, n, 0);                                              \
}                                                                              \
                                                                               \
/* FFT<T>::transform_interleave (fft.h:100-105): in1 -> re, in2 -> im.      */ \
ORC_API void orc_c2c_interleave_##SFX(T *out, const T *in1, const T *in2,     \
                                      int64_t n)                               \
{                                                                              \
    int lg = ilog2(n);                                                         \
    for (int64_t i = 0; i < n; i++) {                                          \
        int64_t r = bitrev(i, lg);                                             \
        out[2 * r] = in1[i];                                                   \
        out[2 * r + 1] = in2[i];                                               \
    }                                                                          \
    dit_inplace_##SFX(out, n, 0);                                              \
}                                                                              \
                                                                               \
/* adjust_DIT_impl<T> (generic/fft_dit_impl_generic.inl:27-61) with the    \
 * DITTwiddle table t[2k] = sin(2*pi*k/N), t[2k+1] = cos(2*pi*k/N)          \
 * (FFTTwiddle.h:116-127).  Z holds N/2 complex values (may alias F).     */ \
ORC_API void orc_unpack_##SFX(T *F, const T *Z, int64_t N, int half)         \
{                                                                              \
    T zr = Z[0], zi = Z[1];                                                    \
    T zeroval = zr + zi;                                                       \
    if (N == 1) { F[0] = zeroval; F[1] = 0; return; }                          \
    T centerval = zr - zi;                                                     \
    if (N == 2) {                                                              \
        F[0] = zeroval; F[1] = 0; F[2] = centerval; F[3] = 0; return;          \
    }                                                                          \
    int64_t halfN = N >> 1;                                                    \
    T *tw = (T *)malloc(sizeof(T) * (N / 2 + 2));                              \
    for (int64_t i = 0; i <= N / 2; i += 2) {                                  \
        tw[i] = (T)sin(M_PI * (double)i / (double)N);                          \
        tw[i + 1] = (T)cos(M_PI * (double)i / (double)N);                      \
    }                                                                          \
    T qr = Z[2 * (halfN / 2)], qi = -Z[2 * (halfN / 2) + 1];                   \
    for (int64_t i = 1, j = halfN - 1; i < halfN / 2; i++, j--) {              \
        T ar = Z[2 * i], ai = Z[2 * i + 1], br = Z[2 * j], bi = Z[2 * j + 1];  \
        T Er = (ar + br) * (T)0.5, Ei = (ai - bi) * (T)0.5;                    \
        T Or = (ar - br) * (T)0.5, Oi = (ai + bi) * (T)0.5;                    \
        T tr = tw[2 * i], ti = tw[2 * i + 1];                                  \
        T pr = tr * Or - ti * Oi, pi = tr * Oi + ti * Or;                      \
        F[2 * i] = Er - pr; F[2 * i + 1] = Ei - pi;                            \
        F[2 * j] = Er + pr; F[2 * j + 1] = -(Ei + pi);                         \
    }                                                                          \
    F[0] = zeroval; F[1] = 0;                                                  \
    F[2 * (halfN / 2)] = qr; F[2 * (halfN / 2) + 1] = qi;                      \
    F[2 * halfN] = centerval; F[2 * halfN + 1] = 0;                            \
    if (!half)                                                                 \
        for (int64_t i = halfN + 1; i < N; i++) {                              \
            F[2 * i] = F[2 * (N - i)];                                         \
            F[2 * i + 1] = -F[2 * (N - i) + 1];                                \
        }                                                                      \
    free(tw);                                                                  \
}                                                                              \
                                                                               \
/* RealFFT<T>::forward(out, in, half) (FFTReal.h:204-213).                  */ \
ORC_API void orc_real_forward_##SFX(T *out, const T *in, int64_t n, int half) \
{                                                                              \
    if (n == 1) { out[0] = in[0]; out[1] = 0; return; }                        \
    orc_c2c_##SFX(out, in, n / 2, 0);                                          \
    orc_unpack_##SFX(out, out, n, half);                                       \
}                                                                              \
                                                                               \
/* separate_2x_real_FFT (FFTReal.h:36-66); out1/out2/in may alias.          */ \
ORC_API void orc_separate_2x_##SFX(T *out1, T *out2, const T *in, int64_t N)  \
{                                                                              \
    T xr = in[0], yr = in[1];                                                  \
    out1[0] = xr; out1[1] = 0; out2[0] = yr; out2[1] = 0;                      \
    for (int64_t i = 1; i <= N / 2; i++) {                                     \
        int64_t k = N - i;                                                     \
        T a0 = in[2 * i], a1 = in[2 * i + 1], b0 = in[2 * k], b1 = in[2*k+1];  \
        T Xr = (a0 + b0) * (T)0.5f, Xi = (a1 - b1) * (T)0.5f;                  \
        T Yr = (b1 + a1) * (T)0.5f, Yi = (b0 - a0) * (T)0.5f;                  \
        out1[2 * i] = Xr; out1[2 * i + 1] = Xi;                                \
        out2[2 * i] = Yr; out2[2 * i + 1] = Yi;                                \
        out1[2 * k] = Xr; out1[2 * k + 1] = -Xi;                               \
        out2[2 * k] = Yr; out2[2 * k + 1] = -Yi;                               \
    }                                                                          \
}                                                                              \
                                                                               \
/* FFTVert<T>::transform<inv>(out, os, in, is, cols) (fft.h:146-150):        \
 * n-point FFT down each of `cols` columns; strides in complex elements.   */ \
ORC_API void orc_vert_##SFX(T *out, int64_t os, const T *in, int64_t is,      \
                            int64_t n, int64_t cols, int inv)                  \
{                                                                              \
    T *col = (T *)malloc(sizeof(T) * 2 * n);                                   \
    T *res = (T *)malloc(sizeof(T) * 2 * n);                                   \
    for (int64_t c = 0; c < cols; c++) {                                       \
        for (int64_t r = 0; r < n; r++) {                                      \
            col[2 * r] = in[2 * (r * is + c)];                                 \
            col[2 * r + 1] = in[2 * (r * is + c) + 1];                         \
        }                                                                      \
        orc_c2c_##SFX(res, col, n, inv);                                       \
        for (int64_t r = 0; r < n; r++) {                                      \
            out[2 * (r * os + c)] = res[2 * r];                                \
            out[2 * (r * os + c) + 1] = res[2 * r + 1];                        \
        }                                                                      \
    }                                                                          \
    free(col);                                                                 \
    free(res);                                                                 \
}                                                                              \
                                                                               \
/* FFT2D<T>(width, height).transform<inv>(out, os, in, is) (fft.h:214-241):  \
 * row FFTs then column FFTs; natural order, unscaled.                     */ \
ORC_API void orc_fft2d_##SFX(T *out, int64_t os, const T *in, int64_t is,     \
                             int64_t width, int64_t height, int inv)           \
{                                                                              \
    for (int64_t r = 0; r < height; r++)                                       \
        orc_c2c_##SFX(out + 2 * r * os, in + 2 * r * is, width, inv);          \
    orc_vert_##SFX(out, os, out, os, height, width, inv);                      \
}                                                                              \
                                                                               \
/* RealFFT2D<T>::forward(out, os, in, is) (FFTReal.h:83-104): real rows ->  \
 * full complex spectrum; os in complex elements, is in scalars.  Rows are  \
 * transformed two at a time as one complex FFT (row 2q in the real part,  \
 * row 2q+1 in the imaginary part) and split with separate_2x, as          \
 * scramble_row_fwd does (FFTReal.h:144-165); a single row (height 1) is  \
 * a plain real-input FFT.                                                 */ \
ORC_API void orc_real2d_forward_##SFX(T *out, int64_t os, const T *in,        \
                                      int64_t is, int64_t width,               \
                                      int64_t height)                          \
{                                                                              \
    const int64_t M = width, N = height;                                       \
    if (N == 1) {                                                              \
        orc_c2c_real_##SFX(out, in, M);                                        \
    } else {                                                                   \
        T *row = (T *)malloc(sizeof(T) * 2 * M);                               \
        for (int64_t q = 0; q < N / 2; q++) {                                  \
            orc_c2c_interleave_##SFX(row, in + (2 * q) * is,                   \
                                     in + (2 * q + 1) * is, M);                \
            orc_separate_2x_##SFX(out + 2 * (2 * q) * os,                      \
                                  out + 2 * (2 * q + 1) * os, row, M);         \
        }                                                                      \
        free(row);                                                             \
    }                                                                          \
    const int64_t vcols = M / 2 + 1;                                           \
    orc_vert_##SFX(out, os, out, os, N, vcols < M ? vcols : M, 0);             \
    if (vcols < M)                                                             \
        for (int64_t i = 0; i < N; i++) {                                      \
            int64_t k = i ? N - i : 0;                                         \
            const T *irow = out + 2 * i * os;                                  \
            T *orow = out + 2 * k * os;                                        \
            for (int64_t j = M - 1; j >= vcols; j--) {                         \
                orow[2 * j] = irow[2 * (M - j)];                               \
                orow[2 * j + 1] = -irow[2 * (M - j) + 1];                      \
            }                                                                  \
        }                                                                      \
}

ORC_DEFINE_GENFFT(float, f32)
ORC_DEFINE_GENFFT(double, f64)

/* Multi-threaded batch driver for the CPU baseline ("port" kind): the batch
 * is strided across nthreads POSIX threads sharing nothing but the input. */
#include <pthread.h>
typedef struct {
    const void *in;
    void *out;
    int64_t n, batch;
    int inv, dbl, tid, nthreads;
} orc_job;

static void *orc_worker(void *p)
{
    orc_job *j = (orc_job *)p;
    for (int64_t b = j->tid; b < j->batch; b += j->nthreads) {
        if (j->dbl)
            orc_c2c_f64((double *)j->out + 2 * b * j->n,
                        (const double *)j->in + 2 * b * j->n, j->n, j->inv);
        else
            orc_c2c_f32((float *)j->out + 2 * b * j->n,
                        (const float *)j->in + 2 * b * j->n, j->n, j->inv);
    }
    return NULL;
}

ORC_API int orc_c2c_batch_mt(void *out, const void *in, int64_t n, int64_t batch,
                             int inv, int dbl, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
    pthread_t th[256];
    orc_job jobs[256];
    if (nthreads > 256)
        nthreads = 256;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (orc_job){in, out, n, batch, inv, dbl, t, nthreads};
        if (pthread_create(&th[t], NULL, orc_worker, &jobs[t]) != 0)
            return -1;
    }
    for (int t = 0; t < nthreads; t++)
        pthread_join(th[t], NULL);
    return 0;
}
