// kern_misc.hip -- element-wise kernels around the Stockham tile passes:
//
//   unpack      real-FFT post-pass, the GPU form of adjust_DIT_impl
//               (generic/fft_dit_impl_generic.inl:27-61): F from the n/2-point
//               FFT Z of the real signal packed as complex.
//   pack        its inverse (half spectrum F -> Z), the c2r pre-pass; the
//               reference has no c2r (include/genFFT/FFTReal.h:127-130).
//   separate    separate_2x_real_FFT (include/genFFT/FFTReal.h:36-66).
//   interleave  transform_interleave's packing (include/genFFT/fft.h:100-105).
//   hfill       RealFFT2D's Hermitian column fill (include/genFFT/FFTReal.h:90-103).
//   copy        the n = 1 transform (identity on each element).
//   repack      rows of 4- or 8-byte elements between two pitches (odd
//               real-side layouts to and from a dense copy in scratch; the
//               length-1 column transform).
//   amax        per-transform power-of-two scale of 16-bit storage input.
//   bitrev      bit-reversal gather of rows (transform_no_scramble input).
//
// All of them are grid-stride loops over (pair, batch) with 64-bit indices
// and read each input element once.
#include "gfft_misc.h"

namespace gfft {

namespace {

// W_N^m (forward sign) from the two-level FP64 table, then i*W = (sin, cos).
__device__ __forceinline__ cx<double> twl(const TwTable &t, uint64_t m)
{
    m &= t.mask;
    const cx<double> *lo = reinterpret_cast<const cx<double> *>(t.lo);
    const cx<double> *hi = reinterpret_cast<const cx<double> *>(t.hi);
    return cmul(hi[m >> t.lo_bits], lo[m & ((1ull << t.lo_bits) - 1ull)]);
}

template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }

template <typename T>
__device__ __forceinline__ cx<T> scale(cx<T> a, T s) { return {a.re * s, a.im * s}; }

template <typename T>
__global__ void __launch_bounds__(256) unpack_kernel(UnpackArgs a)
{
    using C = cx<T>;
    const C *Z = reinterpret_cast<const C *>(a.in);
    C *F = reinterpret_cast<C *>(a.out);
    const int64_t N = a.n, halfN = N >> 1, quarter = N >> 2;
    const int64_t pairs = quarter + 1;  // p = 0 .. N/4
    const int64_t total = pairs * a.batch;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = g / pairs, p = g - b * pairs;
        const C *z = Z + b * a.in_dist;
        C *f = F + b * a.out_dist;
        if (p == 0) {
            const C z0 = z[0];
            f[0] = {z0.re + z0.im, (T)0};
            if (N >= 2) f[halfN] = {z0.re - z0.im, (T)0};
            continue;
        }
        if (p == quarter) {  // N >= 4: F[N/4] = conj(Z[N/4])
            const C zq = z[quarter];
            f[quarter] = conj(zq);
            if (!a.half) f[N - quarter] = zq;
            continue;
        }
        const int64_t i = p, j = halfN - p;
        const C zi = z[i], zj = z[j];
        const C E = scale(zi + conj(zj), (T)0.5);
        const C O = scale(zi - conj(zj), (T)0.5);
        const cx<double> w = twl(a.tw, (uint64_t)i);
        const C t = {(T)(-w.im), (T)w.re};  // i * W_N^i = (sin, cos)
        const C to = cmul(t, O);
        const C fi = E - to;
        const C fj = conj(E + to);
        f[i] = fi;
        f[j] = fj;
        if (!a.half) {
            f[N - i] = conj(fi);
            f[N - j] = conj(fj);
        }
    }
}

// Z[k] = (F[k] + conj F[L-k]) + i W_N^-k (F[k] - conj F[L-k]),  L = N/2.
// The unscaled inverse L-point FFT of Z, read as N reals, is N * x.  As in
// any c2r, the imaginary parts of bins 0 and L are ignored.
template <typename T>
__global__ void __launch_bounds__(256) pack_kernel(UnpackArgs a)
{
    using C = cx<T>;
    const C *F = reinterpret_cast<const C *>(a.in);
    C *Z = reinterpret_cast<C *>(a.out);
    const int64_t N = a.n, L = N >> 1;
    const int64_t total = L * a.batch;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = g / L, k = g - b * L;
        const C *f = F + b * a.in_dist;
        C fk = f[k], fm = f[L - k];
        if (k == 0) {  // bins 0 and L of a Hermitian spectrum are real: ignore their imaginary parts
            fk.im = (T)0;
            fm.im = (T)0;
        }
        const C s = fk + conj(fm);
        const C d = fk - conj(fm);
        const cx<double> w = twl(a.tw, (uint64_t)k);
        const C wc = {(T)w.re, (T)(-w.im)};  // W_N^-k
        const C q = cmul(wc, d);
        Z[b * a.out_dist + k] = {s.re - q.im, s.im + q.re};  // s + i q
    }
}

template <typename T>
__global__ void __launch_bounds__(256) separate_kernel(SeparateArgs a)
{
    using C = cx<T>;
    const C *Z = reinterpret_cast<const C *>(a.in);
    C *X = reinterpret_cast<C *>(a.out1);
    C *Y = reinterpret_cast<C *>(a.out2);
    const int64_t N = a.n, pairs = N / 2 + 1;
    const int64_t total = pairs * a.batch;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = g / pairs, i = g - b * pairs;
        const C *z = Z + b * a.dist;
        C *x = X + b * a.dist;
        C *y = Y + b * a.dist;
        if (i == 0) {
            const C z0 = z[0];
            x[0] = {z0.re, (T)0};
            y[0] = {z0.im, (T)0};
            continue;
        }
        const int64_t k = N - i;
        const C zi = z[i], zk = z[k];
        const C xv = {(zi.re + zk.re) * (T)0.5, (zi.im - zk.im) * (T)0.5};
        const C yv = {(zk.im + zi.im) * (T)0.5, (zk.re - zi.re) * (T)0.5};
        x[i] = xv;
        y[i] = yv;
        x[k] = conj(xv);
        y[k] = conj(yv);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) interleave_kernel(InterleaveArgs a)
{
    using C = cx<T>;
    const T *x = reinterpret_cast<const T *>(a.in1);
    const T *y = reinterpret_cast<const T *>(a.in2);
    C *z = reinterpret_cast<C *>(a.out);
    const int64_t total = a.n * a.batch;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = g / a.n, c = g - r * a.n;
        z[r * a.out_pitch + c] = {x[r * a.in1_pitch + c], y[r * a.in2_pitch + c]};
    }
}

// out[k][j] = conj(out[i][M - j]) for j in [vcols, M), k = (N - i) mod N.
template <typename T>
__global__ void __launch_bounds__(256) hfill_kernel(HFillArgs a)
{
    using C = cx<T>;
    C *out = reinterpret_cast<C *>(a.out);
    const int64_t span = a.M - a.vcols;
    const int64_t per = span * a.N;
    const int64_t total = per * a.batch;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = g / per, r = g - b * per;
        const int64_t i = r / span, j = a.vcols + (r - i * span);
        const int64_t k = i ? a.N - i : 0;
        C *base = out + b * a.dist;
        base[k * a.stride + j] = conj(base[i * a.stride + (a.M - j)]);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) copy_kernel(CopyArgs a)
{
    using C = cx<T>;
    const C *in = reinterpret_cast<const C *>(a.in);
    C *out = reinterpret_cast<C *>(a.out);
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < a.batch;
         g += (int64_t)gridDim.x * blockDim.x) {
        if (a.real_in) {
            const T *r = reinterpret_cast<const T *>(a.in);
            out[g * a.out_dist] = {r[g * a.in_dist], (T)0};
        } else if (a.real_out) {
            T *r = reinterpret_cast<T *>(a.out);
            r[g * a.out_dist] = in[g * a.in_dist].re;
        } else {
            out[g * a.out_dist] = in[g * a.in_dist];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) repack_kernel(RepackArgs a)
{
    const T *in = reinterpret_cast<const T *>(a.in);
    T *out = reinterpret_cast<T *>(a.out);
    const int64_t per = a.width * a.rows, total = per * a.batch;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = g / per, rem = g - b * per;
        const int64_t r = rem / a.width, c = rem - r * a.width;
        out[b * a.out_dist + r * a.out_pitch + c] = in[b * a.in_dist + r * a.in_pitch + c];
    }
}

// ---- per-transform scale of 16-bit storage transforms -------------------
__device__ __forceinline__ float scale_from_max(float m, float gain, float limit)
{
    const float q = m * gain / limit;  // limit is a power of two: exact division
    if (!(q > 1.0f) || !(q <= 3.0e38f)) return 1.0f;  // fits already, or not finite
    const uint32_t u = __float_as_uint(q);
    const int k = (int)((u >> 23) & 255u) - 127 + ((u & 0x7fffffu) ? 1 : 0);  // ceil(log2 q), q normal
    return __uint_as_float((uint32_t)(127 - k) << 23);
}

template <class S>
__device__ __forceinline__ float l1_of(uint32_t w)
{
    const cx<float> v = hdec<S>(w);
    return fabsf(v.re) + fabsf(v.im);
}

// small transforms (rows * width <= 1024): one thread per transform writes
// its scale directly
template <class S>
__global__ void __launch_bounds__(256) amax_small_kernel(AmaxArgs a)
{
    const uint32_t *in = reinterpret_cast<const uint32_t *>(a.in);
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < a.batch;
         b += (int64_t)gridDim.x * blockDim.x) {
        float m = 0.0f;
        for (int64_t r = 0; r < a.rows; r++)
            for (int64_t w = 0; w < a.width; w++) m = fmaxf(m, l1_of<S>(in[b * a.dist + r * a.pitch + w]));
        reinterpret_cast<float *>(a.bits)[b] = scale_from_max(m, a.gain, a.limit);
    }
}

__global__ void __launch_bounds__(256) amax_reset_kernel(AmaxArgs a)
{
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < a.batch;
         b += (int64_t)gridDim.x * blockDim.x)
        a.bits[b] = 0u;
}

// block = `chunk` elements of one transform; max over the block, then one
// atomic max per block on the float bits (non-negative floats order like
// their bit patterns, so the result does not depend on the order)
template <class S>
__global__ void __launch_bounds__(256) amax_kernel(AmaxArgs a)
{
    __shared__ float red[4];
    const uint32_t *in = reinterpret_cast<const uint32_t *>(a.in);
    const int64_t b = blockIdx.x / a.chunks, c = blockIdx.x - b * a.chunks;
    const int64_t E = a.rows * a.width;
    const int64_t e0 = c * a.chunk, e1 = e0 + a.chunk < E ? e0 + a.chunk : E;
    const uint32_t *base = in + b * a.dist;
    float m = 0.0f;
    if (a.vec && (a.rows == 1 || a.pitch == a.width)) {
        // a contiguous transform: 16-byte loads straight from the chunk, no
        // per-load 64-bit row division, several loads in flight per thread
        const uint4 *q = reinterpret_cast<const uint4 *>(base + e0);
        const int nq = (int)((e1 - e0) >> 2);
#pragma unroll 8
        for (int i = threadIdx.x; i < nq; i += 256) {
            const uint4 v = q[i];
            m = fmaxf(m, fmaxf(fmaxf(l1_of<S>(v.x), l1_of<S>(v.y)), fmaxf(l1_of<S>(v.z), l1_of<S>(v.w))));
        }
    } else if (a.vec) {  // width, pitch, dist multiples of 4 and 16-byte aligned input
        for (int64_t e = e0 + 4 * (int64_t)threadIdx.x; e < e1; e += 4 * 256) {
            const int64_t r = e / a.width, w = e - r * a.width;
            const uint4 q = *reinterpret_cast<const uint4 *>(base + r * a.pitch + w);
            m = fmaxf(m, fmaxf(fmaxf(l1_of<S>(q.x), l1_of<S>(q.y)), fmaxf(l1_of<S>(q.z), l1_of<S>(q.w))));
        }
    } else {
        for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
            const int64_t r = e / a.width, w = e - r * a.width;
            m = fmaxf(m, l1_of<S>(base[r * a.pitch + w]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        // the running max only grows: a block whose max is not above a value
        // already seen (possibly stale, never above the true running max)
        // cannot change it and skips the atomic, which the blocks of one
        // transform would otherwise queue on one cache line
        const uint32_t mb = __float_as_uint(m);
        if (mb > __hip_atomic_load(a.bits + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
            atomicMax(a.bits + b, mb);
    }
}

__global__ void __launch_bounds__(256) amax_finish_kernel(AmaxArgs a)
{
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < a.batch;
         b += (int64_t)gridDim.x * blockDim.x)
        reinterpret_cast<float *>(a.bits)[b] = scale_from_max(__uint_as_float(a.bits[b]), a.gain, a.limit);
}

template <typename T>
__global__ void __launch_bounds__(256) bitrev_kernel(BitrevArgs a)
{
    using C = cx<T>;
    const C *in = reinterpret_cast<const C *>(a.in);
    C *out = reinterpret_cast<C *>(a.out);
    const int64_t per = a.n * a.width, total = per * a.batch;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = g / per, rem = g - b * per;
        const int64_t r = rem / a.width, c = rem - r * a.width;
        const int64_t rr = a.lg ? (int64_t)(__brev((uint32_t)r) >> (32 - a.lg)) : 0;
        out[b * a.dist + rr * a.stride + c] = in[b * a.dist + r * a.stride + c];
    }
}

unsigned grid_for(int64_t work)
{
    int64_t blocks = (work + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 65536) blocks = 65536;
    return (unsigned)blocks;
}

}  // namespace

hipError_t launch_amax_scale(AmaxArgs a, hipStream_t s)
{
    const int64_t E = a.rows * a.width;
    const bool f16 = a.store == 1;
    if (E <= 1024) {
        const unsigned g = grid_for(a.batch);
        if (f16) hipLaunchKernelGGL(amax_small_kernel<st_f16>, dim3(g), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(amax_small_kernel<st_bf16>, dim3(g), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    // 64 Ki elements (256 KiB) per block: with 8 Ki, FP16 2^20 x 64 with the
    // automatic scale took 0.450 ms and one 8192 x 8192 image 0.551 ms; with
    // 64 Ki and the skipped atomics (amax_kernel) 0.450 and 0.517 ms (MI355X)
    a.chunk = 65536;
    a.chunks = (E + a.chunk - 1) / a.chunk;
    if (a.batch * a.chunks > 0x7fffffffll) return hipErrorInvalidValue;
    a.vec = (a.width % 4 == 0) && (a.pitch % 4 == 0) && (a.dist % 4 == 0) && (((uintptr_t)a.in & 15) == 0);
    const unsigned g = grid_for(a.batch);
    hipLaunchKernelGGL(amax_reset_kernel, dim3(g), dim3(256), 0, s, a);
    if (f16) hipLaunchKernelGGL(amax_kernel<st_f16>, dim3((unsigned)(a.batch * a.chunks)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(amax_kernel<st_bf16>, dim3((unsigned)(a.batch * a.chunks)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(amax_finish_kernel, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bitrev(int prec, const BitrevArgs &a, hipStream_t s)
{
    const unsigned g = grid_for(a.n * a.width * a.batch);
    if (prec) hipLaunchKernelGGL(bitrev_kernel<double>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(bitrev_kernel<float>, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_unpack(int prec, const UnpackArgs &a, hipStream_t s, hipEvent_t stop)
{
    const unsigned g = grid_for(((a.n >> 2) + 1) * a.batch);
    return launch_stop(prec ? &unpack_kernel<double> : &unpack_kernel<float>, dim3(g), dim3(256), s, a, stop);
}

hipError_t launch_pack(int prec, const UnpackArgs &a, hipStream_t s)
{
    const unsigned g = grid_for((a.n >> 1) * a.batch);
    if (prec) hipLaunchKernelGGL(pack_kernel<double>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(pack_kernel<float>, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_separate(int prec, const SeparateArgs &a, hipStream_t s)
{
    const unsigned g = grid_for((a.n / 2 + 1) * a.batch);
    if (prec) hipLaunchKernelGGL(separate_kernel<double>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(separate_kernel<float>, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_interleave(int prec, const InterleaveArgs &a, hipStream_t s)
{
    const unsigned g = grid_for(a.n * a.batch);
    if (prec) hipLaunchKernelGGL(interleave_kernel<double>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(interleave_kernel<float>, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hfill(int prec, const HFillArgs &a, hipStream_t s, hipEvent_t stop)
{
    const unsigned g = grid_for((a.M - a.vcols) * a.N * a.batch);
    return launch_stop(prec ? &hfill_kernel<double> : &hfill_kernel<float>, dim3(g), dim3(256), s, a, stop);
}

hipError_t launch_repack(const RepackArgs &a, hipStream_t s)
{
    const unsigned g = grid_for(a.width * a.rows * a.batch);
    if (a.elem_bytes == 8) hipLaunchKernelGGL(repack_kernel<double>, dim3(g), dim3(256), 0, s, a);
    else if (a.elem_bytes == 4) hipLaunchKernelGGL(repack_kernel<float>, dim3(g), dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_copy(int prec, const CopyArgs &a, hipStream_t s)
{
    const unsigned g = grid_for(a.batch);
    if (prec) hipLaunchKernelGGL(copy_kernel<double>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(copy_kernel<float>, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gfft
