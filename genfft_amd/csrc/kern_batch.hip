// kern_batch.hip -- persistent mixed-size FFT kernel with a device work queue
// (see gfft_batch.h).  Every length 2 .. MIX_TOT is a sequence-major LDS tile
// of M = min(64, TOT/L) transforms, run with the same Stockham passes as the
// batched tile kernels; length 1 is a copy.
#include "gfft_batch.h"

namespace gfft {
namespace {

template <int... RS>
struct Radices {};

template <typename T, int L>
struct MixShape;
// radix sequences per length (same as the batched row tiles)
template <typename T> struct MixShape<T, 2> { using R = Radices<2>; };
template <typename T> struct MixShape<T, 4> { using R = Radices<4>; };
template <typename T> struct MixShape<T, 8> { using R = Radices<8>; };
template <typename T> struct MixShape<T, 16> { using R = Radices<16>; };
template <typename T> struct MixShape<T, 32> { using R = Radices<8, 4>; };
template <typename T> struct MixShape<T, 64> { using R = Radices<8, 8>; };
template <typename T> struct MixShape<T, 128> { using R = Radices<16, 8>; };
template <typename T> struct MixShape<T, 256> { using R = Radices<16, 16>; };
template <typename T> struct MixShape<T, 512> { using R = Radices<8, 8, 8>; };
template <typename T> struct MixShape<T, 1024> { using R = Radices<16, 16, 4>; };
template <typename T> struct MixShape<T, 2048> { using R = Radices<16, 16, 8>; };
template <typename T> struct MixShape<T, 4096> { using R = Radices<16, 16, 16>; };

template <typename T, int L, int... RS>
__device__ __forceinline__ void mix_tile(cx<T> *lds, const MixArgs &a, const MixTask &t, const int64_t *jin,
                                         const int64_t *jout, const uint8_t *jinv, int tid, Radices<RS...>)
{
    using C = MixCfg<T>;
    constexpr int M = mix_m(C::TOT, L);
    constexpr int TOT = M * L;
    constexpr int NT = C::NT;
    using G = TileGeom<L, M, LAY_SEQ, PAT_ROW, PAT_ROW>;
    const cx<T> *in = reinterpret_cast<const cx<T> *>(a.in);
    cx<T> *out = reinterpret_cast<cx<T> *>(a.out);
    const int nv = t.count;
    if (sizeof(T) == 4 && a.vec && L >= 2) {
#pragma unroll 4
        for (int it = 0; it < (TOT / 2 + NT - 1) / NT; it++) {
            const int e = (tid + it * NT) * 2;
            const int s = e / L, n = e % L;
            if (e < TOT && s < nv) {
                const float4 q = *reinterpret_cast<const float4 *>(in + jin[s] + n);
                cx<T> x0 = {(T)q.x, (T)q.y}, x1 = {(T)q.z, (T)q.w};
                if (jinv[s]) {
                    x0 = swap_ri(x0);
                    x1 = swap_ri(x1);
                }
                lds[G::off(s, n)] = x0;
                lds[G::off(s, n + 1)] = x1;
            }
        }
    } else {
#pragma unroll 4
        for (int it = 0; it < (TOT + NT - 1) / NT; it++) {
            const int e = tid + it * NT;
            const int s = e / L, n = e % L;
            if (e < TOT && s < nv) {
                const cx<T> x = in[jin[s] + n];
                lds[G::off(s, n)] = jinv[s] ? swap_ri(x) : x;
            }
        }
    }
    __syncthreads();
    const TileArgs none{};  // no four-step twiddle
    stockham_passes<T, L, M, NT, G, true, false, 1, RS...>(
        lds, reinterpret_cast<const cx<T> *>(a.pass_tw[ilog2c(L)]), tid, none, 0, 0, Tw4Pre{});
    if (sizeof(T) == 4 && a.vec && L >= 2) {
#pragma unroll 4
        for (int it = 0; it < (TOT / 2 + NT - 1) / NT; it++) {
            const int e = (tid + it * NT) * 2;
            const int s = e / L, k = e % L;
            if (e < TOT && s < nv) {
                cx<T> x0 = lds[G::off(s, k)], x1 = lds[G::off(s, k + 1)];
                if (jinv[s]) {
                    x0 = swap_ri(x0);
                    x1 = swap_ri(x1);
                }
                *reinterpret_cast<float4 *>(out + jout[s] + k) = float4{(float)x0.re, (float)x0.im, (float)x1.re,
                                                                         (float)x1.im};
            }
        }
    } else {
#pragma unroll 4
        for (int it = 0; it < (TOT + NT - 1) / NT; it++) {
            const int e = tid + it * NT;
            const int s = e / L, k = e % L;
            if (e < TOT && s < nv) {
                const cx<T> x = lds[G::off(s, k)];
                out[jout[s] + k] = jinv[s] ? swap_ri(x) : x;
            }
        }
    }
}

template <typename T, int L>
__device__ __forceinline__ void mix_dispatch_one(cx<T> *lds, const MixArgs &a, const MixTask &t, const int64_t *jin,
                                                 const int64_t *jout, const uint8_t *jinv, int tid)
{
    mix_tile<T, L>(lds, a, t, jin, jout, jinv, tid, typename MixShape<T, L>::R{});
}

template <typename T>
constexpr int mix_lds_elems()
{
    using C = MixCfg<T>;
    int m = 0;
    for (int lg = 1; lg <= C::MAXLG; lg++) {
        const int L = 1 << lg, M = mix_m(C::TOT, L);
        const int LP = L + (L >> 4);
        const int e = M * (LP | 1);
        m = e > m ? e : m;
    }
    return m;
}

template <typename T>
__global__ void __launch_bounds__(256) mix_fft_kernel(MixArgs a)
{
    using C = MixCfg<T>;
    __shared__ cx<T> lds[mix_lds_elems<T>()];
    __shared__ int64_t jin[C::MAXM], jout[C::MAXM];
    __shared__ uint8_t jinv[C::MAXM];
    __shared__ int task_id;
    for (;;) {
        // thread index re-materialised per task: otherwise every tid-derived
        // LDS offset of all twelve tile shapes is hoisted and kept live
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (tid == 0) task_id = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int ti = task_id;
        if (ti >= a.n_tasks) break;  // every wave of the block takes this exit once the queue is drained
        const MixTask t = a.tasks[ti];
        if (tid < t.count) {
            jin[tid] = a.job_in[t.first + tid];
            jout[tid] = a.job_out[t.first + tid];
            jinv[tid] = a.job_inv[t.first + tid];
        }
        __syncthreads();
        switch (t.lg) {
        case 0:  // length 1: the transform is the identity
            if (tid < t.count) reinterpret_cast<cx<T> *>(a.out)[jout[tid]] = reinterpret_cast<const cx<T> *>(a.in)[jin[tid]];
            break;
        case 1: mix_dispatch_one<T, 2>(lds, a, t, jin, jout, jinv, tid); break;
        case 2: mix_dispatch_one<T, 4>(lds, a, t, jin, jout, jinv, tid); break;
        case 3: mix_dispatch_one<T, 8>(lds, a, t, jin, jout, jinv, tid); break;
        case 4: mix_dispatch_one<T, 16>(lds, a, t, jin, jout, jinv, tid); break;
        case 5: mix_dispatch_one<T, 32>(lds, a, t, jin, jout, jinv, tid); break;
        case 6: mix_dispatch_one<T, 64>(lds, a, t, jin, jout, jinv, tid); break;
        case 7: mix_dispatch_one<T, 128>(lds, a, t, jin, jout, jinv, tid); break;
        case 8: mix_dispatch_one<T, 256>(lds, a, t, jin, jout, jinv, tid); break;
        case 9: mix_dispatch_one<T, 512>(lds, a, t, jin, jout, jinv, tid); break;
        case 10: mix_dispatch_one<T, 1024>(lds, a, t, jin, jout, jinv, tid); break;
        case 11: mix_dispatch_one<T, 2048>(lds, a, t, jin, jout, jinv, tid); break;
        case 12:
            if constexpr (C::MAXLG >= 12) mix_dispatch_one<T, 4096>(lds, a, t, jin, jout, jinv, tid);
            break;
        default: break;
        }
        __syncthreads();  // LDS and job slots are reused by the next task
    }
}

__global__ void mix_reset_kernel(uint32_t *counter) { *counter = 0u; }

}  // namespace

hipError_t launch_mix(int prec, const MixArgs &a, int grid, hipStream_t s)
{
    if (prec) hipLaunchKernelGGL(mix_fft_kernel<double>, dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(mix_fft_kernel<float>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mix_reset(uint32_t *counter, hipStream_t s)
{
    hipLaunchKernelGGL(mix_reset_kernel, dim3(1), dim3(1), 0, s, counter);
    return hipGetLastError();
}

int mix_grid(int prec, int num_cus)
{
    int per_cu = 1;
    const void *fn = prec ? (const void *)mix_fft_kernel<double> : (const void *)mix_fft_kernel<float>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return per_cu * num_cus;
}

int mix_radices(int prec, int lg, int *rad)
{
    static const int tab[13][4] = {{0}, {2}, {4}, {8}, {16}, {8, 4}, {8, 8}, {16, 8}, {16, 16}, {8, 8, 8},
                                   {16, 16, 4}, {16, 16, 8}, {16, 16, 16}};
    if (lg < 1 || lg > (prec ? 11 : 12)) return 0;
    int n = 0;
    while (n < 4 && tab[lg][n]) {
        rad[n] = tab[lg][n];
        n++;
    }
    return n;
}

}  // namespace gfft
