// pattern_copy.hip -- a measuring probe, not part of the FFT library: copies
// complex64 matrices tile by tile in the HBM access pattern of a four-step
// tile pass, with the same persistent, register-pipelined loop as the FFT
// tile kernels (the next tile is loaded while the current one is stored) and
// no arithmetic.  Its bandwidth is the ceiling of a tile pass with that
// access pattern, so a pass can be judged against what its own pattern
// allows as well as against the HBM peak.
//
// Matrices are n1 rows x n2 columns (row-major, element (r, c) at r*n2 + c);
// one tile is 16384 elements and one access moves 16 bytes (two elements):
//   mode 0  column tile: m columns x n1 rows read and written in place
//           (the column pass; n1 * m = 16384)
//   mode 1  m whole rows read, written transposed into columns m*j .. m*j+m-1
//           of the n2 x n1 result (the transposing row pass; m * n2 = 16384)
//   mode 2  contiguous 16384-element tiles read and written
// Tiles are visited interleaved (workgroup w takes w, w + grid, ...; order 0)
// or in one contiguous run per workgroup (order 1), the two orders of the
// FFT tile kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NT = 1024;
constexpr int TILE = 16384;
constexpr int IT = TILE / 2 / NT;

struct PcArgs {
    const f4 *in;
    f4 *out;
    int32_t ntiles;
    int32_t mode;
    int32_t lg_n1, lg_n2, lg_m;  // log2 of rows, columns, tile width
    int32_t order;               // 0 interleaved, 1 one contiguous run per workgroup
};

// element offset of access e (even) of tile t; `out` selects the write side
template <int MODE>
__device__ __forceinline__ int64_t pc_off(const PcArgs &a, int t, int e, bool out)
{
    const int64_t mat = (int64_t)1 << (a.lg_n1 + a.lg_n2);
    const int m = 1 << a.lg_m;
    if constexpr (MODE == 0) {
        const int groups = 1 << (a.lg_n2 - a.lg_m);
        const int b = t >> (a.lg_n2 - a.lg_m), g = t & (groups - 1);
        return b * mat + ((int64_t)(e >> a.lg_m) << a.lg_n2) + g * m + (e & (m - 1));
    }
    if (MODE == 1 && out) {
        const int per = 1 << (a.lg_n1 - a.lg_m);  // row tiles per matrix
        const int b = t >> (a.lg_n1 - a.lg_m), r0 = (t & (per - 1)) << a.lg_m;
        return b * mat + ((int64_t)(e >> a.lg_m) << a.lg_n1) + r0 + (e & (m - 1));
    }
    return (int64_t)t * TILE + e;
}

template <int MODE>
__global__ void __launch_bounds__(NT) pattern_copy_kernel(PcArgs a)
{
    f4 r[IT], q[IT];
    const int grid = gridDim.x;
    int t = a.order ? (int)((int64_t)blockIdx.x * a.ntiles / grid) : (int)blockIdx.x;
    const int t_end = a.order ? (int)((int64_t)(blockIdx.x + 1) * a.ntiles / grid) : a.ntiles;
    const int step = a.order ? 1 : grid;
    if (t >= t_end) return;
    const int tid = threadIdx.x;
#pragma unroll
    for (int it = 0; it < IT; it++) r[it] = a.in[pc_off<MODE>(a, t, 2 * (tid + it * NT), false) >> 1];
    while (true) {
        int nx = t + step;
        if (nx < t_end) {
#pragma unroll
            for (int it = 0; it < IT; it++) q[it] = a.in[pc_off<MODE>(a, nx, 2 * (tid + it * NT), false) >> 1];
        }
#pragma unroll
        for (int it = 0; it < IT; it++) a.out[pc_off<MODE>(a, t, 2 * (tid + it * NT), true) >> 1] = r[it];
        if (nx >= t_end) break;
        t = nx;
        nx = t + step;
        if (nx < t_end) {
#pragma unroll
            for (int it = 0; it < IT; it++) r[it] = a.in[pc_off<MODE>(a, nx, 2 * (tid + it * NT), false) >> 1];
        }
#pragma unroll
        for (int it = 0; it < IT; it++) a.out[pc_off<MODE>(a, t, 2 * (tid + it * NT), true) >> 1] = q[it];
        if (nx >= t_end) break;
        t = nx;
    }
}

int lg2(int64_t v)
{
    if (v <= 0 || (v & (v - 1))) return -1;
    int l = 0;
    while (((int64_t)1 << l) < v) l++;
    return l;
}

}  // namespace

// Copies `matrices` n1 x n2 complex64 matrices from in to out (both 16-byte
// aligned, non-overlapping) in access pattern `mode` with m-wide tiles and
// tile order `order`, on `stream`.  Returns 0, or -1 for a shape the pattern does not cover, or the
// HIP error of the launch.
extern "C" __attribute__((visibility("default"))) int gfft_pattern_copy(const void *in, void *out,
                                                                        long long matrices, int n1, int n2,
                                                                        int m, int mode, int order, void *stream)
{
    const int l1 = lg2(n1), l2 = lg2(n2), lm = lg2(m);
    if (l1 < 0 || l2 < 0 || lm < 1 || matrices <= 0 || mode < 0 || mode > 2 || order < 0 || order > 1) return -1;
    if (((uintptr_t)in & 15) || ((uintptr_t)out & 15)) return -1;
    if (mode == 0 && ((int64_t)n1 * m != TILE || m > n2)) return -1;
    if (mode == 1 && ((int64_t)m * n2 != TILE || m > n1)) return -1;
    if (mode == 2 && ((int64_t)n1 * n2) % TILE) return -1;
    const int64_t tiles = matrices * (((int64_t)n1 * n2) / TILE);
    if (tiles > INT32_MAX) return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return -1;
    PcArgs a{};
    a.in = static_cast<const f4 *>(in);
    a.out = static_cast<f4 *>(out);
    a.ntiles = (int32_t)tiles;
    a.mode = mode;
    a.lg_n1 = l1;
    a.lg_n2 = l2;
    a.lg_m = lm;
    a.order = order;
    const int grid = (int)(tiles < cus ? tiles : cus);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (mode == 0) hipLaunchKernelGGL(pattern_copy_kernel<0>, dim3(grid), dim3(NT), 0, s, a);
    else if (mode == 1) hipLaunchKernelGGL(pattern_copy_kernel<1>, dim3(grid), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL(pattern_copy_kernel<2>, dim3(grid), dim3(NT), 0, s, a);
    return (int)hipGetLastError();
}
