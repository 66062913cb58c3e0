// kern_r2c.hip -- second four-step pass of a real-input FFT with the unpack
// fused into its epilogue (see gfft_r2c.h for the index algebra).
//
// Tile t of transform b holds rows a0 .. a0+H-1 (a0 = t*H, H = M/2) in
// slots 0 .. H-1 and their mirrors n1-a0 .. n1-a0-H+1 in slots H .. M-1;
// row 0 pairs with itself, so in tile 0 the mirror slot of row 0 holds
// the other self-paired row, n1/2.  After the R-point Stockham passes, bin
// k = u + n1*k_r (u = a0+s, 0 < u < n1/2) and its partner L - k =
// (n1-u) + n1*(R-1-k_r) are both in LDS; each thread computes whole pairs
// and writes both bins.  Lane s = 0 of tile 0 takes the self-paired
// rows: row 0 gives F_0, F_{n/2}, F_{n/4} and the pairs (k_r, R-k_r); row
// n1/2 the pairs (k_r, R-1-k_r).  Writes along s are contiguous H-element
// runs (64 bytes for complex64, 128 for complex128), in ascending order for
// k and descending for L - k.
#include "gfft_r2c.h"

namespace gfft {
namespace {

// Bins of one tile (transform b, upper rows a0 ..) from the row spectra in
// LDS; thread `t` takes slot s = t % H at k_r = t / H + i * (NT / H).
template <typename T, int R, int M, int NT, class G>
__device__ __forceinline__ void unpack_tile(const cx<T> *lds, const R2cArgs &a, int b, int a0, int t)
{
    constexpr int H = M / 2;
    constexpr int P = H * R / NT;  // pairs per thread
    const int n1 = a.n1;
    cx<T> *out = reinterpret_cast<cx<T> *>(a.out) + (int64_t)b * a.out_dist;
    const int64_t L = (int64_t)n1 * R;
    const int s = t % H;
    const int u = a0 + s;
    if (u != 0) {
        // bins k0 + i * step: W_n^k as an FP64 geometric progression from two
        // table lookups, rounded once per bin
        const int64_t step = (int64_t)n1 * (NT / H);
        const int64_t k0 = u + (int64_t)n1 * (t / H);
        cx<double> w = tw_w(a.tw, (uint64_t)k0);
        const cx<double> wr = tw_w(a.tw, (uint64_t)step);
#pragma unroll
        for (int i = 0; i < P; i++) {
            const int kr = t / H + i * (NT / H);
            const int64_t k = k0 + step * i;
            cx<T> fk, fj;
            unpack_pair(lds[G::off(s, kr)], lds[G::off(H + s, R - 1 - kr)], w, fk, fj);
            out[k] = fk;
            out[L - k] = fj;
            w = cmul(w, wr);
        }
        return;
    }
    // lane 0 of workgroup 0 of a transform: the self-paired rows 0 and n1/2
    for (int i = 0; i < P; i++) {
        const int kr = t / H + i * (NT / H);
        cx<T> fk, fj;
        if (kr == 0) {  // row 0: F_0, F_{n/2} and F_{n/4}
            const cx<T> z0 = lds[G::off(0, 0)];
            out[0] = {z0.re + z0.im, (T)0};
            out[L] = {z0.re - z0.im, (T)0};
            out[L / 2] = cj(lds[G::off(0, R / 2)]);
        } else if (kr < R / 2) {  // row 0: bins n1*kr and n1*(R-kr)
            const int64_t k = (int64_t)n1 * kr;
            unpack_pair(lds[G::off(0, kr)], lds[G::off(0, R - kr)], tw_w(a.tw, (uint64_t)k), fk, fj);
            out[k] = fk;
            out[L - k] = fj;
        } else {  // row n1/2 (slot H): bins n1/2 + n1*m and n1/2 + n1*(R-1-m)
            const int m = kr - R / 2;
            const int64_t k = n1 / 2 + (int64_t)n1 * m;
            unpack_pair(lds[G::off(H, m)], lds[G::off(H, R - 1 - m)], tw_w(a.tw, (uint64_t)k), fk, fj);
            out[k] = fk;
            out[L - k] = fj;
        }
    }
}

// One tile per workgroup.  (A persistent form that prefetches the next
// tile's rows measured 0.86 ms against 0.47 ms on MI355X for 2^20 x 256:
// its loop head waits for the previous tile's scattered bin stores, which
// the hardware overlaps with the next workgroup's loads when tiles are
// separate workgroups.)
template <typename T, int R, int M, int NT, int... RS>
__global__ void __launch_bounds__(NT) r2c_row_unpack_kernel(R2cArgs a)
{
    constexpr int H = M / 2;
    constexpr int TOT = M * R;
    static_assert(NT % H == 0 && (H * R) % NT == 0 && TOT % (2 * NT) == 0, "pairs must split evenly over threads");
    using G = TileGeom<R, M, LAY_SEQ, PAT_ROW, PAT_COL>;
    __shared__ cx<T> lds[G::ELEMS];
    const int tid = threadIdx.x;
    const int tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int b = tile / a.tiles_per;
    const int a0 = (tile - b * a.tiles_per) * H;
    const int n1 = a.n1;
    auto row_of = [&](int s) {
        if (s < H) return a0 + s;
        const int m = s - H;
        return a0 + m == 0 ? n1 / 2 : n1 - a0 - m;
    };

    // the M rows, R contiguous complex each: every load is issued into
    // registers before the first LDS write (the input pointer is generic, so
    // interleaved LDS writes would order each load after the previous write)
    const cx<T> *in = reinterpret_cast<const cx<T> *>(a.in) + (int64_t)b * a.in_dist;
    constexpr int E = TOT / NT;
    cx<T> reg[E];
    // The rows are read once, with the streaming (nontemporal) hint: the L2
    // then keeps the bin lines this tile and its neighbours write in parts
    // until they are whole.  Measured at 2^20 x 256 (complex64 bins, odd
    // distance): 1153 -> 1063 MiB written per GiB of bins, 0.489 -> 0.426 ms.
    const bool vec = sizeof(T) == 4 && a.vec_in;
    if (vec) {
        typedef float f4v __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int it = 0; it < E / 2; it++) {
            const int e = (tid + it * NT) * 2;
            const f4v q = __builtin_nontemporal_load(
                reinterpret_cast<const f4v *>(in + (int64_t)row_of(e / R) * R + e % R));
            reg[2 * it] = {(T)q.x, (T)q.y};
            reg[2 * it + 1] = {(T)q.z, (T)q.w};
        }
#pragma unroll
        for (int it = 0; it < E / 2; it++) {
            const int e = (tid + it * NT) * 2;
            lds[G::off(e / R, e % R)] = reg[2 * it];
            lds[G::off(e / R, e % R + 1)] = reg[2 * it + 1];
        }
    } else {
#pragma unroll
        for (int it = 0; it < E; it++) {
            const int e = tid + it * NT;
            reg[it] = in[(int64_t)row_of(e / R) * R + e % R];
        }
#pragma unroll
        for (int it = 0; it < E; it++) {
            const int e = tid + it * NT;
            lds[G::off(e / R, e % R)] = reg[it];
        }
    }
    __syncthreads();
    TileArgs none{};
    none.tw_mode = TW_NONE;
    stockham_passes<T, R, M, NT, G, true, false, 1, RS...>(lds, reinterpret_cast<const cx<T> *>(a.pass_tw), tid, none,
                                                           0, 0, Tw4Pre{});
    unpack_tile<T, R, M, NT, G>(lds, a, b, a0, tid);
}

// Shapes: 16 elements per thread and ~68 KB (complex64) / ~136 KB
// (complex128) of LDS.  For complex64 a 16-row tile (64-byte output runs,
// two workgroups per CU) measured 0.473 ms against 0.483 ms for a 32-row
// tile of 128-byte runs with one workgroup per CU (2^20 x 256).
template <typename T>
struct Fused;
template <>
struct Fused<float> {
    static constexpr KernelEntry entry = entry_shape<float, 512, 16, 512, LAY_SEQ, PAT_ROW, PAT_COL, 8, 8, 8>();
    static constexpr auto kernel = &r2c_row_unpack_kernel<float, 512, 16, 512, 8, 8, 8>;
};
template <>
struct Fused<double> {
    static constexpr KernelEntry entry = entry_shape<double, 512, 16, 512, LAY_SEQ, PAT_ROW, PAT_COL, 8, 8, 8>();
    static constexpr auto kernel = &r2c_row_unpack_kernel<double, 512, 16, 512, 8, 8, 8>;
};

template <typename T>
hipError_t launch_t(const R2cArgs &a, hipStream_t s, hipEvent_t stop)
{
    return launch_stop(Fused<T>::kernel, dim3((unsigned)a.n_tiles), dim3(Fused<T>::entry.NT), s, a, stop);
}

// ---------------------------------------------------------------------------
// c2r: the pack fused into the column pass of the inverse four-step.
// ---------------------------------------------------------------------------
// Z_k and Z_{L-k} in place, written re/im-swapped (the inverse runs as the
// forward transform of the swapped data); W_n^{L-k} = -conj(W_n^k).
template <typename T, class G>
__device__ __forceinline__ void pack_pair(cx<T> *lds, int sk, int ek, int sj, int ej, cx<double> w)
{
    const cx<T> fk = lds[G::off(sk, ek)], fj = lds[G::off(sj, ej)];
    const cx<double> wj = {-w.re, w.im};
    lds[G::off(sk, ek)] = swap_ri(packz(fk, fj, w));
    lds[G::off(sj, ej)] = swap_ri(packz(fj, fk, wj));
}

template <typename T, int N1, int M, int NT, int... RS>
__global__ void __launch_bounds__(NT) c2r_col_pack_kernel(C2rArgs a)
{
    constexpr int H = M / 2;
    constexpr int TOT = M * N1;
    constexpr int E = TOT / NT;
    static_assert(TOT % NT == 0 && (H * N1) % NT == 0 && NT % M == 0, "tile must split evenly over threads");
    using G = TileGeom<N1, M, LAY_ELEM, PAT_COL, PAT_COL>;
    __shared__ cx<T> lds[G::ELEMS];
    __shared__ cx<T> nyq;  // F_L (workgroup 0 of a transform)
    const int tid = threadIdx.x;
    const int tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int b = tile / a.tiles_per;
    const int a0 = (tile - b * a.tiles_per) * H;
    const int n2 = a.n2;
    const int64_t L = (int64_t)N1 * n2;
    auto col_of = [&](int s) {
        if (s < H) return a0 + s;
        const int m = s - H;
        return a0 + m == 0 ? n2 / 2 : n2 - a0 - m;
    };

    // load F along the columns (slot-fast: runs of H bins), all loads first
    const cx<T> *in = reinterpret_cast<const cx<T> *>(a.in) + (int64_t)b * a.in_dist;
    cx<T> reg[E];
#pragma unroll
    for (int it = 0; it < E; it++) {
        const int e = tid + it * NT;
        reg[it] = in[col_of(e % M) + (int64_t)n2 * (e / M)];
    }
    if (a0 == 0 && tid == 0) nyq = in[L];
#pragma unroll
    for (int it = 0; it < E; it++) {
        const int e = tid + it * NT;
        lds[G::off(e % M, e / M)] = reg[it];
    }
    __syncthreads();

    // pack: thread takes upper slot s = tid % H at e = tid / H + i * (NT / H)
    {
        constexpr int P = H * N1 / NT;
        const int s = tid % H;
        const int u = a0 + s;
        if (u != 0) {
            const int64_t step = (int64_t)n2 * (NT / H);
            const int64_t k0 = u + (int64_t)n2 * (tid / H);
            cx<double> w = tw_w(a.tw, (uint64_t)k0);
            const cx<double> wr = tw_w(a.tw, (uint64_t)step);
#pragma unroll
            for (int i = 0; i < P; i++) {
                const int e = tid / H + i * (NT / H);
                pack_pair<T, G>(lds, s, e, H + s, N1 - 1 - e, w);
                w = cmul(w, wr);
            }
        } else {
            // lane 0 of workgroup 0: columns 0 and n2/2 pair with themselves
            for (int i = 0; i < P; i++) {
                const int e = tid / H + i * (NT / H);
                if (e == 0) {
                    // Z_0 from F_0 and F_L (their imaginary parts ignored), and
                    // the self-paired Z_{L/2} (element N1/2 of column 0)
                    cx<T> f0 = lds[G::off(0, 0)], fl = nyq;
                    f0.im = (T)0;
                    fl.im = (T)0;
                    lds[G::off(0, 0)] = swap_ri(packz(f0, fl, cx<double>{1.0, 0.0}));
                    const cx<T> fq = lds[G::off(0, N1 / 2)];
                    lds[G::off(0, N1 / 2)] = swap_ri(packz(fq, fq, tw_w(a.tw, (uint64_t)(L / 2))));
                } else if (e < N1 / 2) {  // column 0: bins n2*e and n2*(N1-e)
                    pack_pair<T, G>(lds, 0, e, 0, N1 - e, tw_w(a.tw, (uint64_t)((int64_t)n2 * e)));
                } else {  // column n2/2 (slot H): bins n2/2 + n2*m and n2/2 + n2*(N1-1-m)
                    const int m = e - N1 / 2;
                    pack_pair<T, G>(lds, H, m, H, N1 - 1 - m, tw_w(a.tw, (uint64_t)(n2 / 2 + (int64_t)n2 * m)));
                }
            }
        }
    }
    __syncthreads();
    TileArgs none{};
    none.tw_mode = TW_NONE;
    stockham_passes<T, N1, M, NT, G, false, false, 1, RS...>(lds, reinterpret_cast<const cx<T> *>(a.pass_tw), tid,
                                                            none, 0, 0, Tw4Pre{});

    // store with the four-step twiddle W_L^(c k_c), thread: slot s = tid % M
    // at k_c = tid / M + i * (NT / M), an FP64 progression per thread
    cx<T> *out = reinterpret_cast<cx<T> *>(a.out) + (int64_t)b * L;
    const int s = tid % M;
    const int c = col_of(s);
    cx<double> w = tw_w(a.tw4, (uint64_t)c * (uint64_t)(tid / M));
    const cx<double> wr = tw_w(a.tw4, (uint64_t)c * (uint64_t)(NT / M));
#pragma unroll
    for (int i = 0; i < E; i++) {
        const int kc = tid / M + i * (NT / M);
        out[c + (int64_t)n2 * kc] = apply_tw(lds[G::off(s, kc)], w);
        w = cmul(w, wr);
    }
}

template <typename T>
struct C2rShape {
    KernelEntry entry;
    void (*fn)(C2rArgs);
};

// One shape per column length: M columns per tile (M/2 and their mirrors),
// 16 elements per thread (8 for n1 = 32, 4 for n1 = 16).  M = 16 reads and
// writes 64-byte row segments (complex64).  M = 32 (128-byte segments) at
// n1 = 512 measured slower, 0.639 against 0.568 ms for 2^20 x 256: its 128 KB
// tile leaves one workgroup per CU, and nothing overlaps its load with the
// previous tile's store.  A persistent form of that tile (the next tile's
// bins prefetched into registers, one run of tiles per workgroup) spilled
// 52 bytes per lane and took 0.785 ms: the reload at the loop head waits for
// every store still in flight.
template <typename T, int N1, int M, int NT, int... RS>
constexpr C2rShape<T> c2r_shape()
{
    return {entry_shape<T, N1, M, NT, LAY_ELEM, PAT_COL, PAT_COL, RS...>(), &c2r_col_pack_kernel<T, N1, M, NT, RS...>};
}

const C2rShape<float> kC2rF32[] = {
    c2r_shape<float, 32, 16, 64, 8, 4>(),     c2r_shape<float, 64, 16, 64, 8, 8>(),
    c2r_shape<float, 128, 16, 128, 16, 8>(),  c2r_shape<float, 256, 16, 256, 16, 16>(),
    c2r_shape<float, 512, 16, 512, 8, 8, 8>(), c2r_shape<float, 1024, 16, 1024, 16, 16, 4>(),
};
const C2rShape<double> kC2rF64[] = {
    c2r_shape<double, 16, 16, 64, 16>(),      c2r_shape<double, 32, 16, 64, 8, 4>(),
    c2r_shape<double, 64, 16, 64, 8, 8>(),    c2r_shape<double, 128, 16, 128, 16, 8>(),
    c2r_shape<double, 256, 16, 256, 16, 16>(), c2r_shape<double, 512, 16, 512, 8, 8, 8>(),
};

template <typename T, size_t K>
const C2rShape<T> *c2r_find(const C2rShape<T> (&tab)[K], int n1)
{
    for (size_t i = 0; i < K; i++)
        if (tab[i].entry.L == n1) return &tab[i];
    return nullptr;
}

}  // namespace

const KernelEntry *r2c_fused_entry(int prec)
{
    return prec ? &Fused<double>::entry : &Fused<float>::entry;
}

hipError_t launch_r2c_fused(int prec, const R2cArgs &a, hipStream_t s, hipEvent_t stop)
{
    return prec ? launch_t<double>(a, s, stop) : launch_t<float>(a, s, stop);
}

const KernelEntry *c2r_fused_entry(int prec, int n1)
{
    if (prec) {
        const auto *e = c2r_find(kC2rF64, n1);
        return e ? &e->entry : nullptr;
    }
    const auto *e = c2r_find(kC2rF32, n1);
    return e ? &e->entry : nullptr;
}

hipError_t launch_c2r_fused(int prec, int n1, const C2rArgs &a, hipStream_t s)
{
    void (*fn)(C2rArgs) = nullptr;
    int nt = 0;
    if (prec) {
        if (const auto *e = c2r_find(kC2rF64, n1)) fn = e->fn, nt = e->entry.NT;
    } else {
        if (const auto *e = c2r_find(kC2rF32, n1)) fn = e->fn, nt = e->entry.NT;
    }
    if (!fn) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3((unsigned)a.n_tiles), dim3(nt), 0, s, a);
    return hipGetLastError();
}

}  // namespace gfft
