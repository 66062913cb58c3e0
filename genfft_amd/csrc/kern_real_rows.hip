// kern_real_rows.hip -- real transforms of n = 2L points with L up to the
// single-pass row length, in one HBM round trip.
//
//   r2c: the L-point FFT of the (even, odd) pairs of each row, then the
//        unpack of adjust_DIT_impl (generic/fft_dit_impl_generic.inl:27-61)
//        from LDS as the epilogue: bins k and L-k of pair p = k, bin 0 and
//        n/2 from Z_0, bin n/4 = conj Z_{L/2}; with a full spectrum also
//        F_{n-k} = conj F_k (FFTReal.h:204-213 with half = false).
//   c2r: the pack Z_k = (F_k + conj F_{L-k}) + i W_n^-k (F_k - conj F_{L-k})
//        as the prologue (both bins of a pair loaded by one thread, all
//        loads issued before the first LDS write), written re/im-swapped,
//        then the row FFT and the swapped store: the unscaled inverse.
//        The imaginary parts of bins 0 and n/2 are ignored, as in any c2r.
// A row tile holds M rows; the shapes are those of the c2c row tiles.
#include "gfft_r2c.h"

namespace gfft {
namespace {

template <typename T, int L, int M, int NT, int MODE, int... RS>
__global__ void __launch_bounds__(NT) real_rows_kernel(RealRowsArgs a)
{
    using G = TileGeom<L, M, LAY_SEQ, PAT_ROW, PAT_ROW>;
    constexpr int PP = L / 2 + 1;  // pairs per row, p = 0 .. L/2
    constexpr int IT = (M * PP + NT - 1) / NT;
    __shared__ cx<T> lds[G::ELEMS];
    const int tid = threadIdx.x;
    const int tile = a.t.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int s0 = tile * M;
    const int nvalid = min(M, a.t.n_inner - s0);
    const cx<T> *ptw = reinterpret_cast<const cx<T> *>(a.t.pass_tw);

    if constexpr (MODE == RR_R2C) {
        load_tile<T, L, M, NT, G, PAT_ROW>(lds, a.t, 0, 0, s0, nvalid, tid);
        __syncthreads();
        stockham_passes<T, L, M, NT, G, true, false, 1, RS...>(lds, ptw, tid, a.t, s0, 0, Tw4Pre{});
        cx<T> *out = reinterpret_cast<cx<T> *>(a.t.out);
#pragma unroll 2
        for (int it = 0; it < IT; it++) {
            const int q = tid + it * NT;
            const int s = q / PP, p = q - s * PP;
            if (q >= M * PP || s >= nvalid) continue;
            cx<T> *f = out + (int64_t)(s0 + s) * a.t.out_inner;
            if (p == 0) {
                const cx<T> z0 = lds[G::off(s, 0)];
                f[0] = {z0.re + z0.im, (T)0};
                f[L] = {z0.re - z0.im, (T)0};
            } else if (p == L / 2) {
                const cx<T> zq = lds[G::off(s, L / 2)];
                f[L / 2] = cj(zq);
                if (!a.half) f[2 * L - L / 2] = zq;
            } else {
                cx<T> fk, fj;
                unpack_pair(lds[G::off(s, p)], lds[G::off(s, L - p)], tw_w(a.tw, (uint64_t)p), fk, fj);
                f[p] = fk;
                f[L - p] = fj;
                if (!a.half) {
                    f[2 * L - p] = cj(fk);
                    f[L + p] = cj(fj);
                }
            }
        }
    } else {
        const cx<T> *in = reinterpret_cast<const cx<T> *>(a.t.in);
        cx<T> fa[IT], fb[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int q = tid + it * NT;
            const int s = q / PP, p = q - s * PP;
            if (q < M * PP && s < nvalid) {
                const cx<T> *f = in + (int64_t)(s0 + s) * a.t.in_inner;
                fa[it] = f[p];
                fb[it] = f[L - p];
            }
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int q = tid + it * NT;
            const int s = q / PP, p = q - s * PP;
            if (q >= M * PP || s >= nvalid) continue;
            cx<T> f = fa[it], g = fb[it];
            if (p == 0) {
                f.im = (T)0;
                g.im = (T)0;
                lds[G::off(s, 0)] = swap_ri(packz(f, g, cx<double>{1.0, 0.0}));
            } else if (p == L / 2) {
                lds[G::off(s, p)] = swap_ri(packz(f, f, tw_w(a.tw, (uint64_t)p)));
            } else {
                const cx<double> w = tw_w(a.tw, (uint64_t)p);
                lds[G::off(s, p)] = swap_ri(packz(f, g, w));
                lds[G::off(s, L - p)] = swap_ri(packz(g, f, cx<double>{-w.re, w.im}));
            }
        }
        __syncthreads();
        stockham_passes<T, L, M, NT, G, true, false, 1, RS...>(lds, ptw, tid, a.t, s0, 0, Tw4Pre{});
        store_tile<T, L, M, NT, G, PAT_ROW>(lds, a.t, 0, 0, s0, nvalid, tid);
    }
}

template <typename T>
struct RowsShape {
    KernelEntry entry;
    void (*fn[2])(RealRowsArgs);
};

template <typename T, int L, int M, int NT, int... RS>
constexpr RowsShape<T> rows_shape()
{
    return {entry_shape<T, L, M, NT, LAY_SEQ, PAT_ROW, PAT_ROW, RS...>(),
            {&real_rows_kernel<T, L, M, NT, RR_R2C, RS...>, &real_rows_kernel<T, L, M, NT, RR_C2R, RS...>}};
}

// the default c2c row-tile shapes of kern_tile_f32.hip / kern_tile_f64.hip;
// from L = 2 (n = 4) on, so short real transforms are one launch as well
// (n = 2 runs as a copy and the unpack)
const RowsShape<float> kRowsF32[] = {
    rows_shape<float, 2, 1024, 128, 2>(),          rows_shape<float, 4, 1024, 256, 4>(),
    rows_shape<float, 8, 512, 256, 8>(),           rows_shape<float, 16, 256, 256, 16>(),
    rows_shape<float, 32, 128, 256, 8, 4>(),
    rows_shape<float, 64, 64, 256, 8, 8>(),        rows_shape<float, 128, 32, 256, 16, 8>(),
    rows_shape<float, 256, 16, 256, 16, 16>(),     rows_shape<float, 512, 8, 256, 8, 8, 8>(),
    rows_shape<float, 1024, 4, 256, 16, 16, 4>(),  rows_shape<float, 2048, 2, 256, 16, 16, 8>(),
    rows_shape<float, 4096, 1, 256, 16, 16, 16>(), rows_shape<float, 8192, 1, 512, 16, 16, 8, 4>(),
    rows_shape<float, 16384, 1, 1024, 16, 16, 16, 4>(),
};
const RowsShape<double> kRowsF64[] = {
    rows_shape<double, 2, 512, 64, 2>(),            rows_shape<double, 4, 512, 128, 4>(),
    rows_shape<double, 8, 256, 128, 8>(),           rows_shape<double, 16, 128, 128, 16>(),
    rows_shape<double, 32, 64, 128, 8, 4>(),
    rows_shape<double, 64, 32, 128, 8, 8>(),        rows_shape<double, 128, 16, 128, 16, 8>(),
    rows_shape<double, 256, 8, 256, 8, 8, 4>(),     rows_shape<double, 512, 4, 256, 8, 8, 8>(),
    rows_shape<double, 1024, 2, 256, 8, 4, 8, 4>(), rows_shape<double, 2048, 1, 256, 8, 8, 8, 4>(),
    rows_shape<double, 4096, 1, 512, 8, 8, 8, 8>(), rows_shape<double, 8192, 1, 1024, 8, 8, 8, 8, 2>(),
};

template <typename T, size_t K>
const RowsShape<T> *rows_find(const RowsShape<T> (&tab)[K], int L)
{
    for (size_t i = 0; i < K; i++)
        if (tab[i].entry.L == L) return &tab[i];
    return nullptr;
}

}  // namespace

const KernelEntry *real_rows_entry(int prec, int L)
{
    if (prec) {
        const auto *e = rows_find(kRowsF64, L);
        return e ? &e->entry : nullptr;
    }
    const auto *e = rows_find(kRowsF32, L);
    return e ? &e->entry : nullptr;
}

hipError_t launch_real_rows(int prec, int L, int mode, const RealRowsArgs &a, int64_t grid, hipStream_t s,
                            hipEvent_t stop)
{
    void (*fn)(RealRowsArgs) = nullptr;
    int nt = 0;
    if (prec) {
        if (const auto *e = rows_find(kRowsF64, L)) fn = e->fn[mode], nt = e->entry.NT;
    } else {
        if (const auto *e = rows_find(kRowsF32, L)) fn = e->fn[mode], nt = e->entry.NT;
    }
    if (!fn) return hipErrorInvalidValue;
    return launch_stop(fn, dim3((unsigned)grid), dim3(nt), s, a, stop);
}

}  // namespace gfft
