// gfft_r2c.h -- the fused last pass of a four-step real-input FFT.
//
// A real signal of n = 2L samples is transformed as the L-point complex FFT
// Z of its (even, odd) pairs, then unpacked into the n/2 + 1 bins F
// (FFTReal.h:204-213 with adjust_DIT_impl, generic/fft_dit_impl_generic.inl:27-61):
//   F_k = E_k - i W_n^k O_k,  F_{L-k} = conj(E_k + i W_n^k O_k),
//   E_k = (Z_k + conj Z_{L-k}) / 2,  O_k = (Z_k - conj Z_{L-k}) / 2.
// For L = n1 * R the four-step column pass leaves row k_c of Z's second
// pass in scratch; bin k = k_c + n1 * k_r comes out of row k_c, element k_r,
// and its partner L - k out of row n1 - k_c, element R - 1 - k_r.  One
// workgroup takes M/2 rows and their M/2 mirror rows, so the unpack pairs
// meet in LDS and the bins are written straight to the output: the
// separate unpack pass over HBM disappears.
#pragma once

#include "gfft_misc.h"
#include "gfft_registry.h"

namespace gfft {

// W_n^m from a two-level FP64 table
__device__ __forceinline__ cx<double> tw_w(const TwTable &t, uint64_t m)
{
    m &= t.mask;
    const cx<double> *lo = reinterpret_cast<const cx<double> *>(t.lo);
    const cx<double> *hi = reinterpret_cast<const cx<double> *>(t.hi);
    return cmul(hi[m >> t.lo_bits], lo[m & ((1ull << t.lo_bits) - 1ull)]);
}

template <typename T>
__device__ __forceinline__ cx<T> cj(cx<T> a) { return {a.re, -a.im}; }

// F_k and F_{L-k} from Z_k, Z_{L-k} (generic/fft_dit_impl_generic.inl:48-52)
template <typename T>
__device__ __forceinline__ void unpack_pair(cx<T> zk, cx<T> zj, cx<double> w, cx<T> &fk, cx<T> &fj)
{
    const cx<T> e = {(T)0.5 * (zk.re + zj.re), (T)0.5 * (zk.im - zj.im)};
    const cx<T> o = {(T)0.5 * (zk.re - zj.re), (T)0.5 * (zk.im + zj.im)};
    const cx<T> t = {(T)(-w.im), (T)w.re};  // i * W_n^k
    const cx<T> to = cmul(t, o);
    fk = e - to;
    fj = cj(e + to);
}

// Z_k from F_k and F_{L-k} with w = W_n^k (pack_kernel's arithmetic,
// kern_misc.hip): s + i q, s = f + conj g, q = conj(w) (f - conj g).
template <typename T>
__device__ __forceinline__ cx<T> packz(cx<T> f, cx<T> g, cx<double> w)
{
    const cx<T> sm = {f.re + g.re, f.im - g.im};
    const cx<T> d = {f.re - g.re, f.im + g.im};
    const cx<T> wc = {(T)w.re, (T)(-w.im)};
    const cx<T> q = cmul(wc, d);
    return {sm.re - q.im, sm.im + q.re};
}

struct R2cArgs {
    const void *in;    // row r of transform b at in + b * in_dist + r * R (complex elements)
    void *out;         // bins 0 .. L of transform b at out + b * out_dist
    int64_t in_dist, out_dist;
    int32_t n1;        // rows per transform (the column pass length)
    int32_t tiles_per; // workgroups per transform: n1 / M
    int32_t n_tiles;   // tiles_per * transforms of this launch
    int32_t vec_in;    // 16-byte loads allowed (complex64, alignment proven on host)
    int32_t xcd_remap;
    int32_t pad;
    const void *pass_tw;  // Stockham pass twiddles of the R-point row FFT
    TwTable tw;           // W_n, n = 2L the real length
};

// Row length R, tile rows M (half of them mirrors), threads and radices of
// the fused kernel for a precision (0 FP32, 1 FP64); the entry has no
// tile-kernel function, only the shape (for the pass twiddle table).
const KernelEntry *r2c_fused_entry(int prec);

hipError_t launch_r2c_fused(int prec, const R2cArgs &a, hipStream_t s, hipEvent_t stop = nullptr);  // one workgroup per tile

// The inverse (c2r) mirror image: the pack Z_k = (F_k + conj F_{L-k})
// + i W_n^-k (F_k - conj F_{L-k}) fused into the first four-step pass of the
// L-point inverse, L = n1 * n2.  Element e of column c is Z_{c + n2 e}, whose
// partner L - k is element n1 - 1 - e of column n2 - c, so a tile of M/2
// columns and their mirrors packs in LDS, runs the n1-point column FFTs
// (inverse by re/im swap) and stores with the four-step twiddle W_L^(c k_c).
struct C2rArgs {
    const void *in;    // bins 0 .. L of transform b at in + b * in_dist
    void *out;         // element k_c of column c of transform b at out + b * L + c + n2 * k_c
    int64_t in_dist;
    int32_t n2;        // columns (the second pass's length)
    int32_t tiles_per; // workgroups per transform: n2 / M
    int32_t n_tiles;
    int32_t xcd_remap;
    const void *pass_tw;  // Stockham pass twiddles of the n1-point column FFT
    TwTable tw;           // W_n, n = 2L (the pack)
    TwTable tw4;          // W_L (the four-step twiddle)
};

// Shape of the fused column kernel for column length n1 (nullptr: none).
const KernelEntry *c2r_fused_entry(int prec, int n1);
hipError_t launch_c2r_fused(int prec, int n1, const C2rArgs &a, hipStream_t s);

// Real transforms whose n/2 points fit one row tile: the r2c is the row
// FFT of the (even, odd) pairs with the unpack as its epilogue, the c2r the
// pack as its prologue and the inverse row FFT; one HBM round trip either way.
struct RealRowsArgs {
    TileArgs t;      // the rows as a ROW -> ROW tile pass (reals read or written as complex pairs)
    TwTable tw;      // W_n
    int32_t half;    // r2c: n/2 + 1 bins per row (else all n, Hermitian-filled)
    int32_t pad;
};

enum RealRowsMode { RR_R2C = 0, RR_C2R = 1 };

// Shape for n/2 = L (the row-tile shape of that length; nullptr: none).
const KernelEntry *real_rows_entry(int prec, int L);
hipError_t launch_real_rows(int prec, int L, int mode, const RealRowsArgs &a, int64_t grid, hipStream_t s,
                            hipEvent_t stop = nullptr);

}  // namespace gfft
