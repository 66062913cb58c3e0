// kern_filter.hip -- fused spectral filter on real rows and its adjoint.
//
//   forward   y  = irfft(H * rfft(x))                         (length-n rows)
//   adjoint   gx = irfft(conj(H) * rfft(gy))
//             gH = sum_rows  w_k/n * rfft(gy)_k * conj(rfft(x)_k)   (w = 1 at k = 0, n/2; else 2)
//
// irfft follows the usual Hermitian convention: the imaginary parts of bins 0
// and n/2 are ignored, and it is normalised (1/n), so y = IFFT(H_full * FFT(x))
// / n with H_full the Hermitian extension of H.  gH is the gradient of a real
// loss with respect to H in the (d/dRe + i d/dIm) convention.
//
// One workgroup owns one row and keeps it in LDS: the n reals are read as
// L = n/2 complex values straight into the registers of the first pass of a
// forward L-point Stockham FFT, which gives Z, the
// real-FFT unpack (the adjust_DIT_impl math, generic/fft_dit_impl_generic.inl:27-61)
// gives the half spectrum, the filter is applied pairwise (bins k and L-k
// belong to one thread), the c2r pack is applied in the same pass, and a
// second L-point FFT (the inverse, by the re/im swap), whose last pass writes
// to HBM, produces the row.  One
// HBM read and one write per row; the adjoint also reads x and writes a row
// of gH partials that a small kernel sums over the rows.
#include "gfft_filter.h"

namespace gfft {
namespace {

template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
// a * i
template <typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) { return {-a.im, a.re}; }

// real-FFT unpack of the pair (k, j = L - k), 0 < k < L/2:  F_k = E - i W^k O,
// F_j = conj(E + i W^k O) with E = (Z_k + conj Z_j)/2, O = (Z_k - conj Z_j)/2
template <typename T>
__device__ __forceinline__ void unpack_pair(cx<T> zk, cx<T> zj, cx<T> w, cx<T> &fk, cx<T> &fj)
{
    const cx<T> E = cscale(zk + cconj(zj), (T)0.5);
    const cx<T> O = cscale(zk - cconj(zj), (T)0.5);
    const cx<T> t = mul_i(cmul(w, O));
    fk = E - t;
    fj = cconj(E + t);
}

// c2r pack of the pair: Z'_k = (Y_k + conj Y_j) + i W^-k (Y_k - conj Y_j),
// Z'_j = (Y_j + conj Y_k) + i W^-j (Y_j - conj Y_k) with W^-j = -W^k.
template <typename T>
__device__ __forceinline__ void pack_pair(cx<T> yk, cx<T> yj, cx<T> w, cx<T> &zk, cx<T> &zj)
{
    zk = (yk + cconj(yj)) + mul_i(cmul(cconj(w), yk - cconj(yj)));
    zj = (yj + cconj(yk)) - mul_i(cmul(w, yj - cconj(yk)));
}

// Spectrum of a row from its packed FFT Z in LDS: bins k and L-k of pair p.
// p = 0 yields the real bins 0 and L (as .re of fk / fj).
template <typename T, class G>
__device__ __forceinline__ void row_bins(const cx<T> *lds, int p, int L, cx<T> w, cx<T> &fk, cx<T> &fj)
{
    if (p == 0) {
        const cx<T> z0 = lds[G::off(0, 0)];
        fk = {z0.re + z0.im, (T)0};
        fj = {z0.re - z0.im, (T)0};
        return;
    }
    unpack_pair(lds[G::off(0, p)], lds[G::off(0, L - p)], w, fk, fj);
}

// The first and last Stockham passes of a row FFT against global memory.
// One workgroup per row and per CU leaves nothing to hide a row's latency,
// so the load is folded into the first pass (each thread reads its R1
// inputs j + r L/R1 straight from the row: lanes run along j, coalesced) and
// the store into the inverse's last pass, saving an LDS round trip and two
// barriers at each end.
//
// Pass 1 (NS = 1, no twiddles) of M sequences, sequence s read from src(s);
// results into LDS for the remaining passes.
template <typename T, int L, int M, int NT, class G, int R, class Src>
__device__ __forceinline__ void first_pass_global(cx<T> *lds, Src src, int tid)
{
    constexpr int LR = L / R;
    constexpr int BF = M * LR;
    constexpr int PER = (BF + NT - 1) / NT;
    cx<T> v[PER][R];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * NT;
        if (BF % NT == 0 || b < BF) {
            const cx<T> *row = src(b / LR);
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = row[b % LR + r * LR];
        }
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * NT;
        if (BF % NT == 0 || b < BF) {
            const int s = b / LR, j = b % LR;
            dft<R>(v[i]);
#pragma unroll
            for (int r = 0; r < R; r++) lds[G::off(s, j * R + r)] = v[i][r];
        }
    }
    __syncthreads();
}

// every pass of RS... but the last, in LDS (NS: the product of earlier radices)
template <typename T, int L, int M, int NT, class G, int NS, int R, int... RS>
__device__ __forceinline__ void passes_but_last(cx<T> *lds, const cx<T> *tw, int tid)
{
    if constexpr (sizeof...(RS) > 0) {
        const TileArgs none{};
        stockham_pass<T, L, M, NT, G, true, NS, R, false, false>(lds, tw, tid, none, 0, 0, Tw4Pre{});
        constexpr int used = NS > 1 ? NS * (R - 1) : 0;
        passes_but_last<T, L, M, NT, G, NS * R, RS...>(lds, tw + used, tid);
    }
}

// The last pass (radix R, NS = L / R) of sequence 0 in LDS, written to `row`
// as s * swap(v): the inverse ran on swapped data.  `tw` is the full pass
// twiddle table of the radix sequence.
template <typename T, int L, int NT, class G, int TWN, int R>
__device__ __forceinline__ void last_pass_global(cx<T> *row, const cx<T> *lds, const cx<T> *tw, T scale, int tid)
{
    constexpr int NS = L / R;
    constexpr int PER = (NS + NT - 1) / NT;
    const cx<T> *twl = tw + (TWN - NS * (R - 1));
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int j = tid + i * NT;
        if (NS % NT == 0 || j < NS) {
            cx<T> v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = lds[G::off(0, j + r * NS)];
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], twl[j * (R - 1) + (r - 1)]);
            dft<R>(v);
#pragma unroll
            for (int r = 0; r < R; r++) row[j + r * NS] = cscale(swap_ri(v[r]), scale);
        }
    }
}

template <typename T, int L, int M, int NT, class G, int R1, int... RS, class Src>
__device__ __forceinline__ void fft_from_global(cx<T> *lds, Src src, const cx<T> *ptw, int tid)
{
    static_assert(sizeof...(RS) > 0, "a row FFT of at least two passes");
    first_pass_global<T, L, M, NT, G, R1>(lds, src, tid);
    const TileArgs none{};
    stockham_passes<T, L, M, NT, G, true, false, R1, RS...>(lds, ptw, tid, none, 0, 0, Tw4Pre{});
}

template <typename T, int L, int NT, class G, int... RS>
__device__ __forceinline__ void fft_to_global(cx<T> *row, cx<T> *lds, const cx<T> *ptw, T scale, int tid)
{
    passes_but_last<T, L, 1, NT, G, 1, RS...>(lds, ptw, tid);
    last_pass_global<T, L, NT, G, pass_tw_count<1, RS...>(), last_radix<RS...>()>(row, lds, ptw, scale, tid);
}

// ---------------------------------------------------------------------------
// The forward FFT's last pass, the per-pair work on the spectrum and the
// inverse FFT's first pass in registers (one LDS round trip instead of three).
//
// The last forward pass (radix R, NS = L/R butterflies) leaves Z[j + r NS] in
// butterfly j.  Bin k = j + r NS pairs with L - k = (NS - j) + (R-1-r) NS, so
// a thread that runs butterflies j and NS - j holds both bins of every pair
// it produces (thread 0 runs j = 0 and j = NS/2, each its own partner set).
// With NS = 2 NT every butterfly is taken exactly once.  The inverse FFT runs
// in the radix order (R, ...): its first pass (NS = 1, no twiddles) reads
// x[j + r L/R] = x[j + r NS], the same registers, so the packed spectrum
// never goes through LDS before that pass.
// ---------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ void pair_butterflies(int t, int &j1, int &j2)
{
    j1 = t;
    j2 = t ? NS - t : NS / 2;
}

// visit(p, zp, zq) once per bin pair of the two butterflies' outputs, zp =
// Z[p], zq = Z[L - p] (references into the registers; p = 0 and p = L/2 pass
// one value twice)
template <typename T, int L, int R, class F>
__device__ __forceinline__ void visit_pairs(cx<T> (&A)[R], cx<T> (&B)[R], int j1, F visit)
{
    constexpr int NS = L / R;
    if (j1 != 0) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int k = j1 + r * NS;  // its partner L - k = j2 + (R-1-r) NS
            if (k < L - k) visit(k, A[r], B[R - 1 - r]);
            else visit(L - k, B[R - 1 - r], A[r]);
        }
        return;
    }
    // j = 0: bins r NS -- 0 alone, L/2 self-paired, r <-> R - r
    visit(0, A[0], A[0]);
    visit(L / 2, A[R / 2], A[R / 2]);
#pragma unroll
    for (int r = 1; r < R / 2; r++) visit(r * NS, A[r], A[R - r]);
    // j = NS/2: bins NS/2 + r NS, r <-> R-1-r
#pragma unroll
    for (int r = 0; r < R / 2; r++) visit(NS / 2 + r * NS, B[r], B[R - 1 - r]);
}

// last forward pass of sequence s (LDS, layout G) for butterflies j1, j2;
// `twl` the pass's twiddles (entry m(R-1) + r-1 = W_L^(r m))
template <typename T, int L, class G, int R>
__device__ __forceinline__ void last_fwd_pass_regs(const cx<T> *lds, int s, const cx<T> *twl, int j1, int j2,
                                                   cx<T> (&A)[R], cx<T> (&B)[R])
{
    constexpr int NS = L / R;
#pragma unroll
    for (int r = 0; r < R; r++) {
        A[r] = lds[G::off(s, j1 + r * NS)];
        B[r] = lds[G::off(s, j2 + r * NS)];
    }
#pragma unroll
    for (int r = 1; r < R; r++) {
        A[r] = cmul(A[r], twl[j1 * (R - 1) + (r - 1)]);
        B[r] = cmul(B[r], twl[j2 * (R - 1) + (r - 1)]);
    }
    dft<R>(A);
    dft<R>(B);
}

// the inverse's first pass (radix R, NS = 1) on the registers, written to
// sequence 0 of layout G once every thread has read the forward data
template <typename T, class G, int R>
__device__ __forceinline__ void first_inv_pass_regs(cx<T> *lds, int j1, int j2, cx<T> (&A)[R], cx<T> (&B)[R])
{
    dft<R>(A);
    dft<R>(B);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        lds[G::off(0, j1 * R + r)] = A[r];
        lds[G::off(0, j2 * R + r)] = B[r];
    }
    __syncthreads();
}

template <int... RS>
struct Radices {};

// the inverse passes after the first (radices RI..., NS from R), the last to HBM
template <typename T, int L, int NT, class G, int R, int... RI>
__device__ __forceinline__ void inverse_rest(cx<T> *row, cx<T> *lds, const cx<T> *ptw_inv, T scale, int tid,
                                             Radices<R, RI...>)
{
    passes_but_last<T, L, 1, NT, G, R, RI...>(lds, ptw_inv, tid);
    last_pass_global<T, L, NT, G, pass_tw_count<1, R, RI...>(), last_radix<RI...>()>(row, lds, ptw_inv, scale, tid);
}

// forward passes but the last (radices R1, RS..., the first from HBM)
template <typename T, int L, int M, int NT, class G, int R1, int... RS, class Src>
__device__ __forceinline__ void forward_but_last(cx<T> *lds, Src src, const cx<T> *ptw, int tid, Radices<R1, RS...>)
{
    first_pass_global<T, L, M, NT, G, R1>(lds, src, tid);
    passes_but_last<T, L, M, NT, G, R1, RS...>(lds, ptw, tid);
}

template <int R, int... RS>
constexpr int first_radix() { return R; }

// visit_pairs over two sequences' butterflies at once: visit(p, gp, gq, xp, xq)
template <typename T, int L, int R, class F>
__device__ __forceinline__ void visit_pairs2(cx<T> (&A)[R], cx<T> (&B)[R], cx<T> (&C)[R], cx<T> (&D)[R], int j1,
                                             F visit)
{
    constexpr int NS = L / R;
    if (j1 != 0) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int k = j1 + r * NS;
            if (k < L - k) visit(k, A[r], B[R - 1 - r], C[r], D[R - 1 - r]);
            else visit(L - k, B[R - 1 - r], A[r], D[R - 1 - r], C[r]);
        }
        return;
    }
    visit(0, A[0], A[0], C[0], C[0]);
    visit(L / 2, A[R / 2], A[R / 2], C[R / 2], C[R / 2]);
#pragma unroll
    for (int r = 1; r < R / 2; r++) visit(r * NS, A[r], A[R - r], C[r], C[R - r]);
#pragma unroll
    for (int r = 0; r < R / 2; r++) visit(NS / 2 + r * NS, B[r], B[R - 1 - r], D[r], D[R - 1 - r]);
}

// The fused filter with the pair pass: forward radices Fwd (last R), inverse
// radices Inv (first R); pass twiddle tables a.pass_tw (Fwd) and
// a.pass_tw_inv (Inv).
template <typename T, int L, int NT, class Fwd, class Inv>
struct PairFilter;
template <typename T, int L_, int NT_, int... RS, int... RI>
struct PairFilter<T, L_, NT_, Radices<RS...>, Radices<RI...>> {
    static constexpr int L = L_, NT = NT_;
    static constexpr int R = last_radix<RS...>();
    static constexpr int NS = L / R;
    static constexpr int TWN = pass_tw_count<1, RS...>();
    static_assert(NS == 2 * NT, "the pair pass runs two butterflies per thread");
    static_assert(first_radix<RI...>() == R, "the inverse starts with the forward's last radix");

    static __device__ __forceinline__ void forward(const FilterArgs &a)
    {
        using G = TileGeom<L, 1, LAY_SEQ, PAT_ROW, PAT_ROW>;
        __shared__ cx<T> lds[G::ELEMS];
        const int tid = threadIdx.x;
        const int64_t row = blockIdx.x;
        const cx<T> *x = reinterpret_cast<const cx<T> *>(a.x) + row * L;
        cx<T> *y = reinterpret_cast<cx<T> *>(a.y) + row * L;
        const cx<T> *H = reinterpret_cast<const cx<T> *>(a.H);
        const cx<T> *tw = reinterpret_cast<const cx<T> *>(a.tw);
        const cx<T> *ptw = reinterpret_cast<const cx<T> *>(a.pass_tw);
        forward_but_last<T, L, 1, NT, G>(lds, [&](int) { return x; }, ptw, tid, Radices<RS...>{});
        int j1, j2;
        pair_butterflies<NS>(tid, j1, j2);
        cx<T> A[R], B[R];
        last_fwd_pass_regs<T, L, G, R>(lds, 0, ptw + (TWN - NS * (R - 1)), j1, j2, A, B);
        // filter each pair and pack it for the inverse (stored swapped: the
        // inverse runs as the forward transform of the swapped data)
        visit_pairs2<T, L, R>(A, B, A, B, j1, [&](int p, cx<T> &zp, cx<T> &zq, cx<T> &, cx<T> &) {
            if (p == 0) {  // bins 0 and L are real; their filter taps' imaginary parts are dropped
                const T y0 = H[0].re * (zp.re + zp.im), yl = H[L].re * (zp.re - zp.im);
                zp = swap_ri(cx<T>{y0 + yl, y0 - yl});
                return;
            }
            const cx<T> w = tw[p];
            cx<T> fk, fj;
            unpack_pair(zp, zq, w, fk, fj);
            const cx<T> hk = H[p], hj = H[L - p];
            const cx<T> yk = cmul(a.conj_h ? cconj(hk) : hk, fk);
            const cx<T> yj = cmul(a.conj_h ? cconj(hj) : hj, fj);
            cx<T> zk, zj;
            pack_pair(yk, yj, w, zk, zj);
            zp = swap_ri(zk);
            if (p != L - p) zq = swap_ri(zj);
        });
        first_inv_pass_regs<T, G, R>(lds, j1, j2, A, B);
        inverse_rest<T, L, NT, G>(y, lds, reinterpret_cast<const cx<T> *>(a.pass_tw_inv), (T)1 / (T)(2 * L), tid,
                                  Radices<RI...>{});
    }

    static __device__ __forceinline__ void adjoint(const FilterArgs &a)
    {
        using G = TileGeom<L, 1, LAY_SEQ, PAT_ROW, PAT_ROW>;
        using G2 = TileGeom<L, 2, LAY_SEQ, PAT_ROW, PAT_ROW>;  // sequence 0 at G's offsets
        __shared__ cx<T> lds[G2::ELEMS];
        const int tid = threadIdx.x;
        const int64_t row = blockIdx.x;
        const cx<T> *gy = reinterpret_cast<const cx<T> *>(a.gy) + row * L;
        const cx<T> *x = reinterpret_cast<const cx<T> *>(a.x) + row * L;
        cx<T> *gx = reinterpret_cast<cx<T> *>(a.y) + row * L;
        cx<T> *part = reinterpret_cast<cx<T> *>(a.partial) + row * a.ppitch;
        const cx<T> *H = reinterpret_cast<const cx<T> *>(a.H);
        const cx<T> *tw = reinterpret_cast<const cx<T> *>(a.tw);
        const cx<T> *ptw = reinterpret_cast<const cx<T> *>(a.pass_tw);
        // rfft(gy) and rfft(x) as one two-sequence pass chain
        forward_but_last<T, L, 2, NT, G2>(lds, [&](int s) { return s ? x : gy; }, ptw, tid, Radices<RS...>{});
        int j1, j2;
        pair_butterflies<NS>(tid, j1, j2);
        cx<T> GA[R], GB[R], XA[R], XB[R];
        const cx<T> *twl = ptw + (TWN - NS * (R - 1));
        last_fwd_pass_regs<T, L, G2, R>(lds, 0, twl, j1, j2, GA, GB);
        last_fwd_pass_regs<T, L, G2, R>(lds, 1, twl, j1, j2, XA, XB);
        const T inv_n = (T)1 / (T)(2 * L);
        visit_pairs2<T, L, R>(GA, GB, XA, XB, j1, [&](int p, cx<T> &gp, cx<T> &gq, cx<T> &xp, cx<T> &xq) {
            if (p == 0) {  // bins 0 and L, real in both spectra; weight 1
                const T g0 = gp.re + gp.im, gl = gp.re - gp.im, x0 = xp.re + xp.im, xl = xp.re - xp.im;
                part[0] = {g0 * x0 * inv_n, (T)0};
                part[L] = {gl * xl * inv_n, (T)0};
                const T y0 = H[0].re * g0, yl = H[L].re * gl;
                gp = swap_ri(cx<T>{y0 + yl, y0 - yl});
                return;
            }
            const cx<T> w = tw[p];
            cx<T> gk, gj, xk, xj;
            unpack_pair(gp, gq, w, gk, gj);
            unpack_pair(xp, xq, w, xk, xj);
            part[p] = cscale(cmul(gk, cconj(xk)), 2 * inv_n);
            if (p != L - p) part[L - p] = cscale(cmul(gj, cconj(xj)), 2 * inv_n);
            const cx<T> yk = cmul(cconj(H[p]), gk), yj = cmul(cconj(H[L - p]), gj);
            cx<T> zk, zj;
            pack_pair(yk, yj, w, zk, zj);
            gp = swap_ri(zk);
            if (p != L - p) gq = swap_ri(zj);
        });
        first_inv_pass_regs<T, G, R>(lds, j1, j2, GA, GB);
        inverse_rest<T, L, NT, G>(gx, lds, reinterpret_cast<const cx<T> *>(a.pass_tw_inv), inv_n, tid,
                                  Radices<RI...>{});
    }
};

template <class P>
__global__ void __launch_bounds__(P::NT) filter_pair_fwd_kernel(FilterArgs a)
{
    P::forward(a);
}

template <class P>
__global__ void __launch_bounds__(P::NT) filter_pair_bwd_kernel(FilterArgs a)
{
    P::adjoint(a);
}

template <typename T, int L, int NT, int... RS>
__global__ void __launch_bounds__(NT) filter_fwd_kernel(FilterArgs a)
{
    using G = TileGeom<L, 1, LAY_SEQ, PAT_ROW, PAT_ROW>;
    __shared__ cx<T> lds[G::ELEMS];
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    const cx<T> *x = reinterpret_cast<const cx<T> *>(a.x) + row * L;
    cx<T> *y = reinterpret_cast<cx<T> *>(a.y) + row * L;
    const cx<T> *H = reinterpret_cast<const cx<T> *>(a.H);
    const cx<T> *tw = reinterpret_cast<const cx<T> *>(a.tw);
    const cx<T> *ptw = reinterpret_cast<const cx<T> *>(a.pass_tw);
    fft_from_global<T, L, 1, NT, G, RS...>(lds, [&](int) { return x; }, ptw, tid);
    // filter each pair of bins and pack for the inverse (stored swapped: the
    // second forward FFT then computes the unscaled inverse)
    for (int p = tid; p <= L / 2; p += NT) {
        const cx<T> w = tw[p];
        cx<T> fk, fj;
        row_bins<T, G>(lds, p, L, w, fk, fj);
        if (p == 0) {
            const T y0 = H[0].re * fk.re, yl = H[L].re * fj.re;  // imaginary parts of bins 0, L dropped
            lds[G::off(0, 0)] = swap_ri(cx<T>{y0 + yl, y0 - yl});
        } else {
            const cx<T> hk = H[p], hj = H[L - p];
            const cx<T> yk = cmul(a.conj_h ? cconj(hk) : hk, fk);
            const cx<T> yj = cmul(a.conj_h ? cconj(hj) : hj, fj);
            cx<T> zk, zj;
            pack_pair(yk, yj, w, zk, zj);
            lds[G::off(0, p)] = swap_ri(zk);
            if (L - p != p) lds[G::off(0, L - p)] = swap_ri(zj);
        }
    }
    __syncthreads();
    fft_to_global<T, L, NT, G, RS...>(y, lds, ptw, (T)1 / (T)(2 * L), tid);
}

template <typename T, int L, int NT, int... RS>
__global__ void __launch_bounds__(NT) filter_bwd_kernel(FilterArgs a)
{
    using G = TileGeom<L, 1, LAY_SEQ, PAT_ROW, PAT_ROW>;
    // gy and x are the two sequences of one tile, so their forward FFTs run
    // as one two-sequence pass chain: half the barriers, and the radix-16
    // passes of an 8192-point row keep all 1024 threads busy.  Sequence 0
    // has the same offsets in both geometries.
    using G2 = TileGeom<L, 2, LAY_SEQ, PAT_ROW, PAT_ROW>;
    __shared__ cx<T> lds[G2::ELEMS];
    cx<T> *gl = lds;           // rfft(gy), then the gx pass
    cx<T> *xl = lds + G2::SS;  // rfft(x)
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    const cx<T> *gy = reinterpret_cast<const cx<T> *>(a.gy) + row * L;
    const cx<T> *x = reinterpret_cast<const cx<T> *>(a.x) + row * L;
    cx<T> *gx = reinterpret_cast<cx<T> *>(a.y) + row * L;
    cx<T> *part = reinterpret_cast<cx<T> *>(a.partial) + row * a.ppitch;
    const cx<T> *H = reinterpret_cast<const cx<T> *>(a.H);
    const cx<T> *tw = reinterpret_cast<const cx<T> *>(a.tw);
    const cx<T> *ptw = reinterpret_cast<const cx<T> *>(a.pass_tw);
    fft_from_global<T, L, 2, NT, G2, RS...>(lds, [&](int s) { return s ? x : gy; }, ptw, tid);
    const T inv_n = (T)1 / (T)(2 * L);
    for (int p = tid; p <= L / 2; p += NT) {
        const cx<T> w = tw[p];
        cx<T> gk, gj, xk, xj;
        row_bins<T, G>(gl, p, L, w, gk, gj);
        row_bins<T, G>(xl, p, L, w, xk, xj);
        if (p == 0) {
            // bins 0 and L are real in both spectra; weight 1
            part[0] = {gk.re * xk.re * inv_n, (T)0};
            part[L] = {gj.re * xj.re * inv_n, (T)0};
            const T y0 = H[0].re * gk.re, yl = H[L].re * gj.re;
            gl[G::off(0, 0)] = swap_ri(cx<T>{y0 + yl, y0 - yl});
        } else {
            part[p] = cscale(cmul(gk, cconj(xk)), 2 * inv_n);
            if (L - p != p) part[L - p] = cscale(cmul(gj, cconj(xj)), 2 * inv_n);
            const cx<T> yk = cmul(cconj(H[p]), gk), yj = cmul(cconj(H[L - p]), gj);
            cx<T> zk, zj;
            pack_pair(yk, yj, w, zk, zj);
            gl[G::off(0, p)] = swap_ri(zk);
            if (L - p != p) gl[G::off(0, L - p)] = swap_ri(zj);
        }
    }
    __syncthreads();
    fft_to_global<T, L, NT, G, RS...>(gx, gl, ptw, inv_n, tid);
}

// dst[y * out_pitch + k] = sum of src[b * pitch + k] over the rows b of
// segment y (blockIdx.y).  A block covers 32 bins with GR row lanes (rows
// b = lane, lane + GR, ...); the lanes are combined in a fixed order, so the
// sum is deterministic.  32 row lanes (1024 threads) keep 16 waves per CU in
// flight over 256 x 8193 partials, where 8 lanes left one wave per SIMD
// waiting on its chain of loads.  The first pass may write in place into the
// first row of its own segment: only this block reads that region, and all
// of its reads precede the barrier.
constexpr int kReduceLanes = 32;

template <typename T>
__global__ void __launch_bounds__(32 * kReduceLanes) filter_reduce_kernel(ReduceArgs r)
{
    constexpr int GR = kReduceLanes;
    __shared__ cx<T> acc[GR][32];
    const int lane = threadIdx.x & 31, grp = threadIdx.x >> 5;
    const int64_t k = blockIdx.x * 32ll + lane;
    const int64_t b0 = (int64_t)blockIdx.y * r.seg;
    const int64_t b1 = b0 + r.seg < r.rows ? b0 + r.seg : r.rows;
    const cx<T> *src = reinterpret_cast<const cx<T> *>(r.src);
    cx<T> s = {(T)0, (T)0};
    if (k < r.bins) {
#pragma unroll 4
        for (int64_t b = b0 + grp; b < b1; b += GR) s = s + src[b * r.pitch + k];
    }
    acc[grp][lane] = s;
    __syncthreads();
    if (grp == 0 && k < r.bins) {
        cx<T> t = acc[0][lane];
#pragma unroll
        for (int g = 1; g < GR; g++) t = t + acc[g][lane];
        reinterpret_cast<cx<T> *>(r.dst)[(int64_t)blockIdx.y * r.out_pitch + k] = t;
    }
}

template <typename T>
struct FilterTable {
    int L, NT, nrad, rad[8];
    int nrad_inv, rad_inv[8];  // the inverse FFT's radices (the same order unless the pair pass runs)
    void (*fwd)(FilterArgs);
    void (*bwd)(FilterArgs);
};

template <typename T, int L, int NT, int... RS>
constexpr FilterTable<T> ft()
{
    FilterTable<T> e{};
    e.L = L;
    e.NT = NT;
    e.nrad = e.nrad_inv = sizeof...(RS);
    const int r[] = {RS...};
    for (int i = 0; i < (int)sizeof...(RS); i++) e.rad[i] = e.rad_inv[i] = r[i];
    e.fwd = &filter_fwd_kernel<T, L, NT, RS...>;
    e.bwd = &filter_bwd_kernel<T, L, NT, RS...>;
    return e;
}

// with the pair pass: forward radices RS..., inverse radices RI...
template <typename T, int L, int NT, class Fwd, class Inv>
struct FtPair;
template <typename T, int L, int NT, int... RS, int... RI>
struct FtPair<T, L, NT, Radices<RS...>, Radices<RI...>> {
    static constexpr FilterTable<T> make()
    {
        using P = PairFilter<T, L, NT, Radices<RS...>, Radices<RI...>>;
        FilterTable<T> e{};
        e.L = L;
        e.NT = NT;
        e.nrad = sizeof...(RS);
        e.nrad_inv = sizeof...(RI);
        const int r[] = {RS...}, ri[] = {RI...};
        for (int i = 0; i < (int)sizeof...(RS); i++) e.rad[i] = r[i];
        for (int i = 0; i < (int)sizeof...(RI); i++) e.rad_inv[i] = ri[i];
        e.fwd = &filter_pair_fwd_kernel<P>;
        e.bwd = &filter_pair_bwd_kernel<P>;
        return e;
    }
};

// The pair pass needs L = 2 NT times the forward's last radix.
const FilterTable<float> kF32[] = {
    ft<float, 32, 64, 8, 4>(),       ft<float, 64, 64, 8, 8>(),        ft<float, 128, 64, 16, 8>(),
    ft<float, 256, 64, 16, 16>(),    ft<float, 512, 64, 8, 8, 8>(),
    FtPair<float, 1024, 64, Radices<16, 8, 8>, Radices<8, 8, 16>>::make(),
    FtPair<float, 2048, 128, Radices<16, 16, 8>, Radices<8, 16, 16>>::make(),
    ft<float, 4096, 256, 16, 16, 16>(),
    // five passes that keep all 1024 threads busy (radix 16 over 8192 points
    // leaves half of them idle): fused forward + adjoint on MI355X 0.0492-0.0496
    // ms against 0.0527 for 16*16*8*4 / 4*8*16*16, 0.0504-0.0511 for
    // 8*8*4*8*4 and 0.0518 for 8*8*8*4*4
    FtPair<float, 8192, 1024, Radices<4, 8, 8, 8, 4>, Radices<4, 8, 8, 8, 4>>::make(),
};
const FilterTable<double> kF64[] = {
    ft<double, 32, 64, 8, 4>(),      ft<double, 64, 64, 8, 8>(),       ft<double, 128, 64, 16, 8>(),
    ft<double, 256, 64, 16, 16>(),   ft<double, 512, 64, 8, 8, 8>(),   ft<double, 1024, 64, 16, 16, 4>(),
    ft<double, 2048, 128, 16, 16, 8>(), ft<double, 4096, 256, 16, 16, 16>(),
};

template <typename T, size_t K>
const FilterTable<T> *lookup(const FilterTable<T> (&tab)[K], int L)
{
    for (size_t i = 0; i < K; i++)
        if (tab[i].L == L) return &tab[i];
    return nullptr;
}

}  // namespace

int filter_shape(int prec, int L, int *nt, int *rad, int inverse)
{
    auto take = [&](const auto *e) {
        if (!e) return 0;
        *nt = e->NT;
        const int n = inverse ? e->nrad_inv : e->nrad;
        for (int i = 0; i < n; i++) rad[i] = inverse ? e->rad_inv[i] : e->rad[i];
        return n;
    };
    return prec ? take(lookup(kF64, L)) : take(lookup(kF32, L));
}

hipError_t launch_filter(int prec, int which, const FilterArgs &a, hipStream_t s, hipEvent_t stop)
{
    const int L = (int)(a.bins - 1);
    void (*fn)(FilterArgs) = nullptr;
    int nt = 0;
    if (which == 2) {  // gH from the per-row partials: one or two reduction passes
        const int64_t bins = a.bins, rows = a.batch;
        const int64_t gx = (bins + 31) / 32;
        // row segments only while the bin columns alone give fewer than ~256
        // blocks: a second pass (its own launch) then adds the segment sums
        int64_t S = std::min<int64_t>((256 + gx - 1) / gx, (rows + 8 * kReduceLanes - 1) / (8 * kReduceLanes));
        if (S < 1) S = 1;
        const int64_t seg = (rows + S - 1) / S;
        S = (rows + seg - 1) / seg;
        ReduceArgs r{};
        r.src = a.partial;
        r.bins = bins;
        r.rows = rows;
        r.pitch = a.ppitch;
        r.seg = seg;
        if (S > 1) {
            r.dst = a.partial;  // segment sums into the first row of each segment
            r.out_pitch = seg * a.ppitch;
            if (prec) hipLaunchKernelGGL(filter_reduce_kernel<double>, dim3((unsigned)gx, (unsigned)S), dim3(32 * kReduceLanes), 0, s, r);
            else hipLaunchKernelGGL(filter_reduce_kernel<float>, dim3((unsigned)gx, (unsigned)S), dim3(32 * kReduceLanes), 0, s, r);
            r.rows = S;
            r.pitch = seg * a.ppitch;
            r.seg = S;
        }
        r.dst = a.y;
        r.out_pitch = 0;
        const void *fr = prec ? (const void *)&filter_reduce_kernel<double> : (const void *)&filter_reduce_kernel<float>;
        void *rargs[] = {&r};
        return hipExtLaunchKernel(fr, dim3((unsigned)gx, 1), dim3(32 * kReduceLanes), rargs, 0, s, nullptr, stop, 0);
    }
    if (prec) {
        const auto *e = lookup(kF64, L);
        if (!e) return hipErrorInvalidValue;
        fn = which ? e->bwd : e->fwd;
        nt = e->NT;
    } else {
        const auto *e = lookup(kF32, L);
        if (!e) return hipErrorInvalidValue;
        fn = which ? e->bwd : e->fwd;
        nt = e->NT;
    }
    FilterArgs args = a;
    void *kargs[] = {&args};
    return hipLaunchKernel((const void *)fn, dim3((unsigned)a.batch), dim3(nt), kargs, 0, s);
}

}  // namespace gfft
