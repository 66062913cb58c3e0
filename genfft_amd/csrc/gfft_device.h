// gfft_device.h -- device-side building blocks of the gfx950 FFT engine.
//
// One kernel template, `tile_fft_kernel`, does every complex pass:
//
//   * a workgroup owns a TILE of M independent sequences of length L;
//   * the tile is read from HBM with wide coalesced loads into LDS
//     (INP = ROW: a sequence's elements are contiguous; INP = COL: the M
//     sequences are contiguous, elements are `in_elem` apart);
//   * log_R(L) self-sorting Stockham radix-R passes run LDS -> registers ->
//     LDS (radix 2/4/8/16 butterflies in registers), so no bit reversal pass
//     exists anywhere;
//   * the last pass may multiply by the four-step twiddle W_Ntw^(c*k)
//     (generated from FP64 tables, rounded once);
//   * the tile is written back with coalesced stores (OUTP = ROW or COL).
//
// Batched 1D rows, the column pass of a four-step decomposition, the
// transposing row pass, and the 2D row / column passes are all instances.
// The inverse transform is the forward one with re/im swapped on the first
// load and the last store: swap(DFT(swap(x))) = unscaled IDFT(x).
//
// The math replaces genFFT's radix-2 DIT stages (generic/fft_impl_generic.h:36-120,
// x86/fft_float_impl_x86.inl:31-221): forward kernel exp(-2*pi*i*j*k/N),
// unscaled inverse, natural-order input and output.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfft {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * (-i)
template <typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a) { return {a.im, -a.re}; }
template <typename T>
__device__ __forceinline__ cx<T> swap_ri(cx<T> a) { return {a.im, a.re}; }

// ---------------------------------------------------------------------------
// In-register DFTs of size 2/4/8/16, forward sign, natural order in and out.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void dft2(cx<T> *v)
{
    cx<T> a = v[0];
    v[0] = a + v[1];
    v[1] = a - v[1];
}

template <typename T>
__device__ __forceinline__ void dft4(cx<T> &x0, cx<T> &x1, cx<T> &x2, cx<T> &x3)
{
    cx<T> t0 = x0 + x2, t1 = x0 - x2, t2 = x1 + x3, t3 = mul_mi(x1 - x3);
    x0 = t0 + t2;
    x2 = t0 - t2;
    x1 = t1 + t3;
    x3 = t1 - t3;
}

template <typename T>
__device__ __forceinline__ void dft4(cx<T> *v) { dft4(v[0], v[1], v[2], v[3]); }

template <typename T>
__device__ __forceinline__ void dft8(cx<T> *v)
{
    const T r = (T)0.70710678118654752440084436210484903928L;
    // decimation in frequency: a_n = x_n + x_{n+4}, b_n = (x_n - x_{n+4}) W8^n
    cx<T> a0 = v[0] + v[4], a1 = v[1] + v[5], a2 = v[2] + v[6], a3 = v[3] + v[7];
    cx<T> b0 = v[0] - v[4], b1 = v[1] - v[5], b2 = v[2] - v[6], b3 = v[3] - v[7];
    b1 = {r * (b1.re + b1.im), r * (b1.im - b1.re)};
    b2 = mul_mi(b2);
    b3 = {r * (b3.im - b3.re), -r * (b3.re + b3.im)};
    dft4(a0, a1, a2, a3);
    dft4(b0, b1, b2, b3);
    v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
    v[1] = b0; v[3] = b1; v[5] = b2; v[7] = b3;
}

template <typename T>
__device__ __forceinline__ void dft16(cx<T> *v)
{
    const T c1 = (T)0.92387953251128675612818318939678828682L;  // cos(pi/8)
    const T s1 = (T)0.38268343236508977172845998403039886676L;  // sin(pi/8)
    const T r = (T)0.70710678118654752440084436210484903928L;
    // n = 4*n1 + n2, k = k1 + 4*k2
    cx<T> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        a[n2][0] = v[n2];
        a[n2][1] = v[n2 + 4];
        a[n2][2] = v[n2 + 8];
        a[n2][3] = v[n2 + 12];
        dft4(a[n2]);
    }
    // a[n2][k1] *= W16^(n2*k1)
    a[1][1] = cmul(a[1][1], cx<T>{c1, -s1});
    a[1][2] = {r * (a[1][2].re + a[1][2].im), r * (a[1][2].im - a[1][2].re)};
    a[1][3] = cmul(a[1][3], cx<T>{s1, -c1});
    a[2][1] = {r * (a[2][1].re + a[2][1].im), r * (a[2][1].im - a[2][1].re)};
    a[2][2] = mul_mi(a[2][2]);
    a[2][3] = {r * (a[2][3].im - a[2][3].re), -r * (a[2][3].re + a[2][3].im)};
    a[3][1] = cmul(a[3][1], cx<T>{s1, -c1});
    a[3][2] = {r * (a[3][2].im - a[3][2].re), -r * (a[3][2].re + a[3][2].im)};
    a[3][3] = cmul(a[3][3], cx<T>{-c1, s1});
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        cx<T> x0 = a[0][k1], x1 = a[1][k1], x2 = a[2][k1], x3 = a[3][k1];
        dft4(x0, x1, x2, x3);
        v[k1] = x0;
        v[k1 + 4] = x1;
        v[k1 + 8] = x2;
        v[k1 + 12] = x3;
    }
}

// 32-point DFT: n = 2*n1 + n2, two 16-point DFTs over n1, the twiddle
// W32^(n2*k1), then radix-2 across n2 (k = k1 + 16*k2).
template <typename T>
__device__ __forceinline__ void dft32(cx<T> *v)
{
    // cos(pi*k/16), sin(pi*k/16), k = 0..8 (the rest by symmetry)
    constexpr long double C[9] = {1.0L,
                                  0.98078528040323044912618223613423903697L,
                                  0.92387953251128675612818318939678828682L,
                                  0.83146961230254523707878837761790575673L,
                                  0.70710678118654752440084436210484903928L,
                                  0.55557023301960222474283081394853287438L,
                                  0.38268343236508977172845998403039886676L,
                                  0.19509032201612826784828486847702224093L,
                                  0.0L};
    cx<T> a[16], b[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++) {
        a[n1] = v[2 * n1];
        b[n1] = v[2 * n1 + 1];
    }
    dft16(a);
    dft16(b);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
        // W32^k1 = (cos(pi k1/16), -sin(pi k1/16)); sin(pi k/16) = cos(pi (8-k)/16)
        const T c = k1 <= 8 ? (T)C[k1] : -(T)C[16 - k1];
        const T sn = k1 <= 8 ? (T)C[8 - k1] : (T)C[k1 - 8];
        cx<T> t = b[k1];
        if (k1 == 8) t = mul_mi(t);
        else if (k1 != 0) t = cmul(t, cx<T>{c, -sn});
        v[k1] = a[k1] + t;
        v[k1 + 16] = a[k1] - t;
    }
}

template <int R, typename T>
__device__ __forceinline__ void dft(cx<T> *v)
{
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 16) dft16(v);
    else if constexpr (R == 32) dft32(v);
    else static_assert(R == 2 || R == 4 || R == 8 || R == 16 || R == 32, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Tile geometry and kernel arguments
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ilog2c(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

enum : int { PAT_ROW = 0, PAT_COL = 1 };
enum : int { LAY_SEQ = 0, LAY_ELEM = 1 };
enum : int { TW_NONE = 0, TW_INNER = 1, TW_OUTER = 2 };

struct TileArgs {
    const void *in;
    void *out;
    // sequence (o2, o1, c) starts at o2*in_outer2 + o1*in_outer + c*in_inner;
    // its element n sits in_elem further on (strides in complex elements)
    int64_t in_outer2, in_outer, in_inner, in_elem;
    int64_t out_outer2, out_outer, out_inner, out_elem;
    int32_t n_inner;      // sequences along the inner index c
    int32_t tiles_inner;  // ceil(n_inner / M)
    int32_t n_outer1;     // range of o1
    int32_t n_tiles;      // tiles_inner * n_outer1 * n_outer2
    int32_t swap_in, swap_out;
    int32_t vec_in, vec_out;  // 16-byte vector access allowed (alignment proven on host)
    int32_t xcd_remap;    // map blocks so consecutive tiles run on one XCD
    // persistent kernels' tile order: 0 interleaved (tile, tile + grid, ...),
    // -1 one contiguous run per workgroup, K > 1 (a power of two) runs of K
    // consecutive tiles dealt out to the workgroups in turn
    int32_t runs;
    int32_t tw_mode;      // TW_NONE / TW_INNER (c * k) / TW_OUTER (o1 * k)
    uint32_t tw_mask;     // Ntw - 1
    int32_t tw_lo_bits;   // two-level split of the exponent
    // Optional job tables (uneven batches): when job_in / job_out is set, the
    // transform o1 of this launch reads from a.in + job_in[job_first + o1]
    // (writes to a.out + job_out[job_first + o1]) instead of o1 * in_outer,
    // and job_inv[job_first + o1] selects the inverse where job_swap_* is set.
    const int64_t *job_in;
    const int64_t *job_out;
    const uint8_t *job_inv;
    int32_t job_first;
    int32_t job_swap_in, job_swap_out;
    const void *pass_tw;  // Stockham pass twiddles, complex T
    const double *tw_lo;  // complex double [1 << tw_lo_bits]: W_Ntw^m
    const double *tw_hi;  // complex double [Ntw >> tw_lo_bits]: W_Ntw^(m << tw_lo_bits)
    // Two-level FP64 table of W_L, b = (log2 L + 1) / 2: gen_lo[m] = W_L^m
    // (m < 2^b), gen_hi[h] = W_L^(h << b), for pipelined kernels whose last
    // pass builds its twiddles instead of reading pass_tw (PassTw::GEN)
    const double *gen_lo;
    const double *gen_hi;
    // Per-transform output scale (16-bit storage kernels): sequence starting
    // at element offset q of this launch's input belongs to transform
    // (scale_off + q) / scale_div, and its stored values are multiplied by
    // scale[] of that transform.  Null: no scaling.
    const float *scale;
    int64_t scale_div, scale_off;
};

// FP64 twiddle W_Ntw^m from the two-level table.
__device__ __forceinline__ cx<double> tw_fp64(const TileArgs &a, uint32_t m)
{
    m &= a.tw_mask;
    const cx<double> *lo = reinterpret_cast<const cx<double> *>(a.tw_lo);
    const cx<double> *hi = reinterpret_cast<const cx<double> *>(a.tw_hi);
    cx<double> l = lo[m & ((1u << a.tw_lo_bits) - 1u)];
    cx<double> h = hi[m >> a.tw_lo_bits];
    return cmul(h, l);
}

template <typename T>
__device__ __forceinline__ cx<T> apply_tw(cx<T> v, cx<double> w)
{
    if constexpr (sizeof(T) == 8) {
        return cmul(v, w);
    } else {
        cx<T> wt = {(T)w.re, (T)w.im};
        return cmul(v, wt);
    }
}

// ---------------------------------------------------------------------------
// One Stockham pass of radix R on every sequence of the tile, in LDS.
//   NS = product of the radices of the previous passes.
//   butterfly (s, j), j < L/R: inputs j + r*L/R, twiddle W_{NS*R}^(r*(j%NS)),
//   outputs (j - j%NS)*R + j%NS + r*NS.
// ---------------------------------------------------------------------------
// Four-step twiddle bases a persistent tile kernel loads ahead of time: the
// last pass then needs W^(c*(j0 + i*NT/M + r*L/R)) = hj * hd^i * hr^r (FP64,
// each base already the product of its two table entries).
struct Tw4Pre {
    cx<double> hj, hd, hr;
};

// GEN (complex64, last pass only, NS * R == L): the pass builds W_L^(r*m)
// from the two-level FP64 table `gen` (lo entries, then hi entries; see
// TileArgs::gen_lo) instead of reading `tw`.  The base W_L^m is one FP64
// product, its powers an FP64 geometric progression (relative error
// ~R * 2^-53), each rounded once to FP32, so the result matches the rounded
// table entry except within ~1e-15 of a rounding boundary.
template <typename T, int L, int M, int NT, class G, bool FASTJ, int NS, int R, bool LAST, bool PRE, bool GEN = false>
__device__ __forceinline__ void stockham_pass(cx<T> *lds, const cx<T> *__restrict__ tw, int tid,
                                              const TileArgs &a, int s0, int o, const Tw4Pre &pre,
                                              const cx<double> *gen = nullptr)
{
    constexpr int LR = L / R;
    constexpr int BF = M * LR;
    constexpr int PER = (BF + NT - 1) / NT;
    cx<T> v[PER][R];
    int sj[PER], jj[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * NT;
        int s, j;
        if constexpr (FASTJ) {
            j = b % LR;
            s = b / LR;
        } else {
            s = b % M;
            j = b / M;
        }
        sj[i] = s;
        jj[i] = j;
        if (BF % NT == 0 || b < BF) {
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = lds[G::off(s, j + r * LR)];
        }
    }
    // GEN: the twiddles are built and applied one at a time after the
    // barrier (no array of R - 1 of them stays live beside the prefetched
    // tile); complex64 computes each butterfly's FP64 base here
    constexpr int LB = (ilog2c(L) + 1) / 2;
    const cx<double> *glo = gen, *ghi = gen + (1 << LB);
    cx<double> gb[GEN ? PER : 1];
    cx<T> w[GEN ? 1 : PER][R > 1 && !GEN ? R - 1 : 1];
    if constexpr (GEN) {
        static_assert(NS > 1 && NS * R == L, "generated twiddles serve the last pass");
        static_assert(sizeof(T) == 4, "generated pass twiddles are a complex64 form");
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int m = jj[i] & (NS - 1);
            gb[i] = cmul(glo[m & ((1 << LB) - 1)], ghi[m >> LB]);
        }
    } else if constexpr (NS > 1) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int m = jj[i] & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; r++) w[i][r - 1] = tw[m * (R - 1) + (r - 1)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * NT;
        if (BF % NT == 0 || b < BF) {
            const int s = sj[i], j = jj[i];
            const int m = j & (NS - 1);
            if constexpr (GEN) {
                cx<double> wk = gb[i];
                v[i][1] = cmul(v[i][1], cx<T>{(T)wk.re, (T)wk.im});
#pragma unroll
                for (int r = 2; r < R; r++) {
                    wk = cmul(wk, gb[i]);
                    v[i][r] = cmul(v[i][r], cx<T>{(T)wk.re, (T)wk.im});
                }
            } else if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[i][r] = cmul(v[i][r], w[i][r - 1]);
            }
            dft<R>(v[i]);
            if constexpr (LAST && PRE) {
                static_assert(!FASTJ && NT % M == 0, "prefetched twiddles need the s-fast mapping");
                if (a.tw_mode != TW_NONE) {
                    // i-th butterfly of this thread: j = j0 + i*NT/M
                    cx<double> wi = pre.hj;
                    const cx<double> wd = pre.hd, wr = pre.hr;
                    for (int q = 0; q < i; q++) wi = cmul(wi, wd);
                    cx<double> wk = wi;
                    v[i][0] = apply_tw(v[i][0], wk);
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        wk = cmul(wk, wr);
                        v[i][r] = apply_tw(v[i][r], wk);
                    }
                }
            }
            if constexpr (LAST && !PRE) {
                // four-step twiddle W_Ntw^(c*k), k = j + r*L/R (here NS*R == L)
                if (a.tw_mode != TW_NONE) {
                    const uint32_t c = a.tw_mode == TW_INNER ? (uint32_t)(s0 + s) : (uint32_t)o;
                    if constexpr (sizeof(T) == 4) {
                        // FP64 geometric progression, rounded once per element
                        cx<double> wk = tw_fp64(a, c * (uint32_t)j);
                        const cx<double> ws = tw_fp64(a, c * (uint32_t)LR);
                        v[i][0] = apply_tw(v[i][0], wk);
#pragma unroll
                        for (int r = 1; r < R; r++) {
                            wk = cmul(wk, ws);
                            v[i][r] = apply_tw(v[i][r], wk);
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < R; r++)
                            v[i][r] = apply_tw(v[i][r], tw_fp64(a, c * (uint32_t)(j + r * LR)));
                    }
                }
            }
            const int base = (j - m) * R + m;
#pragma unroll
            for (int r = 0; r < R; r++) lds[G::off(s, base + r * NS)] = v[i][r];
        }
    }
    __syncthreads();
}

// The passes of a radix sequence; with GEN the last one builds its twiddles
// from `gen` (stockham_pass) and `tw` holds only the earlier passes' tables.
template <typename T, int L, int M, int NT, class G, bool FASTJ, bool PRE, bool GEN, int NS, int R, int... RS>
__device__ __forceinline__ void stockham_passes_g(cx<T> *lds, const cx<T> *__restrict__ tw, int tid,
                                                  const TileArgs &a, int s0, int o, const Tw4Pre &pre,
                                                  const cx<double> *gen)
{
    constexpr bool last = sizeof...(RS) == 0;
    stockham_pass<T, L, M, NT, G, FASTJ, NS, R, last, PRE, last && GEN>(lds, tw, tid, a, s0, o, pre, gen);
    if constexpr (!last) {
        constexpr int used = NS > 1 ? NS * (R - 1) : 0;
        stockham_passes_g<T, L, M, NT, G, FASTJ, PRE, GEN, NS * R, RS...>(lds, tw + used, tid, a, s0, o, pre, gen);
    }
}

template <typename T, int L, int M, int NT, class G, bool FASTJ, bool PRE, int NS, int R, int... RS>
__device__ __forceinline__ void stockham_passes(cx<T> *lds, const cx<T> *__restrict__ tw, int tid,
                                                const TileArgs &a, int s0, int o, const Tw4Pre &pre)
{
    stockham_passes_g<T, L, M, NT, G, FASTJ, PRE, false, NS, R, RS...>(lds, tw, tid, a, s0, o, pre, nullptr);
}

template <int R, int... RS>
constexpr int last_radix()
{
    if constexpr (sizeof...(RS) == 0) return R;
    else return last_radix<RS...>();
}

// number of Stockham pass twiddles of a radix sequence
template <int NS, int R, int... RS>
constexpr int pass_tw_count()
{
    constexpr int used = NS > 1 ? NS * (R - 1) : 0;
    if constexpr (sizeof...(RS) == 0) return used;
    else return used + pass_tw_count<NS * R, RS...>();
}

// ---------------------------------------------------------------------------
// HBM <-> LDS tile movement
// ---------------------------------------------------------------------------
// element offset of transform (o2, o1) and its direction swaps
__device__ __forceinline__ int64_t in_base(const TileArgs &a, int o2, int o)
{
    return a.job_in ? a.job_in[a.job_first + o] : (int64_t)o * a.in_outer + (int64_t)o2 * a.in_outer2;
}
__device__ __forceinline__ int64_t out_base(const TileArgs &a, int o2, int o)
{
    return a.job_out ? a.job_out[a.job_first + o] : (int64_t)o * a.out_outer + (int64_t)o2 * a.out_outer2;
}
__device__ __forceinline__ bool swap_in_of(const TileArgs &a, int o)
{
    return a.job_swap_in ? a.job_inv[a.job_first + o] != 0 : a.swap_in != 0;
}
__device__ __forceinline__ bool swap_out_of(const TileArgs &a, int o)
{
    return a.job_swap_out ? a.job_inv[a.job_first + o] != 0 : a.swap_out != 0;
}

// The 16-byte accesses of a tile: access `it` of thread `tid` covers
// elements e, e + 1 with e = 2 tid + 2 NT it, i.e. element n of sequence s
// (ROW: s = e / L, n = e % L; COL: n = e / M, s = e % M).  With L, M and NT
// powers of two its offset splits into a part that depends on tid alone and
// a part that depends on `it` alone, so every access is one uniform 64-bit
// base (scalar arithmetic) plus the same 32-bit per-thread byte offset: no
// per-access 64-bit multiply.  The host enables the vector path only where
// every in-tile byte offset fits 31 bits.
template <int L, int M, int NT, int PAT>
struct VecWalk {
    static_assert((L & (L - 1)) == 0 && (M & (M - 1)) == 0 && (NT & (NT - 1)) == 0, "powers of two");
    // element offset of (s, n) = s * inner + n (ROW) or n * elem + s (COL)
    static __device__ __forceinline__ uint32_t thread_off(int tid, uint32_t inner, uint32_t elem)
    {
        const uint32_t e = 2u * (uint32_t)tid;
        if constexpr (PAT == PAT_ROW) {
            if constexpr ((2 * NT) % L == 0) return (e / L) * inner + e % L;
            else return e;  // L is a multiple of 2 NT: a thread stays inside one sequence
        } else {
            if constexpr ((2 * NT) % M == 0) return (e / M) * elem + e % M;
            else return e;  // M is a multiple of 2 NT: a thread stays inside one element row
        }
    }
    static __device__ __forceinline__ int64_t iter_off(int it, int64_t inner, int64_t elem)
    {
        const int e = 2 * NT * it;
        if constexpr (PAT == PAT_ROW) return (int64_t)(e / L) * inner + e % L;
        else return (int64_t)(e / M) * elem + e % M;
    }
};

template <typename T, int L, int M, int NT, class G, int INP>
__device__ __forceinline__ void load_tile(cx<T> *lds, const TileArgs &a, int o2, int o, int s0,
                                          int nvalid, int tid)
{
    constexpr int TOT = M * L;
    const bool sw = swap_in_of(a, o);
    const cx<T> *in = reinterpret_cast<const cx<T> *>(a.in) + in_base(a, o2, o);
    if constexpr (INP == PAT_ROW) in += (int64_t)s0 * a.in_inner;
    else in += s0;

    if constexpr (sizeof(T) == 4 && TOT % (2 * NT) == 0 && (INP == PAT_ROW ? L : M) % 2 == 0) {
        if (a.vec_in && nvalid == M) {
            using VW = VecWalk<L, M, NT, INP>;
            const char *b = reinterpret_cast<const char *>(in);
            const uint32_t to = VW::thread_off(tid, (uint32_t)a.in_inner, (uint32_t)a.in_elem) * sizeof(cx<T>);
#pragma unroll
            for (int it = 0; it < TOT / (2 * NT); it++) {
                const int e = (tid + it * NT) * 2;
                const float4 q = *reinterpret_cast<const float4 *>(
                    b + VW::iter_off(it, a.in_inner, a.in_elem) * (int64_t)sizeof(cx<T>) + to);
                const cx<T> x0 = {q.x, q.y}, x1 = {q.z, q.w};
                if constexpr (INP == PAT_ROW) {
                    lds[G::off(e / L, e % L)] = sw ? swap_ri(x0) : x0;
                    lds[G::off(e / L, e % L + 1)] = sw ? swap_ri(x1) : x1;
                } else {
                    lds[G::off(e % M, e / M)] = sw ? swap_ri(x0) : x0;
                    lds[G::off(e % M + 1, e / M)] = sw ? swap_ri(x1) : x1;
                }
            }
            return;
        }
    }
    // Element by element (complex128, odd strides, partial tiles): every
    // load is issued into registers before the first LDS write.  Written
    // interleaved, the compiler orders each load after the previous LDS
    // write (the input pointer is generic) and the loads run one at a time.
    constexpr int IT = (TOT + NT - 1) / NT;
    cx<T> r[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int e = tid + it * NT;
        const int s = INP == PAT_ROW ? e / L : e % M;
        const int n = INP == PAT_ROW ? e % L : e / M;
        if ((TOT % NT == 0 || e < TOT) && s < nvalid)
            r[it] = INP == PAT_ROW ? in[(int64_t)s * a.in_inner + n] : in[(int64_t)n * a.in_elem + s];
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int e = tid + it * NT;
        const int s = INP == PAT_ROW ? e / L : e % M;
        const int n = INP == PAT_ROW ? e % L : e / M;
        if ((TOT % NT == 0 || e < TOT) && s < nvalid) lds[G::off(s, n)] = sw ? swap_ri(r[it]) : r[it];
    }
}

template <typename T, int L, int M, int NT, class G, int OUTP>
__device__ __forceinline__ void store_tile(const cx<T> *lds, const TileArgs &a, int o2, int o,
                                           int s0, int nvalid, int tid)
{
    constexpr int TOT = M * L;
    const bool sw = swap_out_of(a, o);
    cx<T> *out = reinterpret_cast<cx<T> *>(a.out) + out_base(a, o2, o);
    if constexpr (OUTP == PAT_ROW) out += (int64_t)s0 * a.out_inner;
    else out += s0;

    if constexpr (sizeof(T) == 4 && TOT % (2 * NT) == 0 && (OUTP == PAT_ROW ? L : M) % 2 == 0) {
        if (a.vec_out && nvalid == M) {
            using VW = VecWalk<L, M, NT, OUTP>;
            char *b = reinterpret_cast<char *>(out);
            const uint32_t to = VW::thread_off(tid, (uint32_t)a.out_inner, (uint32_t)a.out_elem) * sizeof(cx<T>);
#pragma unroll
            for (int it = 0; it < TOT / (2 * NT); it++) {
                const int e = (tid + it * NT) * 2;
                cx<T> x0, x1;
                if constexpr (OUTP == PAT_ROW) {
                    x0 = lds[G::off(e / L, e % L)];
                    x1 = lds[G::off(e / L, e % L + 1)];
                } else {
                    x0 = lds[G::off(e % M, e / M)];
                    x1 = lds[G::off(e % M + 1, e / M)];
                }
                float4 *p = reinterpret_cast<float4 *>(b + VW::iter_off(it, a.out_inner, a.out_elem) * (int64_t)sizeof(cx<T>) + to);
                if (sw) {
                    x0 = swap_ri(x0);
                    x1 = swap_ri(x1);
                }
                // Streaming (nontemporal) stores where each contiguous run
                // of the tile covers whole 128-byte lines: nothing in this
                // launch reads the output again.  Measured on c64 2^20 x 128
                // (128-byte runs): transposing row pass 0.392 -> 0.380 ms,
                // column pass 0.441 -> 0.435 ms on one MI355X, no change on
                // another.  Narrower runs keep plain stores, so the L2 can
                // merge neighbouring tiles' parts of a line: streamed, the
                // 64-byte runs of c64 2^22 x 32 wrote 15-19% more than their
                // bytes and the 32-byte runs of a 4096-point column pass 54%.
                // (The choice is made at compile time: a runtime choice
                // between the two stores of one value is merged by the
                // compiler into one plain store.)
                typedef float f4v __attribute__((ext_vector_type(4)));
                const f4v q = {x0.re, x0.im, x1.re, x1.im};
                constexpr bool kStream = (OUTP == PAT_ROW ? L : M) * (int)sizeof(cx<T>) >= 128;
                if constexpr (kStream) __builtin_nontemporal_store(q, reinterpret_cast<f4v *>(p));
                else *reinterpret_cast<f4v *>(p) = q;
            }
            return;
        }
    }
#pragma unroll 4
    for (int it = 0; it < (TOT + NT - 1) / NT; it++) {
        const int e = tid + it * NT;
        if (TOT % NT == 0 || e < TOT) {
            int s, k;
            if constexpr (OUTP == PAT_ROW) {
                s = e / L;
                k = e % L;
            } else {
                k = e / M;
                s = e % M;
            }
            if (s < nvalid) {
                cx<T> x = lds[G::off(s, k)];
                cx<T> *p = OUTP == PAT_ROW ? out + (int64_t)s * a.out_inner + k
                                           : out + (int64_t)k * a.out_elem + s;
                *p = sw ? swap_ri(x) : x;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 16-bit storage (FP16 / BF16): a complex element is one 32-bit word, the
// real part in the low half.  Values are widened to FP32 on load, all
// arithmetic runs in FP32 in LDS, and they are rounded once (to nearest
// even) on store.
// ---------------------------------------------------------------------------
struct st_f16 {};
struct st_bf16 {};

template <class S>
struct HalfFmt;
template <>
struct HalfFmt<st_f16> {
    static __device__ __forceinline__ float dec(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)h); }
    static __device__ __forceinline__ uint32_t enc(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
};
template <>
struct HalfFmt<st_bf16> {
    static __device__ __forceinline__ float dec(uint32_t h) { return __uint_as_float(h << 16); }
    // round to nearest even on the 16 dropped bits; a NaN stays a (quiet) NaN
    static __device__ __forceinline__ uint32_t enc(float f)
    {
        const uint32_t u = __float_as_uint(f);
        if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    }
};
template <class S>
__device__ __forceinline__ cx<float> hdec(uint32_t w)
{
    return {HalfFmt<S>::dec(w & 0xffffu), HalfFmt<S>::dec(w >> 16)};
}
template <class S>
__device__ __forceinline__ uint32_t henc(cx<float> v)
{
    return HalfFmt<S>::enc(v.re) | (HalfFmt<S>::enc(v.im) << 16);
}

// load_tile for 16-bit storage: 16-byte vectors carry four elements of one
// sequence (rows) or of four adjacent sequences (columns).
template <class S, int L, int M, int NT, class G, int INP>
__device__ __forceinline__ void load_tile_h(cx<float> *lds, const TileArgs &a, int o2, int o, int s0, int nvalid,
                                            int tid)
{
    constexpr int TOT = M * L;
    const bool sw = swap_in_of(a, o);
    const uint32_t *in = reinterpret_cast<const uint32_t *>(a.in) + in_base(a, o2, o);
    if constexpr (INP == PAT_ROW) in += (int64_t)s0 * a.in_inner;
    else in += s0;
    if constexpr (TOT % (4 * NT) == 0 && (INP == PAT_ROW ? L : M) % 4 == 0) {
        if (a.vec_in && nvalid == M) {
#pragma unroll
            for (int it = 0; it < TOT / (4 * NT); it++) {
                const int e = (tid + it * NT) * 4;
                if constexpr (INP == PAT_ROW) {
                    const int s = e / L, n = e % L;
                    const uint4 q = *reinterpret_cast<const uint4 *>(in + (int64_t)s * a.in_inner + n);
                    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const cx<float> x = hdec<S>(w[v]);
                        lds[G::off(s, n + v)] = sw ? swap_ri(x) : x;
                    }
                } else {
                    const int n = e / M, s = e % M;
                    const uint4 q = *reinterpret_cast<const uint4 *>(in + (int64_t)n * a.in_elem + s);
                    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const cx<float> x = hdec<S>(w[v]);
                        lds[G::off(s + v, n)] = sw ? swap_ri(x) : x;
                    }
                }
            }
            return;
        }
    }
#pragma unroll 4
    for (int it = 0; it < (TOT + NT - 1) / NT; it++) {
        const int e = tid + it * NT;
        if (TOT % NT == 0 || e < TOT) {
            const int s = INP == PAT_ROW ? e / L : e % M;
            const int n = INP == PAT_ROW ? e % L : e / M;
            if (s < nvalid) {
                const cx<float> x =
                    hdec<S>(INP == PAT_ROW ? in[(int64_t)s * a.in_inner + n] : in[(int64_t)n * a.in_elem + s]);
                lds[G::off(s, n)] = sw ? swap_ri(x) : x;
            }
        }
    }
}

// store_tile for 16-bit storage; `scl` (LDS, one factor per sequence of the
// tile) multiplies the values before rounding when non-null.
template <class S, int L, int M, int NT, class G, int OUTP>
__device__ __forceinline__ void store_tile_h(const cx<float> *lds, const TileArgs &a, int o2, int o, int s0,
                                             int nvalid, int tid, const float *scl)
{
    constexpr int TOT = M * L;
    const bool sw = swap_out_of(a, o);
    uint32_t *out = reinterpret_cast<uint32_t *>(a.out) + out_base(a, o2, o);
    if constexpr (OUTP == PAT_ROW) out += (int64_t)s0 * a.out_inner;
    else out += s0;
    auto get = [&](int s, int k) {
        cx<float> x = lds[G::off(s, k)];
        if (scl) x = {x.re * scl[s], x.im * scl[s]};
        return henc<S>(sw ? swap_ri(x) : x);
    };
    if constexpr (TOT % (4 * NT) == 0 && (OUTP == PAT_ROW ? L : M) % 4 == 0) {
        if (a.vec_out && nvalid == M) {
            // plain stores: with the streaming hint each 16-bit pass ran faster
            // alone, the plan did not (FP16 2^20 x 64: the row pass then found
            // the 256 MiB intermediate outside the Infinity Cache)
#pragma unroll
            for (int it = 0; it < TOT / (4 * NT); it++) {
                const int e = (tid + it * NT) * 4;
                if constexpr (OUTP == PAT_ROW) {
                    const int s = e / L, k = e % L;
                    *reinterpret_cast<uint4 *>(out + (int64_t)s * a.out_inner + k) =
                        uint4{get(s, k), get(s, k + 1), get(s, k + 2), get(s, k + 3)};
                } else {
                    const int k = e / M, s = e % M;
                    *reinterpret_cast<uint4 *>(out + (int64_t)k * a.out_elem + s) =
                        uint4{get(s, k), get(s + 1, k), get(s + 2, k), get(s + 3, k)};
                }
            }
            return;
        }
    }
#pragma unroll 4
    for (int it = 0; it < (TOT + NT - 1) / NT; it++) {
        const int e = tid + it * NT;
        if (TOT % NT == 0 || e < TOT) {
            const int s = OUTP == PAT_ROW ? e / L : e % M;
            const int k = OUTP == PAT_ROW ? e % L : e / M;
            if (s < nvalid) {
                const uint32_t w = get(s, k);
                if (OUTP == PAT_ROW) out[(int64_t)s * a.out_inner + k] = w;
                else out[(int64_t)k * a.out_elem + s] = w;
            }
        }
    }
}

// LDS geometry of a tile: element n of sequence s lives at G::off(s, n).
//   LAY_SEQ : sequence-major.  One pad element after every 16 elements of a
//             sequence breaks the stride-R conflicts of the Stockham writes;
//             the sequence pitch is odd so walks across sequences (COL side
//             of a transposing pass) hit distinct banks.
//   LAY_ELEM: element-major (the M sequences of one element index adjacent),
//             pitch padded by one when an HBM side walks along a sequence.
template <int L, int M, int LAYOUT, int INP, int OUTP>
struct TileGeom {
    static constexpr int LP = L + (L >> 4);                       // padded sequence length
    static constexpr int SS = LAYOUT == LAY_SEQ ? (LP | 1) : 1;   // sequence pitch
    static constexpr int MS = M + ((INP == PAT_ROW || OUTP == PAT_ROW) ? 1 : 0);
    static constexpr int ELEMS = LAYOUT == LAY_SEQ ? M * SS : L * MS;
    static __device__ __forceinline__ int off(int s, int n)
    {
        if constexpr (LAYOUT == LAY_SEQ) return s * SS + n + (n >> 4);
        else return n * MS + s;
    }
};

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8),
// each XCD with its own L2.  This permutation hands XCD x a contiguous run
// of tiles, so neighbouring tiles -- which share 128-byte lines when a tile
// is narrower than a line -- are fetched into one L2 at about the same time.
__device__ __forceinline__ int xcd_tile(int b, int grid)
{
    constexpr int XCDS = 8;
    const int full = grid - grid % XCDS;
    if (b >= full) return b;
    return (b % XCDS) * (full / XCDS) + b / XCDS;
}

template <typename T, int L, int M, int NT, int LAYOUT, int INP, int OUTP, int... RS>
__global__ void __launch_bounds__(NT) tile_fft_kernel(TileArgs a)
{
    using G = TileGeom<L, M, LAYOUT, INP, OUTP>;
    __shared__ cx<T> lds[G::ELEMS];
    const int tid = threadIdx.x;
    const int tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int t2 = tile / a.tiles_inner;
    const int ti = tile - t2 * a.tiles_inner;
    const int o2 = t2 / a.n_outer1;
    const int o = t2 - o2 * a.n_outer1;
    const int s0 = ti * M;
    const int nvalid = min(M, a.n_inner - s0);
    load_tile<T, L, M, NT, G, INP>(lds, a, o2, o, s0, nvalid, tid);
    __syncthreads();
    stockham_passes<T, L, M, NT, G, LAYOUT == LAY_SEQ, false, 1, RS...>(
        lds, reinterpret_cast<const cx<T> *>(a.pass_tw), tid, a, s0, o, Tw4Pre{});
    store_tile<T, L, M, NT, G, OUTP>(lds, a, o2, o, s0, nvalid, tid);
}

// tile_fft_kernel over 16-bit storage (FP32 arithmetic).  The launch that
// reads the plan's input applies the per-transform scale on its store.
template <class S, int L, int M, int NT, int LAYOUT, int INP, int OUTP, int... RS>
__global__ void __launch_bounds__(NT) tile_fft_kernel_h(TileArgs a)
{
    using T = float;
    using G = TileGeom<L, M, LAYOUT, INP, OUTP>;
    __shared__ cx<T> lds[G::ELEMS];
    __shared__ float scl[M];
    const int tid = threadIdx.x;
    const int tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int t2 = tile / a.tiles_inner;
    const int ti = tile - t2 * a.tiles_inner;
    const int o2 = t2 / a.n_outer1;
    const int o = t2 - o2 * a.n_outer1;
    const int s0 = ti * M;
    const int nvalid = min(M, a.n_inner - s0);
    load_tile_h<S, L, M, NT, G, INP>(lds, a, o2, o, s0, nvalid, tid);
    if (a.scale) {
        const int64_t base = in_base(a, o2, o);
        for (int s = tid; s < nvalid; s += NT) {
            const int64_t q = base + (INP == PAT_ROW ? (int64_t)(s0 + s) * a.in_inner : (int64_t)(s0 + s));
            scl[s] = a.scale[(a.scale_off + q) / a.scale_div];
        }
    }
    __syncthreads();
    stockham_passes<T, L, M, NT, G, LAYOUT == LAY_SEQ, false, 1, RS...>(
        lds, reinterpret_cast<const cx<T> *>(a.pass_tw), tid, a, s0, o, Tw4Pre{});
    store_tile_h<S, L, M, NT, G, OUTP>(lds, a, o2, o, s0, nvalid, tid, a.scale ? scl : nullptr);
}

// ---------------------------------------------------------------------------
// Persistent, software-pipelined form of tile_fft_kernel for tiles that fill
// most of a CU's LDS (one or two workgroups per CU): while tile t runs its
// Stockham passes and stores, the loads of the workgroup's next tile are
// already in flight into registers.  Pass twiddles live in LDS (no global
// loads inside the passes, so waiting on them never waits on the prefetch),
// and the four-step twiddle bases of tile t are issued before the prefetch
// of tile t+1 for the same reason (vector-memory counters drain in order).
// ---------------------------------------------------------------------------
// Pass twiddles of a pipelined tile are copied to LDS when the tile and the
// table fit a CU's 160 KB together.  Otherwise (a single 16384-point
// complex64 sequence: 136 KB of tile, 128 KB of table, 120 KB of it the last
// pass's) the last pass generates its twiddles (PassTw::GEN) from a two-level
// FP64 table of W_L, 2^b + L/2^b entries (4 KB at L = 16384), kept in LDS with
// the earlier passes' tables.  Reading them from global memory instead, each
// wait on a twiddle load also waits on the next tile's prefetch issued before
// it (the vector-memory counter drains in order), which undoes the pipeline.
template <typename T, int ELEMS, int TWN>
constexpr bool pass_tw_in_lds()
{
    return (ELEMS + (TWN > 0 ? TWN : 1)) * sizeof(cx<T>) <= 160 * 1024;
}

// LDS layout of a pipelined tile's pass twiddles, in complex T units:
// TWL every table; GEN the generation table (16-byte aligned, at offset 0)
// then the tables of every pass but the last.
template <typename T, int L, int ELEMS, int... RS>
struct PassTw {
    static constexpr int TWN = pass_tw_count<1, RS...>();
    static constexpr int RL = last_radix<RS...>();
    static constexpr int LAST = sizeof...(RS) > 1 ? L / RL * (RL - 1) : 0;
    static constexpr int LB = (ilog2c(L) + 1) / 2;
    static constexpr int GEN_D = (1 << LB) + (L >> LB);  // complex double entries
    static constexpr int GEN_U = GEN_D * (int)(sizeof(cx<double>) / sizeof(cx<T>));
    static constexpr bool TWL = pass_tw_in_lds<T, ELEMS, TWN>();
    static constexpr bool GEN = sizeof(T) == 4 && !TWL && LAST > 0 &&
                                (ELEMS + 1 + GEN_U + TWN - LAST) * sizeof(cx<T>) <= 160 * 1024;
    static constexpr int UNITS = TWL ? (TWN > 0 ? TWN : 1) : GEN ? GEN_U + (TWN - LAST > 0 ? TWN - LAST : 1) : 1;
};

// Thread index of a persistent loop, rebuilt on every iteration from the
// wave's index (uniform, kept in an SGPR) and a lane id that v_mbcnt
// computes afresh.  The volatile asm cannot be hoisted, so no VGPR stays
// live across the loop to hold the index and nothing derived from it is
// kept either: with tid itself live, the column kernel of the 2^20
// four-step spilled it, and the scratch reload at the loop head waited for
// every outstanding store and prefetch (vmcnt counts all of them, in order)
// before the tile's four-step twiddle loads were even issued.
__device__ __forceinline__ int fresh_tid(int wave)
{
    int lane;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
    return (wave << 6) + lane;
}

template <typename T, int L, int M, int NT, class G, int INP>
__device__ __forceinline__ bool fetch_tile(cx<T> (&reg)[M * L / NT], const TileArgs &a, int o2, int o, int s0,
                                           int nvalid, int tid)
{
    constexpr int TOT = M * L;
    const cx<T> *in = reinterpret_cast<const cx<T> *>(a.in) + in_base(a, o2, o);
    if constexpr (INP == PAT_ROW) in += (int64_t)s0 * a.in_inner;
    else in += s0;
    if constexpr (sizeof(T) == 4 && TOT % (2 * NT) == 0 && (INP == PAT_ROW ? L : M) % 2 == 0) {
        if (a.vec_in && nvalid == M) {
            using VW = VecWalk<L, M, NT, INP>;
            const char *b = reinterpret_cast<const char *>(in);
            const uint32_t to = VW::thread_off(tid, (uint32_t)a.in_inner, (uint32_t)a.in_elem) * sizeof(cx<T>);
#pragma unroll
            for (int it = 0; it < TOT / (2 * NT); it++) {
                const float4 q = *reinterpret_cast<const float4 *>(
                    b + VW::iter_off(it, a.in_inner, a.in_elem) * (int64_t)sizeof(cx<T>) + to);
                reg[2 * it] = {q.x, q.y};
                reg[2 * it + 1] = {q.z, q.w};
            }
            return true;
        }
    }
#pragma unroll
    for (int it = 0; it < TOT / NT; it++) {
        const int e = tid + it * NT;
        const int s = INP == PAT_ROW ? e / L : e % M;
        const int n = INP == PAT_ROW ? e % L : e / M;
        if (s < nvalid)
            reg[it] = INP == PAT_ROW ? in[(int64_t)s * a.in_inner + n] : in[(int64_t)n * a.in_elem + s];
    }
    return false;
}

template <typename T, int L, int M, int NT, class G, int INP>
__device__ __forceinline__ void put_tile(cx<T> *lds, const cx<T> (&reg)[M * L / NT], bool vec, const TileArgs &a,
                                         int o, int nvalid, int tid)
{
    constexpr int TOT = M * L;
    const bool sw = swap_in_of(a, o);
    if (vec) {
#pragma unroll
        for (int it = 0; it < TOT / (2 * NT); it++) {
            const int e = (tid + it * NT) * 2;
            const cx<T> x0 = sw ? swap_ri(reg[2 * it]) : reg[2 * it];
            const cx<T> x1 = sw ? swap_ri(reg[2 * it + 1]) : reg[2 * it + 1];
            if constexpr (INP == PAT_ROW) {
                lds[G::off(e / L, e % L)] = x0;
                lds[G::off(e / L, e % L + 1)] = x1;
            } else {
                lds[G::off(e % M, e / M)] = x0;
                lds[G::off(e % M + 1, e / M)] = x1;
            }
        }
        return;
    }
#pragma unroll
    for (int it = 0; it < TOT / NT; it++) {
        const int e = tid + it * NT;
        const int s = INP == PAT_ROW ? e / L : e % M;
        const int n = INP == PAT_ROW ? e % L : e / M;
        if (s < nvalid) lds[G::off(s, n)] = sw ? swap_ri(reg[it]) : reg[it];
    }
}

// LDS footprint of a pipelined tile shape: the tile, then the pass twiddles
// (PassTw).
template <typename T, int L, int M, int NT, int LAYOUT, int INP, int OUTP, int... RS>
struct PipeLds {
    using G = TileGeom<L, M, LAYOUT, INP, OUTP>;
    using P = PassTw<T, L, G::ELEMS, RS...>;
    static constexpr int TILE = G::ELEMS;
    static constexpr int TILE_A = (G::ELEMS + 1) & ~1;  // offset of the twiddles in one array (16-byte aligned)
    static constexpr int TW = P::UNITS;
};

// The persistent loop over a launch's tiles, as workgroup `block` of `grid`
// workgroups working on them (a kernel may hand other workgroups other work).
// `lds` holds the tile and `twl` (16-byte aligned) the pass twiddles (PipeLds).
template <typename T, int L, int M, int NT, int LAYOUT, int INP, int OUTP, int... RS>
__device__ __forceinline__ void pipe_tiles(const TileArgs &a, int block, int grid, cx<T> *lds, cx<T> *twl)
{
    using G = TileGeom<L, M, LAYOUT, INP, OUTP>;
    using P = PassTw<T, L, G::ELEMS, RS...>;
    constexpr int E = M * L / NT;
    constexpr int LAST_R = last_radix<RS...>();
    const int tid = threadIdx.x;
    const cx<T> *ptw = reinterpret_cast<const cx<T> *>(a.pass_tw);
    const cx<double> *gen = nullptr;
    if constexpr (P::TWL) {
        for (int i = tid; i < P::TWN; i += NT) twl[i] = ptw[i];
        ptw = twl;
    } else if constexpr (P::GEN) {
        cx<double> *g = reinterpret_cast<cx<double> *>(twl);
        const cx<double> *lo = reinterpret_cast<const cx<double> *>(a.gen_lo);
        const cx<double> *hi = reinterpret_cast<const cx<double> *>(a.gen_hi);
        constexpr int NLO = 1 << P::LB;
        for (int i = tid; i < P::GEN_D; i += NT) g[i] = i < NLO ? lo[i] : hi[i - NLO];
        for (int i = tid; i < P::TWN - P::LAST; i += NT) twl[P::GEN_U + i] = ptw[i];
        ptw = twl + P::GEN_U;
        gen = g;
    }
    const int slot = a.xcd_remap ? xcd_tile(block, grid) : block;
    auto decode = [&](int t, int &o2, int &o, int &s0, int &nv) {
        const int t2 = t / a.tiles_inner;
        const int ti = t - t2 * a.tiles_inner;
        o2 = t2 / a.n_outer1;
        o = t2 - o2 * a.n_outer1;
        s0 = ti * M;
        nv = min(M, a.n_inner - s0);
    };
    cx<T> reg[E];
    // Tile order.  In runs of consecutive tiles, the tile a workgroup
    // prefetches is the neighbour (adjacent sequences of the same transform)
    // of the one it stores and both touch the same HBM pages: complex64 /
    // complex128 column passes (128-byte row segments) run ~5-10% faster.
    // Interleaved, consecutive tiles go to different workgroups and run at
    // the same time, which suits the transposing row pass, whose neighbours
    // then fill whole output rows together.
    const int K = a.runs > 0 ? a.runs : 1;  // run length of the dealt-out order
    int tile = a.runs < 0 ? (int)((int64_t)slot * a.n_tiles / grid) : slot * K;
    const int t_end = a.runs < 0 ? (int)((int64_t)(slot + 1) * a.n_tiles / grid) : a.n_tiles;
    auto advance = [&](int t) { return a.runs < 0 || ((t + 1) & (K - 1)) ? t + 1 : t + 1 + (grid - 1) * K; };
    int o2 = 0, o = 0, s0 = 0, nvalid = 0;
    bool vec = false;
    if (tile < t_end) {
        decode(tile, o2, o, s0, nvalid);
        vec = fetch_tile<T, L, M, NT, G, INP>(reg, a, o2, o, s0, nvalid, tid);
    }
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    while (tile < t_end) {
        // The thread index is rebuilt every iteration: otherwise the compiler
        // hoists every tid-derived LDS offset of every pass out of the loop
        // and keeps them live, which costs ~100 VGPRs.  Its range, which the
        // asm hides, is restated so index divisions stay unsigned shifts.
        const int t = fresh_tid(wave);
        __builtin_assume(t >= 0 && t < NT);
        // four-step twiddle bases of this tile (s-fast mapping: s = t % M):
        // issued first, so their latency hides behind the wait for the
        // prefetched tile data that put_tile needs anyway
        cx<double> hj{}, lj{}, hd{}, ld{}, hr{}, lr{};
        const bool tw = LAYOUT == LAY_ELEM && NT % M == 0 && a.tw_mode != TW_NONE;
        if (tw) {
            const uint32_t c = a.tw_mode == TW_INNER ? (uint32_t)(s0 + t % M) : (uint32_t)o;
            const cx<double> *lo = reinterpret_cast<const cx<double> *>(a.tw_lo);
            const cx<double> *hi = reinterpret_cast<const cx<double> *>(a.tw_hi);
            const uint32_t lm = (1u << a.tw_lo_bits) - 1u;
            const uint32_t mj = (c * (uint32_t)(t / M)) & a.tw_mask;
            const uint32_t md = (c * (uint32_t)(NT / M)) & a.tw_mask;
            const uint32_t mr = (c * (uint32_t)(L / LAST_R)) & a.tw_mask;
            hj = hi[mj >> a.tw_lo_bits];
            lj = lo[mj & lm];
            hd = hi[md >> a.tw_lo_bits];
            ld = lo[md & lm];
            hr = hi[mr >> a.tw_lo_bits];
            lr = lo[mr & lm];
        }
        put_tile<T, L, M, NT, G, INP>(lds, reg, vec, a, o, nvalid, t);
        Tw4Pre pre{};
        if (tw) {
            pre.hj = cmul(hj, lj);
            pre.hd = cmul(hd, ld);
            pre.hr = cmul(hr, lr);
        }
        __syncthreads();
        // prefetch the next tile
        const int next = advance(tile);
        int no2 = 0, no = 0, ns0 = 0, nnv = 0;
        if (next < t_end) {
            decode(next, no2, no, ns0, nnv);
            vec = fetch_tile<T, L, M, NT, G, INP>(reg, a, no2, no, ns0, nnv, t);
        }
        stockham_passes_g<T, L, M, NT, G, LAYOUT == LAY_SEQ, (LAYOUT == LAY_ELEM && NT % M == 0), P::GEN, 1, RS...>(
            lds, ptw, t, a, s0, o, pre, gen);
        store_tile<T, L, M, NT, G, OUTP>(lds, a, o2, o, s0, nvalid, t);
        __syncthreads();
        tile = next;
        o2 = no2;
        o = no;
        s0 = ns0;
        nvalid = nnv;
    }
}

template <typename T, int L, int M, int NT, int LAYOUT, int INP, int OUTP, int... RS>
__global__ void __launch_bounds__(NT) tile_fft_pipe_kernel(TileArgs a)
{
    using S = PipeLds<T, L, M, NT, LAYOUT, INP, OUTP, RS...>;
    __shared__ cx<T> lds[S::TILE];
    __shared__ __align__(16) cx<T> twl[S::TW];
    pipe_tiles<T, L, M, NT, LAYOUT, INP, OUTP, RS...>(a, (int)blockIdx.x, (int)gridDim.x, lds, twl);
}

// A pipelined tile shape as one type (for kernels that combine two).
template <typename T_, int L, int M, int NT_, int LAYOUT, int INP, int OUTP, int... RS>
struct PipeShape {
    using T = T_;
    static constexpr int NT = NT_;
    using S = PipeLds<T_, L, M, NT_, LAYOUT, INP, OUTP, RS...>;
    static constexpr int ELEMS = S::TILE_A + S::TW;
    static __device__ __forceinline__ void run(const TileArgs &a, int block, int grid, cx<T_> *smem)
    {
        pipe_tiles<T_, L, M, NT_, LAYOUT, INP, OUTP, RS...>(a, block, grid, smem, smem + S::TILE_A);
    }
};

// Two tile passes of different shapes in one launch: workgroups [0, ga) run
// pass A's tiles, the rest run pass B's.  The four-step transform runs its
// batch in chunks through a ring of two scratch slots, and each launch
// pairs the column pass of chunk k with the transposing row pass of chunk
// k - 1: the two passes' HBM streams (reads of the input, writes of the
// output, the scratch ring in between) overlap instead of following each
// other.  The choice of role is uniform per workgroup, so the barriers
// inside each role's loop see every thread of the workgroup.
template <class A, class B>
__global__ void __launch_bounds__(A::NT) dual_pipe_kernel(TileArgs a, TileArgs b, int ga)
{
    static_assert(A::NT == B::NT, "both roles run with the kernel's thread count");
    static_assert(sizeof(typename A::T) == sizeof(typename B::T), "one element type");
    using T = typename A::T;
    constexpr int E = A::ELEMS > B::ELEMS ? A::ELEMS : B::ELEMS;
    __shared__ __align__(16) cx<T> smem[E];
    if ((int)blockIdx.x < ga) A::run(a, (int)blockIdx.x, ga, smem);
    else B::run(b, (int)blockIdx.x - ga, (int)gridDim.x - ga, smem);
}

// ---------------------------------------------------------------------------
// Persistent pipelined form over 16-bit storage: the same loop as
// tile_fft_pipe_kernel with the next tile prefetched as packed 32-bit words
// (half the registers of complex64) and the per-transform scale of
// tile_fft_kernel_h.
// ---------------------------------------------------------------------------
template <class S, int L, int M, int NT, class G, int INP>
__device__ __forceinline__ bool fetch_tile_h(uint32_t (&reg)[M * L / NT], const TileArgs &a, int o2, int o, int s0,
                                             int nvalid, int tid)
{
    constexpr int TOT = M * L;
    const uint32_t *in = reinterpret_cast<const uint32_t *>(a.in) + in_base(a, o2, o);
    if constexpr (INP == PAT_ROW) in += (int64_t)s0 * a.in_inner;
    else in += s0;
    if constexpr (TOT % (4 * NT) == 0 && (INP == PAT_ROW ? L : M) % 4 == 0) {
        if (a.vec_in && nvalid == M) {
            // Streaming (nontemporal) loads: each tile is read once, and the
            // hint keeps the Infinity Cache for the intermediate the next
            // pass reads and the L2 for merging the 64-byte output runs.
            // Measured (FP16 2^20 x 64, 256 MiB intermediate): 0.422 ->
            // 0.382-0.388 ms per step, the row pass writing 257.8 MiB per
            // 256 instead of 269.9.  The complex64 tiles with narrow output
            // runs got slower with it (c64 2^22 x 32: 1.080 -> 1.171 ms).
            typedef unsigned u4v __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int it = 0; it < TOT / (4 * NT); it++) {
                const int e = (tid + it * NT) * 4;
                const u4v *p;
                if constexpr (INP == PAT_ROW) p = reinterpret_cast<const u4v *>(in + (int64_t)(e / L) * a.in_inner + e % L);
                else p = reinterpret_cast<const u4v *>(in + (int64_t)(e / M) * a.in_elem + e % M);
                const u4v q = __builtin_nontemporal_load(p);
                reg[4 * it] = q.x;
                reg[4 * it + 1] = q.y;
                reg[4 * it + 2] = q.z;
                reg[4 * it + 3] = q.w;
            }
            return true;
        }
    }
#pragma unroll
    for (int it = 0; it < TOT / NT; it++) {
        const int e = tid + it * NT;
        const int s = INP == PAT_ROW ? e / L : e % M;
        const int n = INP == PAT_ROW ? e % L : e / M;
        if (s < nvalid) reg[it] = INP == PAT_ROW ? in[(int64_t)s * a.in_inner + n] : in[(int64_t)n * a.in_elem + s];
    }
    return false;
}

template <class S, int L, int M, int NT, class G, int INP>
__device__ __forceinline__ void put_tile_h(cx<float> *lds, const uint32_t (&reg)[M * L / NT], bool vec,
                                           const TileArgs &a, int o, int nvalid, int tid)
{
    constexpr int TOT = M * L;
    const bool sw = swap_in_of(a, o);
    if constexpr (TOT % (4 * NT) == 0 && (INP == PAT_ROW ? L : M) % 4 == 0) {
        if (vec) {
#pragma unroll
            for (int it = 0; it < TOT / (4 * NT); it++) {
                const int e = (tid + it * NT) * 4;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const cx<float> x = hdec<S>(reg[4 * it + v]);
                    if constexpr (INP == PAT_ROW) lds[G::off(e / L, e % L + v)] = sw ? swap_ri(x) : x;
                    else lds[G::off(e % M + v, e / M)] = sw ? swap_ri(x) : x;
                }
            }
            return;
        }
    }
#pragma unroll
    for (int it = 0; it < TOT / NT; it++) {
        const int e = tid + it * NT;
        const int s = INP == PAT_ROW ? e / L : e % M;
        const int n = INP == PAT_ROW ? e % L : e / M;
        if (s < nvalid) {
            const cx<float> x = hdec<S>(reg[it]);
            lds[G::off(s, n)] = sw ? swap_ri(x) : x;
        }
    }
}

template <class S, int L, int M, int NT, int LAYOUT, int INP, int OUTP, int... RS>
__global__ void __launch_bounds__(NT) tile_fft_pipe_kernel_h(TileArgs a)
{
    using T = float;
    using G = TileGeom<L, M, LAYOUT, INP, OUTP>;
    constexpr int TWN = pass_tw_count<1, RS...>();
    constexpr int E = M * L / NT;
    constexpr int LAST_R = last_radix<RS...>();
    constexpr bool TWL = pass_tw_in_lds<T, G::ELEMS + M / 2, TWN>();
    __shared__ cx<T> lds[G::ELEMS];
    __shared__ cx<T> twl[TWL && TWN > 0 ? TWN : 1];
    __shared__ float scl[M];
    const int tid = threadIdx.x;
    const cx<T> *ptw = reinterpret_cast<const cx<T> *>(a.pass_tw);
    if constexpr (TWL) {
        for (int i = tid; i < TWN; i += NT) twl[i] = ptw[i];
        ptw = twl;
    }
    const int grid = gridDim.x;
    const int slot = a.xcd_remap ? xcd_tile(blockIdx.x, grid) : (int)blockIdx.x;
    auto decode = [&](int t, int &o2, int &o, int &s0, int &nv) {
        const int t2 = t / a.tiles_inner;
        const int ti = t - t2 * a.tiles_inner;
        o2 = t2 / a.n_outer1;
        o = t2 - o2 * a.n_outer1;
        s0 = ti * M;
        nv = min(M, a.n_inner - s0);
    };
    uint32_t reg[E];
    // Tile order.  In runs of consecutive tiles, the tile a workgroup
    // prefetches is the neighbour (adjacent sequences of the same transform)
    // of the one it stores and both touch the same HBM pages: complex64 /
    // complex128 column passes (128-byte row segments) run ~5-10% faster.
    // Interleaved, consecutive tiles go to different workgroups and run at
    // the same time, which suits the transposing row pass, whose neighbours
    // then fill whole output rows together.
    const int K = a.runs > 0 ? a.runs : 1;  // run length of the dealt-out order
    int tile = a.runs < 0 ? (int)((int64_t)slot * a.n_tiles / grid) : slot * K;
    const int t_end = a.runs < 0 ? (int)((int64_t)(slot + 1) * a.n_tiles / grid) : a.n_tiles;
    auto advance = [&](int t) { return a.runs < 0 || ((t + 1) & (K - 1)) ? t + 1 : t + 1 + (grid - 1) * K; };
    int o2 = 0, o = 0, s0 = 0, nvalid = 0;
    bool vec = false;
    if (tile < t_end) {
        decode(tile, o2, o, s0, nvalid);
        vec = fetch_tile_h<S, L, M, NT, G, INP>(reg, a, o2, o, s0, nvalid, tid);
    }
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    while (tile < t_end) {
        // thread index rebuilt per iteration (see tile_fft_pipe_kernel)
        const int t = fresh_tid(wave);
        __builtin_assume(t >= 0 && t < NT);
        cx<double> hj{}, lj{}, hd{}, ld{}, hr{}, lr{};
        const bool tw = LAYOUT == LAY_ELEM && NT % M == 0 && a.tw_mode != TW_NONE;
        if (tw) {
            const uint32_t c = a.tw_mode == TW_INNER ? (uint32_t)(s0 + t % M) : (uint32_t)o;
            const cx<double> *lo = reinterpret_cast<const cx<double> *>(a.tw_lo);
            const cx<double> *hi = reinterpret_cast<const cx<double> *>(a.tw_hi);
            const uint32_t lm = (1u << a.tw_lo_bits) - 1u;
            const uint32_t mj = (c * (uint32_t)(t / M)) & a.tw_mask;
            const uint32_t md = (c * (uint32_t)(NT / M)) & a.tw_mask;
            const uint32_t mr = (c * (uint32_t)(L / LAST_R)) & a.tw_mask;
            hj = hi[mj >> a.tw_lo_bits];
            lj = lo[mj & lm];
            hd = hi[md >> a.tw_lo_bits];
            ld = lo[md & lm];
            hr = hi[mr >> a.tw_lo_bits];
            lr = lo[mr & lm];
        }
        put_tile_h<S, L, M, NT, G, INP>(lds, reg, vec, a, o, nvalid, t);
        if (a.scale) {
            const int64_t base = in_base(a, o2, o);
            for (int s = t; s < nvalid; s += NT) {
                const int64_t q = base + (INP == PAT_ROW ? (int64_t)(s0 + s) * a.in_inner : (int64_t)(s0 + s));
                scl[s] = a.scale[(a.scale_off + q) / a.scale_div];
            }
        }
        Tw4Pre pre{};
        if (tw) {
            pre.hj = cmul(hj, lj);
            pre.hd = cmul(hd, ld);
            pre.hr = cmul(hr, lr);
        }
        __syncthreads();
        const int next = advance(tile);
        int no2 = 0, no = 0, ns0 = 0, nnv = 0;
        if (next < t_end) {
            decode(next, no2, no, ns0, nnv);
            vec = fetch_tile_h<S, L, M, NT, G, INP>(reg, a, no2, no, ns0, nnv, t);
        }
        stockham_passes<T, L, M, NT, G, LAYOUT == LAY_SEQ, (LAYOUT == LAY_ELEM && NT % M == 0), 1, RS...>(
            lds, ptw, t, a, s0, o, pre);
        store_tile_h<S, L, M, NT, G, OUTP>(lds, a, o2, o, s0, nvalid, t, a.scale ? scl : nullptr);
        __syncthreads();
        tile = next;
        o2 = no2;
        o = no;
        s0 = ns0;
        nvalid = nnv;
    }
}

}  // namespace gfft
