// gfft_wave.h -- 1024-point complex64 FFT held by one wavefront, no LDS.
//
// A sequence lives in the 64 lanes x 16 registers of one wave: element
// n = lane + 64 r starts in register r of lane `lane`.  The three radix
// stages (16, 16, 4) exchange data between lanes with the cross-lane
// instructions of gfx950 instead of LDS round trips and workgroup barriers:
//   * DPP row shifts with bank masks (lane bits 2 and 3),
//   * v_permlane16_swap / v_permlane32_swap (lane bits 4 and 5),
//   * DPP quad permutations (lane bits 0 and 1, as radix-2 butterflies).
//
// With lane = l (6 bits) and the output index k = k1 + 16 c + 256 d:
//   1. DFT-16 over r in registers (r: bits 6..9 of n) -> k1; twiddle W_1024^(l k1);
//   2. swap lane bits 2..5 with register bits 0..3 (four transposes), so the
//      registers hold l_hi = l >> 2 and lane bits 2..5 hold k1;
//      DFT-16 over l_hi in registers -> c; twiddle W_64^(l_lo c), l_lo = l & 3;
//   3. DFT-4 over l_lo across lanes (radix-2 over lane bit 1, W_4 on lanes
//      with l & 3 == 3, radix-2 over lane bit 0) -> d = bit 1 + 2 * bit 0.
// Output X[k1 + 16 c + 256 d] ends in register c of lane
// (k1 << 2) | ((d & 1) << 1) | (d >> 1).
//
// Twiddles are generated in FP64 and rounded once: the bases W_1024^l and
// W_64^(l_lo) = W_1024^(16 l_lo) are products of two entries of the two-level
// FP64 table of W_1024 (TileArgs::gen_lo / gen_hi, 32 + 32 entries), and their
// powers an FP64 geometric progression.  A table of FP32 twiddles read from
// global memory instead (30 loads per lane, each waited on where the compiler
// placed it) ran the kernel 1.5x slower.
#pragma once
#include "gfft_device.h"

namespace gfft {

// Transpose step: swap lane bit LB (mask 1 << LB) with register bit RB for
// every register pair (i, i | 1 << RB).  In lanes whose bit LB is clear,
// register i keeps its value and register i | 1 << RB receives the partner
// lane's register i; in lanes whose bit is set, register i | 1 << RB keeps
// its value and register i receives the partner's register i | 1 << RB.
template <int LB, int RB>
__device__ __forceinline__ void wave_swap_bits(cx<float> (&v)[16])
{
    constexpr int RM = 1 << RB;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (i & RM) continue;
        float *p0 = &v[i].re, *p1 = &v[i | RM].re;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int a = __float_as_int(p0[h]), b = __float_as_int(p1[h]);
            int n0, n1;
            if constexpr (LB == 2 || LB == 3) {
                // DPP within 16-lane rows: row_shr:m reads lane - m, row_shl:m
                // reads lane + m; bank masks (4-lane groups) pick the lanes whose
                // bit LB is set (0xA for bit 2, 0xC for bit 3) or clear
                constexpr int SHR = LB == 2 ? 0x114 : 0x118, SHL = LB == 2 ? 0x104 : 0x108;
                constexpr int SET = LB == 2 ? 0xA : 0xC, CLR = LB == 2 ? 0x5 : 0x3;
                n0 = __builtin_amdgcn_update_dpp(a, b, SHR, 0xF, SET, false);
                n1 = __builtin_amdgcn_update_dpp(b, a, SHL, 0xF, CLR, false);
            } else if constexpr (LB == 4) {
                // odd 16-lane rows of the first operand <-> even rows of the second
                const auto r = __builtin_amdgcn_permlane16_swap((unsigned)a, (unsigned)b, false, false);
                n0 = (int)r[0];
                n1 = (int)r[1];
            } else {
                static_assert(LB == 5, "lane bits 2..5 only");
                // upper 32 lanes of the first operand <-> lower 32 lanes of the second
                const auto r = __builtin_amdgcn_permlane32_swap((unsigned)a, (unsigned)b, false, false);
                n0 = (int)r[0];
                n1 = (int)r[1];
            }
            p0[h] = __int_as_float(n0);
            p1[h] = __int_as_float(n1);
        }
    }
}

// Radix-2 butterfly across lane bit LB (0 or 1) on every register: the lane
// whose bit is clear keeps own + partner, the other partner - own.
template <int LB>
__device__ __forceinline__ void wave_bfly_lanes(cx<float> (&v)[16], float sgn)
{
    constexpr int QP = LB == 0 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float pr = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v[i].re), QP, 0xF, 0xF, false));
        const float pi = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v[i].im), QP, 0xF, 0xF, false));
        v[i].re = fmaf(sgn, v[i].re, pr);
        v[i].im = fmaf(sgn, v[i].im, pi);
    }
}

// v[k] *= b^k for k = 1..15: an FP64 geometric progression (relative error
// ~15 * 2^-53), each power rounded once to FP32
__device__ __forceinline__ void wave_twiddle(cx<float> (&v)[16], cx<double> b)
{
    cx<double> w = b;
#pragma unroll
    for (int k = 1; k < 16; k++) {
        if (k > 1) w = cmul(w, b);
        v[k] = cmul(v[k], cx<float>{(float)w.re, (float)w.im});
    }
}

// The twiddle bases of lane `lane`, W_1024^lane and W_64^(lane & 3) =
// W_1024^(16 (lane & 3)), each the product of two entries of the two-level
// FP64 table of W_1024 (32 + 32 entries).
struct WaveTw {
    cx<double> l1, h1, l2, h2;
};
__device__ __forceinline__ WaveTw wave_tw_load(const double *gen_lo, const double *gen_hi, int lane)
{
    const cx<double> *lo = reinterpret_cast<const cx<double> *>(gen_lo);
    const cx<double> *hi = reinterpret_cast<const cx<double> *>(gen_hi);
    const int m2 = 16 * (lane & 3);
    return WaveTw{lo[lane & 31], hi[lane >> 5], lo[m2 & 31], hi[m2 >> 5]};
}

// Forward 1024-point DFT of the wave's sequence (layout above).
__device__ __forceinline__ void wave_fft1024(cx<float> (&v)[16], int lane, const WaveTw &t)
{
    dft<16>(v);
    wave_twiddle(v, cmul(t.l1, t.h1));
    wave_swap_bits<2, 0>(v);
    wave_swap_bits<3, 1>(v);
    wave_swap_bits<4, 2>(v);
    wave_swap_bits<5, 3>(v);
    dft<16>(v);
    wave_twiddle(v, cmul(t.l2, t.h2));
    // DFT-4 over l_lo = b0 + 2 b1 (lane bits 0, 1): t[b0, d0] = v[b0] + (-1)^d0 v[b0 + 2]
    // leaves d0 in lane bit 1; W_4^(b0 d0) = -i where b0 = d0 = 1; then over b0
    wave_bfly_lanes<1>(v, (lane & 2) ? -1.f : 1.f);
    if ((lane & 3) == 3) {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = cx<float>{v[i].im, -v[i].re};
    }
    wave_bfly_lanes<0>(v, (lane & 1) ? -1.f : 1.f);
}

// Batched 1024-point rows (PAT_ROW -> PAT_ROW), one sequence per wave and
// NT / 64 sequences per workgroup; no LDS, no barrier (a wave past the last
// sequence returns at once).  64 VGPRs, 8 waves per SIMD.  A persistent form
// that kept each lane's 30 FP32 twiddles in registers across sequences
// needed 141 VGPRs (3 waves per SIMD) and ran slower (c64 2^10 x 131072:
// 0.490 against 0.407 ms).
//
// Measured on MI355X against the LDS tile kernels (autotuner, c64 n = 1024):
// faster for a single row (4.4 against 5.3 us), slower from 16 rows on
// (0.407 against 0.366 ms at 131072 rows).  Per wave it executes ~1,060
// VALU instructions against ~870 VALU + ~100 LDS for the LDS kernel: the
// cross-lane moves and the FP64 twiddle progressions are VALU work, while
// LDS exchanges run beside the VALU on their own pipe.
template <int NT>
__global__ void __launch_bounds__(NT) wave_rows1024_kernel(TileArgs a)
{
    constexpr int M = NT / 64;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int t2 = tile / a.tiles_inner;
    const int ti = tile - t2 * a.tiles_inner;
    const int o2 = t2 / a.n_outer1;
    const int o = t2 - o2 * a.n_outer1;
    const int s = ti * M + w;
    if (s >= a.n_inner) return;
    const cx<float> *in = reinterpret_cast<const cx<float> *>(a.in) + in_base(a, o2, o) + (int64_t)s * a.in_inner;
    cx<float> *out = reinterpret_cast<cx<float> *>(a.out) + out_base(a, o2, o) + (int64_t)s * a.out_inner;
    const bool si = swap_in_of(a, o), so = swap_out_of(a, o);
    const WaveTw tw = wave_tw_load(a.gen_lo, a.gen_hi, lane);
    // (the direction swaps as selects: uniform branches around the loads and
    // stores measured slower, 0.501 against 0.407 ms at c64 2^10 x 131072)
    cx<float> v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const cx<float> x = in[lane + 64 * r];
        v[r] = si ? swap_ri(x) : x;
    }
    wave_fft1024(v, lane, tw);
    const int base = ((lane >> 2) | ((lane & 2) << 7) | ((lane & 1) << 9));  // k1 + 256 d
#pragma unroll
    for (int c = 0; c < 16; c++) out[base + 16 * c] = so ? swap_ri(v[c]) : v[c];
}

}  // namespace gfft
