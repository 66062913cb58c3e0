// gfft_registry.h -- the table of compiled tile-kernel shapes.
//
// Every instantiation of gfft::tile_fft_kernel is described by one entry;
// the planner (plan.cpp) picks entries by (precision, L, input pattern,
// output pattern) and launches them through the function pointer.
#pragma once
#include <type_traits>

#include "gfft_device.h"
#include "gfft_wave.h"

namespace gfft {

struct KernelEntry {
    int prec;      // 0 = complex64, 1 = complex128
    int L;         // sequence length
    int M;         // sequences per tile
    int NT;        // threads per workgroup
    int layout;    // LAY_SEQ / LAY_ELEM
    int inp, outp; // PAT_ROW / PAT_COL
    int nrad;
    int rad[8];
    int lds_bytes;
    int persistent;  // 1: tile_fft_pipe_kernel (grid sized to the device, loops over tiles)
    int tune_only;   // 1: an alternative (radix order / tile width / threads) only the autotuner selects
    void (*fn)(TileArgs);
    int store;       // 0: complex T in memory; 1: FP16 pairs; 2: BF16 pairs (FP32 arithmetic)
    int gen_tw;      // 1: the last pass builds its twiddles from TileArgs::gen_lo / gen_hi (W_L)
    int wave;        // 1: one wavefront per row, lanes exchanged without LDS (wave_rows1024_kernel)
    int wide;        // 1: double-width row tile the planner takes for batches past the Infinity Cache
};

// Shape of a tile entry, without its kernel: each make_* below takes the
// address of exactly the kernel it registers, so no translation unit
// instantiates kernels of another table.
template <typename T, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry entry_shape()
{
    static_assert(((RS * ...) == L), "radices must multiply to L");
    static_assert((M * L) % NT == 0, "tile must split evenly over threads");
    static_assert(NT % 64 == 0 && NT <= 1024, "whole wavefronts, at most 1024 threads");
    using G = TileGeom<L, M, LAY, INP, OUTP>;
    static_assert(G::ELEMS * sizeof(cx<T>) <= 160 * 1024, "tile exceeds gfx950 LDS");
    KernelEntry e{};
    e.prec = sizeof(T) == 8 ? 1 : 0;
    e.L = L;
    e.M = M;
    e.NT = NT;
    e.layout = LAY;
    e.inp = INP;
    e.outp = OUTP;
    e.nrad = sizeof...(RS);
    const int r[] = {RS...};
    for (int i = 0; i < (int)sizeof...(RS); i++) e.rad[i] = r[i];
    e.lds_bytes = (int)(G::ELEMS * sizeof(cx<T>));
    e.persistent = 0;
    e.tune_only = 0;
    e.fn = nullptr;
    return e;
}

template <typename T, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry make_entry()
{
    KernelEntry e = entry_shape<T, L, M, NT, LAY, INP, OUTP, RS...>();
    e.fn = &tile_fft_kernel<T, L, M, NT, LAY, INP, OUTP, RS...>;
    return e;
}

template <typename T, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry make_variant()
{
    KernelEntry e = make_entry<T, L, M, NT, LAY, INP, OUTP, RS...>();
    e.tune_only = 1;
    return e;
}

// A tuning alternative the planner also takes by itself for batches larger
// than the Infinity Cache (KernelEntry::wide).
template <typename T, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry make_wide_variant()
{
    KernelEntry e = make_variant<T, L, M, NT, LAY, INP, OUTP, RS...>();
    e.wide = 1;
    return e;
}

// Persistent, software-pipelined variant (same tile math, LDS-resident pass
// twiddles, next tile prefetched into registers).
template <typename T, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry pipe_shape()
{
    KernelEntry e = entry_shape<T, L, M, NT, LAY, INP, OUTP, RS...>();
    using S = PipeLds<T, L, M, NT, LAY, INP, OUTP, RS...>;
    e.lds_bytes = (int)((S::TILE_A + S::TW) * sizeof(cx<T>));
    e.gen_tw = S::P::GEN ? 1 : 0;
    e.persistent = 1;
    return e;
}

// One wavefront per 1024-point complex64 row, no LDS (wave_rows1024_kernel),
// radix stages (16, 16, 4); twiddles generated from the two-level table of W_1024.
template <int NT>
constexpr KernelEntry make_wave_entry()
{
    KernelEntry e = entry_shape<float, 1024, NT / 64, NT, LAY_SEQ, PAT_ROW, PAT_ROW, 16, 16, 4>();
    e.lds_bytes = 0;
    e.tune_only = 1;
    e.gen_tw = 1;
    e.wave = 1;
    e.fn = &wave_rows1024_kernel<NT>;
    return e;
}

template <typename T, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry make_pipe_entry()
{
    KernelEntry e = pipe_shape<T, L, M, NT, LAY, INP, OUTP, RS...>();
    e.fn = &tile_fft_pipe_kernel<T, L, M, NT, LAY, INP, OUTP, RS...>;
    return e;
}

// 16-bit storage form of a complex64 shape (tile_fft_kernel_h).
template <class S, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry make_entry_h()
{
    KernelEntry e = entry_shape<float, L, M, NT, LAY, INP, OUTP, RS...>();
    e.lds_bytes += M * (int)sizeof(float);
    e.store = std::is_same<S, st_f16>::value ? 1 : 2;
    e.fn = &tile_fft_kernel_h<S, L, M, NT, LAY, INP, OUTP, RS...>;
    return e;
}

// Pipelined 16-bit storage form (tile_fft_pipe_kernel_h).
template <class S, int L, int M, int NT, int LAY, int INP, int OUTP, int... RS>
constexpr KernelEntry make_pipe_entry_h()
{
    KernelEntry e = pipe_shape<float, L, M, NT, LAY, INP, OUTP, RS...>();
    // its own LDS: the tile, the pass tables when they fit beside the tile
    // and the scales (else they are read from global memory), the scales
    using G = TileGeom<L, M, LAY, INP, OUTP>;
    constexpr int twn = pass_tw_count<1, RS...>();
    constexpr bool twl = pass_tw_in_lds<float, G::ELEMS + M / 2, twn>();
    e.lds_bytes = (int)((G::ELEMS + (twl && twn > 0 ? twn : 1)) * sizeof(cx<float>)) + M * (int)sizeof(float);
    e.gen_tw = 0;
    e.store = std::is_same<S, st_f16>::value ? 1 : 2;
    e.fn = &tile_fft_pipe_kernel_h<S, L, M, NT, LAY, INP, OUTP, RS...>;
    return e;
}

// A dual_pipe_kernel pairing two pipelined tile passes: role A (the four-step
// column pass of one chunk) and role B (the transposing row pass of the
// chunk before), identified by the single-pass kernels they replace.
struct DualEntry {
    int prec;
    int NT;
    int lds_bytes;
    void (*a_fn)(TileArgs);  // the tile_fft_pipe_kernel whose tiles role A runs
    void (*b_fn)(TileArgs);  // the same for role B
    void (*fn)(TileArgs, TileArgs, int);
};

template <class A, class B>
constexpr DualEntry make_dual(void (*a_fn)(TileArgs), void (*b_fn)(TileArgs))
{
    static_assert(A::NT == B::NT, "one thread count");
    constexpr int e = A::ELEMS > B::ELEMS ? A::ELEMS : B::ELEMS;
    static_assert(e * sizeof(cx<typename A::T>) <= 160 * 1024, "both roles fit gfx950 LDS");
    DualEntry d{};
    d.prec = sizeof(typename A::T) == 8 ? 1 : 0;
    d.NT = A::NT;
    d.lds_bytes = (int)(e * sizeof(cx<typename A::T>));
    d.a_fn = a_fn;
    d.b_fn = b_fn;
    d.fn = &dual_pipe_kernel<A, B>;
    return d;
}

// Defined in kern_tile_f32.hip / kern_tile_f64.hip / kern_tile_h.hip.
const KernelEntry *tile_kernels_f32(int *count);
const KernelEntry *tile_kernels_f64(int *count);
const DualEntry *dual_kernels_f32(int *count);
const DualEntry *dual_kernels_f64(int *count);
const KernelEntry *tile_kernels_f16(int *count);
const KernelEntry *tile_kernels_bf16(int *count);

}  // namespace gfft
