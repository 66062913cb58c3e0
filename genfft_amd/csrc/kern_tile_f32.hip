// kern_tile_f32.hip -- complex64 tile-kernel shapes.
//
// Rows (PAT_ROW -> PAT_ROW, sequence-major LDS): batched 1D transforms and
// the 2D row pass; a tile holds 4096 complex (32 KB) up to L = 4096, then a
// single sequence up to L = 16384 (136 KB of gfx950's 160 KB LDS).
// Columns (PAT_COL -> PAT_COL, element-major LDS): the first pass of a
// four-step transform and the 2D column pass; M adjacent columns give
// >= 64-byte contiguous row segments.
// Transposing rows (PAT_ROW -> PAT_COL): the second four-step pass, writing
// element k of row k1 to k1 + N1*k, M rows wide.
// Every default shape keeps 16 elements per thread, so radix-16 butterflies
// live entirely in registers (the radix-32 alternatives hold 32).
#include "gfft_registry.h"

namespace gfft {
namespace {
using T = float;
constexpr int R_ = PAT_ROW, C_ = PAT_COL, S_ = LAY_SEQ, E_ = LAY_ELEM;

const KernelEntry kTable[] = {
    // --- batched rows ---
    make_entry<T, 2, 1024, 128, S_, R_, R_, 2>(),
    make_entry<T, 4, 1024, 256, S_, R_, R_, 4>(),
    make_entry<T, 8, 512, 256, S_, R_, R_, 8>(),
    make_entry<T, 16, 256, 256, S_, R_, R_, 16>(),
    make_entry<T, 32, 128, 256, S_, R_, R_, 8, 4>(),
    make_entry<T, 64, 64, 256, S_, R_, R_, 8, 8>(),
    make_entry<T, 128, 32, 256, S_, R_, R_, 16, 8>(),
    make_entry<T, 256, 16, 256, S_, R_, R_, 16, 16>(),
    make_entry<T, 512, 8, 256, S_, R_, R_, 8, 8, 8>(),
    make_entry<T, 1024, 4, 256, S_, R_, R_, 16, 16, 4>(),
    make_entry<T, 2048, 2, 256, S_, R_, R_, 16, 16, 8>(),
    make_entry<T, 4096, 1, 256, S_, R_, R_, 16, 16, 16>(),
    make_entry<T, 8192, 1, 512, S_, R_, R_, 16, 16, 8, 4>(),
    make_entry<T, 16384, 1, 1024, S_, R_, R_, 16, 16, 16, 4>(),
    // --- column tiles ---
    make_entry<T, 2, 512, 64, E_, C_, C_, 2>(),
    make_entry<T, 4, 256, 64, E_, C_, C_, 4>(),
    make_entry<T, 8, 128, 64, E_, C_, C_, 8>(),
    make_entry<T, 16, 64, 64, E_, C_, C_, 16>(),
    make_entry<T, 32, 64, 128, E_, C_, C_, 8, 4>(),
    make_entry<T, 64, 64, 256, E_, C_, C_, 8, 8>(),
    make_entry<T, 128, 32, 256, E_, C_, C_, 16, 8>(),
    make_entry<T, 256, 16, 256, E_, C_, C_, 16, 16>(),
    make_entry<T, 512, 16, 512, E_, C_, C_, 8, 8, 8>(),
    make_entry<T, 1024, 16, 1024, E_, C_, C_, 16, 16, 4>(),
    make_entry<T, 1024, 8, 512, E_, C_, C_, 16, 16, 4>(),
    make_entry<T, 2048, 8, 1024, E_, C_, C_, 16, 8, 16>(),  // c64 2^22 column pass 0.550 ms (16*16*8: 0.592, 8*16*16: 0.565)
    make_entry<T, 4096, 4, 1024, E_, C_, C_, 16, 16, 16>(),  // 32-byte rows: column transforms only, not four-step
    // --- transposing rows ---
    make_entry<T, 64, 64, 256, S_, R_, C_, 8, 8>(),
    make_entry<T, 128, 32, 256, S_, R_, C_, 16, 8>(),
    make_entry<T, 256, 16, 256, S_, R_, C_, 16, 16>(),
    make_entry<T, 512, 16, 512, S_, R_, C_, 8, 8, 8>(),
    make_entry<T, 1024, 16, 1024, S_, R_, C_, 16, 16, 4>(),
    make_entry<T, 1024, 8, 512, S_, R_, C_, 16, 16, 4>(),
    make_entry<T, 2048, 8, 1024, S_, R_, C_, 16, 16, 8>(),
    make_entry<T, 4096, 4, 1024, S_, R_, C_, 16, 16, 16>(),
    // --- persistent pipelined forms, kept where they measured faster on
    // MI355X (c64 2^20 x 128: column pass 0.565 -> 0.524 ms, transposing row
    // pass 0.462 -> 0.392 ms); the L = 512 / 2048 and M = 8 forms fall to
    // one workgroup per CU and lose to the plain kernels (c64 2^22 x 32,
    // L = 2048, M = 8: column pass 0.63-0.72 ms pipelined, with or without
    // tile runs, against 0.59; transposing row pass 0.55 against 0.50) ---
    // column pass, radix 16 * 4 * 16: measured on c64 2^20 x 128 against the
    // other orders at 1024 threads, 0.440-0.443 ms (8*8*16: 0.464, 4*16*16:
    // 0.479, 16*16*4: 0.493; the transposing row pass does not move with order)
    make_pipe_entry<T, 1024, 16, 1024, E_, C_, C_, 16, 4, 16>(),
    make_pipe_entry<T, 1024, 16, 1024, S_, R_, C_, 16, 16, 4>(),
    // single 16384-point sequences (tile: 136 KB of LDS); radix 16*4*16*16:
    // c64 2^14 x 8192 0.500 ms with the pass twiddles read from global memory
    // (16*16*16*4: 0.509, 16*16*4*16: 0.502, 4*16*16*16: 0.521), 0.442 ms with
    // the last pass's twiddles generated from a 4 KB table in LDS
    // (PassTw::GEN: the first three passes' 8 KB of tables fit beside it)
    make_pipe_entry<T, 16384, 1, 1024, S_, R_, R_, 16, 4, 16, 16>(),
    // --- autotuner alternatives: radix order, tile width, threads per tile.
    // The double-width row tiles (64 KB, 512 threads; make_wide_variant) are
    // also the planner's choice for batches of >= 512 MiB: measured on MI355X
    // through the autotuner, 1 GiB batches ran 3-5% faster on them (L = 256:
    // 0.385 -> 0.372 ms, 512: 0.385 -> 0.365, 1024: 0.379 -> 0.365, 2048:
    // 0.388 -> 0.367), 256 MiB batches (within the Infinity Cache) 0-12% slower ---
    make_wide_variant<T, 256, 32, 512, S_, R_, R_, 16, 16>(),
    make_variant<T, 512, 8, 256, S_, R_, R_, 16, 8, 4>(),
    make_wide_variant<T, 512, 16, 512, S_, R_, R_, 8, 8, 8>(),
    make_variant<T, 1024, 4, 256, S_, R_, R_, 4, 16, 16>(),
    make_variant<T, 1024, 4, 256, S_, R_, R_, 8, 8, 16>(),
    make_wide_variant<T, 1024, 8, 512, S_, R_, R_, 16, 16, 4>(),
    make_variant<T, 1024, 2, 128, S_, R_, R_, 16, 16, 4>(),
    make_variant<T, 1024, 4, 128, S_, R_, R_, 32, 32>(),   // two radix-32 passes, 32 elements per thread
    make_variant<T, 1024, 8, 256, S_, R_, R_, 32, 32>(),
    make_variant<T, 2048, 2, 256, S_, R_, R_, 8, 16, 16>(),
    make_wide_variant<T, 2048, 4, 512, S_, R_, R_, 16, 16, 8>(),
    make_variant<T, 4096, 2, 512, S_, R_, R_, 16, 16, 16>(),
    make_variant<T, 4096, 1, 512, S_, R_, R_, 8, 8, 8, 8>(),
    make_variant<T, 8192, 1, 512, S_, R_, R_, 16, 16, 16, 2>(),
    make_variant<T, 16384, 1, 1024, S_, R_, R_, 16, 16, 8, 8>(),
    // one wavefront per 1024-point row, lanes exchanged by DPP / permlane swaps
    make_wave_entry<256>(),
};
// The 2^20 four-step's two pipelined passes paired in one launch (column pass
// of chunk k, transposing row pass of chunk k - 1).
using ColPass = PipeShape<T, 1024, 16, 1024, E_, C_, C_, 16, 4, 16>;
using RowPass = PipeShape<T, 1024, 16, 1024, S_, R_, C_, 16, 16, 4>;
const DualEntry kDual[] = {
    make_dual<ColPass, RowPass>(&tile_fft_pipe_kernel<T, 1024, 16, 1024, E_, C_, C_, 16, 4, 16>,
                                &tile_fft_pipe_kernel<T, 1024, 16, 1024, S_, R_, C_, 16, 16, 4>),
};
}  // namespace

const KernelEntry *tile_kernels_f32(int *count)
{
    *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
    return kTable;
}

const DualEntry *dual_kernels_f32(int *count)
{
    *count = (int)(sizeof(kDual) / sizeof(kDual[0]));
    return kDual;
}

}  // namespace gfft
