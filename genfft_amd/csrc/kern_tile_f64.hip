// kern_tile_f64.hip -- complex128 tile-kernel shapes.
//
// Same three families as complex64 (rows, column tiles, transposing rows)
// with half the elements per tile, since each element is 16 bytes: rows hold
// 2048 complex (32 KB) up to L = 2048 and a single sequence up to L = 8192;
// column tiles are >= 8 columns (128-byte segments) where LDS allows.  The
// batched rows use radix-8 passes (FP64 register pressure, see below).
#include "gfft_registry.h"

namespace gfft {
namespace {
using T = double;
constexpr int R_ = PAT_ROW, C_ = PAT_COL, S_ = LAY_SEQ, E_ = LAY_ELEM;

const KernelEntry kTable[] = {
    // --- batched rows ---
    make_entry<T, 2, 512, 64, S_, R_, R_, 2>(),
    make_entry<T, 4, 512, 128, S_, R_, R_, 4>(),
    make_entry<T, 8, 256, 128, S_, R_, R_, 8>(),
    make_entry<T, 16, 128, 128, S_, R_, R_, 16>(),
    make_entry<T, 32, 64, 128, S_, R_, R_, 8, 4>(),
    make_entry<T, 64, 32, 128, S_, R_, R_, 8, 8>(),
    make_entry<T, 128, 16, 128, S_, R_, R_, 16, 8>(),
    // radix <= 8 and 8 elements per thread from L = 256 on: an FP64 radix-16
    // butterfly holds 64 VGPRs of data plus 60 of twiddles (148-190 VGPRs per
    // kernel); the radix-8 forms use 83-116 and run 25-30% faster (MI355X,
    // 1 GiB batches: L = 1024 0.624 -> 0.449 ms, 8192 0.787 -> 0.569 ms)
    make_entry<T, 256, 8, 256, S_, R_, R_, 8, 8, 4>(),
    make_entry<T, 512, 4, 256, S_, R_, R_, 8, 8, 8>(),
    make_entry<T, 1024, 2, 256, S_, R_, R_, 8, 4, 8, 4>(),
    make_entry<T, 2048, 1, 256, S_, R_, R_, 8, 8, 8, 4>(),
    make_entry<T, 4096, 1, 512, S_, R_, R_, 8, 8, 8, 8>(),
    make_entry<T, 8192, 1, 1024, S_, R_, R_, 8, 8, 8, 8, 2>(),
    // --- column tiles ---
    make_entry<T, 2, 256, 64, E_, C_, C_, 2>(),
    make_entry<T, 4, 256, 64, E_, C_, C_, 4>(),
    make_entry<T, 8, 128, 64, E_, C_, C_, 8>(),
    make_entry<T, 16, 64, 64, E_, C_, C_, 16>(),
    make_entry<T, 32, 32, 64, E_, C_, C_, 8, 4>(),
    make_entry<T, 64, 32, 128, E_, C_, C_, 8, 8>(),
    make_entry<T, 128, 16, 128, E_, C_, C_, 16, 8>(),
    make_entry<T, 256, 8, 128, E_, C_, C_, 16, 16>(),
    make_entry<T, 512, 8, 256, E_, C_, C_, 8, 8, 8>(),
    make_entry<T, 1024, 8, 512, E_, C_, C_, 16, 16, 4>(),
    make_entry<T, 2048, 4, 512, E_, C_, C_, 16, 8, 16>(),  // c128 2^22 column pass 0.62-0.63 ms (16*16*8: 0.64)
    // --- transposing rows ---
    make_entry<T, 64, 32, 128, S_, R_, C_, 8, 8>(),
    make_entry<T, 128, 16, 128, S_, R_, C_, 16, 8>(),
    make_entry<T, 256, 8, 128, S_, R_, C_, 16, 16>(),
    make_entry<T, 512, 8, 256, S_, R_, C_, 8, 8, 8>(),
    make_entry<T, 1024, 8, 512, S_, R_, C_, 16, 16, 4>(),
    make_entry<T, 2048, 4, 512, S_, R_, C_, 16, 16, 8>(),
    make_entry<T, 4096, 2, 512, S_, R_, C_, 16, 16, 16>(),
    // --- persistent pipelined forms of the 2^20 four-step passes ---
    make_pipe_entry<T, 1024, 8, 512, E_, C_, C_, 16, 16, 4>(),
    make_pipe_entry<T, 1024, 8, 512, S_, R_, C_, 16, 16, 4>(),
    // --- autotuner alternatives ---
    make_variant<T, 1024, 4, 256, S_, R_, R_, 16, 16, 4>(),
    make_variant<T, 1024, 2, 128, S_, R_, R_, 4, 16, 16>(),
    make_variant<T, 2048, 2, 256, S_, R_, R_, 16, 16, 8>(),
    make_variant<T, 1024, 4, 512, S_, R_, R_, 8, 4, 8, 4>(),
    make_variant<T, 2048, 2, 512, S_, R_, R_, 8, 8, 8, 4>(),
    // the radix-16 forms
    make_variant<T, 256, 8, 128, S_, R_, R_, 16, 16>(),
    make_variant<T, 512, 4, 128, S_, R_, R_, 8, 8, 8>(),
    make_variant<T, 1024, 2, 128, S_, R_, R_, 16, 16, 4>(),
    make_variant<T, 2048, 1, 128, S_, R_, R_, 16, 16, 8>(),
    make_variant<T, 4096, 1, 256, S_, R_, R_, 16, 16, 16>(),
    make_variant<T, 8192, 1, 512, S_, R_, R_, 16, 16, 8, 4>(),
};
// The 2^20 four-step's two pipelined passes paired in one launch (column pass
// of chunk k, transposing row pass of chunk k - 1).
using ColPass = PipeShape<T, 1024, 8, 512, E_, C_, C_, 16, 16, 4>;
using RowPass = PipeShape<T, 1024, 8, 512, S_, R_, C_, 16, 16, 4>;
const DualEntry kDual[] = {
    make_dual<ColPass, RowPass>(&tile_fft_pipe_kernel<T, 1024, 8, 512, E_, C_, C_, 16, 16, 4>,
                                &tile_fft_pipe_kernel<T, 1024, 8, 512, S_, R_, C_, 16, 16, 4>),
};
}  // namespace

const KernelEntry *tile_kernels_f64(int *count)
{
    *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
    return kTable;
}

const DualEntry *dual_kernels_f64(int *count)
{
    *count = (int)(sizeof(kDual) / sizeof(kDual[0]));
    return kDual;
}

}  // namespace gfft
