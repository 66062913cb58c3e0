// kern_tile_h.hip -- 16-bit storage tile kernels (FP16 and BF16 pairs, FP32
// arithmetic in LDS).
//
// The shapes are the complex64 ones (rows up to L = 16384, column tiles up
// to L = 2048, transposing rows up to L = 4096): the LDS tile is FP32 in both
// cases, so the tile sizes that fit are the same; only the HBM side narrows
// to four bytes per element.
#include "gfft_registry.h"

namespace gfft {
namespace {
constexpr int R_ = PAT_ROW, C_ = PAT_COL, S_ = LAY_SEQ, E_ = LAY_ELEM;

template <class S>
struct HalfTable {
    static constexpr KernelEntry tab[] = {
        // --- batched rows ---
        make_entry_h<S, 2, 1024, 128, S_, R_, R_, 2>(),
        make_entry_h<S, 4, 1024, 256, S_, R_, R_, 4>(),
        make_entry_h<S, 8, 512, 256, S_, R_, R_, 8>(),
        make_entry_h<S, 16, 256, 256, S_, R_, R_, 16>(),
        make_entry_h<S, 32, 128, 256, S_, R_, R_, 8, 4>(),
        make_entry_h<S, 64, 64, 256, S_, R_, R_, 8, 8>(),
        make_entry_h<S, 128, 32, 256, S_, R_, R_, 16, 8>(),
        make_entry_h<S, 256, 16, 256, S_, R_, R_, 16, 16>(),
        make_entry_h<S, 512, 8, 256, S_, R_, R_, 8, 8, 8>(),
        make_entry_h<S, 1024, 4, 256, S_, R_, R_, 16, 16, 4>(),
        make_entry_h<S, 2048, 2, 256, S_, R_, R_, 16, 16, 8>(),
        make_entry_h<S, 4096, 1, 256, S_, R_, R_, 16, 16, 16>(),
        make_entry_h<S, 8192, 1, 512, S_, R_, R_, 16, 16, 8, 4>(),
        make_entry_h<S, 16384, 1, 1024, S_, R_, R_, 16, 16, 16, 4>(),
        // --- column tiles ---
        make_entry_h<S, 2, 512, 64, E_, C_, C_, 2>(),
        make_entry_h<S, 4, 256, 64, E_, C_, C_, 4>(),
        make_entry_h<S, 8, 128, 64, E_, C_, C_, 8>(),
        make_entry_h<S, 16, 64, 64, E_, C_, C_, 16>(),
        make_entry_h<S, 32, 64, 128, E_, C_, C_, 8, 4>(),
        make_entry_h<S, 64, 64, 256, E_, C_, C_, 8, 8>(),
        make_entry_h<S, 128, 32, 256, E_, C_, C_, 16, 8>(),
        make_entry_h<S, 256, 16, 256, E_, C_, C_, 16, 16>(),
        make_entry_h<S, 512, 16, 512, E_, C_, C_, 8, 8, 8>(),
        make_entry_h<S, 1024, 16, 1024, E_, C_, C_, 16, 16, 4>(),
        make_entry_h<S, 2048, 8, 1024, E_, C_, C_, 16, 16, 8>(),
        // --- transposing rows ---
        make_entry_h<S, 64, 64, 256, S_, R_, C_, 8, 8>(),
        make_entry_h<S, 128, 32, 256, S_, R_, C_, 16, 8>(),
        make_entry_h<S, 256, 16, 256, S_, R_, C_, 16, 16>(),
        make_entry_h<S, 512, 16, 512, S_, R_, C_, 8, 8, 8>(),
        make_entry_h<S, 1024, 16, 1024, S_, R_, C_, 16, 16, 4>(),
        make_entry_h<S, 2048, 8, 1024, S_, R_, C_, 16, 16, 8>(),
        make_entry_h<S, 4096, 4, 1024, S_, R_, C_, 16, 16, 16>(),
        // --- persistent pipelined forms of the 2^20 four-step passes ---
        make_pipe_entry_h<S, 1024, 16, 1024, E_, C_, C_, 8, 8, 16>(),
        make_pipe_entry_h<S, 1024, 16, 1024, S_, R_, C_, 16, 16, 4>(),
    };
};
}  // namespace

const KernelEntry *tile_kernels_f16(int *count)
{
    *count = (int)(sizeof(HalfTable<st_f16>::tab) / sizeof(KernelEntry));
    return HalfTable<st_f16>::tab;
}

const KernelEntry *tile_kernels_bf16(int *count)
{
    *count = (int)(sizeof(HalfTable<st_bf16>::tab) / sizeof(KernelEntry));
    return HalfTable<st_bf16>::tab;
}

}  // namespace gfft
