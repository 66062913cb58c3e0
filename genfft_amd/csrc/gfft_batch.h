// gfft_batch.h -- the persistent mixed-size kernel of uneven batches.
//
// One launch runs every transform of length 1 .. MIX_TOT (4096 complex64 /
// 2048 complex128) of an uneven batch.  The work is a device-resident task
// list -- a task is up to 64 transforms of one length that share a tile --
// which resident workgroups drain through one atomic counter, largest
// lengths first.  Each transform has its own input / output offset and
// direction (job tables).  A separate one-thread kernel clears the counter
// (no memset, so the pair can be captured in a HIP graph and replayed).
#pragma once
#include "gfft_device.h"

namespace gfft {

struct MixTask {
    int32_t lg;     // log2 of the transform length
    int32_t first;  // first job (index into the job tables)
    int32_t count;  // transforms in this task (<= mix_m(L))
    int32_t pad;
};

struct MixArgs {
    const void *in;
    void *out;
    const int64_t *job_in;   // element offsets of each job's input
    const int64_t *job_out;  // element offsets of each job's output
    const uint8_t *job_inv;  // 1 = inverse (unscaled)
    const MixTask *tasks;
    uint32_t *counter;       // work-queue head, zeroed by mix_reset_kernel
    int32_t n_tasks;
    int32_t vec;             // every offset even and both buffers 16-byte aligned
    const void *pass_tw[13]; // Stockham twiddles per log2 length (complex T)
};

// Tile budget of the mixed kernel: elements per tile and threads.
template <typename T>
struct MixCfg {
    static constexpr int TOT = sizeof(T) == 4 ? 4096 : 2048;
    static constexpr int NT = 256;
    static constexpr int MAXM = 64;
    static constexpr int MAXLG = sizeof(T) == 4 ? 12 : 11;
};

constexpr int mix_m(int tot, int L) { return tot / L < 64 ? tot / L : 64; }

hipError_t launch_mix(int prec, const MixArgs &a, int grid, hipStream_t s);
hipError_t launch_mix_reset(uint32_t *counter, hipStream_t s);
int mix_grid(int prec, int num_cus);
// radix sequence of the mixed kernel's tile for length 2^lg (for twiddle tables)
int mix_radices(int prec, int lg, int *rad);

}  // namespace gfft
