// gfft_filter.h -- fused spectral filter (kern_filter.hip).
#pragma once
#include <algorithm>

#include "gfft_device.h"

namespace gfft {

struct FilterArgs {
    const void *x;        // real rows (read as complex pairs), batch x n
    const void *gy;       // adjoint: upstream gradient rows
    void *y;              // forward: y rows; adjoint: gx rows; reduce: gH
    const void *H;        // n/2 + 1 complex bins
    void *partial;        // adjoint: per-row gH contributions, batch x (n/2 + 1)
    const void *tw;       // W_n^k, k < n/2 (complex T, FP64-generated, rounded once)
    const void *pass_tw;  // Stockham twiddles of the n/2-point forward FFT
    const void *pass_tw_inv;  // the same for the inverse FFT's radix order
    int64_t batch;
    int64_t bins;         // n/2 + 1
    int64_t ppitch;       // complex elements between partial rows: bins rounded up to 128 bytes
    int32_t conj_h;       // forward kernel with conj(H) (the x-adjoint alone)
    int32_t pad;
};

// Partial-row pitch: whole 128-byte lines per row, so the reduction's
// 32-bin column segments never straddle an extra line.
inline int64_t filter_partial_pitch(int64_t bins, int prec)
{
    const int64_t per_line = prec ? 8 : 16;
    return (bins + per_line - 1) / per_line * per_line;
}

struct ReduceArgs {
    const void *src;      // rows x pitch complex partials
    void *dst;            // one sum row per segment, out_pitch apart
    int64_t bins, rows, pitch, seg, out_pitch;
};

// which: 0 forward, 1 adjoint (gx + partials), 2 partial reduction (gH);
// `stop`, if given, completes with the launch's last kernel
hipError_t launch_filter(int prec, int which, const FilterArgs &a, hipStream_t s, hipEvent_t stop = nullptr);
// threads and radices (of the forward FFT, or with `inverse` of the inverse
// FFT) of the fused kernel for L = n/2; 0 if unsupported
int filter_shape(int prec, int L, int *nt, int *rad, int inverse);

}  // namespace gfft
