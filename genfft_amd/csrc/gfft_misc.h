// gfft_misc.h -- argument blocks and launchers of the element-wise kernels
// (kern_misc.hip).  Strides and distances are in elements of the buffer's own
// type (complex, or real where a field says so).
#pragma once
#include "gfft_device.h"

namespace gfft {

// A one-argument kernel launch whose completion also completes `stop` (when
// given): the event needs no marker packet of its own after the kernel.
template <class A>
inline hipError_t launch_stop(void (*fn)(A), dim3 grid, dim3 block, hipStream_t s, A a, hipEvent_t stop)
{
    void *args[] = {&a};
    return hipExtLaunchKernel((const void *)fn, grid, block, args, 0, s, nullptr, stop, 0);
}


// Two-level FP64 table of W_N^m = exp(-2*pi*i*m/N):
//   W_N^m = hi[m >> lo_bits] * lo[m & ((1 << lo_bits) - 1)]
struct TwTable {
    const double *lo;
    const double *hi;
    uint64_t mask;  // N - 1
    int32_t lo_bits;
    int32_t pad;
};

struct UnpackArgs {
    const void *in;
    void *out;
    int64_t n;         // real transform length N
    int64_t batch;
    int64_t in_dist;   // complex elements
    int64_t out_dist;  // complex elements
    int32_t half;      // write N/2+1 bins only
    int32_t pad;
    TwTable tw;        // W_N
};

struct SeparateArgs {
    const void *in;
    void *out1;
    void *out2;
    int64_t n, batch, dist;
};

// out row r = in1 row r + i * in2 row r, n elements per row; row pitches in
// real elements (in1, in2) and complex elements (out)
struct InterleaveArgs {
    const void *in1;
    const void *in2;
    void *out;
    int64_t n, batch;
    int64_t in1_pitch, in2_pitch, out_pitch;
};

struct HFillArgs {
    void *out;
    int64_t M, N;      // columns (width), rows (height)
    int64_t vcols;     // columns already computed (M/2 + 1)
    int64_t stride;    // row pitch, complex elements
    int64_t dist;      // batch distance, complex elements
    int64_t batch;
};

struct CopyArgs {
    const void *in;
    void *out;
    int64_t batch, in_dist, out_dist;
    int32_t real_in, real_out;
};

// Rows of `width` elements of elem_bytes (4 or 8) copied between two layouts:
// row r of image b starts at b * in_dist + r * in_pitch in the input and
// b * out_dist + r * out_pitch in the output (in elements).  Moves a real side
// whose pitch or distance is odd to or from a dense copy the tile kernels
// can read as complex pairs, and runs the length-1 column transform (a copy).
struct RepackArgs {
    const void *in;
    void *out;
    int64_t width, rows, batch;
    int64_t in_pitch, in_dist, out_pitch, out_dist;
    int32_t elem_bytes;
    int32_t pad;
};

// Per-transform scale of a 16-bit storage transform: scale[b] = 2^-k, the
// largest power of two <= 1 with gain * max|x_b|_1 * 2^-k <= limit, where
// max|x|_1 is the largest |re| + |im| of transform b's input (rows x width
// elements, row pitch `pitch`, transform b at b * dist).  Every partial and
// final DFT value of the transform is bounded by gain * max|x|_1.
struct AmaxArgs {
    const void *in;   // 16-bit pairs
    uint32_t *bits;   // batch words: running max (float bits), then the scale (float)
    int32_t store;    // 1 FP16, 2 BF16
    int32_t vec;      // 16-byte loads (set by the launcher)
    int64_t batch, rows, width, pitch, dist;
    int64_t chunk, chunks;  // elements per block and blocks per transform (set by the launcher)
    float gain, limit;
};

hipError_t launch_amax_scale(AmaxArgs a, hipStream_t s);

// Bit-reversal gather of rows: for each transform b and row r < n,
//   out[b*dist + rev(r)*stride + c] = in[b*dist + r*stride + c],  c < width
// (a 1D transform is rows of one element).  n is a power of two.
struct BitrevArgs {
    const void *in;
    void *out;
    int64_t n, width, stride, dist, batch;
    int32_t lg, pad;
};
hipError_t launch_bitrev(int prec, const BitrevArgs &a, hipStream_t s);

hipError_t launch_unpack(int prec, const UnpackArgs &a, hipStream_t s, hipEvent_t stop = nullptr);
hipError_t launch_pack(int prec, const UnpackArgs &a, hipStream_t s);
hipError_t launch_separate(int prec, const SeparateArgs &a, hipStream_t s);
hipError_t launch_interleave(int prec, const InterleaveArgs &a, hipStream_t s);
hipError_t launch_hfill(int prec, const HFillArgs &a, hipStream_t s, hipEvent_t stop = nullptr);
hipError_t launch_copy(int prec, const CopyArgs &a, hipStream_t s);
hipError_t launch_repack(const RepackArgs &a, hipStream_t s);

}  // namespace gfft
