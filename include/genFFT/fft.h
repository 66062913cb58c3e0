// Source-compatible entry point for code written against genFFT: keeps the
// reference's include line (#include <genFFT/fft.h>) working and routes it to
// the gfx950 engine's genfft-compatible classes.
#ifndef GENFFT_AMD_COMPAT_FFT_H
#define GENFFT_AMD_COMPAT_FFT_H
#include "../genfft_amd/fft.h"
#endif
