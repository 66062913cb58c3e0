// Source-compatible entry point: the real-input classes live in <genfft_amd/fft.h>.
#ifndef GENFFT_AMD_COMPAT_FFTREAL_H
#define GENFFT_AMD_COMPAT_FFTREAL_H
#include "../genfft_amd/fft.h"
#endif
