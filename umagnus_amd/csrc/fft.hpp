// LDS-staged Stockham FFT engine for the demag convolution (gfx950).
//
// A work-group holds `nl` lines of length `n` in LDS and runs one radix pass
// per step: every thread pulls its butterflies' inputs into registers, the
// group synchronises, then each thread twiddles, transforms and writes its
// outputs back to the Stockham (autosort) positions.  Radices 2,4,8,16 cover
// the powers of two; 3, 5, 7 the other factors that zero-padding produces
// (2N-1 for N<=5, mag/demagkernel.go:404-420, and 3*2^k grids).
#pragma once
#include "common.hpp"

namespace um {

constexpr int FFT_NT = 256;      // threads per FFT work-group
constexpr int FFT_MAXPASS = 12;

struct FftLen {
  int n;                         // transform length
  int np;                        // number of radix passes
  int rad[FFT_MAXPASS];          // radices in pass order
  const void *tw;                // exp(-2 pi i m / n), m in [0, n), precision T
};

// Elements per thread: bounds n*nl per work-group (8192 float2 / 4096 double2 = 64 KiB LDS).
template <typename T> struct FftEpt;
template <> struct FftEpt<float> { static constexpr int value = 32; };
template <> struct FftEpt<double> { static constexpr int value = 16; };

// Host side: radix plan + twiddle table for length n (throws UM_EUNSUP if n
// has a prime factor above 7 or exceeds the LDS budget).
std::vector<int> fft_radices(int n);
bool fft_supported(int n);

}  // namespace um
