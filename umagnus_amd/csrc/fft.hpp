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

// Unsigned division by a run-time constant with a multiply-high (round-up
// method of Granlund & Montgomery); exact for numerators below 2^31.
struct FastDiv {
  uint32_t d = 1, m = 1, s = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t dd) { init(dd); }
  __host__ void init(uint32_t dd) {
    d = dd;
    s = 0;
    while ((1u << s) < d) s++;
    m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << s) - d)) / d + 1);
  }
  __device__ __forceinline__ uint32_t div(uint32_t x) const { return (__umulhi(x, m) + x) >> s; }
  __device__ __forceinline__ uint32_t mod(uint32_t x) const { return x - div(x) * d; }
};

struct FftLen {
  int n;                         // transform length
  int np;                        // number of radix passes
  int rad[FFT_MAXPASS];          // radices in pass order
  FastDiv fp[FFT_MAXPASS];       // p = product of the radices before pass s
  FastDiv fnr[FFT_MAXPASS];      // n / rad[s]
  FastDiv fn;                    // n
  const void *tw;                // exp(-2 pi i m / n), m in [0, n), precision T
};

// Elements per thread: bounds n*nl per work-group (8192 float2 / 4096 double2 = 64 KiB LDS).
template <typename T> struct FftEpt;
template <> struct FftEpt<float> { static constexpr int value = 32; };
template <> struct FftEpt<double> { static constexpr int value = 16; };

// Host side: radix plan for length n under radix schedule `sched` (throws
// UM_EUNSUP if n has a prime factor above 7).
std::vector<int> fft_radices(int n, int sched = 0);
bool fft_supported(int n);

}  // namespace um
