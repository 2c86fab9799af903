// Compile-time specialised LDS Stockham FFT: length N and line count NL are
// template constants, so every LDS address of a butterfly is a base register
// plus an immediate offset and all index divisions are shifts or constant
// multiplies.  The twiddle table lives in LDS (read once per work-group).
#pragma once
#include "fft_device.hpp"

namespace um {
namespace ct {

struct Plan {
  int np;
  int r[16];
};
// Radix plan: 7s, 5s, 3s, then the power of two split into near-equal radices <= 16.
constexpr Plan plan(int n) {
  Plan p{0, {}};
  int m = n;
  while (m % 7 == 0) { p.r[p.np++] = 7; m /= 7; }
  while (m % 5 == 0) { p.r[p.np++] = 5; m /= 5; }
  while (m % 3 == 0) { p.r[p.np++] = 3; m /= 3; }
  int e2 = 0;
  while (m % 2 == 0) { m /= 2; e2++; }
  if (e2 > 0) {
    const int np = (e2 + 3) / 4, base = e2 / np, rem = e2 % np;
    for (int i = 0; i < np; i++) p.r[p.np++] = 1 << (base + (i < rem ? 1 : 0));
  }
  return p;
}
constexpr int npass(int n) { return plan(n).np; }
constexpr int radix(int n, int s) { return plan(n).r[s]; }

}  // namespace ct

// One pass.  Element (l, e) at lds[l*LS + e*ES]; twiddle W_N^m = tw[m*TWS].
template <typename T, int N, int NL, int R, int P, int DIR, int MAPEM, int TWS>
__device__ __forceinline__ void ct_pass(C2<T> *lds, const C2<T> *tw) {
  constexpr int NR = N / R, TOTAL = NL * NR, S = (TOTAL + FFT_NT - 1) / FFT_NT, TSTEP = N / (P * R);
  constexpr int LS = MAPEM ? 1 : N, ES = MAPEM ? NL : 1;
  const int tid = threadIdx.x;
  C2<T> v[S][R];
  int lb[S], ii[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    const int g = tid + s * FFT_NT;
    if (TOTAL % FFT_NT == 0 || g < TOTAL) {
      int l, i;
      if (MAPEM) { i = g / NL; l = g - i * NL; }
      else { l = g / NR; i = g - l * NR; }
      lb[s] = l * LS;
      ii[s] = i;
      const C2<T> *src = lds + lb[s] + i * ES;
#pragma unroll
      for (int j = 0; j < R; j++) v[s][j] = src[j * NR * ES];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; s++) {
    const int g = tid + s * FFT_NT;
    if (TOTAL % FFT_NT == 0 || g < TOTAL) {
      const int i = ii[s];
      const int k = (P == 1) ? 0 : i % P;
      if (P > 1) {
#pragma unroll
        for (int j = 1; j < R; j++) {
          C2<T> w = tw[j * k * TSTEP * TWS];
          if (DIR > 0) w.y = -w.y;
          v[s][j] = cmul(v[s][j], w);
        }
      }
      dft<R, DIR>(v[s]);
      C2<T> *dst = lds + lb[s] + ((i - k) * R + k) * ES;
#pragma unroll
      for (int j = 0; j < R; j++) dst[j * P * ES] = v[s][j];
    }
  }
  __syncthreads();
}

template <typename T, int N, int NL, int DIR, int MAPEM, int TWS, int S, int P>
__device__ __forceinline__ void ct_passes(C2<T> *lds, const C2<T> *tw) {
  if constexpr (S < ct::npass(N)) {
    constexpr int R = ct::radix(N, S);
    ct_pass<T, N, NL, R, P, DIR, MAPEM, TWS>(lds, tw);
    ct_passes<T, N, NL, DIR, MAPEM, TWS, S + 1, P * R>(lds, tw);
  }
}

template <typename T, int N, int NL, int DIR, int MAPEM, int TWS>
__device__ __forceinline__ void ct_fft(C2<T> *lds, const C2<T> *tw) {
  ct_passes<T, N, NL, DIR, MAPEM, TWS, 0, 1>(lds, tw);
}

}  // namespace um
