// Device side of the LDS Stockham FFT engine (included by demag.hip only).
#pragma once
#include "fft.hpp"

namespace um {

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) { return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename C>
__device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }
template <typename C, typename S>
__device__ __forceinline__ C cscale(C a, S s) { return C{a.x * s, a.y * s}; }
// a * (DIR * i): forward (DIR=-1) multiplies by -i, inverse by +i
template <int DIR, typename C>
__device__ __forceinline__ C crot(C a) {
  if (DIR < 0) return C{a.y, -a.x};
  return C{-a.y, a.x};
}
// multiply by exp(DIR * i * angle) given (cos, sin) of angle
template <int DIR, typename C, typename S>
__device__ __forceinline__ C cmul_cs(C a, S c, S s) {
  const S sn = DIR < 0 ? -s : s;
  return C{a.x * c - a.y * sn, a.x * sn + a.y * c};
}

// ---- small DFTs --------------------------------------------------------------
template <int DIR, typename C>
__device__ __forceinline__ void dft2(C &a0, C &a1) {
  C t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}
template <int DIR, typename C>
__device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3) {
  C t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = crot<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}
template <int DIR, typename C>
__device__ __forceinline__ void dft8(C *a) {
  using S = decltype(a[0].x);
  C e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  C o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  const S h = S(0.70710678118654752440);
  o1 = cmul_cs<DIR>(o1, h, h);
  o2 = crot<DIR>(o2);
  o3 = cmul_cs<DIR>(o3, -h, h);
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}
template <int DIR, typename C>
__device__ __forceinline__ void dft16(C *a) {
  using S = decltype(a[0].x);
  C e[8], o[8];
#pragma unroll
  for (int j = 0; j < 8; j++) { e[j] = a[2 * j]; o[j] = a[2 * j + 1]; }
  dft8<DIR>(e);
  dft8<DIR>(o);
  const S c1 = S(0.92387953251128675613), s1 = S(0.38268343236508977173);
  const S h = S(0.70710678118654752440);
  o[1] = cmul_cs<DIR>(o[1], c1, s1);
  o[2] = cmul_cs<DIR>(o[2], h, h);
  o[3] = cmul_cs<DIR>(o[3], s1, c1);
  o[4] = crot<DIR>(o[4]);
  o[5] = cmul_cs<DIR>(o[5], -s1, c1);
  o[6] = cmul_cs<DIR>(o[6], -h, h);
  o[7] = cmul_cs<DIR>(o[7], -c1, s1);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    a[k] = cadd(e[k], o[k]);
    a[k + 8] = csub(e[k], o[k]);
  }
}
template <int DIR, typename C>
__device__ __forceinline__ void dft32(C *a) {
  using S = decltype(a[0].x);
  C e[16], o[16];
#pragma unroll
  for (int j = 0; j < 16; j++) { e[j] = a[2 * j]; o[j] = a[2 * j + 1]; }
  dft16<DIR>(e);
  dft16<DIR>(o);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    if (k == 8) {
      o[k] = crot<DIR>(o[k]);
    } else if (k > 0) {
      // W_32^k, folded to constants once the loop is unrolled
      const double ang = 2.0 * M_PI * k / 32.0;
      o[k] = cmul_cs<DIR>(o[k], S(__builtin_cos(ang)), S(__builtin_sin(ang)));
    }
    a[k] = cadd(e[k], o[k]);
    a[k + 16] = csub(e[k], o[k]);
  }
}
// cos / sin of 2 pi m / R for the odd radices
template <int R> struct OddTw;
template <> struct OddTw<3> {
  static __device__ __forceinline__ double c(int m) { return m == 0 ? 1.0 : -0.5; }
  static __device__ __forceinline__ double s(int m) { return m == 0 ? 0.0 : (m == 1 ? 0.86602540378443864676 : -0.86602540378443864676); }
};
template <> struct OddTw<5> {
  static __device__ __forceinline__ double c(int m) {
    return m == 0 ? 1.0 : (m == 1 || m == 4) ? 0.30901699437494742410 : -0.80901699437494742410;
  }
  static __device__ __forceinline__ double s(int m) {
    return m == 0 ? 0.0 : m == 1 ? 0.95105651629515357212 : m == 2 ? 0.58778525229247312917
         : m == 3 ? -0.58778525229247312917 : -0.95105651629515357212;
  }
};
template <> struct OddTw<7> {
  static __device__ __forceinline__ double c(int m) {
    return m == 0 ? 1.0 : (m == 1 || m == 6) ? 0.62348980185873353053
         : (m == 2 || m == 5) ? -0.22252093395631440429 : -0.90096886790241912624;
  }
  static __device__ __forceinline__ double s(int m) {
    return m == 0 ? 0.0 : m == 1 ? 0.78183148246802980871 : m == 2 ? 0.97492791218182360702
         : m == 3 ? 0.43388373911755812048 : m == 4 ? -0.43388373911755812048
         : m == 5 ? -0.97492791218182360702 : -0.78183148246802980871;
  }
};
// Odd prime radix: pair q with R-q using sums/differences of a_j and a_{R-j}.
template <int R, int DIR, typename C>
__device__ __forceinline__ void dft_odd(C *a) {
  using S = decltype(a[0].x);
  constexpr int H = (R - 1) / 2;
  C s[H], d[H];
#pragma unroll
  for (int j = 1; j <= H; j++) {
    s[j - 1] = cadd(a[j], a[R - j]);
    d[j - 1] = csub(a[j], a[R - j]);
  }
  C y0 = a[0];
#pragma unroll
  for (int j = 0; j < H; j++) y0 = cadd(y0, s[j]);
  C out[R];
  out[0] = y0;
#pragma unroll
  for (int q = 1; q <= H; q++) {
    C ca = a[0], cb = C{S(0), S(0)};
#pragma unroll
    for (int j = 1; j <= H; j++) {
      const int m = (j * q) % R;
      const S cc = S(OddTw<R>::c(m)), ss = S(OddTw<R>::s(m));
      ca = C{ca.x + cc * s[j - 1].x, ca.y + cc * s[j - 1].y};
      cb = C{cb.x + ss * d[j - 1].x, cb.y + ss * d[j - 1].y};
    }
    C r = crot<DIR>(cb);  // DIR * i * sum sin * d
    out[q] = cadd(ca, r);
    out[R - q] = csub(ca, r);
  }
#pragma unroll
  for (int j = 0; j < R; j++) a[j] = out[j];
}
template <int R, int DIR, typename C>
__device__ __forceinline__ void dft(C *a) {
  if constexpr (R == 2) dft2<DIR>(a[0], a[1]);
  else if constexpr (R == 4) dft4<DIR>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 8) dft8<DIR>(a);
  else if constexpr (R == 16) dft16<DIR>(a);
  else if constexpr (R == 32) dft32<DIR>(a);
  else dft_odd<R, DIR>(a);
}

// ---- register-level helpers ------------------------------------------------
// base + byte offset: a uniform base and a zero-extended 32-bit lane offset
// map onto the scalar-base + vector-offset form of the global loads
// (the base passes through readfirstlane so that it stays in scalar registers
// instead of being re-associated with the lane offset into 64-bit VGPR pairs)
template <typename V, typename B>
__device__ __forceinline__ __attribute__((address_space(1))) V *at_off(B *base, uint32_t off) {
  using Ch = typename std::conditional<std::is_const<V>::value, const char, char>::type;
  using GCh = __attribute__((address_space(1))) Ch;
  const uint64_t u = (uint64_t)base;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (__attribute__((address_space(1))) V *)((GCh *)(((uint64_t)hi << 32) | lo) + off);
}

// Materialise an unrolled array at this point: its computation may neither
// sink past the call (e.g. into a guarded store block) nor interleave with
// the code after it, which bounds the register peak of long unrolled FFTs.
template <typename CT, int N>
__device__ __forceinline__ void pin(CT (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; i++) asm volatile("" : "+v"(v[i].x), "+v"(v[i].y));
  __builtin_amdgcn_sched_barrier(0);
}

// ---- one Stockham pass over nl lines in LDS ---------------------------------
// Element e of line l lives at lds[l*ls + e*es].  MAPEM=1 assigns consecutive
// threads to consecutive lines (element-major LDS, ls=1), MAPEM=0 to consecutive
// butterflies of one line (line-major LDS, es=1); both keep LDS access unit-stride
// across a wavefront when loading.
template <typename T, int R, int DIR, int MAPEM>
__device__ __forceinline__ void fft_pass(C2<T> *lds, int nl, const FastDiv &fnl, int n, int p, const FastDiv &fp,
                                         const FastDiv &fnR, int ls, int es, const C2<T> *__restrict__ tw) {
  constexpr int EPT = FftEpt<T>::value;
  constexpr int S = (EPT + R - 1) / R;
  const int tid = threadIdx.x;
  const int nR = n / R;
  const int total = nl * nR;
  const int tstep = n / (p * R);
  C2<T> v[S][R];
  int li[S], ii[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    const int g = tid + s * FFT_NT;
    li[s] = -1;
    if (g < total) {
      int l, i;
      if (MAPEM) { i = (int)fnl.div(g); l = g - i * nl; }
      else { l = (int)fnR.div(g); i = g - l * nR; }
      li[s] = l;
      ii[s] = i;
#pragma unroll
      for (int j = 0; j < R; j++) v[s][j] = lds[l * ls + (i + j * nR) * es];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; s++) {
    if (li[s] >= 0) {
      const int l = li[s], i = ii[s];
      const int k = (int)fp.mod(i);
      if (k != 0) {
#pragma unroll
        for (int j = 1; j < R; j++) {
          C2<T> w = tw[j * k * tstep];
          if (DIR > 0) w.y = -w.y;
          v[s][j] = cmul(v[s][j], w);
        }
      }
      dft<R, DIR>(v[s]);
      const int j0 = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; j++) lds[l * ls + (j0 + j * p) * es] = v[s][j];
    }
  }
  __syncthreads();
}

template <typename T, int DIR, int MAPEM>
__device__ __forceinline__ void lds_fft(C2<T> *lds, int nl, const FastDiv &fnl, int ls, int es, const FftLen &f) {
  const C2<T> *tw = (const C2<T> *)f.tw;
  int p = 1;
  for (int s = 0; s < f.np; s++) {
    const int R = f.rad[s];
    switch (R) {
      case 2: fft_pass<T, 2, DIR, MAPEM>(lds, nl, fnl, f.n, p, f.fp[s], f.fnr[s], ls, es, tw); break;
      case 3: fft_pass<T, 3, DIR, MAPEM>(lds, nl, fnl, f.n, p, f.fp[s], f.fnr[s], ls, es, tw); break;
      case 4: fft_pass<T, 4, DIR, MAPEM>(lds, nl, fnl, f.n, p, f.fp[s], f.fnr[s], ls, es, tw); break;
      case 5: fft_pass<T, 5, DIR, MAPEM>(lds, nl, fnl, f.n, p, f.fp[s], f.fnr[s], ls, es, tw); break;
      case 7: fft_pass<T, 7, DIR, MAPEM>(lds, nl, fnl, f.n, p, f.fp[s], f.fnr[s], ls, es, tw); break;
      case 8: fft_pass<T, 8, DIR, MAPEM>(lds, nl, fnl, f.n, p, f.fp[s], f.fnr[s], ls, es, tw); break;
      default: fft_pass<T, 16, DIR, MAPEM>(lds, nl, fnl, f.n, p, f.fp[s], f.fnr[s], ls, es, tw); break;
    }
    p *= R;
  }
}

}  // namespace um
