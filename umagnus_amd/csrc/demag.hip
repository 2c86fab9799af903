// Demag convolution on gfx950: B = IFFT( K^ . FFT( pad(mu0 Ms vol m) ) ).
//
// Reference path replaced: opencl/conv_demag.go:43-85 (zero1_async + copypadmul2
// + VkFFT R2C per component, kernmulRSymm3D/2D, VkFFT C2R + copyunpad), the
// tensor generation mag/demagkernel.go:115-336 and DemagConvolution.init
// (opencl/conv_demag.go:131-202, conv_common.go:27-50).
//
// Layout in HBM.  Spectra are kept per component as [z][y][kx] complex rows of
// pitch Hp >= H = Px/2+1 (128-byte aligned rows).  Only the z-planes that hold
// data (z < Nz) are ever stored: the X and Y passes never write the padded
// half, the Z pass zero-extends in LDS, multiplies by the tensor and
// truncates again, and the inverse Y/X passes only emit y < Ny, x < Nx.
//
// Passes (3D): X-fwd R2C (reads m directly, fused mu0*Ms*vol scaling and
// zero padding) -> Y-fwd -> Z (fwd FFT, 6-component symmetric tensor
// multiply, inverse FFT, all in LDS) -> Y-inv -> X-inv C2R (writes the
// unpadded field).  2D (Nz==1): X-fwd -> Y (fwd, 4-component multiply, inv) -> X-inv.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>

#include "batch.hpp"
#include "fft_ct.hpp"
#include "field.hpp"
#include "sim.hpp"

namespace um {

static constexpr double MU0 = 4 * M_PI * 1e-7;

// Data movement between HBM and LDS uses fixed-trip, fully unrolled slots
// (slot s of thread t handles flat index g = t + s*FFT_NT, g < nl*n <= FFT_NT*EPT)
// so every global load of a work-group is in flight before the first LDS store.
//
// Every pass kernel is a template on (CN, CNL): CN > 0 instantiates it for one
// transform length and line count (compile-time LDS offsets, twiddles copied
// into LDS); CN == 0 is the run-time generic version used for other sizes.

template <int C>
__device__ __forceinline__ int cdiv(int x, const FastDiv &f) {
  if constexpr (C > 0) return x / C;
  else return (int)f.div(x);
}

// FFT over the LDS lines: compile-time engine when the size is fixed.
template <typename T, int CN, int CNL, int DIR, int MAPEM, int TWS>
__device__ __forceinline__ void run_fft(C2<T> *lds, const C2<T> *twl, int nl, const FastDiv &fnl, const FftLen &f) {
  if constexpr (CN > 0) {
    ct_fft<T, CN, CNL, DIR, MAPEM, TWS>(lds, twl);
  } else {
    if (MAPEM) lds_fft<T, DIR, 1>(lds, nl, fnl, 1, nl, f);
    else lds_fft<T, DIR, 0>(lds, nl, f.fn, f.n, 1, f);
  }
}

// ---------------------------------------------------------------------------
// X forward: rows of real data -> half spectrum (packed two-for-one R2C).
// ---------------------------------------------------------------------------
template <typename T>
struct XArgs {
  FftLen f;               // complex length n = Px/2
  const C2<T> *tw2;       // exp(-2 pi i k / Px), k in [0, Px)
  int mode;               // 0: mu0*Ms*vol*m rows of Nx valid reals; 1: raw real rows of Px
  int nl;                 // lines per work-group
  FastDiv fhalf;          // n/2 + 1
  int Nx;                 // valid reals per input row
  int rows;               // input rows per component
  FastDiv fRy;            // input rows per z-plane (Ny, or Py for mode 1)
  int Py;                 // spectrum rows per z-plane
  int Hp;                 // spectrum row pitch
  long in_pitch;          // input row pitch (reals)
  const T *src[3];
  Par<T> Ms;
  const T *vol;
  C2<T> *out[3];
};

template <typename T, int CN, int CNL>
struct BXfwd {
  using Args = XArgs<T>;
  static constexpr int NT = FFT_NT, MINB = 1;
  static __device__ __forceinline__ void run(const XArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    constexpr int EPT = FftEpt<T>::value;
    extern __shared__ __align__(16) unsigned char smem_[];
    C2<T> *lds = (C2<T> *)smem_;
    const int c = blockIdx.y;
    const int n = CN > 0 ? CN : a.f.n, nl = CN > 0 ? CNL : a.nl, total = nl * n;
    C2<T> *twl = lds + nl * n;   // W_{2n}^k, k < 2n (compile-time path only)
    if constexpr (CN > 0) {
      for (int g = threadIdx.x; g < 2 * CN; g += FFT_NT) twl[g] = a.tw2[g];
    }
    const int row0 = blockIdx.x * nl;
    const T *src = a.src[c];
    const int Nx = a.Nx;
    // uniform Ms and no geometry: one factor mu0*Ms (copypadmul2.cl:15-19 with vol = 1)
    const bool uniform = (a.mode != 0) || (a.Ms.arr == nullptr && a.vol == nullptr);
    const T fac = (a.mode != 0) ? T(1) : T(MU0) * a.Ms.mul;
    const bool pairs = ((Nx & 1) == 0) || a.mode != 0;   // rows start 8-byte aligned
    constexpr int SL = CN > 0 ? (CN * CNL + FFT_NT - 1) / FFT_NT : EPT;
    constexpr int BATCH = SL > EPT / 2 ? EPT / 2 : SL;
  #pragma unroll
    for (int s0 = 0; s0 < SL; s0 += BATCH) {
      C2<T> v[BATCH];
  #pragma unroll
      for (int s = 0; s < BATCH; s++) {
        const int g = threadIdx.x + (s0 + s) * FFT_NT;
        v[s] = C2<T>{T(0), T(0)};
        if (s0 + s < SL && g < total) {
          const int l = cdiv<CN>(g, a.f.fn), e = g - l * n;
          const int r = row0 + l;
          if (r < a.rows) {
            const long i0 = (long)r * a.in_pitch + 2 * e;
            if (pairs && 2 * e + 1 < (a.mode != 0 ? 2 * n : Nx)) {
              v[s] = *(const C2<T> *)(src + i0);
            } else {
              if (a.mode != 0 || 2 * e < Nx) v[s].x = src[i0];
              if (a.mode != 0 || 2 * e + 1 < Nx) v[s].y = src[i0 + 1];
            }
          }
        }
      }
  #pragma unroll
      for (int s = 0; s < BATCH; s++) {
        const int g = threadIdx.x + (s0 + s) * FFT_NT;
        if (s0 + s < SL && g < total) {
          C2<T> w = v[s];
          if (uniform) {
            w.x = fac * w.x;
            w.y = fac * w.y;
          } else {
            const int l = cdiv<CN>(g, a.f.fn), e = g - l * n;
            const int r = row0 + l;
            if (r < a.rows) {
              const long i0 = (long)r * a.in_pitch + 2 * e;
              if (2 * e < Nx) w.x = (T(MU0) * a.Ms.at(i0)) * (a.vol ? a.vol[i0] : T(1)) * w.x;
              if (2 * e + 1 < Nx) w.y = (T(MU0) * a.Ms.at(i0 + 1)) * (a.vol ? a.vol[i0 + 1] : T(1)) * w.y;
            }
          }
          lds[g] = w;
        }
      }
    }
    __syncthreads();
    run_fft<T, CN, CNL, -1, 0, 2>(lds, twl, nl, a.f.fn, a.f);
    // two-for-one post-processing: X[k] = E[k] + W^k O[k], X[n-k] = conj(E[k] - W^k O[k])
    const C2<T> *tw2 = CN > 0 ? twl : a.tw2;
    const int half = n / 2 + 1;
    for (int g = threadIdx.x; g < nl * half; g += FFT_NT) {
      const int l = cdiv<(CN > 0 ? CN / 2 + 1 : 0)>(g, a.fhalf), k = g - l * half;
      const int r = row0 + l;
      if (r >= a.rows) continue;
      const C2<T> A = lds[l * n + k];
      const C2<T> Bc = cconj(lds[l * n + (k == 0 ? 0 : n - k)]);
      const C2<T> E = C2<T>{T(0.5) * (A.x + Bc.x), T(0.5) * (A.y + Bc.y)};
      const C2<T> O = C2<T>{T(0.5) * (A.y - Bc.y), T(-0.5) * (A.x - Bc.x)};
      const C2<T> t = cmul(tw2[k], O);
      const int z = (int)a.fRy.div(r), y = r - z * (int)a.fRy.d;
      C2<T> *orow = a.out[c] + ((long)z * a.Py + y) * a.Hp;
      orow[k] = cadd(E, t);
      if (n - k != k) orow[n - k] = cconj(csub(E, t));
    }
  }
};
template <typename T, int CN, int CNL>
__global__ __launch_bounds__(FFT_NT) void k_xfwd(XArgs<T> a) { BXfwd<T, CN, CNL>::run(a, hw_bid(), hw_grid()); }

// ---------------------------------------------------------------------------
// X inverse: half spectrum rows -> real rows (C2R), keeping x < Nx.
// ---------------------------------------------------------------------------
template <typename T>
struct XiArgs {
  FftLen f;
  const C2<T> *tw2;
  int nl, Nx, rows, Py, Hp;
  FastDiv fhalf, fRy;
  const C2<T> *in[3];
  T *dst[3];
};

template <typename T, int CN, int CNL>
struct BXinv {
  using Args = XiArgs<T>;
  static constexpr int NT = FFT_NT, MINB = 1;
  static __device__ __forceinline__ void run(const XiArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    constexpr int EPT = FftEpt<T>::value;
    extern __shared__ __align__(16) unsigned char smem_[];
    C2<T> *lds = (C2<T> *)smem_;
    const int c = blockIdx.y;
    const int n = CN > 0 ? CN : a.f.n, nl = CN > 0 ? CNL : a.nl;
    C2<T> *twl = lds + nl * n;
    if constexpr (CN > 0) {
      for (int g = threadIdx.x; g < 2 * CN; g += FFT_NT) twl[g] = a.tw2[g];
    }
    const C2<T> *tw2 = a.tw2;   // pre-processing reads the global table (once per element)
    const int row0 = blockIdx.x * nl;
    const int half = n / 2 + 1;
    const int total = nl * half;
    constexpr int S = CN > 0 ? (CNL * (CN / 2 + 1) + FFT_NT - 1) / FFT_NT : EPT / 2 + 1;
    for (int g0 = 0; g0 < total; g0 += S * FFT_NT) {
      C2<T> xk[S], xn[S];
  #pragma unroll
      for (int s = 0; s < S; s++) {
        const int g = g0 + threadIdx.x + s * FFT_NT;
        xk[s] = C2<T>{T(0), T(0)};
        xn[s] = xk[s];
        if (g < total) {
          const int l = cdiv<(CN > 0 ? CN / 2 + 1 : 0)>(g, a.fhalf), k = g - l * half;
          const int r = row0 + l;
          if (r < a.rows) {
            const int z = (int)a.fRy.div(r), y = r - z * (int)a.fRy.d;
            const C2<T> *irow = a.in[c] + ((long)z * a.Py + y) * a.Hp;
            xk[s] = irow[k];
            xn[s] = irow[n - k];
          }
        }
      }
  #pragma unroll
      for (int s = 0; s < S; s++) {
        const int g = g0 + threadIdx.x + s * FFT_NT;
        if (g < total) {
          const int l = cdiv<(CN > 0 ? CN / 2 + 1 : 0)>(g, a.fhalf), k = g - l * half;
          const C2<T> Xk = xk[s], Xnk = xn[s];
          // Z'[k] = (X[k] + conj X[n-k]) + i W^-k (X[k] - conj X[n-k])
          C2<T> sm = cadd(Xk, cconj(Xnk)), d = csub(Xk, cconj(Xnk));
          C2<T> t = cmul(cconj(tw2[k]), d);
          lds[l * n + k] = C2<T>{sm.x - t.y, sm.y + t.x};
          if (k != 0 && n - k != k) {
            sm = cadd(Xnk, cconj(Xk));
            d = csub(Xnk, cconj(Xk));
            t = cmul(cconj(tw2[n - k]), d);
            lds[l * n + n - k] = C2<T>{sm.x - t.y, sm.y + t.x};
          }
        }
      }
    }
    __syncthreads();
    run_fft<T, CN, CNL, +1, 0, 2>(lds, twl, nl, a.f.fn, a.f);
    T *dst = a.dst[c];
    const int Nx = a.Nx;
    constexpr int SL = CN > 0 ? (CN * CNL + FFT_NT - 1) / FFT_NT : EPT;
  #pragma unroll
    for (int s = 0; s < SL; s++) {
      const int g = threadIdx.x + s * FFT_NT;
      if (g < nl * n) {
        const int l = cdiv<CN>(g, a.f.fn), e = g - l * n;
        const int r = row0 + l;
        if (r < a.rows) {
          const C2<T> v = lds[g];
          const long base = (long)r * Nx;
          if (2 * e + 1 < Nx && (Nx & 1) == 0) {
            *(C2<T> *)(dst + base + 2 * e) = v;
          } else {
            if (2 * e < Nx) dst[base + 2 * e] = v.x;
            if (2 * e + 1 < Nx) dst[base + 2 * e + 1] = v.y;
          }
        }
      }
    }
  }
};
template <typename T, int CN, int CNL>
__global__ __launch_bounds__(FFT_NT) void k_xinv(XiArgs<T> a) { BXinv<T, CN, CNL>::run(a, hw_bid(), hw_grid()); }

// ---------------------------------------------------------------------------
// X axis of odd length (periodic along x with odd Nx, so Px = Nx is odd and
// the two-for-one packing does not apply): each row is a full complex
// transform of length Px with zero imaginary parts, of which the H = Px/2+1
// non-redundant outputs are stored; the inverse rebuilds the Hermitian
// spectrum X[Px-k] = conj X[k] and keeps the real parts.  Small grids only
// (nl*Px <= FFT_NT*EPT per work-group), generic radix passes.
// a.f: FftLen of length Px; a.fhalf: H; a.nl lines per work-group.
// ---------------------------------------------------------------------------
template <typename T>
struct BXfwdOdd {
  using Args = XArgs<T>;
  static constexpr int NT = FFT_NT, MINB = 1;
  static __device__ __forceinline__ void run(const XArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    extern __shared__ __align__(16) unsigned char smem_[];
    C2<T> *lds = (C2<T> *)smem_;
    const int c = blockIdx.y;
    const int n = a.f.n, nl = a.nl, half = n / 2 + 1;
    const int row0 = blockIdx.x * nl;
    const T *src = a.src[c];
    for (int g = threadIdx.x; g < nl * n; g += FFT_NT) {
      const int l = (int)a.f.fn.div(g), e = g - l * n;
      const int r = row0 + l;
      T v = T(0);
      if (r < a.rows && e < a.Nx) {
        const long i = (long)r * a.in_pitch + e;
        v = src[i];
        if (a.mode == 0) v = (T(MU0) * a.Ms.at(i)) * (a.vol ? a.vol[i] : T(1)) * v;
      }
      lds[g] = C2<T>{v, T(0)};
    }
    __syncthreads();
    lds_fft<T, -1, 0>(lds, nl, a.f.fn, n, 1, a.f);
    for (int g = threadIdx.x; g < nl * half; g += FFT_NT) {
      const int l = (int)a.fhalf.div(g), k = g - l * half;
      const int r = row0 + l;
      if (r >= a.rows) continue;
      const int z = (int)a.fRy.div(r), y = r - z * (int)a.fRy.d;
      a.out[c][((long)z * a.Py + y) * a.Hp + k] = lds[l * n + k];
    }
  }
};
template <typename T>
__global__ __launch_bounds__(FFT_NT) void k_xfwd_odd(XArgs<T> a) { BXfwdOdd<T>::run(a, hw_bid(), hw_grid()); }

template <typename T>
struct BXinvOdd {
  using Args = XiArgs<T>;
  static constexpr int NT = FFT_NT, MINB = 1;
  static __device__ __forceinline__ void run(const XiArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    extern __shared__ __align__(16) unsigned char smem_[];
    C2<T> *lds = (C2<T> *)smem_;
    const int c = blockIdx.y;
    const int n = a.f.n, nl = a.nl, half = n / 2 + 1;
    const int row0 = blockIdx.x * nl;
    for (int g = threadIdx.x; g < nl * n; g += FFT_NT) {
      const int l = (int)a.f.fn.div(g), k = g - l * n;
      const int r = row0 + l;
      C2<T> v = C2<T>{T(0), T(0)};
      if (r < a.rows) {
        const int z = (int)a.fRy.div(r), y = r - z * (int)a.fRy.d;
        const C2<T> *irow = a.in[c] + ((long)z * a.Py + y) * a.Hp;
        v = k < half ? irow[k] : cconj(irow[n - k]);
      }
      lds[g] = v;
    }
    __syncthreads();
    lds_fft<T, +1, 0>(lds, nl, a.f.fn, n, 1, a.f);
    for (int g = threadIdx.x; g < nl * n; g += FFT_NT) {
      const int l = (int)a.f.fn.div(g), e = g - l * n;
      const int r = row0 + l;
      if (r < a.rows && e < a.Nx) a.dst[c][(long)r * a.Nx + e] = lds[g].x;
    }
  }
};
template <typename T>
__global__ __launch_bounds__(FFT_NT) void k_xinv_odd(XiArgs<T> a) { BXinvOdd<T>::run(a, hw_bid(), hw_grid()); }

// ---------------------------------------------------------------------------
// Register four-step X inverse (FP32, zero-padded rows: Px = 2 Nx, n = Px/2
// complex = A*B with B = A or 2A).  A group of B lanes owns G = B/A rows.
// Step 1: lane (h, n1) transforms z[n1 + A n2] over n2 (B points) in
// registers, twiddles by W_n^(n1 k2) and writes to LDS; step 2: lane k2
// transforms over n1 (A points) for each of the G rows, giving Z[k2 + B k1].
// The LDS transpose rows are padded to A+1 entries, so both the writes and
// the stride-(A+1) reads are bank-conflict free.
// ---------------------------------------------------------------------------
constexpr int XF4_NT = 128;

// Twiddle table of the step between the two register transforms of a
// four-step pass: W_n^(r k2) (r < R, k2 < K) at [r][k2], rows padded by one
// entry.  Loading and storing are split so that a kernel issues the table's
// global loads after its data loads and writes the table once both arrived:
// no memory round trip of its own ahead of the data.
template <typename T, int R, int K, int NT>
struct TwStage {
  static constexpr int Q = (R * K + NT - 1) / NT;
  C2<T> v[Q];
  __device__ __forceinline__ void load(const C2<T> *tw, int step) {   // tw[step r k2] = W_n^(r k2)
  #pragma unroll
    for (int q = 0; q < Q; q++) {
      const int i = (int)threadIdx.x + q * NT;
      if ((R * K) % NT == 0 || i < R * K) {
        const int r = i / K, k2 = i - r * K;
        v[q] = tw[step * r * k2];
      }
    }
  }
  __device__ __forceinline__ void store(C2<T> *lds) const {
  #pragma unroll
    for (int q = 0; q < Q; q++) {
      const int i = (int)threadIdx.x + q * NT;
      if ((R * K) % NT == 0 || i < R * K) {
        const int r = i / K, k2 = i - r * K;
        lds[r * (K + 1) + k2] = v[q];
      }
    }
  }
};

// exp(dir * 2 pi i j / n); constant once the caller's loops are unrolled
template <typename T>
__device__ __forceinline__ C2<T> cexpi(int j, int n, int dir) {
  const double ang = 2.0 * M_PI * (double)j / (double)n;
  return C2<T>{T(__builtin_cos(ang)), T(dir * __builtin_sin(ang))};
}

template <int N> struct X4Shape;
template <> struct X4Shape<64> { static constexpr int A = 8, B = 8; };
template <> struct X4Shape<128> { static constexpr int A = 8, B = 16; };
template <> struct X4Shape<256> { static constexpr int A = 16, B = 16; };
template <> struct X4Shape<512> { static constexpr int A = 16, B = 32; };
template <> struct X4Shape<1024> { static constexpr int A = 32, B = 32; };

// inverse: half spectrum X[k], k = 0..n -> real rows, x < Nx = n:
// Z'[k] = (X[k] + conj X[n-k]) + i W_2n^-k (X[k] - conj X[n-k]), z = IDFT_n(Z'),
// x[2j] = Re z[j], x[2j+1] = Im z[j] for j < n/2
template <typename T, int N>
struct BXi4 {
  using Args = XiArgs<T>;
  static constexpr int NT = XF4_NT, MINB = 1;
  static __device__ __forceinline__ void run(const XiArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    constexpr int A = X4Shape<N>::A, B = X4Shape<N>::B, G = B / A;
    constexpr int GROUPS = XF4_NT / B, ROWS = GROUPS * G, TS = B * (A + 1);
    __shared__ C2<T> lds[ROWS * TS];
    // twiddles W_n^(n1 k2) between the transforms, [n1][k2] (rows padded by one)
    __shared__ C2<T> twb[A * (B + 1)];
    TwStage<T, A, B, XF4_NT> twt;
    const int c = blockIdx.y;
    const int grp = threadIdx.x / B, l = threadIdx.x % B;
    const int row0 = blockIdx.x * ROWS + grp * G;
    {
      const int h = l / A, n1 = l % A, r = row0 + h;
      const C2<T> w1 = a.tw2[n1];   // W_2n^n1: W_2n^k = W_2n^n1 W_2n^(A n2) with the second factor constant
      C2<T> v[B];
  #pragma unroll
      for (int n2 = 0; n2 < B; n2++) v[n2] = C2<T>{T(0), T(0)};
      if (r < a.rows) {
        const int z = (int)a.fRy.div(r), y = r - z * (int)a.fRy.d;
        const C2<T> *irow = a.in[c] + ((long)z * a.Py + y) * a.Hp;
        C2<T> xk[B], xn[B];
  #pragma unroll
        for (int n2 = 0; n2 < B; n2++) {
          const int k = n1 + A * n2;
          xk[n2] = irow[k];
          xn[n2] = irow[N - k];
        }
        twt.load(a.tw2, 2);   // after the data loads
        __builtin_amdgcn_sched_barrier(0);
  #pragma unroll
        for (int n2 = 0; n2 < B; n2++) {
          const C2<T> sm = cadd(xk[n2], cconj(xn[n2])), d = csub(xk[n2], cconj(xn[n2]));
          const C2<T> wk = n2 == 0 ? w1 : cmul(w1, cexpi<T>(A * n2, 2 * N, -1));
          const C2<T> t = cmul(cconj(wk), d);
          v[n2] = C2<T>{sm.x - t.y, sm.y + t.x};
        }
      } else {
        twt.load(a.tw2, 2);
      }
      twt.store(twb);
      dft<B, +1>(v);
      __syncthreads();   // the twiddle table is complete
  #pragma unroll
      for (int k2 = 1; k2 < B; k2++) v[k2] = cmul(v[k2], cconj(twb[n1 * (B + 1) + k2]));   // W_n^-(n1 k2)
      C2<T> *t = lds + (grp * G + h) * TS + n1;
  #pragma unroll
      for (int k2 = 0; k2 < B; k2++) t[k2 * (A + 1)] = v[k2];
    }
    __syncthreads();
    T *dst = a.dst[c];
  #pragma unroll
    for (int h = 0; h < G; h++) {
      C2<T> u[A];
      const C2<T> *t = lds + (grp * G + h) * TS + l * (A + 1);
  #pragma unroll
      for (int n1 = 0; n1 < A; n1++) u[n1] = t[n1];
      dft<A, +1>(u);
      const int r = row0 + h;
      if (r < a.rows) {
        T *orow = dst + (long)r * N;
  #pragma unroll
        for (int k1 = 0; k1 < A / 2; k1++) *(float2 *)(orow + 2 * (l + B * k1)) = float2{u[k1].x, u[k1].y};
      }
    }
  }
};
template <typename T, int N>
__global__ __launch_bounds__(XF4_NT) void k_xi4(XiArgs<T> a) { BXi4<T, N>::run(a, hw_bid(), hw_grid()); }

// Register four-step X forward (FP32, Px = 2 Nx, n = Px/2 = Nx = A*B): the
// two-for-one packed row z[j] = x[2j] + i x[2j+1] is zero for j >= n/2, i.e.
// for n2 >= B/2, so step 1 loads half its points.  Step 1: lane (h, n1)
// transforms z[n1 + A n2] over n2 in registers and twiddles by W_n^(n1 k2);
// one LDS transpose (real, then imaginary parts) hands lane k2 the A values
// over n1; step 2 gives Z[k2 + B k1].  The unpacking
//   X[k] = E[k] + W_2n^k O[k],  E = (Z[k] + conj Z[n-k]) / 2,  O = (Z[k] - conj Z[n-k]) / 2i
// takes Z[n-k] from the group's lane B - k2 (lane 0 holds its own partner)
// through a cross-lane permute, and each lane stores the X[k] it holds
// (lane 0 also X[n]).
template <typename T, int N>
struct BXf4 {
  using Args = XArgs<T>;
  static constexpr int NT = XF4_NT, MINB = 1;
  static __device__ __forceinline__ void run(const XArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    constexpr int A = X4Shape<N>::A, B = X4Shape<N>::B, G = B / A;
    constexpr int GROUPS = XF4_NT / B, ROWS = GROUPS * G, TS = B * (A + 1), HB = B / 2;
    __shared__ T ldr[ROWS * TS];
    // twiddles W_n^(n1 k2) between the transforms, [n1][k2] (rows padded by one)
    __shared__ C2<T> twb[A * (B + 1)];
    TwStage<T, A, B, XF4_NT> twt;
    const int c = blockIdx.y;
    const int grp = threadIdx.x / B, l = threadIdx.x % B;
    const int row0 = blockIdx.x * ROWS + grp * G;
    const bool uniform = a.Ms.arr == nullptr && a.vol == nullptr;
    const T fac = T(MU0) * a.Ms.mul;
    const int h1 = l / A, n1 = l % A;
    {
      const int r = row0 + h1;
      C2<T> v[B];
  #pragma unroll
      for (int n2 = 0; n2 < B; n2++) v[n2] = C2<T>{T(0), T(0)};
      if (r < a.rows) {
        const long i0 = (long)r * a.in_pitch;
        const T *src = a.src[c] + i0;
  #pragma unroll
        for (int n2 = 0; n2 < HB; n2++) v[n2] = *(const C2<T> *)(src + 2 * (n1 + A * n2));
        twt.load(a.tw2, 2);   // after the data loads
        __builtin_amdgcn_sched_barrier(0);
        if (uniform) {
  #pragma unroll
          for (int n2 = 0; n2 < HB; n2++) v[n2] = C2<T>{fac * v[n2].x, fac * v[n2].y};
        } else {
  #pragma unroll
          for (int n2 = 0; n2 < HB; n2++) {
            const long i = i0 + 2 * (n1 + A * n2);
            v[n2].x = (T(MU0) * a.Ms.at(i)) * (a.vol ? a.vol[i] : T(1)) * v[n2].x;
            v[n2].y = (T(MU0) * a.Ms.at(i + 1)) * (a.vol ? a.vol[i + 1] : T(1)) * v[n2].y;
          }
        }
      } else {
        twt.load(a.tw2, 2);
      }
      twt.store(twb);
      dft<B, -1>(v);
      __syncthreads();   // the twiddle table is complete
  #pragma unroll
      for (int k2 = 1; k2 < B; k2++) v[k2] = cmul(v[k2], twb[n1 * (B + 1) + k2]);   // W_n^(n1 k2)
      T *tw = ldr + (grp * G + h1) * TS + n1;
  #pragma unroll
      for (int k2 = 0; k2 < B; k2++) tw[k2 * (A + 1)] = v[k2].x;
      __syncthreads();
      C2<T> u[G][A];
  #pragma unroll
      for (int h = 0; h < G; h++) {
        const T *t = ldr + (grp * G + h) * TS + l * (A + 1);
  #pragma unroll
        for (int q = 0; q < A; q++) u[h][q].x = t[q];
      }
      __syncthreads();
  #pragma unroll
      for (int k2 = 0; k2 < B; k2++) tw[k2 * (A + 1)] = v[k2].y;
      __syncthreads();
  #pragma unroll
      for (int h = 0; h < G; h++) {
        const T *t = ldr + (grp * G + h) * TS + l * (A + 1);
  #pragma unroll
        for (int q = 0; q < A; q++) u[h][q].y = t[q];
      }
      // unpack and store: lane l = k2 holds Z[k2 + B k1]; W_2n^k = W_2n^l W_2n^(B k1)
      const C2<T> wl = a.tw2[l];
      const int base = threadIdx.x - l;                       // first lane of the group
      const int src_lane = base + ((B - l) & (B - 1));        // holder of Z[n - k]
  #pragma unroll
      for (int h = 0; h < G; h++) {
        dft<A, -1>(u[h]);
        const int rr = row0 + h;
        const int z = (int)a.fRy.div(rr < a.rows ? rr : 0), y = (rr < a.rows ? rr : 0) - z * (int)a.fRy.d;
        C2<T> *orow = a.out[c] + ((long)z * a.Py + y) * a.Hp;
  #pragma unroll
        for (int k1 = 0; k1 < A; k1++) {
          // partner element: A-1-k1 of lane B-l, or (A-k1) mod A of lane 0 itself
          const C2<T> q = u[h][A - 1 - k1];
          C2<T> Zp{__shfl(q.x, src_lane, 64), __shfl(q.y, src_lane, 64)};
          if (l == 0) Zp = u[h][(A - k1) & (A - 1)];
          const C2<T> Zk = u[h][k1];
          const C2<T> Bc = cconj(Zp);
          const C2<T> E = C2<T>{T(0.5) * (Zk.x + Bc.x), T(0.5) * (Zk.y + Bc.y)};
          const C2<T> O = C2<T>{T(0.5) * (Zk.y - Bc.y), T(-0.5) * (Zk.x - Bc.x)};
          const int k = l + B * k1;
          const C2<T> t = cmul(k1 == 0 ? wl : cmul(wl, cexpi<T>(B * k1, 2 * N, -1)), O);
          if (rr < a.rows) {
            orow[k] = cadd(E, t);
            if (k == 0) orow[N] = cconj(csub(E, t));
          }
        }
      }
    }
  }
};
template <typename T, int N>
__global__ __launch_bounds__(XF4_NT) void k_xf4(XArgs<T> a) { BXf4<T, N>::run(a, hw_bid(), hw_grid()); }

// ---------------------------------------------------------------------------
// Column pass along y (or z): nl adjacent kx columns per work-group.
// ---------------------------------------------------------------------------
template <typename T>
struct ColArgs {
  FftLen f;
  int nl;          // columns per work-group
  FastDiv fnl;
  int H;           // valid kx columns
  int nvalid;      // input elements that may be non-zero
  int nout;        // output elements stored
  long estride;    // element stride (complex) along the transform axis
  long pstride;    // stride between blockIdx.y planes (complex)
  C2<T> *buf[3];
};

template <typename T, int DIR, int CN, int CNL>
struct BColfft {
  using Args = ColArgs<T>;
  static constexpr int NT = FFT_NT, MINB = 1;
  static __device__ __forceinline__ void run(const ColArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    constexpr int EPT = FftEpt<T>::value;
    extern __shared__ __align__(16) unsigned char smem_[];
    C2<T> *lds = (C2<T> *)smem_;
    const int n = CN > 0 ? CN : a.f.n, nl = CN > 0 ? CNL : a.nl, total = nl * n;
    C2<T> *twl = lds + nl * n;
    if constexpr (CN > 0) {
      const C2<T> *tw = (const C2<T> *)a.f.tw;
      for (int g = threadIdx.x; g < CN; g += FFT_NT) twl[g] = tw[g];
    }
    const int kx0 = blockIdx.x * nl;
    C2<T> *base = a.buf[blockIdx.z] + (long)blockIdx.y * a.pstride;
    constexpr int SL = CN > 0 ? (CN * CNL + FFT_NT - 1) / FFT_NT : EPT;
    C2<T> v[SL];
  #pragma unroll
    for (int s = 0; s < SL; s++) {
      const int g = threadIdx.x + s * FFT_NT;
      v[s] = C2<T>{T(0), T(0)};
      if (g < total) {
        const int e = cdiv<CNL>(g, a.fnl), l = g - e * nl;
        const int kx = kx0 + l;
        if (kx < a.H && e < a.nvalid) v[s] = base[(long)e * a.estride + kx];
      }
    }
  #pragma unroll
    for (int s = 0; s < SL; s++) {
      const int g = threadIdx.x + s * FFT_NT;
      if (g < total) lds[g] = v[s];   // element-major: g = e*nl + l
    }
    __syncthreads();
    run_fft<T, CN, CNL, DIR, 1, 1>(lds, twl, nl, a.fnl, a.f);
  #pragma unroll
    for (int s = 0; s < SL; s++) {
      const int g = threadIdx.x + s * FFT_NT;
      if (g < total) {
        const int e = cdiv<CNL>(g, a.fnl), l = g - e * nl;
        const int kx = kx0 + l;
        if (kx < a.H && e < a.nout) base[(long)e * a.estride + kx] = lds[g];
      }
    }
  }
};
template <typename T, int DIR, int CN, int CNL>
__global__ __launch_bounds__(FFT_NT) void k_colfft(ColArgs<T> a) { BColfft<T, DIR, CN, CNL>::run(a, hw_bid(), hw_grid()); }

// ---------------------------------------------------------------------------
// Register four-step column pass for P = A*A (A in 8, 16, 32): C adjacent kx
// columns per work-group, A lanes per column.  With n = n1 + A n2 and
// k = k2 + A k1,
//   X[k2 + A k1] = sum_n1 W_A^(n1 k1) W_P^(n1 k2) sum_n2 x[n1 + A n2] W_A^(n2 k2):
// lane n1 transforms its strided A elements in registers, twiddles them,
// one LDS transpose hands lane k2 the A values over n1, and a second
// register transform gives the outputs k2 + A k1.  HALF: the forward input is
// zero for n >= P/2 (n2 >= A/2, never loaded) and the inverse keeps only
// k < P/2 (k1 < A/2).  One LDS round trip per element instead of one per
// radix pass.
// ---------------------------------------------------------------------------
// columns per work-group of the P = 1024 pass (build-time tunable)
#ifndef UM_COLY_C32
#define UM_COLY_C32 16
#endif
// z planes per work-group of the FP32 Y passes (build-time tunable)
#ifndef UM_COLY_ZB
#define UM_COLY_ZB 1
#endif

// Work-group index permuted so that the tiles one XCD runs (linear ids w,
// w+8, w+16, ... are dispatched to the same XCD) are adjacent in memory and
// share their 128-byte lines in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int w, int total) {
  const int q = total >> 3, full = q << 3;
  return w >= full ? w : (w & 7) * q + (w >> 3);
}

template <typename T, int A, int DIR, int C, bool HALF, int ZB>
struct BColy {
  using Args = ColArgs<T>;
  static constexpr int NT = A * C, MINB = (HALF && sizeof(C2<T>) * A * (A * C + C) > 65536) ? 4 : 1;
  static __device__ __forceinline__ void run(const ColArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    // transpose row stride A*C + C: a wave's 64 b64 accesses then share each
    // LDS bank pair exactly twice (the minimum for 8-byte lanes)
    constexpr int P = A * A, S = A * C + C;
    // SPLIT: a transpose of A x (A*C) complex values would hold 135 KB of LDS
    // at A = 32 (one work-group per CU); exchanging the real and then the
    // imaginary parts halves it, so two work-groups (4 waves per SIMD, hence
    // the launch bound of the half-column passes; the full-column inverse
    // would spill under it) overlap their load and arithmetic phases, without
    // extra registers (each half is read back into the slots just written out)
    constexpr bool SPLIT = sizeof(C2<T>) * A * S > 65536;
    extern __shared__ __align__(16) unsigned char smem_[];
    C2<T> *lds = (C2<T> *)smem_;
    T *ldr = (T *)smem_;
    // twiddles W_P^(i k2) of the step between the transforms, [i][k2] with
    // rows padded by one entry (the A rows of a wave fall on distinct banks),
    // after the transpose space: LDS reads instead of global-memory reads in
    // the middle of the transform
    C2<T> *twt = (C2<T> *)(smem_ + (SPLIT ? sizeof(T) : sizeof(C2<T>)) * A * S);
    TwStage<T, A, A, A * C> twv;   // loaded after the first plane's data
    const int ngx = gridDim.x, ngy = gridDim.y;
    const int w = xcd_tile(blockIdx.x + ngx * (blockIdx.y + ngy * blockIdx.z), ngx * ngy * gridDim.z);
    const int gx = w % ngx, gyz = w / ngx, gy = gyz % ngy, gz = gyz / ngy;
    const int col = threadIdx.x % C, i = threadIdx.x / C;
    const int kx = gx * C + col;
    const bool valid = kx < a.H;
    // uniform plane base + one 32-bit lane offset (row i, column kx)
    const uint32_t loff = (uint32_t)(((long)i * a.estride + min(kx, a.H - 1)) * sizeof(C2<T>));
    const long rstep = (long)A * a.estride;   // rows n2 -> n2 + 1
    constexpr int NIN = (HALF && DIR < 0) ? A / 2 : A;
    constexpr int NOUT = (HALF && DIR > 0) ? A / 2 : A;

    // ZB planes per work-group; the loads of plane p+1 are issued before plane p
    // is transformed, so they overlap its arithmetic and LDS exchange
    C2<T> nxt[NIN];
    auto load = [&](int plane) {
      C2<T> *ub = a.buf[gz] + (long)plane * a.pstride;
  #pragma unroll
      for (int n2 = 0; n2 < NIN; n2++) {
        const auto *q = at_off<const T>(ub + n2 * rstep, loff);
        nxt[n2] = C2<T>{q[0], q[1]};
      }
    };
    load(gy * ZB);
    twv.load((const C2<T> *)a.f.tw, 1);   // W_P^m, m < P
    __builtin_amdgcn_sched_barrier(0);
    twv.store(twt);
  #pragma unroll
    for (int p = 0; p < ZB; p++) {
      C2<T> v[A];
  #pragma unroll
      for (int n2 = 0; n2 < A; n2++) v[n2] = n2 < NIN ? nxt[n2] : C2<T>{T(0), T(0)};
      if (p + 1 < ZB) load(gy * ZB + p + 1);
      dft<A, DIR>(v);
      pin(v);
      if (p == 0) __syncthreads();   // the twiddle table is complete
  #pragma unroll
      for (int k2 = 1; k2 < A; k2++) {
        C2<T> t = twt[i * (A + 1) + k2];
        if (DIR > 0) t.y = -t.y;
        v[k2] = cmul(v[k2], t);
      }
      if (p > 0) __syncthreads();   // the previous plane's transpose reads are done
      if constexpr (SPLIT) {
  #pragma unroll
        for (int k2 = 0; k2 < A; k2++) ldr[k2 * S + threadIdx.x] = v[k2].x;
        __syncthreads();
  #pragma unroll
        for (int n1 = 0; n1 < A; n1++) v[n1].x = ldr[i * S + n1 * C + col];
        __syncthreads();
  #pragma unroll
        for (int k2 = 0; k2 < A; k2++) ldr[k2 * S + threadIdx.x] = v[k2].y;
        __syncthreads();
  #pragma unroll
        for (int n1 = 0; n1 < A; n1++) v[n1].y = ldr[i * S + n1 * C + col];
      } else {
  #pragma unroll
        for (int k2 = 0; k2 < A; k2++) lds[k2 * S + threadIdx.x] = v[k2];
        __syncthreads();
  #pragma unroll
        for (int n1 = 0; n1 < A; n1++) v[n1] = lds[i * S + n1 * C + col];
      }
      dft<A, DIR>(v);
      if (valid) {
        C2<T> *ub = a.buf[gz] + (long)(gy * ZB + p) * a.pstride;
  #pragma unroll
        for (int k1 = 0; k1 < NOUT; k1++) {
          auto *q = at_off<T>(ub + k1 * rstep, loff);
          q[0] = v[k1].x;
          q[1] = v[k1].y;
        }
      }
    }
  }
};
template <typename T, int A, int DIR, int C, bool HALF, int ZB>
__global__ __launch_bounds__(A * C, (HALF && sizeof(C2<T>) * A * (A * C + C) > 65536) ? 4 : 1) void k_coly(ColArgs<T> a) { BColy<T, A, DIR, C, HALF, ZB>::run(a, hw_bid(), hw_grid()); }

// ---------------------------------------------------------------------------
// Fused convolution pass along the last axis: forward FFT of the three
// components, symmetric tensor multiply, inverse FFT (kernmulrsymm3d.cl:27-91,
// kernmulrsymm2dxy.cl / 2dz.cl for 2D).
// ---------------------------------------------------------------------------
template <typename T>
struct ConvArgs {
  FftLen f;        // length along the convolved axis (Pz in 3D, Py in 2D)
  int B;           // kx columns per work-group (lines = 3B)
  FastDiv fB, f3B;
  int H;           // valid kx columns
  int nvalid;      // non-zero input elements along the axis (Nz or Ny)
  long estride;    // element stride along the axis (complex)
  long pstride;    // stride of blockIdx.y (the other k index: ky in 3D, unused in 2D)
  int is2d;
  int Py, Pz;      // logical padded sizes for mirror indexing
  int KHy;         // Py/2+1
  int Hk;          // tensor row pitch (H rounded up to a 128-byte line)
  const T *K[6];   // xx yy zz yz xz xy, each Hk x (Py/2+1) x (Pz/2+1)
  C2<T> *buf[3];
};

constexpr int CONV_QK = 6;   // tensor prefetch slots per thread: B*(n/2+1) <= FFT_NT*CONV_QK
// a pass launching fewer work-groups than this (3 per CU on 256 CUs) switches
// to a variant with fewer lines per work-group when one is compiled in
constexpr long SMALL_GRID_WG = 768;

template <typename T, int CN, int CB>
struct BConv {
  using Args = ConvArgs<T>;
  static constexpr int NT = FFT_NT, MINB = 1;
  static __device__ __forceinline__ void run(const ConvArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    constexpr int EPT = FftEpt<T>::value;
    extern __shared__ __align__(16) unsigned char smem_[];
    C2<T> *lds = (C2<T> *)smem_;
    const int n = CN > 0 ? CN : a.f.n, B = CN > 0 ? CB : a.B, nl = 3 * B, total = nl * n;
    C2<T> *twl = lds + nl * n;
    if constexpr (CN > 0) {
      const C2<T> *tw = (const C2<T> *)a.f.tw;
      for (int g = threadIdx.x; g < CN; g += FFT_NT) twl[g] = tw[g];
    }
    const int kx0 = blockIdx.x * B;
    // 3D: this work-group handles ky and its mirror Py-ky, which use the same
    // tensor values (kernmulrsymm3d.cl:62-66 flips the yz and xy signs); 2D: one tile
    const int kya = blockIdx.y;
    const int kyb = a.Py - kya;
    const int ntiles = (!a.is2d && kya != 0 && kyb != kya) ? 2 : 1;
    const int nh = n / 2 + 1;   // unique points along the axis: k and n-k share |K|
    // tensor values of frequency e, column kx (2D: yz and xz are not used)
    auto tensor = [&](int e, int kx, T *k) {
  #pragma unroll
      for (int j = 0; j < 6; j++) k[j] = T(0);
      if (!a.is2d) {
        const long I = ((long)e * a.KHy + kya) * a.Hk + kx;
        k[0] = a.K[0][I]; k[1] = a.K[1][I]; k[2] = a.K[2][I];
        k[3] = a.K[3][I]; k[4] = a.K[4][I]; k[5] = a.K[5][I];
      } else {
        const long I = (long)e * a.Hk + kx;
        k[0] = a.K[0][I]; k[1] = a.K[1][I]; k[2] = a.K[2][I]; k[5] = a.K[5][I];
      }
    };
    // 1. tensor prefetch (independent of the data, so it overlaps the loads
    // below); the run-time-length kernel reads it at the multiply instead, to
    // keep its register count down
    constexpr int QK = CN > 0 ? (CB * (CN / 2 + 1) + FFT_NT - 1) / FFT_NT : 1;
    T kv[QK][6];
    if constexpr (CN > 0) {
  #pragma unroll
      for (int q = 0; q < QK; q++) {
        const int g = threadIdx.x + q * FFT_NT;
  #pragma unroll
        for (int j = 0; j < 6; j++) kv[q][j] = T(0);
        if (g < B * nh) {
          const int e = cdiv<CB>(g, a.fB), b = g - e * B;
          if (kx0 + b < a.H) tensor(e, kx0 + b, kv[q]);
        }
      }
    }
    // 2. spectrum loads of a tile
    // (fixed-length kernels: every load of the tile in flight at once; the
    // run-time-length kernel streams them to keep its register count down)
    constexpr int SL = CN > 0 ? (3 * CB * CN + FFT_NT - 1) / FFT_NT : 1;
    (void)EPT;
    C2<T> v[SL];
    auto load = [&](int ky) {
      const long off = (long)ky * a.pstride;
  #pragma unroll
      for (int s = 0; s < SL; s++) {
        const int g = threadIdx.x + s * FFT_NT;
        v[s] = C2<T>{T(0), T(0)};
        if (g < total) {
          const int e = cdiv<3 * CB>(g, a.f3B), l = g - e * nl;
          const int c = cdiv<CB>(l, a.fB), b = l - c * B;
          const int kx = kx0 + b;
          if (kx < a.H && e < a.nvalid) v[s] = a.buf[c][off + (long)e * a.estride + kx];
        }
      }
    };
    auto mul = [&](C2<T> *p, T xx, T yy, T zz, T yz, T xz, T xy) {
      const C2<T> M0 = p[0], M1 = p[B], M2 = p[2 * B];
      if (!a.is2d) {
        p[0] = C2<T>{M0.x * xx + M1.x * xy + M2.x * xz, M0.y * xx + M1.y * xy + M2.y * xz};
        p[B] = C2<T>{M0.x * xy + M1.x * yy + M2.x * yz, M0.y * xy + M1.y * yy + M2.y * yz};
        p[2 * B] = C2<T>{M0.x * xz + M1.x * yz + M2.x * zz, M0.y * xz + M1.y * yz + M2.y * zz};
      } else {
        p[0] = C2<T>{M0.x * xx + M1.x * xy, M0.y * xx + M1.y * xy};
        p[B] = C2<T>{M0.x * xy + M1.x * yy, M0.y * xy + M1.y * yy};
        p[2 * B] = C2<T>{M2.x * zz, M2.y * zz};
      }
    };
    for (int t = 0; t < ntiles; t++) {
      const int ky = (t == 0) ? kya : kyb;
      if constexpr (CN > 0) {
        load(ky);
  #pragma unroll
        for (int s = 0; s < SL; s++) {
          const int g = threadIdx.x + s * FFT_NT;
          if (g < total) lds[g] = v[s];
        }
      } else {
        const long off = (long)ky * a.pstride;
  #pragma unroll 4
        for (int g = threadIdx.x; g < total; g += FFT_NT) {
          const int e = cdiv<3 * CB>(g, a.f3B), l = g - e * nl;
          const int c = cdiv<CB>(l, a.fB), b = l - c * B;
          const int kx = kx0 + b;
          C2<T> w{T(0), T(0)};
          if (kx < a.H && e < a.nvalid) w = a.buf[c][off + (long)e * a.estride + kx];
          lds[g] = w;
        }
      }
      __syncthreads();
      run_fft<T, CN, 3 * CB, -1, 1, 1>(lds, twl, nl, a.f3B, a.f);
      // 3. tensor multiply at k and at its mirror n-k (signs of the odd components flip)
      const T sy = (t == 0) ? T(1) : T(-1);   // ky mirror: yz and xy change sign
      auto multiply = [&](int e, int b, const T *k) {
        const T yz = sy * k[3], xy = sy * k[5];
        mul(lds + e * nl + b, k[0], k[1], k[2], yz, k[4], xy);
        const int e2 = n - e;
        if (e != 0 && e2 != e) {
          // 3D mirror in z flips yz and xz (kernmulrsymm3d.cl:67-71); 2D mirror in y flips xy
          if (!a.is2d) mul(lds + e2 * nl + b, k[0], k[1], k[2], -yz, -k[4], xy);
          else mul(lds + e2 * nl + b, k[0], k[1], k[2], k[3], k[4], -k[5]);
        }
      };
      if constexpr (CN > 0) {
  #pragma unroll
        for (int q = 0; q < QK; q++) {
          const int g = threadIdx.x + q * FFT_NT;
          if (g < B * nh) {
            const int e = cdiv<CB>(g, a.fB), b = g - e * B;
            if (kx0 + b < a.H) multiply(e, b, kv[q]);
          }
        }
      } else {
  #pragma unroll 1
        for (int g = threadIdx.x; g < B * nh; g += FFT_NT) {
          const int e = cdiv<CB>(g, a.fB), b = g - e * B;
          if (kx0 + b < a.H) {
            T k[6];
            tensor(e, kx0 + b, k);
            multiply(e, b, k);
          }
        }
      }
      __syncthreads();
      run_fft<T, CN, 3 * CB, +1, 1, 1>(lds, twl, nl, a.f3B, a.f);
      const long off = (long)ky * a.pstride;
      auto store = [&](int g) {
        const int e = cdiv<3 * CB>(g, a.f3B), l = g - e * nl;
        const int c = cdiv<CB>(l, a.fB), b = l - c * B;
        const int kx = kx0 + b;
        if (kx < a.H && e < a.nvalid) a.buf[c][off + (long)e * a.estride + kx] = lds[g];
      };
      if constexpr (CN > 0) {
  #pragma unroll
        for (int s = 0; s < SL; s++) {
          const int g = threadIdx.x + s * FFT_NT;
          if (g < total) store(g);
        }
      } else {
  #pragma unroll 4
        for (int g = threadIdx.x; g < total; g += FFT_NT) store(g);
      }
      __syncthreads();
    }
  }
};
template <typename T, int CN, int CB>
__global__ __launch_bounds__(FFT_NT) void k_conv(ConvArgs<T> a) { BConv<T, CN, CB>::run(a, hw_bid(), hw_grid()); }

// ---------------------------------------------------------------------------
// Register Z pass (3D with Pz = 2 Nz, Nz in {8, 16, 32}): one lane per
// (component, kx, ky) column.  The Nz non-zero inputs give the even and odd
// halves of the Pz-point spectrum,
//   X[2m] = DFT_Nz(x)[m],   X[2m+1] = DFT_Nz(x W_Pz^n)[m],
// both transformed in registers with compile-time twiddles.  The three
// components of a column meet in LDS only for the tensor multiply, FR
// frequencies per round; the inverse runs in registers again,
//   y[n] = IDFT_Nz(Y_even)[n] + W_Pz^-n IDFT_Nz(Y_odd)[n],   n < Nz,
// and only the Nz planes that hold data are stored.  A work-group is three
// waves, one per component; lanes 0-31 take ky and lanes 32-63 its mirror
// Py-ky, which read the same tensor values (kernmulrsymm3d.cl:62-66).
// When H = 32 q + 1 (H = Px/2 + 1 with Px a multiple of 64), the last kx,
// the Nyquist column, would take a work-group of its own for every ky with
// one useful lane in 32; instead the blocks of the last grid column hold the
// Nyquist column of 32 ky pairs each (lane j: ky pair 32 blockIdx.y + j) and
// the remaining blocks of that column return at once.
// ---------------------------------------------------------------------------
template <typename T>
struct ConvZArgs {
  int H;           // valid kx columns
  int nyq;         // 1: the last grid column holds the Nyquist column (H = 32 q + 1)
  int Py, KHy;     // padded y size and Py/2+1
  int Hk;          // tensor row pitch (H rounded up to a 128-byte line)
  long zstride;    // complex elements between z planes of a component
  long ystride;    // complex elements between ky rows
  const T *K;      // xx yy zz yz xz xy as [ky][kz][element][kx] (row pitch Hk): one ky's slab contiguous
  C2<T> *buf[3];
};

constexpr int CONVZ_NT = 192;
constexpr int CONVZ_COLS = 32;


// Visiting order of the Pz frequencies: 0, Nz, then the mirror pairs (p, Pz-p)
// next to each other so a pair shares its tensor loads.
template <int NZ>
__device__ __forceinline__ constexpr int zorder(int i) {
  return i == 0 ? 0 : i == 1 ? NZ : (i & 1) ? 2 * NZ - i / 2 : i / 2;
}

template <typename T, int NZ, int FR, int G>
struct BConvz {
  using Args = ConvZArgs<T>;
  static constexpr int NT = CONVZ_NT, MINB = 1;
  static __device__ __forceinline__ void run(const ConvZArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    constexpr int PZ = 2 * NZ, KZ = NZ + 1;
    static_assert(PZ % FR == 0 && FR % G == 0, "rounds must tile the spectrum");
    __shared__ C2<T> xch[FR * 3 * 64];             // exchange: [freq][component][lane]
    __shared__ T tens[KZ * 6 * CONVZ_COLS];        // tensor slab: [kz][element][kx]
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    const int t = l >> 5;
    // Nyquist blocks: one kx (H - 1), 32 ky pairs across the lanes
    const bool nyqb = a.nyq && blockIdx.x == gridDim.x - 1;
    if (nyqb && (int)blockIdx.y * CONVZ_COLS >= a.KHy) return;   // uniform: the whole work-group
    const int kx0 = nyqb ? a.H - 1 : blockIdx.x * CONVZ_COLS;
    const int kx = nyqb ? kx0 : kx0 + (l & 31);
    const int kyq = nyqb ? (int)blockIdx.y * CONVZ_COLS + (l & 31) : (int)blockIdx.y;
    const int kya = min(kyq, a.KHy - 1);
    const bool mirror = kya != 0 && 2 * kya != a.Py;
    const bool valid = kx < a.H && kyq < a.KHy && (t == 0 || mirror);
    // every lane loads in-bounds data (clamped column / row); only valid lanes store
    const int kxc = min(kx, a.H - 1);
    const int ky = (t && mirror) ? a.Py - kya : kya;
    // Addresses are a wave-uniform base (plane n, scalar registers) plus one
    // per-lane 32-bit byte offset shared by every load and store of the lane.
    C2<T> *plane0 = a.buf[c];
    const uint32_t loff = (uint32_t)(((long)ky * a.ystride + kxc) * sizeof(C2<T>));

    // 0. tensor slab of this work-group (32 kx x 6 elements x Nz+1 kz) and the
    // lane's column: every global load is issued before the first use of any
    // (fixed trip counts, unrolled, then a scheduling barrier).  Thread t
    // loads element e = t / 32 of slab column t % 32 for every kz.  The
    // tensor is stored [ky][kz][element][kx], so the slab of one ky pair is
    // one contiguous range (no per-lane read of a pointer argument).
    static_assert(CONVZ_NT == 6 * CONVZ_COLS, "one slab element per 32 threads");
    const int j = threadIdx.x & (CONVZ_COLS - 1), eld = threadIdx.x / CONVZ_COLS;
    T tv[KZ];
    {
      const long kzs = 6L * a.Hk;
      // slab column j: kx0 + j at this ky, or (Nyquist block) kx0 at ky pair 32 blockIdx.y + j
      const int x = nyqb ? kx0 : min(kx0 + j, a.H - 1);
      const long ky_ = nyqb ? min((int)blockIdx.y * CONVZ_COLS + j, a.KHy - 1) : kya;
      const T *Ke = a.K + ((ky_ * KZ) * 6 + eld) * a.Hk + x;
  #pragma unroll
      for (int kz = 0; kz < KZ; kz++) tv[kz] = Ke[kz * kzs];
    }

    // 1. forward: even and odd halves of the spectrum
    C2<T> E[NZ], O[NZ];
  #pragma unroll
    for (int n = 0; n < NZ; n++) {
      const auto *p = at_off<const T>(plane0 + n * a.zstride, loff);
      E[n] = C2<T>{p[0], p[1]};
    }
    __builtin_amdgcn_sched_barrier(0);
  #pragma unroll
    for (int kz = 0; kz < KZ; kz++) tens[(kz * 6 + eld) * CONVZ_COLS + j] = tv[kz];
  #pragma unroll
    for (int n = 0; n < NZ; n++) O[n] = n == 0 ? E[0] : cmul(E[n], cexpi<T>(n, PZ, -1));
    // one transform at a time: interleaving the two raises the register peak
    dft<NZ, -1>(E);
    pin(E);
    dft<NZ, -1>(O);
    pin(O);

    // 2. tensor multiply: component c of K . (X0, X1, X2) for each frequency.
    // Row c of the symmetric tensor: x: (xx, xy, xz), y: (xy, yy, yz), z: (xz, yz, zz).
    const int e0 = c == 0 ? 0 : c == 1 ? 5 : 4;
    const int e1 = c == 0 ? 5 : c == 1 ? 1 : 3;
    const int e2 = c == 0 ? 4 : c == 1 ? 3 : 2;
    // the ky mirror flips xy and yz; the kz mirror flips xz and yz
    const T sy = t ? T(-1) : T(1);
    auto fy = [&](int e) { return (e == 5 || e == 3) ? sy : T(1); };
    auto fz = [&](int e) { return (e == 4 || e == 3) ? T(-1) : T(1); };
    const T y0 = fy(e0), y1 = fy(e1), y2 = fy(e2);
    const T z0 = y0 * fz(e0), z1 = y1 * fz(e1), z2 = y2 * fz(e2);
    const T *tl = tens + (l & (CONVZ_COLS - 1));
    C2<T> *xw = xch + c * 64 + l;
  #pragma unroll
    for (int r = 0; r < PZ / FR; r++) {
  #pragma unroll
      for (int j = 0; j < FR; j++) {
        const int k = zorder<NZ>(r * FR + j);
        xw[j * 192] = (k & 1) ? O[k >> 1] : E[k >> 1];
      }
      __syncthreads();
  #pragma unroll
      for (int j0 = 0; j0 < FR; j0 += G) {
        // groups of G frequencies: bounded look-ahead keeps the live registers low
        __builtin_amdgcn_sched_barrier(0);
  #pragma unroll
        for (int j = j0; j < j0 + G; j++) {
          const int k = zorder<NZ>(r * FR + j);
          const int kz = k <= NZ ? k : PZ - k;
          const C2<T> X0 = xch[(j * 3 + 0) * 64 + l];
          const C2<T> X1 = xch[(j * 3 + 1) * 64 + l];
          const C2<T> X2 = xch[(j * 3 + 2) * 64 + l];
          const T *tk = tl + kz * 6 * CONVZ_COLS;
          const T k0 = tk[e0 * CONVZ_COLS] * (k <= NZ ? y0 : z0);
          const T k1 = tk[e1 * CONVZ_COLS] * (k <= NZ ? y1 : z1);
          const T k2 = tk[e2 * CONVZ_COLS] * (k <= NZ ? y2 : z2);
          C2<T> Y{X0.x * k0 + X1.x * k1 + X2.x * k2, X0.y * k0 + X1.y * k1 + X2.y * k2};
          // keep the product here: otherwise it sinks into the guarded store
          // block below, and all LDS reads of the kernel stay live at once
          asm volatile("" : "+v"(Y.x), "+v"(Y.y));
          if (k & 1) O[k >> 1] = Y;
          else E[k >> 1] = Y;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();
    }

    // 3. inverse and store of the Nz data planes
    dft<NZ, +1>(E);
    pin(E);
    dft<NZ, +1>(O);
    pin(O);
    if (valid) {
  #pragma unroll
      for (int n = 0; n < NZ; n++) {
        const C2<T> o = n == 0 ? O[0] : cmul(O[n], cexpi<T>(n, PZ, +1));
        const C2<T> y = cadd(E[n], o);
        auto *p = at_off<T>(plane0 + n * a.zstride, loff);
        p[0] = y.x;
        p[1] = y.y;
      }
    }
  }
};
template <typename T, int NZ, int FR, int G>
__global__ __launch_bounds__(CONVZ_NT) void k_convz(ConvZArgs<T> a) { BConvz<T, NZ, FR, G>::run(a, hw_bid(), hw_grid()); }

template <typename T> struct ConvZTable;
// frequencies per LDS exchange round and per scheduling group (build-time tunables)
#ifndef UM_CONVZ_FR
#define UM_CONVZ_FR 8
#endif
#ifndef UM_CONVZ_G
#define UM_CONVZ_G 4
#endif
template <> struct ConvZTable<float> {
  static constexpr int FR = UM_CONVZ_FR, G = UM_CONVZ_G;
  static bool has(int nz) { return nz == 8 || nz == 16 || nz == 32; }
};
template <> struct ConvZTable<double> {
  static constexpr int FR = 4, G = 2;
  static bool has(int nz) { return nz == 8 || nz == 16; }
};

template <typename T>
static void launch_convz(const ConvZArgs<T> &a, int Nz, hipStream_t s) {
  constexpr int FR = ConvZTable<T>::FR, G = ConvZTable<T>::G;
  const dim3 grid(a.nyq ? (a.H - 1) / CONVZ_COLS + 1 : (a.H + CONVZ_COLS - 1) / CONVZ_COLS, a.Py / 2 + 1);
  switch (Nz) {
    case 8: launch_b<BConvz<T, 8, FR, G>>(k_convz<T, 8, FR, G>, grid, dim3(CONVZ_NT), 0, s, a); break;
    case 16: launch_b<BConvz<T, 16, FR, G>>(k_convz<T, 16, FR, G>, grid, dim3(CONVZ_NT), 0, s, a); break;
    case 32:
      if constexpr (std::is_same<T, float>::value) {
        launch_b<BConvz<T, 32, FR, G>>(k_convz<T, 32, FR, G>, grid, dim3(CONVZ_NT), 0, s, a);
        break;
      }
      [[fallthrough]];
    default: throw Error(UM_EUNSUP, "register Z pass: unsupported Nz");
  }
  UM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Tensor generation in FP64: brute-force face-charge integration,
// mag/demagkernel.go:115-336.  One thread per wrapped destination element
// (xw, yw, zw) with every coordinate <= P/2 and per source direction; the
// thread visits the periodic images of that element (every offset of the
// kernelRanges box that wraps onto it, :339-352) in the reference loop order,
// z outer, y, x inner, each ascending, and sums their FP64 fields in that
// order.  Without PBC an element has at most one image.  FP contraction is
// disabled so the FP64 values match a sequential evaluation.
//
// The integration grid shrinks as (accuracy / distance)^5, so the few offsets
// next to the source hold nearly all the points (31104 each at accuracy 6,
// 3.3e8 at accuracy 29 for 2.5 x 2.5 x 5 nm cells, against 4 far away).  The
// offsets inside a small "near box" (|x| <= R[0], |y| <= R[1], |z| <= R[2],
// chosen on the host so that every offset outside it has at most
// BF_NEAR_POINTS points) are therefore taken out of the per-element threads
// and integrated by whole wavefronts, BF_CHUNK points per wave, each wave's
// partial sum reduced across its lanes; a last kernel adds the near images of
// every element, in the reference image order, to the far sum.
// ---------------------------------------------------------------------------
struct BfArgs {
  int P[3];
  int r1[3], r2[3];      // image offsets per axis: r1 <= x <= r2
  int H[3];              // wrapped destinations per axis: 0..P/2
  int R[3];              // near box, integrated by k_demag_bf_near (-1: none)
  double cs[3], L, accuracy;
  double *K[6];          // xx yy zz yz xz xy
};

constexpr long long BF_NEAR_POINTS = 4096;  // offsets above this go to the near kernel
constexpr int BF_CHUNK_ITERS = 256;         // points per lane per chunk
constexpr long long BF_CHUNK = 64LL * BF_CHUNK_ITERS;

// One near-box offset and source direction, with its integration grid
// (nv, nw already doubled as in the reference).
struct BfTask {
  int x, y, z, s;
  int n[5];              // nv nw nx ny nz
  long long npts;
  int chunk0, nchunk;    // its chunks in the partial-sum array
};
// One destination element and direction, with its near tasks [t0, t1) in
// image order.
struct BfGroup {
  long long I;
  int s, t0, t1;
};

__device__ __forceinline__ double bf_delta(int d) {
  if (d < 0) d = -d;
  if (d > 0) d -= 1;
  return (double)d;
}

// Field at offset (x, y, z) of a source cell magnetised along s: two face
// charge sheets of nv x nw points against an nx x ny x nz destination grid
// (:186-262).
__device__ void bf_image(const BfArgs &a, int s, int x, int y, int z, double B[3]) {
#pragma clang fp contract(off)
  const int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
  const double *cs = a.cs;
  const double R[3] = {(double)x * cs[0], (double)y * cs[1], (double)z * cs[2]};
  const double dx = bf_delta(x) * cs[0], dy = bf_delta(y) * cs[1], dz = bf_delta(z) * cs[2];
  double d = sqrt(dx * dx + dy * dy + dz * dz);
  if (d == 0) d = a.L;
  const double maxSize = d / a.accuracy;
  int nv = (int)(fmax(cs[v] / maxSize, 1.0) + 0.5);
  int nw = (int)(fmax(cs[w] / maxSize, 1.0) + 0.5);
  const int nx = (int)(fmax(cs[0] / maxSize, 1.0) + 0.5);
  const int ny = (int)(fmax(cs[1] / maxSize, 1.0) + 0.5);
  const int nz = (int)(fmax(cs[2] / maxSize, 1.0) + 0.5);
  nv *= 2;
  nw *= 2;
  // 64-bit count as Go's int: 32 bits overflow at DemagAccuracy ~40
  const double scale = 1 / (double)((long long)nv * nw * nx * ny * nz);
  const double surface = cs[v] * cs[w];
  const double charge = surface * scale;
  const double pu1 = cs[u] / 2.;
  const double pu2 = -pu1;
  double B0 = 0, B1 = 0, B2 = 0;
  double pole[3];
  for (int i = 0; i < nv; i++) {
    pole[v] = -(cs[v] / 2.) + cs[v] / (double)(2 * nv) + (double)i * (cs[v] / (double)nv);
    for (int j = 0; j < nw; j++) {
      pole[w] = -(cs[w] / 2.) + cs[w] / (double)(2 * nw) + (double)j * (cs[w] / (double)nw);
      for (int al = 0; al < nx; al++) {
        const double rx = R[0] - cs[0] / 2 + cs[0] / (double)(2 * nx) + (cs[0] / (double)nx) * (double)al;
        for (int be = 0; be < ny; be++) {
          const double ry = R[1] - cs[1] / 2 + cs[1] / (double)(2 * ny) + (cs[1] / (double)ny) * (double)be;
          for (int ga = 0; ga < nz; ga++) {
            const double rz = R[2] - cs[2] / 2 + cs[2] / (double)(2 * nz) + (cs[2] / (double)nz) * (double)ga;
            pole[u] = pu1;
            double X = rx - pole[0], Y = ry - pole[1], Z = rz - pole[2];
            double r = sqrt(X * X + Y * Y + Z * Z);
            double qr = charge / (4 * M_PI * r * r * r);
            const double bx = X * qr, by = Y * qr, bz = Z * qr;
            pole[u] = pu2;
            X = rx - pole[0]; Y = ry - pole[1]; Z = rz - pole[2];
            r = sqrt(X * X + Y * Y + Z * Z);
            qr = -charge / (4 * M_PI * r * r * r);
            B0 += (bx + X * qr);
            B1 += (by + Y * qr);
            B2 += (bz + Z * qr);
          }
        }
      }
    }
  }
  B[0] = B0; B[1] = B1; B[2] = B2;
}

__global__ void k_demag_bf(BfArgs a) {
#pragma clang fp contract(off)
  const long total = (long)a.H[0] * a.H[1] * a.H[2] * 3;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int s = (int)(t % 3);
    long q = t / 3;
    const int xw = (int)(q % a.H[0]);
    q /= a.H[0];
    const int yw = (int)(q % a.H[1]);
    const int zw = (int)(q / a.H[1]);
    double S0 = 0, S1 = 0, S2 = 0;
    for (int z = image_first(zw, a.P[2], a.r1[2]); z <= a.r2[2]; z += a.P[2])
      for (int y = image_first(yw, a.P[1], a.r1[1]); y <= a.r2[1]; y += a.P[1])
        for (int x = image_first(xw, a.P[0], a.r1[0]); x <= a.r2[0]; x += a.P[0]) {
          if (abs(x) <= a.R[0] && abs(y) <= a.R[1] && abs(z) <= a.R[2]) continue;   // near kernel
          double B[3];
          bf_image(a, s, x, y, z, B);
          S0 += B[0];
          S1 += B[1];
          S2 += B[2];
        }
    const long I = ((long)zw * a.P[1] + yw) * a.P[0] + xw;
    // array[s][d] for d >= s: s=0 -> xx, xy, xz; s=1 -> yy, yz; s=2 -> zz
    if (s == 0) { a.K[0][I] = S0; a.K[5][I] = S1; a.K[4][I] = S2; }
    else if (s == 1) { a.K[1][I] = S1; a.K[3][I] = S2; }
    else { a.K[2][I] = S2; }
  }
}

// Near-box integration: one wavefront per chunk of BF_CHUNK consecutive points
// of one task, in the order of the reference's loops (nv, nw, nx, ny, nz with
// nz fastest); lane l takes points l, l + 64, ...  The positions are formed
// with the same expressions as bf_image.
__global__ void __launch_bounds__(256) k_demag_bf_near(BfArgs a, const BfTask *tasks, const int *chunk_task,
                                                       int nchunk, double *part) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= nchunk) return;   // wave-uniform
  const BfTask t = tasks[chunk_task[c]];
  const int s = t.s, u = s, v = (s + 1) % 3, w = (s + 2) % 3;
  const double *cs = a.cs;
  const int nv = t.n[0], nw = t.n[1], nx = t.n[2], ny = t.n[3], nz = t.n[4];
  const double scale = 1 / (double)((long long)nv * nw * nx * ny * nz);
  const double charge = (cs[v] * cs[w]) * scale;
  const double pu1 = cs[u] / 2., pu2 = -pu1;
  const double v0 = -(cs[v] / 2.) + cs[v] / (double)(2 * nv), vs = cs[v] / (double)nv;
  const double w0 = -(cs[w] / 2.) + cs[w] / (double)(2 * nw), ws = cs[w] / (double)nw;
  const double x0 = (double)t.x * cs[0] - cs[0] / 2 + cs[0] / (double)(2 * nx), xs = cs[0] / (double)nx;
  const double y0 = (double)t.y * cs[1] - cs[1] / 2 + cs[1] / (double)(2 * ny), ys = cs[1] / (double)ny;
  const double z0 = (double)t.z * cs[2] - cs[2] / 2 + cs[2] / (double)(2 * nz), zs = cs[2] / (double)nz;
  const long long first = (long long)(c - t.chunk0) * BF_CHUNK;
  const long long end = first + BF_CHUNK < t.npts ? first + BF_CHUNK : t.npts;
  long long idx = first + lane;
  long long q = idx;
  int ga = (int)(q % nz); q /= nz;
  int be = (int)(q % ny); q /= ny;
  int al = (int)(q % nx); q /= nx;
  int j = (int)(q % nw);
  int i = (int)(q / nw);
  double B0 = 0, B1 = 0, B2 = 0;
  for (; idx < end; idx += 64) {
    const double pv = v0 + (double)i * vs, pw = w0 + (double)j * ws;
    const double rx = x0 + xs * (double)al, ry = y0 + ys * (double)be, rz = z0 + zs * (double)ga;
    // pole[u] = pu, pole[v] = pv, pole[w] = pw
    double p0 = u == 0 ? pu1 : (v == 0 ? pv : pw);
    double p1 = u == 1 ? pu1 : (v == 1 ? pv : pw);
    double p2 = u == 2 ? pu1 : (v == 2 ? pv : pw);
    double X = rx - p0, Y = ry - p1, Z = rz - p2;
    double r = sqrt(X * X + Y * Y + Z * Z);
    double qr = charge / (4 * M_PI * r * r * r);
    const double bx = X * qr, by = Y * qr, bz = Z * qr;
    p0 = u == 0 ? pu2 : p0;
    p1 = u == 1 ? pu2 : p1;
    p2 = u == 2 ? pu2 : p2;
    X = rx - p0; Y = ry - p1; Z = rz - p2;
    r = sqrt(X * X + Y * Y + Z * Z);
    qr = -charge / (4 * M_PI * r * r * r);
    B0 += (bx + X * qr);
    B1 += (by + Y * qr);
    B2 += (bz + Z * qr);
    // advance 64 points in the mixed radix (nv, nw, nx, ny, nz)
    ga += 64;
    int k = ga / nz; ga -= k * nz; be += k;
    k = be / ny; be -= k * ny; al += k;
    k = al / nx; al -= k * nx; j += k;
    k = j / nw; j -= k * nw; i += k;
  }
  for (int o = 32; o > 0; o >>= 1) {
    B0 += __shfl_xor(B0, o, 64);
    B1 += __shfl_xor(B1, o, 64);
    B2 += __shfl_xor(B2, o, 64);
  }
  if (lane == 0) {
    part[3 * (long)c] = B0;
    part[3 * (long)c + 1] = B1;
    part[3 * (long)c + 2] = B2;
  }
}

// Adds each element's near images, in image order (each image the ordered sum
// of its chunks), to the far sum written by k_demag_bf.
__global__ void k_demag_bf_near_add(BfArgs a, const BfGroup *groups, int ngroup, const BfTask *tasks,
                                    const double *part) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroup) return;
  const BfGroup G = groups[g];
  double S0 = 0, S1 = 0, S2 = 0;
  for (int ti = G.t0; ti < G.t1; ti++) {
    const BfTask t = tasks[ti];
    double T0 = 0, T1 = 0, T2 = 0;
    for (int c = t.chunk0; c < t.chunk0 + t.nchunk; c++) {
      T0 += part[3 * (long)c];
      T1 += part[3 * (long)c + 1];
      T2 += part[3 * (long)c + 2];
    }
    S0 += T0;
    S1 += T1;
    S2 += T2;
  }
  const long long I = G.I;
  if (G.s == 0) { a.K[0][I] += S0; a.K[5][I] += S1; a.K[4][I] += S2; }
  else if (G.s == 1) { a.K[1][I] += S1; a.K[3][I] += S2; }
  else { a.K[2][I] += S2; }
}

// Mirror reconstruction, mag/demagkernel.go:280-324: every element with a
// coordinate above P/2 copies its mirror image with the component's parity.
__global__ void k_demag_mirror(BfArgs a) {
  const long total = (long)a.P[0] * a.P[1] * a.P[2];
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int x = (int)(t % a.P[0]);
    const int y = (int)((t / a.P[0]) % a.P[1]);
    const int z = (int)(t / ((long)a.P[0] * a.P[1]));
    const bool mx = x > a.P[0] / 2, my = y > a.P[1] / 2, mz = z > a.P[2] / 2;
    if (!(mx || my || mz)) continue;
    const int x2 = mx ? a.P[0] - x : x, y2 = my ? a.P[1] - y : y, z2 = mz ? a.P[2] - z : z;
    const long S = ((long)z2 * a.P[1] + y2) * a.P[0] + x2;
    // parity per component (xx yy zz yz xz xy): odd axes flip the sign
    const double sx = mx ? -1.0 : 1.0, sy = my ? -1.0 : 1.0, sz = mz ? -1.0 : 1.0;
    a.K[0][t] = a.K[0][S];
    a.K[1][t] = a.K[1][S];
    a.K[2][t] = a.K[2][S];
    a.K[3][t] = sy * sz * a.K[3][S];
    a.K[4][t] = sx * sz * a.K[4][S];
    a.K[5][t] = sx * sy * a.K[5][S];
  }
}

// Non-redundant real parts of the tensor spectrum, scaled by 1/(Px Py Pz)
// (scaleRealParts, opencl/conv_common.go:27-50); also the max |imag| seen.
template <typename T>
__global__ void k_extract_kspec(const double2 *__restrict__ spec, T *__restrict__ out, int H, int Hk, int Hp,
                                int Py, int KHy, int KHz, double scale,
                                unsigned long long *__restrict__ maximg_bits, int zlay, int comp) {
  const long total = (long)Hk * KHy * KHz;
  double mi = 0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int kx = (int)(t % Hk);
    const int ky = (int)((t / Hk) % KHy);
    const int kz = (int)(t / ((long)Hk * KHy));
    // zlay: [ky][kz][element][kx], the Z pass's slab of one ky contiguous;
    // otherwise [element][kz][ky][kx] (out already offset by the element)
    const long o = zlay ? (((long)ky * KHz + kz) * 6 + comp) * Hk + kx : t;
    if (kx >= H) {   // row padding
      out[o] = T(0);
      continue;
    }
    const double2 v = spec[((long)kz * Py + ky) * Hp + kx];
    out[o] = (T)(v.x * scale);
    mi = fmax(mi, fabs(v.y));
  }
  atomicMax(maximg_bits, (unsigned long long)__double_as_longlong(mi));
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Radix schedule of the run-time (generic) Stockham engine: the order and
// size of the passes.  sched 0: odd radices first, then the power of two in
// balanced passes of at most 16; 1: the power-of-two passes first; 2 and 3:
// odd first, power-of-two passes of at most 8 / 4 (more passes, fewer values
// per thread).  The compiled fixed-length kernels have their own schedules.
constexpr int FFT_NSCHED = 4;
std::vector<int> fft_radices(int n, int sched) {
  std::vector<int> odd;
  int m = n;
  for (int p : {7, 5, 3}) {
    while (m % p == 0) { odd.push_back(p); m /= p; }
  }
  int e2 = 0;
  while (m % 2 == 0) { m /= 2; e2++; }
  if (m != 1) throw Error(UM_EUNSUP, "FFT length " + std::to_string(n) + " has a prime factor above 7");
  const int maxlog = sched == 2 ? 3 : sched == 3 ? 2 : 4;
  std::vector<int> two;
  if (e2 > 0) {
    const int np = (e2 + maxlog - 1) / maxlog;
    const int base = e2 / np, rem = e2 % np;
    for (int i = 0; i < np; i++) two.push_back(1 << (base + (i < rem ? 1 : 0)));
  }
  std::vector<int> r = sched == 1 ? two : odd;
  for (int v : (sched == 1 ? odd : two)) r.push_back(v);
  if (r.empty()) r.push_back(1);
  return r;
}
bool fft_supported(int n) {
  try { fft_radices(n); return true; } catch (...) { return false; }
}

template <typename T>
struct TwTable {
  DevBuf buf;
  FftLen len;
  int sched = 0;
  // pass order of the generic engine (the twiddle table does not depend on it)
  void schedule(int sc) {
    const int n = len.n;
    std::vector<int> r = (n > 1) ? fft_radices(n, sc) : std::vector<int>{};
    UM_REQUIRE((int)r.size() <= FFT_MAXPASS, UM_EUNSUP, "FFT length too long");
    sched = sc;
    len.np = (n > 1) ? (int)r.size() : 0;
    int pp = 1;
    for (int i = 0; i < len.np; i++) {
      len.rad[i] = r[i];
      len.fp[i].init(pp);
      len.fnr[i].init(n / r[i]);
      pp *= r[i];
    }
  }
  void make(int n, hipStream_t s) {
    len.n = n;
    len.fn.init(n > 0 ? n : 1);
    schedule(0);
    std::vector<C2<T>> h(n > 0 ? n : 1);
    for (int m = 0; m < n; m++) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
      h[m] = C2<T>{(T)cosl(ang), (T)sinl(ang)};
    }
    buf.alloc(sizeof(C2<T>) * h.size());
    UM_HIP(hipMemcpyAsync(buf.p, h.data(), sizeof(C2<T>) * h.size(), hipMemcpyHostToDevice, s));
    UM_HIP(hipStreamSynchronize(s));
    len.tw = buf.p;
  }
};

// ---- launch dispatch: compile-time instantiations for the hot sizes ------------
// Lines per work-group of each specialised length (LDS budget ~40-72 KiB so two
// or more work-groups share a CU); other lengths use the generic kernels.
template <typename T> struct CtTable;
// The 32/48/96/192-point X, 96/192/384-point Y and Pz = 3/5/7 Z entries serve
// small grids with Nx, Ny in {32, 48, 96, 192} and Nz <= 4 (ensembles).
template <> struct CtTable<float> {
  static int x(int n) {
    switch (n) { case 32: return 32; case 48: return 32; case 64: return 32; case 96: return 16; case 128: return 16;
                 case 192: return 16; case 256: return 16; case 512: return 8; case 1024: return 4; }
    return 0;
  }
  static int col(int n) {
    switch (n) { case 64: return 32; case 96: return 16; case 128: return 16; case 192: return 8; case 256: return 16;
                 case 384: return 8; case 512: return 8; case 1024: return 8; case 2048: return 4; }
    return 0;
  }
  static int conv3(int n) { switch (n) { case 3: return 32; case 5: return 32; case 7: return 32; case 16: return 16; case 32: return 16; case 64: return 16; } return 0; }
  static int conv2(int n) {
    switch (n) { case 64: return 16; case 96: return 16; case 128: return 16; case 192: return 8; case 256: return 8;
                 case 384: return 4; case 512: return 4; }
    return 0;
  }
};
template <> struct CtTable<double> {
  static int x(int n) { switch (n) { case 256: return 8; case 512: return 4; } return 0; }
  static int col(int n) { switch (n) { case 512: return 4; case 1024: return 4; } return 0; }
  static int conv3(int n) { switch (n) { case 16: return 8; case 64: return 8; } return 0; }
  static int conv2(int n) { switch (n) { case 256: return 4; case 512: return 2; } return 0; }
};

template <typename B> struct BTag { using type = B; };
// a parenthesised type as a macro argument: Unparen<void (X<a, b>)>::type is X<a, b>
template <typename> struct Unparen;
template <typename X> struct Unparen<void(X)> { using type = X; };
template <typename K>
static void allow_lds(K kern, size_t bytes) {
  if (bytes > 65536) UM_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

#define UM_CT_CASE(BODY, KERN, N, NL, LDSB)                                  \
  if (n == N && nl == NL) {                                                  \
    allow_lds(KERN, LDSB);                                                   \
    launch_b<typename Unparen<void BODY>::type>(KERN, grid, dim3(FFT_NT), LDSB, s, a); \
    UM_HIP(hipGetLastError());                                               \
    return;                                                                  \
  }

// four-step X inverse: FP32, n in 64..1024, rows of exactly Nx = n reals (Px = 2 Nx)
template <typename T>
static bool x4_ok(int n, int Nx) {
  return std::is_same<T, float>::value && Nx == n && (n == 64 || n == 128 || n == 256 || n == 512 || n == 1024);
}
template <typename T>
static bool launch_xi4(const XiArgs<T> &a, int n, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (!x4_ok<T>(n, a.Nx)) return false;
    auto go = [&](auto tag, auto kern, int rows_per_wg) {
      using K = typename decltype(tag)::type;
      launch_b<K>(kern, dim3((a.rows + rows_per_wg - 1) / rows_per_wg, 3), dim3(XF4_NT), 0, s, a);
      UM_HIP(hipGetLastError());
      return true;
    };
#define UM_XI4(NN) \
    if (n == NN) return go(BTag<BXi4<T, NN>>{}, k_xi4<T, NN>, XF4_NT / X4Shape<NN>::B * (X4Shape<NN>::B / X4Shape<NN>::A));
    UM_XI4(64) UM_XI4(128) UM_XI4(256) UM_XI4(512) UM_XI4(1024)
#undef UM_XI4
  }
  return false;
}

template <typename T>
static bool launch_xf4(const XArgs<T> &a, int n, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (a.mode != 0 || !x4_ok<T>(n, a.Nx) || a.in_pitch != a.Nx) return false;
    auto go = [&](auto tag, auto kern, int rows_per_wg) {
      using K = typename decltype(tag)::type;
      launch_b<K>(kern, dim3((a.rows + rows_per_wg - 1) / rows_per_wg, 3), dim3(XF4_NT), 0, s, a);
      UM_HIP(hipGetLastError());
      return true;
    };
#define UM_XF4(NN) \
    if (n == NN) return go(BTag<BXf4<T, NN>>{}, k_xf4<T, NN>, XF4_NT / X4Shape<NN>::B * (X4Shape<NN>::B / X4Shape<NN>::A));
    UM_XF4(64) UM_XF4(128) UM_XF4(256) UM_XF4(512) UM_XF4(1024)
#undef UM_XF4
  }
  return false;
}

template <typename T>
static void launch_xfwd(const XArgs<T> &a, int n, int nl, dim3 grid, hipStream_t s) {
  const size_t ct = sizeof(C2<T>) * ((size_t)nl * n + 2 * n);
  if constexpr (std::is_same<T, float>::value) {
    UM_CT_CASE((BXfwd<float, 64, 32>), (k_xfwd<float, 64, 32>), 64, 32, ct)
    UM_CT_CASE((BXfwd<float, 128, 16>), (k_xfwd<float, 128, 16>), 128, 16, ct)
    UM_CT_CASE((BXfwd<float, 256, 16>), (k_xfwd<float, 256, 16>), 256, 16, ct)
    UM_CT_CASE((BXfwd<float, 512, 8>), (k_xfwd<float, 512, 8>), 512, 8, ct)
    UM_CT_CASE((BXfwd<float, 1024, 4>), (k_xfwd<float, 1024, 4>), 1024, 4, ct)
    // few rows (small grids): fewer rows per work-group, more work-groups
    UM_CT_CASE((BXfwd<float, 64, 8>), (k_xfwd<float, 64, 8>), 64, 8, ct)
    UM_CT_CASE((BXfwd<float, 64, 1>), (k_xfwd<float, 64, 1>), 64, 1, ct)
    UM_CT_CASE((BXfwd<float, 128, 4>), (k_xfwd<float, 128, 4>), 128, 4, ct)
    UM_CT_CASE((BXfwd<float, 128, 1>), (k_xfwd<float, 128, 1>), 128, 1, ct)
    UM_CT_CASE((BXfwd<float, 256, 4>), (k_xfwd<float, 256, 4>), 256, 4, ct)
    UM_CT_CASE((BXfwd<float, 256, 1>), (k_xfwd<float, 256, 1>), 256, 1, ct)
    UM_CT_CASE((BXfwd<float, 512, 2>), (k_xfwd<float, 512, 2>), 512, 2, ct)
    UM_CT_CASE((BXfwd<float, 512, 1>), (k_xfwd<float, 512, 1>), 512, 1, ct)
    UM_CT_CASE((BXfwd<float, 1024, 1>), (k_xfwd<float, 1024, 1>), 1024, 1, ct)
    // small grids of the other X lengths
    UM_CT_CASE((BXfwd<float, 32, 32>), (k_xfwd<float, 32, 32>), 32, 32, ct)
    UM_CT_CASE((BXfwd<float, 32, 8>), (k_xfwd<float, 32, 8>), 32, 8, ct)
    UM_CT_CASE((BXfwd<float, 32, 1>), (k_xfwd<float, 32, 1>), 32, 1, ct)
    UM_CT_CASE((BXfwd<float, 48, 32>), (k_xfwd<float, 48, 32>), 48, 32, ct)
    UM_CT_CASE((BXfwd<float, 48, 8>), (k_xfwd<float, 48, 8>), 48, 8, ct)
    UM_CT_CASE((BXfwd<float, 48, 1>), (k_xfwd<float, 48, 1>), 48, 1, ct)
    UM_CT_CASE((BXfwd<float, 96, 16>), (k_xfwd<float, 96, 16>), 96, 16, ct)
    UM_CT_CASE((BXfwd<float, 96, 4>), (k_xfwd<float, 96, 4>), 96, 4, ct)
    UM_CT_CASE((BXfwd<float, 96, 1>), (k_xfwd<float, 96, 1>), 96, 1, ct)
    UM_CT_CASE((BXfwd<float, 192, 16>), (k_xfwd<float, 192, 16>), 192, 16, ct)
    UM_CT_CASE((BXfwd<float, 192, 4>), (k_xfwd<float, 192, 4>), 192, 4, ct)
    UM_CT_CASE((BXfwd<float, 192, 1>), (k_xfwd<float, 192, 1>), 192, 1, ct)
  } else {
    UM_CT_CASE((BXfwd<double, 256, 8>), (k_xfwd<double, 256, 8>), 256, 8, ct)
    UM_CT_CASE((BXfwd<double, 512, 4>), (k_xfwd<double, 512, 4>), 512, 4, ct)
  }
  launch_b<BXfwd<T, 0, 0>>(k_xfwd<T, 0, 0>, grid, dim3(FFT_NT), sizeof(C2<T>) * nl * n, s, a);
  UM_HIP(hipGetLastError());
}
template <typename T>
static void launch_xinv(const XiArgs<T> &a, int n, int nl, dim3 grid, hipStream_t s, bool four_step = true) {
  if (four_step && launch_xi4<T>(a, n, s)) return;
  const size_t ct = sizeof(C2<T>) * ((size_t)nl * n + 2 * n);
  if constexpr (std::is_same<T, float>::value) {
    UM_CT_CASE((BXinv<float, 64, 32>), (k_xinv<float, 64, 32>), 64, 32, ct)
    UM_CT_CASE((BXinv<float, 128, 16>), (k_xinv<float, 128, 16>), 128, 16, ct)
    UM_CT_CASE((BXinv<float, 256, 16>), (k_xinv<float, 256, 16>), 256, 16, ct)
    UM_CT_CASE((BXinv<float, 512, 8>), (k_xinv<float, 512, 8>), 512, 8, ct)
    UM_CT_CASE((BXinv<float, 1024, 4>), (k_xinv<float, 1024, 4>), 1024, 4, ct)
    UM_CT_CASE((BXinv<float, 32, 32>), (k_xinv<float, 32, 32>), 32, 32, ct)
    UM_CT_CASE((BXinv<float, 32, 8>), (k_xinv<float, 32, 8>), 32, 8, ct)
    UM_CT_CASE((BXinv<float, 32, 1>), (k_xinv<float, 32, 1>), 32, 1, ct)
    UM_CT_CASE((BXinv<float, 48, 32>), (k_xinv<float, 48, 32>), 48, 32, ct)
    UM_CT_CASE((BXinv<float, 48, 8>), (k_xinv<float, 48, 8>), 48, 8, ct)
    UM_CT_CASE((BXinv<float, 48, 1>), (k_xinv<float, 48, 1>), 48, 1, ct)
    UM_CT_CASE((BXinv<float, 96, 16>), (k_xinv<float, 96, 16>), 96, 16, ct)
    UM_CT_CASE((BXinv<float, 96, 4>), (k_xinv<float, 96, 4>), 96, 4, ct)
    UM_CT_CASE((BXinv<float, 96, 1>), (k_xinv<float, 96, 1>), 96, 1, ct)
    UM_CT_CASE((BXinv<float, 192, 16>), (k_xinv<float, 192, 16>), 192, 16, ct)
    UM_CT_CASE((BXinv<float, 192, 4>), (k_xinv<float, 192, 4>), 192, 4, ct)
    UM_CT_CASE((BXinv<float, 192, 1>), (k_xinv<float, 192, 1>), 192, 1, ct)
  } else {
    UM_CT_CASE((BXinv<double, 256, 8>), (k_xinv<double, 256, 8>), 256, 8, ct)
    UM_CT_CASE((BXinv<double, 512, 4>), (k_xinv<double, 512, 4>), 512, 4, ct)
  }
  launch_b<BXinv<T, 0, 0>>(k_xinv<T, 0, 0>, grid, dim3(FFT_NT), sizeof(C2<T>) * nl * n, s, a);
  UM_HIP(hipGetLastError());
}
template <typename T, int DIR>
static void launch_col(const ColArgs<T> &a, int n, int nl, dim3 grid, hipStream_t s) {
  const size_t ct = sizeof(C2<T>) * ((size_t)nl * n + n);
  if constexpr (std::is_same<T, float>::value) {
    UM_CT_CASE((BColfft<float, DIR, 64, 32>), (k_colfft<float, DIR, 64, 32>), 64, 32, ct)
    UM_CT_CASE((BColfft<float, DIR, 128, 16>), (k_colfft<float, DIR, 128, 16>), 128, 16, ct)
    UM_CT_CASE((BColfft<float, DIR, 256, 16>), (k_colfft<float, DIR, 256, 16>), 256, 16, ct)
    UM_CT_CASE((BColfft<float, DIR, 512, 8>), (k_colfft<float, DIR, 512, 8>), 512, 8, ct)
    UM_CT_CASE((BColfft<float, DIR, 1024, 8>), (k_colfft<float, DIR, 1024, 8>), 1024, 8, ct)
    UM_CT_CASE((BColfft<float, DIR, 2048, 4>), (k_colfft<float, DIR, 2048, 4>), 2048, 4, ct)
    UM_CT_CASE((BColfft<float, DIR, 96, 16>), (k_colfft<float, DIR, 96, 16>), 96, 16, ct)
    UM_CT_CASE((BColfft<float, DIR, 192, 8>), (k_colfft<float, DIR, 192, 8>), 192, 8, ct)
    UM_CT_CASE((BColfft<float, DIR, 384, 8>), (k_colfft<float, DIR, 384, 8>), 384, 8, ct)
  } else {
    UM_CT_CASE((BColfft<double, DIR, 512, 4>), (k_colfft<double, DIR, 512, 4>), 512, 4, ct)
    UM_CT_CASE((BColfft<double, DIR, 1024, 4>), (k_colfft<double, DIR, 1024, 4>), 1024, 4, ct)
  }
  launch_b<BColfft<T, DIR, 0, 0>>(k_colfft<T, DIR, 0, 0>, grid, dim3(FFT_NT), sizeof(C2<T>) * nl * n, s, a);
  UM_HIP(hipGetLastError());
}
// Register four-step Y pass when P = A*A (A = 8, 16, 32 in FP32; 8, 16 in FP64)
// and the column is either half zero-padded or full; false: use launch_col.
template <typename T, int DIR>
static bool launch_coly(const ColArgs<T> &a, int P, int nz, hipStream_t s) {
  const int nin = a.nvalid, nout = a.nout;
  int half;
  if (DIR < 0 && nin * 2 == P && nout == P) half = 1;
  else if (DIR > 0 && nin == P && nout * 2 == P) half = 1;
  else if (nin == P && nout == P) half = 0;
  else return false;
  // planes per work-group (software-pipelined): FP32 forward pass only (the
  // inverse loads twice as many rows and would spill)
  const int zb = (sizeof(T) == 4 && DIR < 0 && nz % UM_COLY_ZB == 0) ? UM_COLY_ZB : 1;
  auto go = [&](auto tag, auto kern, int A, int C, int ZB) {
    using K = typename decltype(tag)::type;
    // k_coly's SPLIT: real and imaginary parts exchanged one after the other
    size_t lds = sizeof(C2<T>) * (size_t)A * (A * C + C);
    if (lds > 65536) lds /= 2;
    lds += sizeof(C2<T>) * (size_t)A * (A + 1);   // the twiddle table
    allow_lds(kern, lds);
    const dim3 grid((a.H + C - 1) / C, nz / ZB, 3);
    launch_b<K>(kern, grid, dim3(A * C), lds, s, a);
    UM_HIP(hipGetLastError());
    return true;
  };
#define UM_COLY(A, C)                                                                                         \
  if (P == A * A) {                                                                                           \
    if constexpr (DIR < 0 && UM_COLY_ZB > 1)                                                                  \
      if (half && zb == UM_COLY_ZB) return go(BTag<BColy<T, A, DIR, C, true, UM_COLY_ZB>>{}, k_coly<T, A, DIR, C, true, UM_COLY_ZB>, A, C, UM_COLY_ZB);     \
    return half ? go(BTag<BColy<T, A, DIR, C, true, 1>>{}, k_coly<T, A, DIR, C, true, 1>, A, C, 1) : go(BTag<BColy<T, A, DIR, C, false, 1>>{}, k_coly<T, A, DIR, C, false, 1>, A, C, 1);  \
  }
  if constexpr (std::is_same<T, float>::value) {
    UM_COLY(32, UM_COLY_C32)
    UM_COLY(16, 16)
    UM_COLY(8, 32)
  } else {
    UM_COLY(16, 16)
    UM_COLY(8, 32)
  }
#undef UM_COLY
  return false;
}

template <typename T>
static void launch_conv(const ConvArgs<T> &a, int n, int B, dim3 grid, hipStream_t s) {
  const int nl = B;  // UM_CT_CASE compares (n, nl)
  const size_t ct = sizeof(C2<T>) * ((size_t)3 * B * n + n);
  if (!a.is2d) {
    if constexpr (std::is_same<T, float>::value) {
      UM_CT_CASE((BConv<float, 16, 16>), (k_conv<float, 16, 16>), 16, 16, ct)
      UM_CT_CASE((BConv<float, 32, 16>), (k_conv<float, 32, 16>), 32, 16, ct)
      UM_CT_CASE((BConv<float, 64, 16>), (k_conv<float, 64, 16>), 64, 16, ct)
      UM_CT_CASE((BConv<float, 3, 32>), (k_conv<float, 3, 32>), 3, 32, ct)     // Pz = 2 Nz - 1 for Nz = 2, 3, 4
      UM_CT_CASE((BConv<float, 5, 32>), (k_conv<float, 5, 32>), 5, 32, ct)
      UM_CT_CASE((BConv<float, 7, 32>), (k_conv<float, 7, 32>), 7, 32, ct)
    } else {
      UM_CT_CASE((BConv<double, 16, 8>), (k_conv<double, 16, 8>), 16, 8, ct)
      UM_CT_CASE((BConv<double, 64, 8>), (k_conv<double, 64, 8>), 64, 8, ct)
    }
  } else {
    if constexpr (std::is_same<T, float>::value) {
      UM_CT_CASE((BConv<float, 64, 16>), (k_conv<float, 64, 16>), 64, 16, ct)
      UM_CT_CASE((BConv<float, 128, 16>), (k_conv<float, 128, 16>), 128, 16, ct)
      UM_CT_CASE((BConv<float, 256, 8>), (k_conv<float, 256, 8>), 256, 8, ct)
      UM_CT_CASE((BConv<float, 512, 4>), (k_conv<float, 512, 4>), 512, 4, ct)
      // narrow spectra (small grids): fewer columns per work-group
      UM_CT_CASE((BConv<float, 64, 4>), (k_conv<float, 64, 4>), 64, 4, ct)
      UM_CT_CASE((BConv<float, 64, 1>), (k_conv<float, 64, 1>), 64, 1, ct)
      UM_CT_CASE((BConv<float, 128, 4>), (k_conv<float, 128, 4>), 128, 4, ct)
      UM_CT_CASE((BConv<float, 128, 1>), (k_conv<float, 128, 1>), 128, 1, ct)
      UM_CT_CASE((BConv<float, 256, 2>), (k_conv<float, 256, 2>), 256, 2, ct)
      UM_CT_CASE((BConv<float, 256, 1>), (k_conv<float, 256, 1>), 256, 1, ct)
      UM_CT_CASE((BConv<float, 512, 1>), (k_conv<float, 512, 1>), 512, 1, ct)
      UM_CT_CASE((BConv<float, 96, 16>), (k_conv<float, 96, 16>), 96, 16, ct)
      UM_CT_CASE((BConv<float, 96, 4>), (k_conv<float, 96, 4>), 96, 4, ct)
      UM_CT_CASE((BConv<float, 96, 1>), (k_conv<float, 96, 1>), 96, 1, ct)
      UM_CT_CASE((BConv<float, 192, 8>), (k_conv<float, 192, 8>), 192, 8, ct)
      UM_CT_CASE((BConv<float, 192, 2>), (k_conv<float, 192, 2>), 192, 2, ct)
      UM_CT_CASE((BConv<float, 192, 1>), (k_conv<float, 192, 1>), 192, 1, ct)
      UM_CT_CASE((BConv<float, 384, 4>), (k_conv<float, 384, 4>), 384, 4, ct)
      UM_CT_CASE((BConv<float, 384, 1>), (k_conv<float, 384, 1>), 384, 1, ct)
    } else {
      UM_CT_CASE((BConv<double, 256, 4>), (k_conv<double, 256, 4>), 256, 4, ct)
      UM_CT_CASE((BConv<double, 512, 2>), (k_conv<double, 512, 2>), 512, 2, ct)
    }
  }
  launch_b<BConv<T, 0, 0>>(k_conv<T, 0, 0>, grid, dim3(FFT_NT), sizeof(C2<T>) * 3 * B * n, s, a);
  UM_HIP(hipGetLastError());
}
#undef UM_CT_CASE

// ---------------------------------------------------------------------------
// X inverse fused with the four-cell effective field + torque (FP32).  A
// work-group of XF_NT threads inverts R rows (y0 .. y0+R-1 of one z plane) of
// all three components with the register four-step of k_xi4 (3R lines, G per
// 32-lane group), leaves the real B_demag lines in LDS, then evaluates the
// effective field and torque of those R x Nx cells: B_demag never touches HBM.
// Work-groups run z fastest (tile w -> y block w / Nz, plane w % Nz), so the
// z +- 1 rows of m that the stencil reads were just read by the neighbouring
// work-groups of the same XCD and are still in its L2.
// ---------------------------------------------------------------------------
constexpr int XF_NT = 192;

template <int N> struct XFShape {
  static constexpr int A = X4Shape<N>::A, B = X4Shape<N>::B, G = B / A;
  static constexpr int R = 64 * G / B;    // rows per work-group
  static constexpr int L = 3 * R;         // lines (component-major: line = c R + j)
  static_assert(L / G * B == XF_NT, "the lines of R rows fill the work-group");
};

template <typename T>
struct XiFieldArgs {
  XiArgs<T> x;
  FieldArgs<T> f;
};
// TH = false (no thermal term, checked on the host) at N = 512: at most 128
// VGPRs without spills (the launch bound's second argument is the minimum
// number of waves per SIMD with hipcc), so five work-groups share a CU
// instead of four; the other lengths would spill under that bound
template <typename T, int N, bool REG, int EX, bool UNI, bool TH>
struct BXi4Field {
  using Args = XiFieldArgs<T>;
  static constexpr int NT = XF_NT, MINB = (!TH && N == 512) ? 4 : 1;
  static __device__ __forceinline__ void run(const XiFieldArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    const XiArgs<T> &x = a.x;
    const FieldArgs<T> &f = a.f;
    using S = XFShape<N>;
    constexpr int A = S::A, B = S::B, G = S::G, R = S::R, TS = B * (A + 1), HB = B / 2;
    static_assert(S::L * N <= S::L * TS, "B_demag lines fit the transpose space");
    // the transpose moves the real, then the imaginary parts (half the LDS of
    // complex entries: more work-groups per CU); afterwards line (c, j) holds
    // B_demag_c of row j, x < N
    __shared__ __align__(16) T ldr[S::L * TS];
    // twiddles of the step between the two transforms, W_N^(n1 k2) at
    // [n1][k2] (rows padded by one entry: the A rows fall on distinct banks);
    // read from LDS instead of global memory, their latency stays out of the
    // transform
    __shared__ C2<T> twb[A * (B + 1)];
    for (int i = threadIdx.x; i < A * B; i += XF_NT) {
      const int r = i / B, k2 = i - r * B;
      twb[r * (B + 1) + k2] = x.tw2[2 * r * k2];
    }
    const int Ny = (int)x.fRy.d, Nz = x.rows / Ny;
    const int w = xcd_tile(blockIdx.x, gridDim.x);
    const int yb = w / Nz, z = w - yb * Nz, y0 = yb * R;
    const int grp = threadIdx.x / B, l = threadIdx.x % B;
    C2<T> v[B];
    const int h1 = l / A, n1 = l % A, line1 = grp * G + h1;
    {
      const int c = line1 / R, y = y0 + (line1 - c * R);
  #pragma unroll
      for (int n2 = 0; n2 < B; n2++) v[n2] = C2<T>{T(0), T(0)};
      if (y < Ny) {
        // component base selected from the uniform arguments (an indexed read
        // of x.in would be a per-lane load ahead of every spectrum load)
        const C2<T> *inc = c == 0 ? x.in[0] : c == 1 ? x.in[1] : x.in[2];
        const C2<T> *irow = inc + ((long)z * x.Py + y) * x.Hp;
        // two halves of loads: half the staging registers live at a time
  #pragma unroll
        for (int hb = 0; hb < 2; hb++) {
          C2<T> xk[HB], xn[HB];
  #pragma unroll
          for (int q = 0; q < HB; q++) {
            const int k = n1 + A * (hb * HB + q);
            xk[q] = irow[k];
            xn[q] = irow[N - k];
          }
  #pragma unroll
          for (int q = 0; q < HB; q++) {
            const int n2 = hb * HB + q, k = n1 + A * n2;
            const C2<T> sm = cadd(xk[q], cconj(xn[q])), d = csub(xk[q], cconj(xn[q]));
            const C2<T> t = cmul(cconj(x.tw2[k]), d);
            v[n2] = C2<T>{sm.x - t.y, sm.y + t.x};
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      dft<B, +1>(v);
      __syncthreads();   // the twiddle table is complete
  #pragma unroll
      for (int k2 = 1; k2 < B; k2++) v[k2] = cmul(v[k2], cconj(twb[n1 * (B + 1) + k2]));   // W_n^-(n1 k2)
    }
    C2<T> u[G][A];
    T *tw = ldr + line1 * TS + n1;
  #pragma unroll
    for (int k2 = 0; k2 < B; k2++) tw[k2 * (A + 1)] = v[k2].x;
    __syncthreads();
  #pragma unroll
    for (int h = 0; h < G; h++) {
      const T *t = ldr + (grp * G + h) * TS + l * (A + 1);
  #pragma unroll
      for (int q = 0; q < A; q++) u[h][q].x = t[q];
    }
    __syncthreads();
  #pragma unroll
    for (int k2 = 0; k2 < B; k2++) tw[k2 * (A + 1)] = v[k2].y;
    __syncthreads();
  #pragma unroll
    for (int h = 0; h < G; h++) {
      const T *t = ldr + (grp * G + h) * TS + l * (A + 1);
  #pragma unroll
      for (int q = 0; q < A; q++) u[h][q].y = t[q];
    }
    T *bl = ldr;
    __syncthreads();   // every transpose read is done: the space now takes the real lines
  #pragma unroll
    for (int h = 0; h < G; h++) {
      dft<A, +1>(u[h]);
      T *o = bl + (grp * G + h) * N;
  #pragma unroll
      for (int k1 = 0; k1 < A / 2; k1++) *(float2 *)(o + 2 * (l + B * k1)) = float2{u[h][k1].x, u[h][k1].y};
    }
    __syncthreads();
    constexpr int NGX = N / 4, NG = R * NGX;
    for (int g = threadIdx.x; g < NG; g += XF_NT) {
      const int j = g / NGX, gx = g - j * NGX, y = y0 + j;
      if (y >= Ny) continue;
      T bdv[3][4];
  #pragma unroll
      for (int c = 0; c < 3; c++) {
        const float4 b = *(const float4 *)(bl + (c * R + j) * N + 4 * gx);
        bdv[c][0] = b.x; bdv[c][1] = b.y; bdv[c][2] = b.z; bdv[c][3] = b.w;
      }
      field_group4<T, REG, EX, false, UNI, false, TH>(f, 4 * gx, y, z, bdv);   // registers over instructions
    }
  }
};
template <typename T, int N, bool REG, int EX, bool UNI, bool TH>
__global__ __launch_bounds__(XF_NT, (BXi4Field<T, N, REG, EX, UNI, TH>::MINB)) void k_xi4_field(XiFieldArgs<T> a) {
  BXi4Field<T, N, REG, EX, UNI, TH>::run(a, hw_bid(), hw_grid());
}

// whether the fused kernel covers these field terms and buffers: the
// four-cell kernel's terms without spin-transfer torques or bulk DMI, every
// per-cell array aligned to four cells
template <typename T>
static bool xi4_field_ok(const FieldArgs<T> &a, int n) {
  if (!std::is_same<T, float>::value || a.g.Nx != n || a.g.Nx % 4) return false;
  if (n != 64 && n != 128 && n != 256 && n != 512 && n != 1024) return false;
  if (a.exch_mode == 3 || a.has_zl || a.has_sl) return false;
  auto al = [](const void *p) { return ((uintptr_t)p & (4 * sizeof(T) - 1)) == 0; };
  for (int c = 0; c < 3; c++) {
    if (!al(a.m[c])) return false;
    if (a.tau[0] && !al(a.tau[c])) return false;
    if (a.beff[0] && !al(a.beff[c])) return false;
  }
  return !a.regions || ((uintptr_t)a.regions & 3) == 0;
}

template <typename T>
static void launch_xi4_field(const XiArgs<T> &a, const FieldArgs<T> &f, int n, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    const int Ny = a.fRy.d, Nz = a.rows / Ny;
    const bool reg = f.regions != nullptr, uni = field_uniform(f), th = f.has_therm != 0;
    auto go = [&](auto tag, auto kern, int R) {
      using K = typename decltype(tag)::type;
      const dim3 grid((unsigned)(((Ny + R - 1) / R) * Nz));
      launch_b<K>(kern, grid, dim3(XF_NT), 0, s, XiFieldArgs<T>{a, f});
      UM_HIP(hipGetLastError());
    };
// the thermal-free variants only for uniform parameters (the hot case);
// per-cell parameters without a thermal term take the TH = true kernel
#define UM_XF2(CN, RG, E)                                                                                      \
  { if (uni && !th) go(BTag<BXi4Field<T, CN, RG, E, true, false>>{}, k_xi4_field<T, CN, RG, E, true, false>,    \
                       XFShape<CN>::R);                                                                          \
    else if (uni) go(BTag<BXi4Field<T, CN, RG, E, true, true>>{}, k_xi4_field<T, CN, RG, E, true, true>,         \
                     XFShape<CN>::R);                                                                            \
    else go(BTag<BXi4Field<T, CN, RG, E, false, true>>{}, k_xi4_field<T, CN, RG, E, false, true>, XFShape<CN>::R); }
#define UM_XF(CN)                                                                                    \
  if (n == CN) {                                                                                     \
    if (f.exch_mode == 0) { if (reg) UM_XF2(CN, true, 0) else UM_XF2(CN, false, 0) }                 \
    else if (f.exch_mode == 1) { if (reg) UM_XF2(CN, true, 1) else UM_XF2(CN, false, 1) }            \
    else { if (reg) UM_XF2(CN, true, 2) else UM_XF2(CN, false, 2) }                                  \
    return;                                                                                          \
  }
    UM_XF(64) UM_XF(128) UM_XF(256) UM_XF(512) UM_XF(1024)
#undef UM_XF
#undef UM_XF2
  }
  throw Error(UM_EUNSUP, "fused X inverse: unsupported length");
}

static bool fused_field_default() {
  const char *e = std::getenv("UMHIP_FUSED_FIELD");
  return !(e && e[0] == '0');
}

struct DemagBase {
  virtual ~DemagBase() {}
  int prec = UM_F32;
  int N[3], P[3];
  virtual void exec(void *const B[3], const void *const m[3], const void *vol, const void *Ms,
                    double Ms_mul, hipStream_t s) = 0;
  virtual void kspec(int c, void *host) = 0;
  virtual bool exec_field(const void *const m[3], const void *vol, const void *Ms, double Ms_mul, const void *fa,
                          hipStream_t s) = 0;
  virtual void schedule(int out[DEMAG_NKNOB]) const = 0;
  virtual void tune(hipStream_t s) = 0;
};

// ---- autotuner ---------------------------------------------------------------
// Launch schedules of a plan are picked per shape by timing the candidates of
// each pass on the plan's own buffers (HIP events, best of a few launches),
// pass by pass in pipeline order.  Results are kept per (precision, N, P)
// until the next set_autotune call and, with a tuning file, appended to it and
// read back by later calls and processes.  Off by default: the static tables are then used, and results do
// not depend on timing noise.  Mode 2 applies known results (from the
// file) without timing anything.
static std::mutex g_tune_mu;
static int g_tune_mode = 0;
static std::string g_tune_file;
static std::map<std::string, std::vector<int>> g_tune_memo;

void set_autotune(int mode, const char *file) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_mode = mode;
  g_tune_file = file ? file : "";
  g_tune_memo.clear();   // the known results are those of the file
  if (g_tune_file.empty()) return;
  FILE *f = std::fopen(g_tune_file.c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    char *colon = std::strchr(line, ':');
    if (!colon) continue;
    std::vector<int> v;
    char *q = colon + 1;
    for (;;) {
      char *end = nullptr;
      const long x = std::strtol(q, &end, 10);
      if (end == q) break;
      v.push_back((int)x);
      q = end;
    }
    if ((int)v.size() == DEMAG_NKNOB) g_tune_memo[std::string(line, colon)] = v;
  }
  std::fclose(f);
}
static bool tune_lookup(const std::string &key, std::vector<int> &v, int *mode) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  *mode = g_tune_mode;
  auto it = g_tune_memo.find(key);
  if (it == g_tune_memo.end()) return false;
  v = it->second;
  return true;
}
static void tune_store(const std::string &key, const std::vector<int> &v) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_memo[key] = v;
  if (g_tune_file.empty()) return;
  if (FILE *f = std::fopen(g_tune_file.c_str(), "a")) {
    std::fprintf(f, "%s:", key.c_str());
    for (int x : v) std::fprintf(f, " %d", x);
    std::fprintf(f, "\n");
    std::fclose(f);
  }
}

template <typename T>
struct DemagPlan : DemagBase {
  int H, Hp, Hk, is2d, KHy, KHz;   // Hk: tensor row pitch (H rounded up to 128 bytes)
  TwTable<T> tx, ty, tz, tx2;   // tx: length Px/2, tx2: length Px (post-processing), ty, tz
  DevBuf spec[3];               // per component: Nz*Py*Hp complex
  DevBuf kern;                  // 6 x H*KHy*KHz of T
  // launch schedule (static tables, or the autotuner's choice): X forward and
  // inverse rows per work-group, four-step X inverse on/off, generic Y
  // lines, convolution columns, and the generic engine's radix schedules
  int nl_x = 1, nl_xi = 1, nl_y = 1, B_conv = 1;
  bool xi4 = true;
  // X inverse fused with the field + torque kernel where it applies
  // (environment UMHIP_FUSED_FIELD=0 keeps the two kernels apart)
  bool fused_field = fused_field_default();
  bool zreg = false;            // register Z pass (Pz = 2 Nz, Nz in ConvZTable)
  bool xodd = false;            // odd Px (periodic along x, odd Nx): full-length X passes
  long kcount = 0;

  // The generic engine takes any lines per work-group: halve a choice while
  // its pass (units lines, times mult in the other grid dimensions) would
  // launch fewer than target work-groups, e.g. 100 x 100 x 2 (200 X rows, 81
  // per group at most) would otherwise run on 3 x 3 work-groups.
  static int spread(int nl, long units, long mult, long target = SMALL_GRID_WG) {
    while (nl > 1 && mult * ((units + nl - 1) / nl) < target) nl = (nl + 1) / 2;
    return nl;
  }

  static int lines_for(int n) {
    const int cap = FFT_NT * FftEpt<T>::value;
    UM_REQUIRE(n <= cap, UM_EUNSUP, "FFT length " + std::to_string(n) + " exceeds the LDS budget");
    return cap / n;
  }

  void setup(const int Nin[3], const int Pin[3], hipStream_t s) {
    for (int i = 0; i < 3; i++) { N[i] = Nin[i]; P[i] = Pin[i]; }
    prec = sizeof(T) == 4 ? UM_F32 : UM_F64;
    xodd = (P[0] % 2) != 0;
    is2d = (N[2] == 1);
    UM_REQUIRE(!is2d || P[2] == 1, UM_EUNSUP, "2D grid with Z padding");
    H = P[0] / 2 + 1;
    const int al = 128 / (int)sizeof(C2<T>), al2 = 128 / (int)sizeof(T);
    Hp = (H + al - 1) / al * al;
    KHy = P[1] / 2 + 1;
    KHz = P[2] / 2 + 1;
    tx.make(P[0] / 2, s);
    tx2.make(P[0], s);
    ty.make(P[1], s);
    tz.make(P[2], s);
    nl_x = xodd ? lines_for(P[0]) : CtTable<T>::x(P[0] / 2);
    if (!nl_x) nl_x = lines_for(P[0] / 2);
    // small grids: trade rows per work-group for work-groups while the X pass
    // would launch fewer than ~3 per CU (the smaller variants are compiled in)
    if (std::is_same<T, float>::value && !xodd && CtTable<T>::x(P[0] / 2)) {
      const int rows = N[1] * N[2];
      for (int nl : {nl_x / 4, 1}) {
        if (nl < 1 || 3L * ((rows + nl_x - 1) / nl_x) >= SMALL_GRID_WG) break;
        nl_x = nl;
      }
    }
    if (xodd || !CtTable<T>::x(P[0] / 2)) nl_x = spread(nl_x, (long)N[1] * N[2], 3);
    nl_xi = nl_x;
    nl_y = CtTable<T>::col(P[1]);
    // the Y columns of a 3D grid run Nz planes deep: fuller groups pay there
    if (!nl_y) nl_y = spread(std::min(lines_for(P[1]), 64), H, 3L * N[2], SMALL_GRID_WG / 6);

    zreg = !is2d && P[2] == 2 * N[2] && ConvZTable<T>::has(N[2]);
    const int convlen = is2d ? P[1] : P[2];
    B_conv = is2d ? CtTable<T>::conv2(convlen) : CtTable<T>::conv3(convlen);
    if (is2d && B_conv && std::is_same<T, float>::value) {
      for (int b : {B_conv / 4, 1}) {   // narrow spectra: more work-groups, as for the X pass
        if (b < 1 || (H + B_conv - 1) / B_conv >= SMALL_GRID_WG) break;
        B_conv = b;
      }
    }
    if (!B_conv) {
      const int bmax = std::min(lines_for(convlen) / 3, 32);
      UM_REQUIRE(bmax >= 1, UM_EUNSUP, "convolution axis too long");
      B_conv = 1;
      while (B_conv * 2 <= bmax) B_conv *= 2;
      B_conv = spread(B_conv, H, is2d ? 1 : P[1] / 2 + 1);
    }
    UM_REQUIRE(B_conv * (convlen / 2 + 1) <= FFT_NT * CONV_QK, UM_EUNSUP, "convolution prefetch slots exceeded");
    UM_REQUIRE(3 * B_conv * convlen <= FFT_NT * FftEpt<T>::value, UM_EUNSUP, "convolution tile exceeds LDS budget");
    const size_t sc = is2d ? (size_t)N[1] * Hp : (size_t)N[2] * P[1] * Hp;
    for (int c = 0; c < 3; c++) spec[c].alloc(sizeof(C2<T>) * sc);
    Hk = (H + al2 - 1) / al2 * al2;
    kcount = (long)Hk * KHy * KHz;
    kern.alloc(sizeof(T) * 6 * kcount);
  }

  // Transform the FP64 real-space tensor (device, 6 x P) into the stored
  // half spectrum of precision T.
  void load_kernel(double *const Kdev[6], hipStream_t s) {
    // FP64 FFT plan over the full padded box
    TwTable<double> dx, dx2, dy, dz;
    dx.make(P[0] / 2, s);
    dx2.make(P[0], s);
    dy.make(P[1], s);
    dz.make(P[2], s);
    const int Hpd = (H + 7) / 8 * 8;
    DevBuf sp(sizeof(double2) * (size_t)P[2] * P[1] * Hpd);
    DevBuf mx(sizeof(unsigned long long));
    UM_HIP(hipMemsetAsync(mx.p, 0, 8, s));
    const int capd = FFT_NT * FftEpt<double>::value;
    UM_REQUIRE((xodd ? P[0] : P[0] / 2) <= capd && P[1] <= capd && P[2] <= capd, UM_EUNSUP,
               "tensor FFT too long for FP64 LDS budget");
    const double scale = 1.0 / ((double)P[0] * P[1] * P[2]);
    for (int c = 0; c < 6; c++) {
      T *out = kern.as<T>() + c * kcount;
      if (is2d && (c == 3 || c == 4)) {
        fill_value<T>(out, T(0), kcount, s);
        continue;
      }
      XArgs<double> xa{};
      xa.f = dx.len; xa.tw2 = (const double2 *)dx2.buf.p; xa.mode = 1;
      xa.nl = capd / (P[0] / 2); xa.fhalf.init(P[0] / 4 + 1); xa.Nx = P[0]; xa.rows = P[1] * P[2];
      xa.fRy.init(P[1]); xa.Py = P[1]; xa.Hp = Hpd;
      xa.in_pitch = P[0];
      xa.src[0] = Kdev[c]; xa.Ms = Par<double>{nullptr, 1.0}; xa.vol = nullptr;
      xa.out[0] = (double2 *)sp.p;
      if (xodd) {
        xa.f = dx2.len; xa.nl = capd / P[0]; xa.fhalf.init(H);
        dim3 gx((xa.rows + xa.nl - 1) / xa.nl, 1);
        UM_LAUNCH(k_xfwd_odd<double>, gx, dim3(FFT_NT), sizeof(double2) * xa.nl * P[0], s, xa);
      } else {
        dim3 gx((xa.rows + xa.nl - 1) / xa.nl, 1);
        UM_LAUNCH((k_xfwd<double, 0, 0>), gx, dim3(FFT_NT), sizeof(double2) * xa.nl * (P[0] / 2), s, xa);
      }
      UM_HIP(hipGetLastError());
      if (P[1] > 1) {
        ColArgs<double> ca{};
        ca.f = dy.len; ca.nl = std::min(capd / P[1], 64); ca.fnl.init(ca.nl); ca.H = H; ca.nvalid = P[1]; ca.nout = P[1];
        ca.estride = Hpd; ca.pstride = (long)P[1] * Hpd; ca.buf[0] = (double2 *)sp.p;
        dim3 g((H + ca.nl - 1) / ca.nl, P[2], 1);
        UM_LAUNCH((k_colfft<double, -1, 0, 0>), g, dim3(FFT_NT), sizeof(double2) * ca.nl * P[1], s, ca);
        UM_HIP(hipGetLastError());
      }
      if (P[2] > 1) {
        ColArgs<double> ca{};
        ca.f = dz.len; ca.nl = std::min(capd / P[2], 64); ca.fnl.init(ca.nl); ca.H = H; ca.nvalid = P[2]; ca.nout = P[2];
        ca.estride = (long)P[1] * Hpd; ca.pstride = Hpd; ca.buf[0] = (double2 *)sp.p;
        dim3 g((H + ca.nl - 1) / ca.nl, P[1], 1);
        UM_LAUNCH((k_colfft<double, -1, 0, 0>), g, dim3(FFT_NT), sizeof(double2) * ca.nl * P[2], s, ca);
        UM_HIP(hipGetLastError());
      }
      UM_LAUNCH(k_extract_kspec<T>, dim3(blocks_for(kcount, 256, 4096)), dim3(256), 0, s,
                         (const double2 *)sp.p, zreg ? kern.as<T>() : out, H, Hk, Hpd, P[1], KHy, KHz, scale,
                         (unsigned long long *)mx.p, zreg ? 1 : 0, c);
      UM_HIP(hipGetLastError());
    }
    unsigned long long bits = 0;
    UM_HIP(hipMemcpyAsync(&bits, mx.p, 8, hipMemcpyDeviceToHost, s));
    UM_HIP(hipStreamSynchronize(s));
    double maximg;
    std::memcpy(&maximg, &bits, 8);
    // conv_common.go:43-49: after one FFT the normalisation is sqrt(N)
    if (maximg * sqrt(scale) > 1e-6)
      fprintf(stderr, "umhip: demag kernel imaginary part %g (tensor lacks symmetry)\n", maximg * sqrt(scale));
  }

  // X-fwd through Y-inv (3D) or the 2D Y pass: leaves the convolved half
  // spectrum (rows y < Ny) in spec[] for the X inverse.
  void pass_xfwd(const void *const mv[3], const void *vol, const void *Msa, double Ms_mul, hipStream_t s) {
    const int Nx = N[0], Ny = N[1], Nz = N[2];
    XArgs<T> a{};
    a.f = tx.len; a.tw2 = (const C2<T> *)tx2.buf.p; a.mode = 0; a.nl = nl_x; a.fhalf.init(P[0] / 4 + 1);
    a.Nx = Nx; a.rows = Ny * Nz; a.fRy.init(Ny); a.Py = is2d ? Ny : P[1]; a.Hp = Hp; a.in_pitch = Nx;
    for (int c = 0; c < 3; c++) { a.src[c] = (const T *)mv[c]; a.out[c] = spec[c].as<C2<T>>(); }
    a.Ms = Par<T>{(const T *)Msa, (T)Ms_mul}; a.vol = (const T *)vol;
    dim3 g((a.rows + nl_x - 1) / nl_x, 3);
    if (xodd) {
      a.f = tx2.len; a.fhalf.init(H);
      launch_b<BXfwdOdd<T>>(k_xfwd_odd<T>, g, dim3(FFT_NT), sizeof(C2<T>) * nl_x * P[0], s, a);
      UM_HIP(hipGetLastError());
    } else if (!(xi4 && launch_xf4<T>(a, P[0] / 2, s))) {
      launch_xfwd<T>(a, P[0] / 2, nl_x, g, s);
    }
  }
  // Y forward (dir < 0) or inverse (dir > 0), 3D only
  void pass_y(int dir, hipStream_t s) {
    ColArgs<T> ya{};
    ya.f = ty.len; ya.nl = nl_y; ya.fnl.init(nl_y); ya.H = H; ya.estride = Hp; ya.pstride = (long)P[1] * Hp;
    for (int c = 0; c < 3; c++) ya.buf[c] = spec[c].as<C2<T>>();
    dim3 gy((H + nl_y - 1) / nl_y, N[2], 3);
    if (dir < 0) {
      ya.nvalid = N[1]; ya.nout = P[1];
      if (!launch_coly<T, -1>(ya, P[1], N[2], s)) launch_col<T, -1>(ya, P[1], nl_y, gy, s);
    } else {
      ya.nvalid = P[1]; ya.nout = N[1];
      if (!launch_coly<T, +1>(ya, P[1], N[2], s)) launch_col<T, +1>(ya, P[1], nl_y, gy, s);
    }
  }
  // the convolution pass: Z (3D) or Y (2D) forward, tensor multiply, inverse
  void pass_conv(hipStream_t s) {
    const T *K[6];
    for (int c = 0; c < 6; c++) K[c] = kern.as<T>() + c * kcount;
    if (!is2d && zreg) {
      ConvZArgs<T> za{};
      za.H = H; za.Py = P[1]; za.KHy = KHy; za.Hk = Hk; za.zstride = (long)P[1] * Hp; za.ystride = Hp;
      za.nyq = H > 1 && (H - 1) % CONVZ_COLS == 0;
      za.K = kern.as<T>();
      for (int c = 0; c < 3; c++) za.buf[c] = spec[c].as<C2<T>>();
      launch_convz<T>(za, N[2], s);
      return;
    }
    ConvArgs<T> ca{};
    ca.f = is2d ? ty.len : tz.len; ca.B = B_conv; ca.fB.init(B_conv); ca.f3B.init(3 * B_conv); ca.H = H;
    ca.nvalid = is2d ? N[1] : N[2];
    ca.estride = is2d ? Hp : (long)P[1] * Hp; ca.pstride = is2d ? 0 : Hp;
    ca.is2d = is2d; ca.Py = P[1]; ca.Pz = is2d ? 1 : P[2]; ca.KHy = KHy; ca.Hk = Hk;
    for (int c = 0; c < 6; c++) ca.K[c] = K[c];
    for (int c = 0; c < 3; c++) ca.buf[c] = spec[c].as<C2<T>>();
    dim3 gc((H + B_conv - 1) / B_conv, is2d ? 1 : P[1] / 2 + 1);   // 3D: ky and Py-ky per work-group
    launch_conv<T>(ca, is2d ? P[1] : P[2], B_conv, gc, s);
  }
  // X forward through Y inverse (3D) or the 2D Y pass: leaves the convolved
  // half spectrum (rows y < Ny) in spec[] for the X inverse.
  void spectral(const void *const mv[3], const void *vol, const void *Msa, double Ms_mul, hipStream_t s) {
    pass_xfwd(mv, vol, Msa, Ms_mul, s);
    if (!is2d) pass_y(-1, s);
    pass_conv(s);
    if (!is2d) pass_y(+1, s);
  }
  void pass_xinv(void *const Bv[3], hipStream_t s) {
    XiArgs<T> a = xinv_args(nl_xi);
    for (int c = 0; c < 3; c++) a.dst[c] = (T *)Bv[c];
    dim3 g((a.rows + nl_xi - 1) / nl_xi, 3);
    if (xodd) {
      a.f = tx2.len; a.fhalf.init(H);
      launch_b<BXinvOdd<T>>(k_xinv_odd<T>, g, dim3(FFT_NT), sizeof(C2<T>) * nl_xi * P[0], s, a);
      UM_HIP(hipGetLastError());
    } else {
      launch_xinv<T>(a, P[0] / 2, nl_xi, g, s, xi4);
    }
  }

  XiArgs<T> xinv_args(int nl) const {
    XiArgs<T> a{};
    a.f = tx.len; a.tw2 = (const C2<T> *)tx2.buf.p; a.nl = nl; a.Nx = N[0]; a.rows = N[1] * N[2];
    a.fhalf.init(P[0] / 4 + 1); a.fRy.init(N[1]); a.Py = is2d ? N[1] : P[1]; a.Hp = Hp;
    for (int c = 0; c < 3; c++) a.in[c] = spec[c].as<C2<T>>();
    return a;
  }

  void exec(void *const Bv[3], const void *const mv[3], const void *vol, const void *Msa,
            double Ms_mul, hipStream_t s) override {
    spectral(mv, vol, Msa, Ms_mul, s);
    pass_xinv(Bv, s);
  }

  // Demag convolution fused with the effective field and torque: the X
  // inverse of all three components of a few rows stays in LDS and feeds the
  // four-cell field kernel, so B_demag never goes to HBM.  false: not
  // available for this size / precision / set of terms (the caller runs exec
  // and the field kernel).
  bool exec_field(const void *const mv[3], const void *vol, const void *Msa, double Ms_mul, const void *fav,
                  hipStream_t s) override {
    if constexpr (!std::is_same<T, float>::value) {
      return false;
    } else {
      const FieldArgs<T> &f = *(const FieldArgs<T> *)fav;
      const int n = P[0] / 2;
      if (!fused_field || xodd || !xi4 || !x4_ok<T>(n, N[0]) || !xi4_field_ok(f, n)) return false;
      spectral(mv, vol, Msa, Ms_mul, s);
      launch_xi4_field(xinv_args(1), f, n, s);
      return true;
    }
  }

  // ---- schedule knobs: nl_x nl_xi xi4 nl_y B_conv sched(x) sched(y) sched(z)
  TwTable<T> &xtab() { return xodd ? tx2 : tx; }
  void schedule(int out[DEMAG_NKNOB]) const override {
    const int v[DEMAG_NKNOB] = {nl_x, nl_xi, xi4 ? 1 : 0, nl_y, B_conv, xodd ? tx2.sched : tx.sched, ty.sched, tz.sched};
    for (int i = 0; i < DEMAG_NKNOB; i++) out[i] = v[i];
  }
  void apply(const std::vector<int> &v) {
    const int nx = xodd ? P[0] : P[0] / 2;
    bool ok = (int)v.size() == DEMAG_NKNOB && v[0] >= 1 && v[0] <= lines_for(nx) && v[1] >= 1 &&
              v[1] <= lines_for(nx) && v[3] >= 1 && v[3] <= std::min(lines_for(P[1]), 64) && conv_ok(v[4]);
    for (int i = 5; ok && i < DEMAG_NKNOB; i++) ok = v[i] >= 0 && v[i] < FFT_NSCHED;
    UM_REQUIRE(ok, UM_EINVAL, "invalid demag launch schedule");
    nl_x = v[0]; nl_xi = v[1]; xi4 = v[2] != 0; nl_y = v[3]; B_conv = v[4];
    xtab().schedule(v[5]); ty.schedule(v[6]); tz.schedule(v[7]);
  }
  std::string tune_key() const {
    char b[160];
    std::snprintf(b, sizeof b, "%s N=%d,%d,%d P=%d,%d,%d", sizeof(T) == 4 ? "f32" : "f64", N[0], N[1], N[2], P[0],
                  P[1], P[2]);
    return b;
  }
  static std::vector<int> halvings(int L) {
    std::vector<int> r;
    for (int v = L; v >= 1; v /= 2) r.push_back(v);
    if (r.empty() || r.back() != 1) r.push_back(1);
    return r;
  }
  static void add(std::vector<int> &v, int x) {
    if (x >= 1 && std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
  }
  bool conv_ok(int B) const {
    const int convlen = is2d ? P[1] : P[2];
    return B >= 1 && B * (convlen / 2 + 1) <= FFT_NT * CONV_QK && 3 * B * convlen <= FFT_NT * FftEpt<T>::value;
  }

  void tune(hipStream_t s) override {
    const std::string key = tune_key();
    std::vector<int> v;
    int mode = 0;
    const bool known = tune_lookup(key, v, &mode);
    if (mode == 0) return;   // static tables
    if (known) {
      std::vector<int> keep(DEMAG_NKNOB);
      schedule(keep.data());
      try { apply(v); return; } catch (const Error &) { apply(keep); }
    }
    if (mode != 1) return;   // mode 2: known results only, never timed
    const size_t n = (size_t)N[0] * N[1] * N[2];
    DevBuf mb(sizeof(T) * 3 * n), bb(sizeof(T) * 3 * n);
    fill_value<T>(mb.as<T>(), T(0), 3 * n, s);
    const void *mv[3] = {mb.as<T>(), mb.as<T>() + n, mb.as<T>() + 2 * n};
    void *bv[3] = {bb.as<T>(), bb.as<T>() + n, bb.as<T>() + 2 * n};
    hipEvent_t e0, e1;
    UM_HIP(hipEventCreate(&e0));
    UM_HIP(hipEventCreate(&e1));
    // best of 3 timed launches after one warm-up; a candidate that fails to
    // launch (LDS limits) is skipped
    auto timed = [&](auto &&run) -> float {
      try {
        run();
        float best = 1e30f;
        for (int r = 0; r < 3; r++) {
          UM_HIP(hipEventRecord(e0, s));
          run();
          UM_HIP(hipEventRecord(e1, s));
          UM_HIP(hipEventSynchronize(e1));
          float ms = 0;
          UM_HIP(hipEventElapsedTime(&ms, e0, e1));
          best = std::min(best, ms);
        }
        return best;
      } catch (const Error &) {
        (void)hipGetLastError();
        return 1e30f;
      }
    };
    try {
      spectral(mv, nullptr, nullptr, 1.0, s);   // valid spectra in the buffers
      // 1. X forward: rows per work-group x radix schedule
      {
        std::vector<int> nls;
        const int nx = xodd ? P[0] : P[0] / 2;
        add(nls, nl_x);
        if (!xodd && std::is_same<T, float>::value && CtTable<T>::x(nx)) {
          add(nls, CtTable<T>::x(nx));
          add(nls, CtTable<T>::x(nx) / 4);
          add(nls, 1);
        }
        for (int x : halvings(lines_for(nx))) add(nls, x);
        float best = 1e30f;
        int bn = nl_x, bs = xtab().sched;
        for (int nl : nls)
          for (int sc = 0; sc < FFT_NSCHED; sc++) {
            nl_x = nl;
            try { xtab().schedule(sc); } catch (const Error &) { continue; }   // too many passes
            const float t = timed([&] { pass_xfwd(mv, nullptr, nullptr, 1.0, s); });
            if (t < best) { best = t; bn = nl; bs = sc; }
          }
        nl_x = bn;
        xtab().schedule(bs);
      }
      // 2. X inverse: four-step kernel or the Stockham one with rows per work-group
      {
        const int nx = xodd ? P[0] : P[0] / 2;
        std::vector<int> nls;
        add(nls, nl_xi);
        if (!xodd && CtTable<T>::x(nx)) add(nls, CtTable<T>::x(nx));
        for (int x : halvings(lines_for(nx))) add(nls, x);
        float best = 1e30f;
        int bn = nl_xi;
        bool b4 = xi4;
        for (int four : {1, 0})
          for (int nl : nls) {
            if (four && nl != nls[0]) continue;   // the four-step kernel has its own rows
            xi4 = four != 0;
            nl_xi = nl;
            const float t = timed([&] { pass_xinv(bv, s); });
            if (t < best) { best = t; bn = nl; b4 = xi4; }
          }
        nl_xi = bn;
        xi4 = b4;
      }
      // 3. Y passes (3D): generic lines per work-group x radix schedule
      if (!is2d) {
        std::vector<int> nls;
        add(nls, nl_y);
        add(nls, CtTable<T>::col(P[1]));
        for (int x : halvings(std::min(lines_for(P[1]), 64))) add(nls, x);
        float best = 1e30f;
        int bn = nl_y, bs = ty.sched;
        for (int nl : nls)
          for (int sc = 0; sc < FFT_NSCHED; sc++) {
            nl_y = nl;
            try { ty.schedule(sc); } catch (const Error &) { continue; }
            const float t = timed([&] { pass_y(-1, s); pass_y(+1, s); });
            if (t < best) { best = t; bn = nl; bs = sc; }
          }
        nl_y = bn;
        ty.schedule(bs);
      }
      // 4. convolution pass: columns per work-group x radix schedule
      if (is2d || !zreg) {
        TwTable<T> &tc = is2d ? ty : tz;
        const int convlen = is2d ? P[1] : P[2];
        std::vector<int> bs_;
        add(bs_, B_conv);
        const int ct = is2d ? CtTable<T>::conv2(convlen) : CtTable<T>::conv3(convlen);
        add(bs_, ct);
        add(bs_, ct / 4);
        for (int x : halvings(std::min(lines_for(convlen) / 3, 32))) add(bs_, x);
        float best = 1e30f;
        int bb_ = B_conv, bsc = tc.sched;
        for (int B : bs_) {
          if (!conv_ok(B)) continue;
          for (int sc = 0; sc < FFT_NSCHED; sc++) {
            B_conv = B;
            try { tc.schedule(sc); } catch (const Error &) { continue; }
            const float t = timed([&] { pass_conv(s); });
            if (t < best) { best = t; bb_ = B; bsc = sc; }
          }
        }
        B_conv = bb_;
        tc.schedule(bsc);
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    UM_HIP(hipEventDestroy(e0));
    UM_HIP(hipEventDestroy(e1));
    std::vector<int> out(DEMAG_NKNOB);
    schedule(out.data());
    tune_store(key, out);
  }

  void kspec(int c, void *host) override {
    UM_REQUIRE(c >= 0 && c < 6, UM_EINVAL, "component out of range");
    UM_HIP(hipDeviceSynchronize());
    // rows of H values out of the padded pitch Hk, returned as [kz][ky][kx]
    if (zreg) {   // stored [ky][kz][element][kx]
      for (int ky = 0; ky < KHy; ky++)
        UM_HIP(hipMemcpy2D((T *)host + (size_t)ky * H, sizeof(T) * H * KHy,
                           kern.as<T>() + ((size_t)ky * KHz * 6 + c) * Hk, sizeof(T) * Hk * 6, sizeof(T) * H,
                           (size_t)KHz, hipMemcpyDeviceToHost));
    } else {
      UM_HIP(hipMemcpy2D(host, sizeof(T) * H, kern.as<T>() + c * kcount, sizeof(T) * Hk, sizeof(T) * H,
                         (size_t)KHy * KHz, hipMemcpyDeviceToHost));
    }
  }
};

// padSize, mag/demagkernel.go:404-420
void pad_size(const int N[3], const int pbc[3], int P[3]) {
  for (int i = 0; i < 3; i++) {
    if (pbc[i] != 0) { P[i] = N[i]; continue; }
    P[i] = (i != 2 || N[i] > 5) ? 2 * N[i] : 2 * N[i] - 1;
  }
}

void demag_kernel_mirror(const int P[3], double *const K[6], hipStream_t s);

// kernelRanges, mag/demagkernel.go:339-352 (P: the padded size)
void kernel_ranges(const int P[3], const int pbc[3], int r1[3], int r2[3]) {
  for (int c = 0; c < 3; c++) {
    if (pbc[c] == 0) {
      r1[c] = -(P[c] - 1) / 2;
      r2[c] = (P[c] - 1) / 2;
    } else {
      r1[c] = -(P[c] * pbc[c] - 1);
      r2[c] = P[c] * pbc[c] - 1;
    }
  }
  if (P[2] == 1 && pbc[2] == 0) r2[2] = 0;
}

// Brute-force tensor on the GPU into device arrays Kdev[6] of size P (FP64).
// The integration grid of bf_image for offset (x, y, z) and direction s_, on
// the host: nv, nw (doubled), nx, ny, nz.
static void bf_grid(const double cs[3], double L, double accuracy, int s_, int x, int y, int z, int n[5]) {
  auto delta = [](int d) { d = std::abs(d); return (double)(d > 0 ? d - 1 : 0); };
  const double dx = delta(x) * cs[0], dy = delta(y) * cs[1], dz = delta(z) * cs[2];
  double d = std::sqrt(dx * dx + dy * dy + dz * dz);
  if (d == 0) d = L;
  const double maxSize = d / accuracy;
  auto cnt = [&](double c) { return (int)(std::fmax(c / maxSize, 1.0) + 0.5); };
  n[0] = 2 * cnt(cs[(s_ + 1) % 3]);
  n[1] = 2 * cnt(cs[(s_ + 2) % 3]);
  n[2] = cnt(cs[0]);
  n[3] = cnt(cs[1]);
  n[4] = cnt(cs[2]);
}

void demag_kernel_bf_device(const int N[3], const int pbc[3], const double cs[3], double accuracy,
                            double *const Kdev[6], int P[3], hipStream_t s) {
  for (int i = 0; i < 3; i++)
    UM_REQUIRE(pbc[i] >= 0 && pbc[i] <= (1 << 16), UM_EINVAL, "PBC repetitions out of range");
  UM_REQUIRE(cs[0] > 0 && cs[1] > 0 && cs[2] > 0 && accuracy > 0, UM_EINVAL, "bad cell size / accuracy");
  pad_size(N, pbc, P);
  BfArgs a{};
  kernel_ranges(P, pbc, a.r1, a.r2);
  for (int i = 0; i < 3; i++) {
    a.P[i] = P[i];
    a.H[i] = P[i] / 2 + 1;
    UM_REQUIRE((long)P[i] * pbc[i] < (1L << 30), UM_EINVAL, "PBC image range too large");
    a.cs[i] = cs[i];
  }
  a.L = std::min(cs[0], std::min(cs[1], cs[2]));
  a.accuracy = accuracy;
  const size_t np = (size_t)P[0] * P[1] * P[2];
  for (int c = 0; c < 6; c++) {
    a.K[c] = Kdev[c];
    fill_value<double>(Kdev[c], 0.0, np, s);
  }
  // near box: the point count only falls with the distance, which only grows
  // with |offset| per axis, so R[c] = the last heavy offset along axis c
  for (int c = 0; c < 3; c++) a.R[c] = -1;
  auto points = [&](int s_, int x, int y, int z, int n[5]) {
    bf_grid(a.cs, a.L, accuracy, s_, x, y, z, n);
    return (long long)n[0] * n[1] * n[2] * n[3] * n[4];
  };
  auto heaviest = [&](int x, int y, int z) {
    long long m = 0;
    int n[5];
    for (int s_ = 0; s_ < 3; s_++) m = std::max(m, points(s_, x, y, z, n));
    return m;
  };
  if (heaviest(0, 0, 0) > BF_NEAR_POINTS)
    for (int c = 0; c < 3; c++) {
      const int lim = std::max(-a.r1[c], a.r2[c]);
      int r = 0;
      for (; r < lim; r++) {
        int o[3] = {0, 0, 0};
        o[c] = r + 1;
        if (heaviest(o[0], o[1], o[2]) <= BF_NEAR_POINTS) break;
      }
      a.R[c] = r;
    }
  // an extreme DemagAccuracy (a near box of millions of offsets) keeps every
  // offset in the per-element kernel rather than building huge task tables
  if ((2L * a.R[0] + 1) * (2L * a.R[1] + 1) * (2L * a.R[2] + 1) > (1L << 22))
    for (int c = 0; c < 3; c++) a.R[c] = -1;
  std::vector<BfTask> tasks;
  std::vector<long long> elem;
  if (a.R[0] >= 0)
    for (int z = std::max(a.r1[2], -a.R[2]); z <= std::min(a.r2[2], a.R[2]); z++)
      for (int y = std::max(a.r1[1], -a.R[1]); y <= std::min(a.r2[1], a.R[1]); y++)
        for (int x = std::max(a.r1[0], -a.R[0]); x <= std::min(a.r2[0], a.R[0]); x++) {
          const int o[3] = {x, y, z};
          int wr[3];
          bool in = true;
          for (int c = 0; c < 3; c++) {
            wr[c] = ((o[c] % P[c]) + P[c]) % P[c];
            in = in && wr[c] < a.H[c];
          }
          if (!in) continue;   // its element is filled by the mirror
          for (int s_ = 0; s_ < 3; s_++) {
            BfTask t{};
            t.x = x; t.y = y; t.z = z; t.s = s_;
            t.npts = points(s_, x, y, z, t.n);
            t.nchunk = (int)((t.npts + BF_CHUNK - 1) / BF_CHUNK);
            tasks.push_back(t);
            elem.push_back((((long long)wr[2] * P[1] + wr[1]) * P[0] + wr[0]) * 3 + s_);
          }
        }
  // group by (element, direction), keeping the image order within a group
  std::vector<int> order(tasks.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return elem[p] < elem[q]; });
  std::vector<BfTask> sorted;
  std::vector<BfGroup> groups;
  std::vector<int> chunk_task;
  for (int i : order) {
    BfTask t = tasks[i];
    t.chunk0 = (int)chunk_task.size();
    UM_REQUIRE((long long)t.chunk0 + t.nchunk < (1LL << 28), UM_ENOMEM, "DemagAccuracy too high for the near-field split");
    const int ti = (int)sorted.size();
    for (int k = 0; k < t.nchunk; k++) chunk_task.push_back(ti);
    sorted.push_back(t);
    if (groups.empty() || groups.back().I * 3 + groups.back().s != elem[i])
      groups.push_back(BfGroup{elem[i] / 3, (int)(elem[i] % 3), ti, ti});
    groups.back().t1 = ti + 1;
  }
  const long total = (long)a.H[0] * a.H[1] * a.H[2] * 3;
  UM_LAUNCH(k_demag_bf, dim3(blocks_for(total, 64, 1 << 16)), dim3(64), 0, s, a);
  UM_HIP(hipGetLastError());
  if (!groups.empty()) {
    const int nchunk = (int)chunk_task.size();
    DevBuf dt(sizeof(BfTask) * sorted.size()), dg(sizeof(BfGroup) * groups.size()),
        dc(sizeof(int) * chunk_task.size()), part(sizeof(double) * 3 * (size_t)nchunk);
    UM_HIP(hipMemcpyAsync(dt.p, sorted.data(), dt.bytes, hipMemcpyHostToDevice, s));
    UM_HIP(hipMemcpyAsync(dg.p, groups.data(), dg.bytes, hipMemcpyHostToDevice, s));
    UM_HIP(hipMemcpyAsync(dc.p, chunk_task.data(), dc.bytes, hipMemcpyHostToDevice, s));
    UM_LAUNCH(k_demag_bf_near, dim3((nchunk + 3) / 4), dim3(256), 0, s, a, (const BfTask *)dt.p,
                       (const int *)dc.p, nchunk, (double *)part.p);
    UM_HIP(hipGetLastError());
    UM_LAUNCH(k_demag_bf_near_add, dim3(((int)groups.size() + 63) / 64), dim3(64), 0, s, a,
                       (const BfGroup *)dg.p, (int)groups.size(), (const BfTask *)dt.p, (const double *)part.p);
    UM_HIP(hipGetLastError());
    UM_HIP(hipStreamSynchronize(s));   // before the host tables and device buffers go
  }
  demag_kernel_mirror(P, Kdev, s);
  if (P[2] == 1) {
    fill_value<double>(Kdev[3], 0.0, np, s);
    fill_value<double>(Kdev[4], 0.0, np, s);
  }
}

// Fill the octants with a coordinate above P/2 from their mirror images.
void demag_kernel_mirror(const int P[3], double *const K[6], hipStream_t s) {
  BfArgs a{};
  for (int i = 0; i < 3; i++) a.P[i] = P[i];
  for (int c = 0; c < 6; c++) a.K[c] = K[c];
  const size_t np = (size_t)P[0] * P[1] * P[2];
  UM_LAUNCH(k_demag_mirror, dim3(blocks_for(np, 256, 1 << 16)), dim3(256), 0, s, a);
  UM_HIP(hipGetLastError());
}

void demag_exec(DemagBase *d, void *const B[3], const void *const m[3], const void *vol, const void *Ms,
                double Ms_mul, hipStream_t s) {
  d->exec(B, m, vol, Ms, Ms_mul, s);
}
bool demag_exec_field(DemagBase *d, const void *const m[3], const void *vol, const void *Ms, double Ms_mul,
                      const void *fa, hipStream_t s) {
  return d->exec_field(m, vol, Ms, Ms_mul, fa, s);
}
void demag_destroy(DemagBase *d) { delete d; }
void demag_kspec(DemagBase *d, int c, void *host) { d->kspec(c, host); }
void demag_schedule(DemagBase *d, int out[DEMAG_NKNOB]) { d->schedule(out); }
void demag_info(DemagBase *d, int P[3], int *prec) {
  for (int i = 0; i < 3; i++) P[i] = d->P[i];
  *prec = d->prec;
}

DemagBase *demag_create(int prec, const int N[3], const int P[3], double *const Kdev[6], hipStream_t s) {
  DemagBase *d = nullptr;
  if (prec == UM_F32) {
    auto *p = new DemagPlan<float>();
    try { p->setup(N, P, s); p->load_kernel(Kdev, s); p->tune(s); } catch (...) { delete p; throw; }
    d = p;
  } else {
    auto *p = new DemagPlan<double>();
    try { p->setup(N, P, s); p->load_kernel(Kdev, s); p->tune(s); } catch (...) { delete p; throw; }
    d = p;
  }
  return d;
}

}  // namespace um
