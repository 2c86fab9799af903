// Vector arithmetic (opencl/madd.go, normalize.go, vecnorm.go), RK stage and
// error kernels (engine/rk45dp.go, heun.go, rk4.go, stepcontroller.go) and
// reductions (opencl/reduce.go) accumulated in FP64.
#include "batch.hpp"
#include "vec.hpp"

namespace um {

// ---- fills (kernel based, so they may sit inside captured graphs) ----------
struct FillArgs {
  unsigned char *dst;
  size_t n;   // words (k_fill_bytes) or bytes (k_fill_tail)
};
struct BFillBytes {
  using Args = FillArgs;
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const FillArgs &a, uint3 blockIdx, uint3 gridDim) {
    uint32_t *dst = (uint32_t *)a.dst;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < a.n; i += (size_t)gridDim.x * blockDim.x)
      dst[i] = 0u;
  }
};
__global__ __launch_bounds__(256) void k_fill_bytes(FillArgs a) { BFillBytes::run(a, hw_bid(), hw_grid()); }
struct BFillTail {
  using Args = FillArgs;
  static constexpr int NT = 64, MINB = 1;
  static __device__ __forceinline__ void run(const FillArgs &a, uint3, uint3) {
    if (threadIdx.x < a.n) a.dst[threadIdx.x] = 0;
  }
};
__global__ __launch_bounds__(64) void k_fill_tail(FillArgs a) { BFillTail::run(a, hw_bid(), hw_grid()); }
void fill_bytes(void *dst, size_t bytes, hipStream_t s) {
  const size_t nw = bytes / 4;
  if (nw) {
    launch_b<BFillBytes>(k_fill_bytes, dim3(blocks_for((long)nw, 256, 8192)), dim3(256), 0, s,
                (FillArgs{(unsigned char *)dst, nw}));
    UM_HIP(hipGetLastError());
  }
  if (bytes % 4) {
    launch_b<BFillTail>(k_fill_tail, dim3(1), dim3(64), 0, s, (FillArgs{(unsigned char *)dst + nw * 4, bytes % 4}));
    UM_HIP(hipGetLastError());
  }
}
template <typename T>
__global__ void k_fill_value(T *dst, T v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = v;
}
template <typename T>
void fill_value(T *dst, T v, size_t n, hipStream_t s) {
  if (!n) return;
  UM_LAUNCH(k_fill_value<T>, dim3(blocks_for((long)n, 256, 8192)), dim3(256), 0, s, dst, v, n);
  UM_HIP(hipGetLastError());
}
template void fill_value<float>(float *, float, size_t, hipStream_t);
template void fill_value<double>(double *, double, size_t, hipStream_t);
template void fill_value<unsigned long long>(unsigned long long *, unsigned long long, size_t, hipStream_t);

// ---- stage: madd + normalize ---------------------------------------------------
template <typename T>
struct BStage {
  using Args = StageArgs<T>;
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const StageArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    if (a.skip && *a.skip) return;
    const T h = a.dev_dt ? (T)(a.dev_dt[0] * a.gam) : T(1);
    auto fac = [&](int k) { return (a.dev_dt && k > 0) ? a.fac[k] * h : a.fac[k]; };
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.N; i += (long)gridDim.x * blockDim.x) {
      T r[3] = {T(0), T(0), T(0)};
      const int nc = a.ncomp;
      for (int c = 0; c < nc; c++) {
        T acc;
        if (a.group3 && a.nsrc == 3) {
          acc = (fac(0) * a.src[0][c][i]) + (fac(1) * a.src[1][c][i] + fac(2) * a.src[2][c][i]);
        } else {
          acc = fac(0) * a.src[0][c][i];
          for (int k = 1; k < a.nsrc; k++) acc = acc + fac(k) * a.src[k][c][i];
        }
        r[c] = acc;
      }
      if (a.normalize && nc == 3) {
        // normalize2.cl:8-23
        const T v = a.vol ? a.vol[i] : T(1);
        T X = v * r[0], Y = v * r[1], Z = v * r[2];
        const T l = sqrt(X * X + Y * Y + Z * Z);
        const T inv = (l != T(0)) ? T(1) / l : T(0);
        if (v == T(0)) {
          r[0] = r[1] = r[2] = T(0);
        } else {
          r[0] = X * inv;
          r[1] = Y * inv;
          r[2] = Z * inv;
        }
        if (a.nrm) a.nrm[i] = l;
      }
      for (int c = 0; c < nc; c++) a.dst[c][i] = r[c];
    }
  }
};
template <typename T>
__global__ __launch_bounds__(256) void k_stage(StageArgs<T> a) { BStage<T>::run(a, hw_bid(), hw_grid()); }
// Same arithmetic, 16 bytes per access: each thread owns W = 16/sizeof(T)
// consecutive cells of every component, and the source count, component
// count, madd3 grouping and normalisation are compile-time.
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float __attribute__((ext_vector_type(4))); static constexpr int n = 4; };
template <> struct Vec16<double> { using type = double __attribute__((ext_vector_type(2))); static constexpr int n = 2; };
// Streaming accesses: every element is touched once per kernel, so the loads
// and stores carry the nontemporal hint (measured ~12% faster on the 7-vector
// RK stages than cached accesses).
template <typename V>
__device__ __forceinline__ V ld_nt(const void *base, long g) { return __builtin_nontemporal_load((const V *)base + g); }
template <typename V>
__device__ __forceinline__ void st_nt(void *base, long g, V v) { __builtin_nontemporal_store(v, (V *)base + g); }

template <typename T, int NS, int NC, bool G3, bool NORM>
struct BStageV {
  using Args = StageArgs<T>;
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const StageArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    using V = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::n;
    const long g = blockIdx.x * 256L + threadIdx.x;
    if (g * W >= a.N) return;
    if (a.skip && *a.skip) return;
    T fac[8];
  #pragma unroll
    for (int k = 0; k < 8; k++) fac[k] = a.fac[k];
    if (a.dev_dt) {
      const T h = (T)(a.dev_dt[0] * a.gam);
  #pragma unroll
      for (int k = 1; k < 8; k++) fac[k] = a.fac[k] * h;
    }
    T r[3][W];
  #pragma unroll
    for (int c = 0; c < NC; c++) {
      T x[NS][W];
  #pragma unroll
      for (int k = 0; k < NS; k++) {
        const V v = ld_nt<V>(a.src[k][c], g);
        __builtin_memcpy(x[k], &v, sizeof(V));
      }
  #pragma unroll
      for (int w = 0; w < W; w++) {
        T acc;
        if (G3) {
          acc = (fac[0] * x[0][w]) + (fac[1] * x[1][w] + fac[2] * x[2][w]);
        } else {
          acc = fac[0] * x[0][w];
  #pragma unroll
          for (int k = 1; k < NS; k++) acc = acc + fac[k] * x[k][w];
        }
        r[c][w] = acc;
      }
    }
    if (NORM) {
      T vv[W];
      if (a.vol) {
        const V v = ld_nt<V>(a.vol, g);
        __builtin_memcpy(vv, &v, sizeof(V));
      } else {
  #pragma unroll
        for (int w = 0; w < W; w++) vv[w] = T(1);
      }
      T ln[W];
  #pragma unroll
      for (int w = 0; w < W; w++) {
        const T v = vv[w];
        const T X = v * r[0][w], Y = v * r[1][w], Z = v * r[2][w];
        const T l = sqrt(X * X + Y * Y + Z * Z);
        ln[w] = l;
        const T inv = (l != T(0)) ? T(1) / l : T(0);
        if (v == T(0)) {
          r[0][w] = r[1][w] = r[2][w] = T(0);
        } else {
          r[0][w] = X * inv;
          r[1][w] = Y * inv;
          r[2][w] = Z * inv;
        }
      }
      if (a.nrm) {
        V v;
        __builtin_memcpy(&v, ln, sizeof(V));
        st_nt<V>(a.nrm, g, v);
      }
    }
  #pragma unroll
    for (int c = 0; c < NC; c++) {
      V v;
      __builtin_memcpy(&v, r[c], sizeof(V));
      st_nt<V>(a.dst[c], g, v);
    }
  }
};
template <typename T, int NS, int NC, bool G3, bool NORM>
__global__ __launch_bounds__(256) void k_stage_v(StageArgs<T> a) { BStageV<T, NS, NC, G3, NORM>::run(a, hw_bid(), hw_grid()); }

template <typename T>
static bool stage_vectorisable(const StageArgs<T> &a) {
  constexpr int W = Vec16<T>::n;
  if (a.N % W) return false;
  auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
  for (int c = 0; c < a.ncomp; c++) {
    if (!al(a.dst[c])) return false;
    for (int k = 0; k < a.nsrc; k++)
      if (!al(a.src[k][c])) return false;
  }
  if (a.normalize && a.nrm && !al(a.nrm)) return false;
  return !(a.normalize && a.vol && !al(a.vol));
}

template <typename T>
void launch_stage(const StageArgs<T> &a, hipStream_t s) {
  UM_REQUIRE(a.nsrc >= 1 && a.nsrc <= 8, UM_EINVAL, "madd: 1..8 sources");
  UM_REQUIRE(a.ncomp == 1 || a.ncomp == 3, UM_EINVAL, "stage: ncomp must be 1 or 3");
  // 16-byte accesses pay on large arrays; small ones keep one cell per thread
  // for four times more work-groups
  if (a.N >= 512L * 256 * Vec16<T>::n && stage_vectorisable(a)) {
    const long groups = a.N / Vec16<T>::n;
    const dim3 grid((unsigned)((groups + 255) / 256));
    const bool g3 = a.group3 && a.nsrc == 3, nrm = a.normalize && a.ncomp == 3;
#define UM_STAGE(NS)                                                                          \
  case NS:                                                                                    \
    if (a.ncomp == 1) launch_b<BStageV<T, NS, 1, false, false>>(k_stage_v<T, NS, 1, false, false>, grid, dim3(256), 0, s, a); \
    else if (nrm) launch_b<BStageV<T, NS, 3, false, true>>(k_stage_v<T, NS, 3, false, true>, grid, dim3(256), 0, s, a);      \
    else launch_b<BStageV<T, NS, 3, false, false>>(k_stage_v<T, NS, 3, false, false>, grid, dim3(256), 0, s, a);             \
    break;
    if (g3) {
      if (a.ncomp == 1) launch_b<BStageV<T, 3, 1, true, false>>(k_stage_v<T, 3, 1, true, false>, grid, dim3(256), 0, s, a);
      else if (nrm) launch_b<BStageV<T, 3, 3, true, true>>(k_stage_v<T, 3, 3, true, true>, grid, dim3(256), 0, s, a);
      else launch_b<BStageV<T, 3, 3, true, false>>(k_stage_v<T, 3, 3, true, false>, grid, dim3(256), 0, s, a);
    } else {
      switch (a.nsrc) {
        UM_STAGE(1) UM_STAGE(2) UM_STAGE(3) UM_STAGE(4) UM_STAGE(5) UM_STAGE(6) UM_STAGE(7) UM_STAGE(8)
      }
    }
#undef UM_STAGE
    UM_HIP(hipGetLastError());
    return;
  }
  launch_b<BStage<T>>(k_stage<T>, dim3(blocks_for(a.N, 256, 16384)), dim3(256), 0, s, a);
  UM_HIP(hipGetLastError());
}
template void launch_stage<float>(const StageArgs<float> &, hipStream_t);
template void launch_stage<double>(const StageArgs<double> &, hipStream_t);

// ---- block max helper -------------------------------------------------------------
// max that keeps NaN (fmax would drop it): a NaN error must reject the step
__device__ __forceinline__ double nmax(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// max-combine non-negative doubles by their bit pattern (NaN propagates as the largest)
__device__ __forceinline__ void block_max_to(double v, unsigned long long *dst) {
  __shared__ double sh[16];
  v = wave_max(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int i = 1; i < (int)(blockDim.x + 63) / 64; i++) m = nmax(m, sh[i]);
    if (m != m) m = __longlong_as_double(0x7ff8000000000000LL);
    atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
  __syncthreads();
}

// ---- fused error reductions --------------------------------------------------------
template <typename T>
struct BErr {
  using Args = ErrArgs<T>;
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const ErrArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    double m0 = 0, m1 = 0, m2 = 0;
    bool nan = false;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.N; i += (long)gridDim.x * blockDim.x) {
      if (a.mode == 0) {
        T E[3], K7[3];
  #pragma unroll
        for (int c = 0; c < 3; c++) {
          // Madd6(Err, k1,k3,k4,k5,k6,k7, e...), engine/rk45dp.go:95
          T acc = a.e[0] * a.k[0][c][i];
  #pragma unroll
          for (int j = 1; j < 6; j++) acc = acc + a.e[j] * a.k[j][c][i];
          E[c] = acc;
          K7[c] = a.k[5][c][i];
        }
        const double e2 = (double)E[0] * E[0] + (double)E[1] * E[1] + (double)E[2] * E[2];
        const double k2 = (double)K7[0] * K7[0] + (double)K7[1] * K7[1] + (double)K7[2] * K7[2];
        // VecNorm(errnorm)/VecNorm(ddtnorm), 0 where |k7| = 0 (div.cl:9)
        const double rel = (k2 == 0.0) ? 0.0 : sqrt(e2) / sqrt(k2);
        nan |= (e2 != e2);
        m0 = fmax(m0, e2);
        m1 = fmax(m1, k2);
        m2 = fmax(m2, rel);
      } else if (a.mode == 1) {
        const double dx = (double)(a.k[0][0][i] - a.k[1][0][i]);
        const double dy = (double)(a.k[0][1][i] - a.k[1][1][i]);
        const double dz = (double)(a.k[0][2][i] - a.k[1][2][i]);
        const double d2 = dx * dx + dy * dy + dz * dz;
        nan |= (d2 != d2);
        m0 = fmax(m0, d2);
      } else {
        const double x = a.k[0][0][i], y = a.k[0][1][i], z = a.k[0][2][i];
        const double d2 = x * x + y * y + z * z;
        nan |= (d2 != d2);
        m0 = fmax(m0, d2);
      }
    }
    if (nan) m0 = __longlong_as_double(0x7ff8000000000000LL);
    block_max_to(m0, a.res);
    if (a.mode == 0) {
      block_max_to(m1, a.res + 1);
      block_max_to(m2, a.res + 2);
    }
  }
};
template <typename T>
__global__ __launch_bounds__(256) void k_err(ErrArgs<T> a) { BErr<T>::run(a, hw_bid(), hw_grid()); }
// Same reductions with 16-byte accesses (W cells per access), mode compile-time.
template <typename T, int MODE>
struct BErrV {
  using Args = ErrArgs<T>;
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const ErrArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    using V = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::n, NK = MODE == 0 ? 6 : MODE == 1 ? 2 : 1;
    double m0 = 0, m1 = 0, m2 = 0;
    bool nan = false;
    const long groups = a.N / W;
  #pragma unroll 2
    for (long g = blockIdx.x * 256L + threadIdx.x; g < groups; g += (long)gridDim.x * 256) {
      T x[NK][3][W];
  #pragma unroll
      for (int j = 0; j < NK; j++)
  #pragma unroll
        for (int c = 0; c < 3; c++) {
          const V v = ld_nt<V>(a.k[j][c], g);
          __builtin_memcpy(x[j][c], &v, sizeof(V));
        }
  #pragma unroll
      for (int w = 0; w < W; w++) {
        if (MODE == 0) {
          T E[3];
  #pragma unroll
          for (int c = 0; c < 3; c++) {
            T acc = a.e[0] * x[0][c][w];
  #pragma unroll
            for (int j = 1; j < 6; j++) acc = acc + a.e[j] * x[j][c][w];
            E[c] = acc;
          }
          const double e2 = (double)E[0] * E[0] + (double)E[1] * E[1] + (double)E[2] * E[2];
          const double k2 = (double)x[5][0][w] * x[5][0][w] + (double)x[5][1][w] * x[5][1][w] +
                            (double)x[5][2][w] * x[5][2][w];
          const double rel = (k2 == 0.0) ? 0.0 : sqrt(e2) / sqrt(k2);
          nan |= (e2 != e2);
          m0 = fmax(m0, e2);
          m1 = fmax(m1, k2);
          m2 = fmax(m2, rel);
        } else if (MODE == 1) {
          const double dx = (double)(x[0][0][w] - x[1][0][w]);
          const double dy = (double)(x[0][1][w] - x[1][1][w]);
          const double dz = (double)(x[0][2][w] - x[1][2][w]);
          const double d2 = dx * dx + dy * dy + dz * dz;
          nan |= (d2 != d2);
          m0 = fmax(m0, d2);
        } else {
          const double X = x[0][0][w], Y = x[0][1][w], Z = x[0][2][w];
          const double d2 = X * X + Y * Y + Z * Z;
          nan |= (d2 != d2);
          m0 = fmax(m0, d2);
        }
      }
    }
    if (nan) m0 = __longlong_as_double(0x7ff8000000000000LL);
    block_max_to(m0, a.res);
    if (MODE == 0) {
      block_max_to(m1, a.res + 1);
      block_max_to(m2, a.res + 2);
    }
  }
};
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_err_v(ErrArgs<T> a) { BErrV<T, MODE>::run(a, hw_bid(), hw_grid()); }

template <typename T>
void launch_err(const ErrArgs<T> &a, hipStream_t s) {
  constexpr int W = Vec16<T>::n;
  const int nk = a.mode == 0 ? 6 : a.mode == 1 ? 2 : 1;
  bool vec = a.N > 0 && a.N % W == 0;
  for (int j = 0; j < nk && vec; j++)
    for (int c = 0; c < 3; c++) vec = vec && ((uintptr_t)a.k[j][c] & 15) == 0;
  if (vec) {
    // at most 512 work-groups: each ends with three device-scope atomic maxima
    // on the same three slots, which serialise (2048 work-groups: 124 us at
    // 8.4M cells, 512: 114 us)
    const dim3 grid(blocks_for(a.N / W, 256, 512));
    if (a.mode == 0) launch_b<BErrV<T, 0>>(k_err_v<T, 0>, grid, dim3(256), 0, s, a);
    else if (a.mode == 1) launch_b<BErrV<T, 1>>(k_err_v<T, 1>, grid, dim3(256), 0, s, a);
    else launch_b<BErrV<T, 2>>(k_err_v<T, 2>, grid, dim3(256), 0, s, a);
    UM_HIP(hipGetLastError());
    return;
  }
  launch_b<BErr<T>>(k_err<T>, dim3(blocks_for(a.N, 256, 2048)), dim3(256), 0, s, a);
  UM_HIP(hipGetLastError());
}
template void launch_err<float>(const ErrArgs<float> &, hipStream_t);
template void launch_err<double>(const ErrArgs<double> &, hipStream_t);

// ---- step statistics on the device ----------------------------------------------------
// Folds one reduction of a captured step into gs (last error, peak error,
// last torque, step counter as u64 bits in gs[3]) and clears the 4 reduction
// slots for the next reduction, so a captured step needs no fill launches;
// count: this is the step's last statistics launch.
struct StatsArgs {
  unsigned long long *res;
  double *gs;
  double h;
  int mode, count;
};
struct BStepStats {
  using Args = StatsArgs;
  static constexpr int NT = 64, MINB = 1;
  static __device__ __forceinline__ void run(const StatsArgs &a, uint3, uint3) {
    if (threadIdx.x != 0) return;
    unsigned long long *res = a.res;
    double *gs = a.gs;
    const int mode = a.mode;
    const double r0 = __longlong_as_double((long long)res[0]);
    const double r1 = __longlong_as_double((long long)res[1]);
    for (int i = 0; i < 4; i++) res[i] = 0ull;
    if (a.count) *(unsigned long long *)(gs + 3) += 1ull;
    if (mode == 2) {
      gs[2] = sqrt(r0);
      return;
    }
    const double err = sqrt(mode == 3 ? 3.0 * r0 : r0) * a.h;
    gs[0] = err;
    if (err > gs[1]) gs[1] = err;
    if (mode == 0) gs[2] = sqrt(r1);
  }
};
__global__ __launch_bounds__(64) void k_step_stats(StatsArgs a) { BStepStats::run(a, hw_bid(), hw_grid()); }
struct BStepCount {
  using Args = StatsArgs;
  static constexpr int NT = 64, MINB = 1;
  static __device__ __forceinline__ void run(const StatsArgs &a, uint3, uint3) {
    if (threadIdx.x == 0) *a.res += 1ull;
  }
};
__global__ __launch_bounds__(64) void k_step_count(StatsArgs a) { BStepCount::run(a, hw_bid(), hw_grid()); }
void launch_step_stats(unsigned long long *res, double *gs, double h, int mode, int count, hipStream_t s) {
  launch_b<BStepStats>(k_step_stats, dim3(1), dim3(64), 0, s, (StatsArgs{res, gs, h, mode, count}));
  UM_HIP(hipGetLastError());
}
void launch_step_count(unsigned long long *counter, hipStream_t s) {
  launch_b<BStepCount>(k_step_count, dim3(1), dim3(64), 0, s, (StatsArgs{counter, nullptr, 0.0, 0, 0}));
  UM_HIP(hipGetLastError());
}

// ---- device-side control of adaptive RK45DP steps -------------------------------------
// integralController (engine/stepcontroller.go:35-97) and adaptDt
// (engine/run.go:150-187) with FixDt = 0 and the alarm at the run's end,
// restated from Sim::step_rk45dp / Sim::adapt: one thread.
struct BAdapt {
  struct Args {
    unsigned long long *res;
    AdaptState *st;
    AdaptPar p;
  };
  static constexpr int NT = 64, MINB = 1;
  static __device__ __forceinline__ void run(const Args &a, uint3, uint3) {
    if (threadIdx.x != 0) return;
    AdaptState *st = a.st;
    const AdaptPar &p = a.p;
    const double r0 = __longlong_as_double((long long)a.res[0]);
    const double r1 = __longlong_as_double((long long)a.res[1]);
    const double r2 = __longlong_as_double((long long)a.res[2]);
    for (int i = 0; i < 4; i++) a.res[i] = 0ull;
    st->accepted = 0;
    if (st->done) return;
    st->nevals += p.evals_per_step;
    const double dt = st->dt, t0 = st->time;
    const double h = p.h_f32 ? (double)(float)(dt * p.gam) : dt * p.gam;
    const double err = sqrt(r0) * h, maxdm = sqrt(r1);
    double corr;
    if (err < p.maxerr || dt <= p.mindt) {
      int fail = 0;
      double rlerr = 0;
      if (!(maxdm < p.minslope)) { rlerr = r2; fail = 1; }
      if (fail == 0 || p.relerr <= 0.0 || rlerr < p.relerr || dt <= p.mindt) {
        st->last_err = err;
        if (err > st->peak_err) st->peak_err = err;
        st->last_torque = maxdm;
        st->nsteps++;
        st->time = t0 + dt;
        st->accepted = 1;
        corr = fail == 0 ? pow(p.maxerr / err, 1. / 5) : pow(p.relerr / rlerr, 1. / 5);
      } else {
        st->nundone++;
        corr = pow(p.relerr / rlerr, 1. / 6);
      }
    } else {
      st->nundone++;
      corr = pow(p.maxerr / err, 1. / 6);
    }
    // adaptDt
    if (isnan(corr)) corr = 1;
    if (corr == 0) { st->error = 1; st->done = 1; return; }
    corr *= p.headroom;
    if (corr > 2) corr = 2;
    if (corr < 0.5) corr = 0.5;
    double ndt = dt * corr;
    if (p.mindt != 0 && ndt < p.mindt) ndt = p.mindt;
    if (p.maxdt != 0 && ndt > p.maxdt) ndt = p.maxdt;
    if (ndt == 0) { st->error = 1; st->done = 1; return; }
    if (st->time < st->stop && st->time + ndt > st->stop) ndt = st->stop - st->time;
    if (!(ndt > 0)) { st->error = 1; st->done = 1; return; }
    st->dt = ndt;
    if (st->accepted && !(st->time < st->stop)) st->done = 1;
  }
};
__global__ __launch_bounds__(64) void k_adapt(BAdapt::Args a) { BAdapt::run(a, hw_bid(), hw_grid()); }

// k1 <- k7 after an accepted step, m <- m0 otherwise (rejected, or past the end)
template <typename T>
struct BAdaptCopy {
  struct Args {
    const AdaptState *st;
    T *m;
    const T *m0;
    T *k1;
    const T *k7;
    long n3;
  };
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const Args &a, uint3 blockIdx, uint3 gridDim) {
    const bool acc = a.st->accepted != 0;
    T *dst = acc ? a.k1 : a.m;
    const T *src = acc ? a.k7 : a.m0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.n3; i += (long)gridDim.x * blockDim.x)
      dst[i] = src[i];
  }
};
template <typename T>
__global__ __launch_bounds__(256) void k_adapt_copy(typename BAdaptCopy<T>::Args a) {
  BAdaptCopy<T>::run(a, hw_bid(), hw_grid());
}

template <typename T>
void launch_adapt(unsigned long long *res, AdaptState *st, const AdaptPar &p, T *m, const T *m0, T *k1,
                  const T *k7, long n3, hipStream_t s) {
  launch_b<BAdapt>(k_adapt, dim3(1), dim3(64), 0, s, BAdapt::Args{res, st, p});
  launch_b<BAdaptCopy<T>>(k_adapt_copy<T>, dim3(blocks_for(n3, 256, 8192)), dim3(256), 0, s,
                          typename BAdaptCopy<T>::Args{st, m, m0, k1, k7, n3});
}
template void launch_adapt<float>(unsigned long long *, AdaptState *, const AdaptPar &, float *, const float *,
                                  float *, const float *, long, hipStream_t);
template void launch_adapt<double>(unsigned long long *, AdaptState *, const AdaptPar &, double *, const double *,
                                   double *, const double *, long, hipStream_t);

// ---- deterministic weighted dot sums ------------------------------------------------
template <typename T>
struct WdotArgs {
  const T *a[3];
  const T *b[3];
  int ncomp;
  const T *w;
  T wmul;
  long N;
  double *partial;
};
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_wdot(WdotArgs<T> a) {
  double acc = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.N; i += (long)gridDim.x * blockDim.x) {
    double d = 0;
    for (int c = 0; c < a.ncomp; c++) d += (double)a.a[c][i] * (a.b[c] ? (double)a.b[c][i] : 1.0);
    const double w = a.w ? (double)a.wmul * (double)a.w[i] : (double)a.wmul;
    acc += w * d;
  }
  __shared__ double sh[RED_THREADS / 64];
  acc = wave_sum(acc);
  if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int k = 0; k < RED_THREADS / 64; k++) t += sh[k];
    a.partial[blockIdx.x] = t;
  }
}
// ---- steepest-descent minimizer ------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_min_step(T *m, T *m0, const T *k, T h, long N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    T v[3], t[3];
    for (int c = 0; c < 3; c++) { v[c] = m[c * N + i]; t[c] = k[c * N + i]; }
    const T t2 = h * h * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    const T div = T(4.0) + t2;
    for (int c = 0; c < 3; c++) {
      m0[c * N + i] = v[c];
      m[c * N + i] = ((T(4.0) - t2) * v[c] + T(4.0) * h * t[c]) / div;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_min_stats(const T *m, const T *m0, const T *k, const T *k0,
                                                          long N, double *partial) {
  // v: max|dm|^2, sum dm.dm, sum dm.dk, sum dk.dk, max|k|^2
  double v[5] = {0, 0, 0, 0, 0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double n2 = 0, k2 = 0;
    for (int c = 0; c < 3; c++) {
      const double dm = (double)(m[c * N + i] - m0[c * N + i]);
      const double dk = (double)(k0[c * N + i] - k[c * N + i]);
      const double kk = (double)k[c * N + i];
      n2 += dm * dm;
      v[2] += dm * dk;
      v[3] += dk * dk;
      k2 += kk * kk;
    }
    v[1] += n2;
    v[0] = fmax(v[0], n2);
    v[4] = fmax(v[4], k2);
  }
  __shared__ double sh[5][RED_THREADS / 64];
#pragma unroll
  for (int j = 0; j < 5; j++) {
    double x = v[j];
    const bool mx = (j == 0 || j == 4);
    for (int off = 32; off >= 1; off >>= 1) {
      const double o = __shfl_xor(x, off, 64);
      x = mx ? fmax(x, o) : x + o;
    }
    if (threadIdx.x % 64 == 0) sh[j][threadIdx.x / 64] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < 5; j++) {
      const bool mx = (j == 0 || j == 4);
      double r = 0;
      for (int w = 0; w < RED_THREADS / 64; w++) r = mx ? fmax(r, sh[j][w]) : r + sh[j][w];
      partial[5 * blockIdx.x + j] = r;
    }
  }
}
__global__ void k_min_final(const double *partial, int n, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double mx = 0, mk = 0, s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
  for (int i = 0; i < n; i++) {
    mx = fmax(mx, partial[5 * i]);
    mk = fmax(mk, partial[5 * i + 4]);
    for (int j = 0; j < 3; j++) {   // Kahan
      const double y = partial[5 * i + 1 + j] - c[j];
      const double t = s[j] + y;
      c[j] = (t - s[j]) - y;
      s[j] = t;
    }
  }
  out[0] = mx;
  for (int j = 0; j < 3; j++) out[1 + j] = s[j];
  out[4] = mk;
}
template <typename T>
void launch_min_step(T *m, T *m0, const T *k, T h, long N, hipStream_t s) {
  UM_LAUNCH(k_min_step<T>, dim3(blocks_for(N, 256, 16384)), dim3(256), 0, s, m, m0, k, h, N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void launch_min_stats(const T *m, const T *m0, const T *k, const T *k0, long N, double *partial, double *out,
                      hipStream_t s) {
  UM_LAUNCH(k_min_stats<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, m, m0, k, k0, N, partial);
  UM_HIP(hipGetLastError());
  UM_LAUNCH(k_min_final, dim3(1), dim3(64), 0, s, (const double *)partial, RED_BLOCKS, out);
  UM_HIP(hipGetLastError());
}
template void launch_min_step<float>(float *, float *, const float *, float, long, hipStream_t);
template void launch_min_step<double>(double *, double *, const double *, double, long, hipStream_t);
template void launch_min_stats<float>(const float *, const float *, const float *, const float *, long, double *,
                                      double *, hipStream_t);
template void launch_min_stats<double>(const double *, const double *, const double *, const double *, long,
                                       double *, double *, hipStream_t);

// Kahan-compensated final sum over the fixed partials, one thread order.
__global__ void k_final_sum(const double *partial, int n, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0, c = 0;
  for (int i = 0; i < n; i++) {
    const double y = partial[i] - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  *out = s;
}
template <typename T>
void launch_wdot(const T *const *a, const T *const *b, int ncomp, const T *w, T wmul, long N, double *partial,
                 double *out, hipStream_t s) {
  WdotArgs<T> args{};
  for (int c = 0; c < ncomp; c++) {
    args.a[c] = a[c];
    args.b[c] = b ? b[c] : nullptr;
  }
  args.ncomp = ncomp;
  args.w = w;
  args.wmul = wmul;
  args.N = N;
  args.partial = partial;
  UM_LAUNCH(k_wdot<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, args);
  UM_HIP(hipGetLastError());
  UM_LAUNCH(k_final_sum, dim3(1), dim3(64), 0, s, (const double *)partial, RED_BLOCKS, out);
  UM_HIP(hipGetLastError());
}
template void launch_wdot<float>(const float *const *, const float *const *, int, const float *, float, long,
                                 double *, double *, hipStream_t);
template void launch_wdot<double>(const double *const *, const double *const *, int, const double *, double, long,
                                  double *, double *, hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void k_maxabs(const T *x, long N, unsigned long long *res) {
  double m = 0;
  bool nan = false;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double v = fabs((double)x[i]);
    nan |= (v != v);
    m = fmax(m, v);
  }
  if (nan) m = __longlong_as_double(0x7ff8000000000000LL);
  block_max_to(m, res);
}
template <typename T>
void launch_maxabs(const T *x, long N, unsigned long long *res, hipStream_t s) {
  UM_LAUNCH(k_maxabs<T>, dim3(blocks_for(N, 256, 2048)), dim3(256), 0, s, x, N, res);
  UM_HIP(hipGetLastError());
}
template void launch_maxabs<float>(const float *, long, unsigned long long *, hipStream_t);
template void launch_maxabs<double>(const double *, long, unsigned long long *, hipStream_t);

// ---- pointwise helpers (vecnorm.cl, div.cl, mul.cl) ---------------------------------
template <typename T>
__global__ void k_vecnorm(T *dst, const T *ax, const T *ay, const T *az, long N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    dst[i] = sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
}
template <typename T>
__global__ void k_div(T *dst, const T *a, const T *b, long N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    dst[i] = (b[i] == T(0)) ? T(0) : (a[i] / b[i]);
}
template <typename T>
__global__ void k_mul(T *dst, const T *a, const T *b, long N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    dst[i] = a[i] * b[i];
}
template <typename T>
void launch_vecnorm(T *dst, const T *const a[3], long N, hipStream_t s) {
  UM_LAUNCH(k_vecnorm<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, dst, a[0], a[1], a[2], N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void launch_div(T *dst, const T *a, const T *b, long N, hipStream_t s) {
  UM_LAUNCH(k_div<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, dst, a, b, N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void launch_mul(T *dst, const T *a, const T *b, long N, hipStream_t s) {
  UM_LAUNCH(k_mul<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, dst, a, b, N);
  UM_HIP(hipGetLastError());
}
template void launch_vecnorm<float>(float *, const float *const[3], long, hipStream_t);
template void launch_vecnorm<double>(double *, const double *const[3], long, hipStream_t);
template void launch_div<float>(float *, const float *, const float *, long, hipStream_t);
template void launch_div<double>(double *, const double *, const double *, long, hipStream_t);
template void launch_mul<float>(float *, const float *, const float *, long, hipStream_t);
template void launch_mul<double>(double *, const double *, const double *, long, hipStream_t);

}  // namespace um
