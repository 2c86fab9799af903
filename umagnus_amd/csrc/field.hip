// Field-term kernels: unfused entry points (opencl/exchange.go, dmi.go,
// anisotropy.go, region.go, zeromask.go, lltorque.go, temperature.go) and the
// fused effective-field + torque pass.
#include "batch.hpp"
#include "field.hpp"

namespace um {

// ---- unfused stencil terms ---------------------------------------------------
template <typename T>
__global__ void k_addexchange(T *Bx, T *By, T *Bz, const T *mx, const T *my, const T *mz, Par<T> Ms,
                              const T *aLUT, const uint8_t *regions, T wx, T wy, T wz, Mesh g) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= g.Nx || iy >= g.Ny) return;
  const T *m[3] = {mx, my, mz};
  const long I = g.idx(ix, iy, iz);
  const V3<T> m0 = load3(m, I);
  if (is0(m0)) return;
  const V3<T> B = exchange_cell(g, ix, iy, iz, I, m0, m, regions, aLUT, wx, wy, wz);
  const T invMs = inv_msat(Ms, I);
  Bx[I] += B.x * invMs;
  By[I] += B.y * invMs;
  Bz[I] += B.z * invMs;
}

template <typename T>
__global__ void k_adddmi(T *Bx, T *By, T *Bz, const T *mx, const T *my, const T *mz, Par<T> Ms,
                         const T *aLUT, const T *dLUT, const uint8_t *regions, T cx, T cy, T cz, Mesh g,
                         int openbc) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= g.Nx || iy >= g.Ny) return;
  const T *m[3] = {mx, my, mz};
  const long I = g.idx(ix, iy, iz);
  const V3<T> m0 = load3(m, I);
  if (is0(m0)) return;
  const V3<T> h = dmi_cell(g, ix, iy, iz, I, m0, m, regions, aLUT, dLUT, cx, cy, cz, openbc);
  const T invMs = inv_msat(Ms, I);
  Bx[I] += h.x * invMs;
  By[I] += h.y * invMs;
  Bz[I] += h.z * invMs;
}

template <typename T>
__global__ void k_adddmibulk(T *Bx, T *By, T *Bz, const T *mx, const T *my, const T *mz, Par<T> Ms,
                             const T *aLUT, const T *dLUT, const uint8_t *regions, T cx, T cy, T cz, Mesh g,
                             int openbc) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= g.Nx || iy >= g.Ny) return;
  const T *m[3] = {mx, my, mz};
  const long I = g.idx(ix, iy, iz);
  const V3<T> m0 = load3(m, I);
  if (is0(m0)) return;
  const V3<T> h = dmibulk_cell(g, ix, iy, iz, I, m0, m, regions, aLUT, dLUT, cx, cy, cz, openbc);
  const T invMs = inv_msat(Ms, I);
  Bx[I] += h.x * invMs;
  By[I] += h.y * invMs;
  Bz[I] += h.z * invMs;
}

// tau += Zhang-Li torque (zhangli2.cl) and/or Slonczewski torque (slonczewski2.cl)
template <typename T>
__global__ void k_addstt(T *tx, T *ty, T *tz, const T *mx, const T *my, const T *mz, Par<T> Ms, Par<T> alpha,
                         SttPar<T> p, T cx, T cy, T cz, Mesh g, int zl, int sl) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= g.Nx || iy >= g.Ny) return;
  const T *m[3] = {mx, my, mz};
  const long I = g.idx(ix, iy, iz);
  const V3<T> m0 = load3(m, I);
  V3<T> t = v3(T(0), T(0), T(0));
  if (zl) t = zhangli_cell(g, ix, iy, iz, I, m0, m, Ms, alpha, p, cx, cy, cz);
  if (sl) {
    const V3<T> z = slonczewski_cell(I, m0, Ms, alpha, p);
    t.x += z.x; t.y += z.y; t.z += z.z;
  }
  tx[I] += t.x;
  ty[I] += t.y;
  tz[I] += t.z;
}

template <typename T>
__global__ void k_adduniaxial2(T *Bx, T *By, T *Bz, const T *mx, const T *my, const T *mz, Par<T> Ms,
                               Par<T> K1, Par<T> K2, Par<T> ux, Par<T> uy, Par<T> uz, int N) {
  const Par<T> u[3] = {ux, uy, uz};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const V3<T> B = uniaxial_cell(i, v3(mx[i], my[i], mz[i]), Ms, K1, K2, u);
    Bx[i] += B.x;
    By[i] += B.y;
    Bz[i] += B.z;
  }
}

template <typename T>
struct CubicArgs {
  T *B[3];
  const T *m[3];
  Par<T> Ms, K1, K2, K3, c1[3], c2[3];
  int N;
};
template <typename T>
__global__ void k_addcubic2(CubicArgs<T> a) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.N; i += (long)gridDim.x * blockDim.x) {
    const V3<T> B = cubic_cell(i, load3(a.m, i), a.Ms, a.K1, a.K2, a.K3, a.c1, a.c2);
    a.B[0][i] += B.x;
    a.B[1][i] += B.y;
    a.B[2][i] += B.z;
  }
}

template <typename T>
__global__ void k_regionaddv(T *dx, T *dy, T *dz, const T *lx, const T *ly, const T *lz, const uint8_t *regions,
                             int N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const int r = regions ? regions[i] : 0;
    dx[i] += lx[r];
    dy[i] += ly[r];
    dz[i] += lz[r];
  }
}

template <typename T>
__global__ void k_zeromask(T *dst, const T *maskLUT, const uint8_t *regions, int N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    if (maskLUT[regions ? regions[i] : 0] != T(0)) dst[i] = T(0);
}

template <typename T>
__global__ void k_lltorque(T *tx, T *ty, T *tz, const T *mx, const T *my, const T *mz, const T *hx, const T *hy,
                           const T *hz, Par<T> alpha, int N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const V3<T> t = lltorque_cell(v3(mx[i], my[i], mz[i]), v3(hx[i], hy[i], hz[i]), alpha.at(i));
    tx[i] = t.x;
    ty[i] = t.y;
    tz[i] = t.z;
  }
}

template <typename T>
__global__ void k_llnoprecess(T *tx, T *ty, T *tz, const T *mx, const T *my, const T *mz, const T *hx,
                              const T *hy, const T *hz, int N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const V3<T> m = v3(mx[i], my[i], mz[i]);
    const V3<T> t = vcross(m, vcross(m, v3(hx[i], hy[i], hz[i])));
    tx[i] = -t.x;
    ty[i] = -t.y;
    tz[i] = -t.z;
  }
}

template <typename T>
__global__ void k_thermal(T *Bx, T *By, T *Bz, uint64_t seed, uint64_t step, uint32_t stream, T k, Par<T> Ms,
                          Par<T> temp, Par<T> alpha, int N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const V3<T> b = thermal_cell(i, seed, step, stream, k, Ms, temp, alpha);
    Bx[i] = b.x;
    By[i] = b.y;
    Bz[i] = b.z;
  }
}

// ---- fused effective field + torque ------------------------------------------------
// B = B_demag + B_exch(+DMI) + B_anis + B_ext + B_therm in the order of
// SetEffectiveField (engine/effectivefield.go:14-24), then tau = LLTorque(m, B)
// (lltorque2.cl) and FreezeSpins (engine/torque.go:348-352), one pass per cell.
template <typename T>
struct BFieldTorque {
  using Args = FieldArgs<T>;
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const FieldArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    const int iz = blockIdx.z;
    const Mesh &g = a.g;
    if (ix >= g.Nx || iy >= g.Ny) return;
    const long I = g.idx(ix, iy, iz);
    const V3<T> m0 = load3(a.m, I);
    V3<T> B = a.bd[0] ? load3(a.bd, I) : v3(T(0), T(0), T(0));
    if (a.exch_mode != 0 && !is0(m0)) {
      const V3<T> h =
          (a.exch_mode == 1)   ? exchange_cell(g, ix, iy, iz, I, m0, a.m, a.regions, a.aLUT, a.wx, a.wy, a.wz)
          : (a.exch_mode == 2) ? dmi_cell(g, ix, iy, iz, I, m0, a.m, a.regions, a.aLUT, a.dLUT, a.cx, a.cy, a.cz, a.openbc)
                               : dmibulk_cell(g, ix, iy, iz, I, m0, a.m, a.regions, a.aLUT, a.dLUT, a.cx, a.cy, a.cz,
                                              a.openbc);
      const T invMs = inv_msat(a.Ms, I);
      B.x += h.x * invMs;
      B.y += h.y * invMs;
      B.z += h.z * invMs;
    }
    if (a.has_uniax) {
      const V3<T> h = uniaxial_cell(I, m0, a.Ms, a.Ku1, a.Ku2, a.u);
      B.x += h.x; B.y += h.y; B.z += h.z;
    }
    if (a.has_cubic) {
      const V3<T> h = cubic_cell(I, m0, a.Ms, a.Kc1, a.Kc2, a.Kc3, a.c1, a.c2);
      B.x += h.x; B.y += h.y; B.z += h.z;
    }
    if (a.bextLUT[0]) {
      const int r = a.regions ? a.regions[I] : 0;
      B.x += a.bextLUT[0][r];
      B.y += a.bextLUT[1][r];
      B.z += a.bextLUT[2][r];
    }
    if (a.has_bext) {
      B.x += a.bext[0].at(I);
      B.y += a.bext[1].at(I);
      B.z += a.bext[2].at(I);
    }
    if (a.has_therm) {
      const uint64_t stp = a.step_dev ? (uint64_t)*a.step_dev : a.step;
      const V3<T> h = thermal_cell(I, a.seed, stp, a.stream, a.kth, a.Ms, a.temp, a.alpha);
      B.x += h.x; B.y += h.y; B.z += h.z;
    }
    if (a.badd[0]) {   // AddCustomField (engine/effectivefield.go:23)
      const V3<T> h = load3(a.badd, I);
      B.x += h.x; B.y += h.y; B.z += h.z;
    }
    if (a.beff[0]) {
      a.beff[0][I] = B.x;
      a.beff[1][I] = B.y;
      a.beff[2][I] = B.z;
    }
    if (a.tau[0]) {
      V3<T> t;
      if (a.precess) {
        t = lltorque_cell(m0, B, a.alpha.at(I));
      } else {
        const V3<T> c = vcross(m0, vcross(m0, B));
        t = v3(-c.x, -c.y, -c.z);
      }
      // AddSTTorque (engine/torque.go:147-196), before FreezeSpins
      if (a.has_zl) {
        const V3<T> z = zhangli_cell(g, ix, iy, iz, I, m0, a.m, a.Ms, a.alpha, a.stt, a.cx, a.cy, a.cz);
        t.x += z.x; t.y += z.y; t.z += z.z;
      }
      if (a.has_sl) {
        const V3<T> z = slonczewski_cell(I, m0, a.Ms, a.alpha, a.stt);
        t.x += z.x; t.y += z.y; t.z += z.z;
      }
      if (a.tadd[0]) {   // AddCustomTorques (engine/torque.go:129)
        const V3<T> z = load3(a.tadd, I);
        t.x += z.x; t.y += z.y; t.z += z.z;
      }
      if (a.frozenLUT && a.frozenLUT[a.regions ? a.regions[I] : 0] != T(0)) t = v3(T(0), T(0), T(0));
      a.tau[0][I] = t.x;
      a.tau[1][I] = t.y;
      a.tau[2][I] = t.z;
    }
  }
};
template <typename T>
__global__ __launch_bounds__(256) void k_field_torque(FieldArgs<T> a) { BFieldTorque<T>::run(a, hw_bid(), hw_grid()); }

// four x cells per thread; consecutive threads take consecutive x groups
template <typename T, bool REG, int EX, bool STT, bool UNI>
struct BFieldTorque4 {
  using Args = FieldArgs<T>;
  static constexpr int NT = 256, MINB = 1;
  static __device__ __forceinline__ void run(const FieldArgs<T> &a, uint3 blockIdx, uint3 gridDim) {
    // 32-bit indices: the launcher takes this kernel only below 2^31 groups
    const uint32_t ngx = (uint32_t)a.g.Nx >> 2;
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t rows = (uint32_t)a.g.Ny * (uint32_t)a.g.Nz;
    if (gid >= rows * ngx) return;
    const uint32_t row = gid / ngx;
    const int gx = (int)(gid - row * ngx);
    const int iz = (int)(row / (uint32_t)a.g.Ny), iy = (int)(row - (uint32_t)iz * (uint32_t)a.g.Ny);
    const long I0 = a.g.idx(gx * 4, iy, iz);
    T bdv[3][4];
  #pragma unroll
    for (int c = 0; c < 3; c++) {
      if (a.bd[0]) ld4(a.bd[c] + I0, bdv[c]);
      else bdv[c][0] = bdv[c][1] = bdv[c][2] = bdv[c][3] = T(0);
    }
    field_group4<T, REG, EX, STT, UNI>(a, gx * 4, iy, iz, bdv);
  }
};
template <typename T, bool REG, int EX, bool STT, bool UNI>
__global__ __launch_bounds__(256) void k_field_torque4(FieldArgs<T> a) { BFieldTorque4<T, REG, EX, STT, UNI>::run(a, hw_bid(), hw_grid()); }

// whether the four-cell kernel applies: FP32 (in FP64 the four cells need
// twice the registers and the one-cell kernel keeps a higher occupancy),
// Nx % 4 == 0 and every per-cell array aligned to four cells
template <typename T>
static bool field4_ok(const FieldArgs<T> &a) {
  if (sizeof(T) != 4 || a.g.Nx % 4) return false;
  auto al = [](const void *p) { return ((uintptr_t)p & (4 * sizeof(T) - 1)) == 0; };
  for (int c = 0; c < 3; c++) {
    if (!al(a.m[c])) return false;
    if (a.bd[0] && !al(a.bd[c])) return false;
    if (a.tau[0] && !al(a.tau[c])) return false;
    if (a.beff[0] && !al(a.beff[c])) return false;
  }
  return !a.regions || ((uintptr_t)a.regions & 3) == 0;
}

template <typename T>
void launch_field_torque(const FieldArgs<T> &a, hipStream_t s) {
  // four cells per thread pays on large grids; below ~512 work-groups the
  // one-cell kernel's four times more work-groups hide the latency better
  const long ncell = (long)a.g.Nx * a.g.Ny * a.g.Nz;
  if (ncell >= 512L * 256 * 4 && ncell < (1L << 31) && field4_ok(a)) {
    const long n4 = (long)a.g.Nx / 4 * a.g.Ny * a.g.Nz;
    const dim3 grid((unsigned)((n4 + 255) / 256));
    const bool reg = a.regions != nullptr;
    const bool stt = a.has_zl || a.has_sl;
    const bool uni = field_uniform(a);
#define UM_F4(R, E)                                                                          \
  do {                                                                                       \
    if (stt) launch_b<BFieldTorque4<T, R, E, true, false>>(k_field_torque4<T, R, E, true, false>, grid, dim3(256), 0, s, a);  \
    else if (uni) launch_b<BFieldTorque4<T, R, E, false, true>>(k_field_torque4<T, R, E, false, true>, grid, dim3(256), 0, s, a); \
    else launch_b<BFieldTorque4<T, R, E, false, false>>(k_field_torque4<T, R, E, false, false>, grid, dim3(256), 0, s, a);     \
  } while (0)
    if constexpr (sizeof(T) == 4) {
      if (a.exch_mode == 0) { if (reg) UM_F4(true, 0); else UM_F4(false, 0); }
      else if (a.exch_mode == 1) { if (reg) UM_F4(true, 1); else UM_F4(false, 1); }
      else if (a.exch_mode == 2) { if (reg) UM_F4(true, 2); else UM_F4(false, 2); }
      else { if (reg) UM_F4(true, 3); else UM_F4(false, 3); }
    }
#undef UM_F4
    UM_HIP(hipGetLastError());
    return;
  }
  dim3 blk(64, 4, 1);
  dim3 grd((a.g.Nx + 63) / 64, (a.g.Ny + 3) / 4, a.g.Nz);
  launch_b<BFieldTorque<T>>(k_field_torque<T>, grd, blk, 0, s, a);
  UM_HIP(hipGetLastError());
}
template void launch_field_torque<float>(const FieldArgs<float> &, hipStream_t);
template void launch_field_torque<double>(const FieldArgs<double> &, hipStream_t);

// ---- host launchers for the unfused entry points --------------------------------------
static inline dim3 grid3(const Mesh &g) { return dim3((g.Nx + 63) / 64, (g.Ny + 3) / 4, g.Nz); }

template <typename T>
void add_exchange(void *const B[3], const void *const m[3], Par<T> Ms, const T *aLUT, const uint8_t *regions, T wx,
                  T wy, T wz, Mesh g, hipStream_t s) {
  UM_LAUNCH(k_addexchange<T>, grid3(g), dim3(64, 4), 0, s, (T *)B[0], (T *)B[1], (T *)B[2],
                     (const T *)m[0], (const T *)m[1], (const T *)m[2], Ms, aLUT, regions, wx, wy, wz, g);
  UM_HIP(hipGetLastError());
}
template <typename T>
void add_dmi(void *const B[3], const void *const m[3], Par<T> Ms, const T *aLUT, const T *dLUT,
             const uint8_t *regions, T cx, T cy, T cz, Mesh g, int openbc, hipStream_t s) {
  UM_LAUNCH(k_adddmi<T>, grid3(g), dim3(64, 4), 0, s, (T *)B[0], (T *)B[1], (T *)B[2], (const T *)m[0],
                     (const T *)m[1], (const T *)m[2], Ms, aLUT, dLUT, regions, cx, cy, cz, g, openbc);
  UM_HIP(hipGetLastError());
}
template <typename T>
void add_dmi_bulk(void *const B[3], const void *const m[3], Par<T> Ms, const T *aLUT, const T *dLUT,
                  const uint8_t *regions, T cx, T cy, T cz, Mesh g, int openbc, hipStream_t s) {
  UM_LAUNCH(k_adddmibulk<T>, grid3(g), dim3(64, 4), 0, s, (T *)B[0], (T *)B[1], (T *)B[2],
                     (const T *)m[0], (const T *)m[1], (const T *)m[2], Ms, aLUT, dLUT, regions, cx, cy, cz, g,
                     openbc);
  UM_HIP(hipGetLastError());
}
template <typename T>
void add_stt(void *const t[3], const void *const m[3], Par<T> Ms, Par<T> alpha, const SttPar<T> &p, T cx, T cy,
             T cz, Mesh g, int zl, int sl, hipStream_t s) {
  UM_LAUNCH(k_addstt<T>, grid3(g), dim3(64, 4), 0, s, (T *)t[0], (T *)t[1], (T *)t[2], (const T *)m[0],
                     (const T *)m[1], (const T *)m[2], Ms, alpha, p, cx, cy, cz, g, zl, sl);
  UM_HIP(hipGetLastError());
}
template <typename T>
void add_uniaxial2(void *const B[3], const void *const m[3], Par<T> Ms, Par<T> K1, Par<T> K2, const Par<T> u[3],
                   int N, hipStream_t s) {
  UM_LAUNCH(k_adduniaxial2<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, (T *)B[0], (T *)B[1],
                     (T *)B[2], (const T *)m[0], (const T *)m[1], (const T *)m[2], Ms, K1, K2, u[0], u[1], u[2], N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void add_cubic2(void *const B[3], const void *const m[3], Par<T> Ms, Par<T> K1, Par<T> K2, Par<T> K3,
                const Par<T> c1[3], const Par<T> c2[3], int N, hipStream_t s) {
  CubicArgs<T> a{};
  for (int c = 0; c < 3; c++) {
    a.B[c] = (T *)B[c];
    a.m[c] = (const T *)m[c];
    a.c1[c] = c1[c];
    a.c2[c] = c2[c];
  }
  a.Ms = Ms; a.K1 = K1; a.K2 = K2; a.K3 = K3; a.N = N;
  UM_LAUNCH(k_addcubic2<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, a);
  UM_HIP(hipGetLastError());
}
template <typename T>
void region_addv(void *const d[3], const void *const l[3], const uint8_t *regions, int N, hipStream_t s) {
  UM_LAUNCH(k_regionaddv<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, (T *)d[0], (T *)d[1],
                     (T *)d[2], (const T *)l[0], (const T *)l[1], (const T *)l[2], regions, N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void zeromask(void *dst, const void *lut, const uint8_t *regions, int N, hipStream_t s) {
  UM_LAUNCH(k_zeromask<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, (T *)dst, (const T *)lut,
                     regions, N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void lltorque(void *const t[3], const void *const m[3], const void *const B[3], Par<T> alpha, int N,
              hipStream_t s) {
  UM_LAUNCH(k_lltorque<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, (T *)t[0], (T *)t[1],
                     (T *)t[2], (const T *)m[0], (const T *)m[1], (const T *)m[2], (const T *)B[0],
                     (const T *)B[1], (const T *)B[2], alpha, N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void llnoprecess(void *const t[3], const void *const m[3], const void *const B[3], int N, hipStream_t s) {
  UM_LAUNCH(k_llnoprecess<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, (T *)t[0], (T *)t[1],
                     (T *)t[2], (const T *)m[0], (const T *)m[1], (const T *)m[2], (const T *)B[0],
                     (const T *)B[1], (const T *)B[2], N);
  UM_HIP(hipGetLastError());
}
template <typename T>
void thermal_field(void *const B[3], uint64_t seed, uint64_t step, uint32_t stream, T k, Par<T> Ms, Par<T> temp,
                   Par<T> alpha, int N, hipStream_t s) {
  UM_LAUNCH(k_thermal<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, (T *)B[0], (T *)B[1],
                     (T *)B[2], seed, step, stream, k, Ms, temp, alpha, N);
  UM_HIP(hipGetLastError());
}

#define UM_INST(T)                                                                                            \
  template void add_exchange<T>(void *const[3], const void *const[3], Par<T>, const T *, const uint8_t *, T, T, \
                                T, Mesh, hipStream_t);                                                        \
  template void add_dmi<T>(void *const[3], const void *const[3], Par<T>, const T *, const T *, const uint8_t *, \
                           T, T, T, Mesh, int, hipStream_t);                                                  \
  template void add_dmi_bulk<T>(void *const[3], const void *const[3], Par<T>, const T *, const T *,           \
                                const uint8_t *, T, T, T, Mesh, int, hipStream_t);                             \
  template void add_stt<T>(void *const[3], const void *const[3], Par<T>, Par<T>, const SttPar<T> &, T, T, T,   \
                           Mesh, int, int, hipStream_t);                                                      \
  template void add_uniaxial2<T>(void *const[3], const void *const[3], Par<T>, Par<T>, Par<T>, const Par<T>[3], \
                                 int, hipStream_t);                                                           \
  template void add_cubic2<T>(void *const[3], const void *const[3], Par<T>, Par<T>, Par<T>, Par<T>,            \
                              const Par<T>[3], const Par<T>[3], int, hipStream_t);                            \
  template void region_addv<T>(void *const[3], const void *const[3], const uint8_t *, int, hipStream_t);       \
  template void zeromask<T>(void *, const void *, const uint8_t *, int, hipStream_t);                         \
  template void lltorque<T>(void *const[3], const void *const[3], const void *const[3], Par<T>, int,          \
                            hipStream_t);                                                                     \
  template void llnoprecess<T>(void *const[3], const void *const[3], const void *const[3], int, hipStream_t); \
  template void thermal_field<T>(void *const[3], uint64_t, uint64_t, uint32_t, T, Par<T>, Par<T>, Par<T>, int, \
                                 hipStream_t);
UM_INST(float)
UM_INST(double)

}  // namespace um
