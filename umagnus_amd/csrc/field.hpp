// Per-cell effective-field terms and the LL torque, shared by the unfused
// entry points (one term per launch, as the reference's opencl package) and
// the fused field+torque kernel used by the simulation handle.
#pragma once
#include "common.hpp"

namespace um {

template <typename T> struct V3 { T x, y, z; };
template <typename T> __device__ __forceinline__ V3<T> v3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <typename T> __device__ __forceinline__ T vdot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> __device__ __forceinline__ V3<T> vcross(V3<T> a, V3<T> b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> __device__ __forceinline__ bool is0(V3<T> m) { return vdot(m, m) == T(0); }
template <typename T> __device__ __forceinline__ V3<T> load3(const T *const *p, long i) {
  return V3<T>{p[0][i], p[1][i], p[2][i]};
}
// normalized(), kernels_src/clh/float3.clh:18-21
template <typename T> __device__ __forceinline__ V3<T> vnormalized(V3<T> a) {
  const T l = sqrt(vdot(a, a));
  const T s = (l != T(0)) ? T(1) / l : T(0);
  return V3<T>{s * a.x, s * a.y, s * a.z};
}
// inv_Msat, kernels_src/clh/amul.clh:18-25
template <typename T> __device__ __forceinline__ T inv_msat(const Par<T> &Ms, long i) {
  const T ms = Ms.at(i);
  return ms == T(0) ? T(0) : T(1) / ms;
}
__device__ __forceinline__ int symidx(int i, int j) {
  return (j <= i) ? (i * (i + 1) / 2 + j) : (j * (j + 1) / 2 + i);
}
__device__ __forceinline__ int modp(int n, int M) { return ((n % M) + M) % M; }

struct Mesh {
  int Nx, Ny, Nz, pbc;
  __device__ __forceinline__ long idx(int ix, int iy, int iz) const {
    return ((long)iz * Ny + iy) * Nx + ix;
  }
  __device__ __forceinline__ int hx(int i) const { return (pbc & 1) ? modp(i, Nx) : min(i, Nx - 1); }
  __device__ __forceinline__ int lx(int i) const { return (pbc & 1) ? modp(i, Nx) : max(i, 0); }
  __device__ __forceinline__ int hy(int i) const { return (pbc & 2) ? modp(i, Ny) : min(i, Ny - 1); }
  __device__ __forceinline__ int ly(int i) const { return (pbc & 2) ? modp(i, Ny) : max(i, 0); }
  __device__ __forceinline__ int hz(int i) const { return (pbc & 4) ? modp(i, Nz) : min(i, Nz - 1); }
  __device__ __forceinline__ int lz(int i) const { return (pbc & 4) ? modp(i, Nz) : max(i, 0); }
};

// ---- exchange, kernels_src/cl/exchange.cl:5-87 -----------------------------
template <typename T>
__device__ __forceinline__ V3<T> exchange_cell(const Mesh &g, int ix, int iy, int iz, long I, V3<T> m0,
                                               const T *const *m, const uint8_t *regions,
                                               const T *aLUT, T wx, T wy, T wz) {
  const int r0 = regions ? regions[I] : 0;
  V3<T> B = v3(T(0), T(0), T(0));
  auto term = [&](long i_, T w) {
    V3<T> mn = load3(m, i_);
    if (is0(mn)) mn = m0;
    const T a = aLUT[symidx(r0, regions ? regions[i_] : 0)];
    const T s = w * a;
    B.x += s * (mn.x - m0.x);
    B.y += s * (mn.y - m0.y);
    B.z += s * (mn.z - m0.z);
  };
  term(g.idx(g.lx(ix - 1), iy, iz), wx);
  term(g.idx(g.hx(ix + 1), iy, iz), wx);
  term(g.idx(ix, g.ly(iy - 1), iz), wy);
  term(g.idx(ix, g.hy(iy + 1), iz), wy);
  if (g.Nz != 1) {
    term(g.idx(ix, iy, g.lz(iz - 1)), wz);
    term(g.idx(ix, iy, g.hz(iz + 1)), wz);
  }
  return B;
}

// ---- exchange + interfacial DMI, kernels_src/cl/dmi.cl:8-148 ----------------
template <typename T>
__device__ __forceinline__ V3<T> dmi_cell(const Mesh &g, int ix, int iy, int iz, long I, V3<T> m0,
                                          const T *const *m, const uint8_t *regions, const T *aLUT,
                                          const T *dLUT, T cx, T cy, T cz, int openbc) {
  const int r0 = regions ? regions[I] : 0;
  V3<T> h = v3(T(0), T(0), T(0));
  // in-plane neighbour along axis 0 (x) or 1 (y); sgn = -1 lower, +1 upper
  auto inplane = [&](long i_, bool inside, T c, int axis, T sgn) {
    V3<T> mn = v3(T(0), T(0), T(0));
    if (inside) mn = load3(m, i_);
    const int rn = is0(mn) ? r0 : (regions ? regions[i_] : 0);
    const T A = aLUT[symidx(r0, rn)];
    const T D = dLUT[symidx(r0, rn)];
    if (!is0(mn) || !openbc) {
      if (is0(mn)) {
        const T f = sgn * c * (T(0.5) * D / A);
        if (axis == 0) {
          mn.x = m0.x - (f * m0.z);
          mn.y = m0.y;
          mn.z = m0.z + (f * m0.x);
        } else {
          mn.x = m0.x;
          mn.y = m0.y - (f * m0.z);
          mn.z = m0.z + (f * m0.y);
        }
      }
      const T e = (T(2.0) * A / (c * c));
      h.x += e * (mn.x - m0.x);
      h.y += e * (mn.y - m0.y);
      h.z += e * (mn.z - m0.z);
      const T dc = D / c;
      if (axis == 0) {
        h.x += dc * (sgn * mn.z);
        h.z -= dc * (sgn * mn.x);
      } else {
        h.y += dc * (sgn * mn.z);
        h.z -= dc * (sgn * mn.y);
      }
    }
  };
  inplane(g.idx(g.lx(ix - 1), iy, iz), (ix - 1 >= 0) || (g.pbc & 1), cx, 0, T(-1));
  inplane(g.idx(g.hx(ix + 1), iy, iz), (ix + 1 < g.Nx) || (g.pbc & 1), cx, 0, T(1));
  inplane(g.idx(ix, g.ly(iy - 1), iz), (iy - 1 >= 0) || (g.pbc & 2), cy, 1, T(-1));
  inplane(g.idx(ix, g.hy(iy + 1), iz), (iy + 1 < g.Ny) || (g.pbc & 2), cy, 1, T(1));
  if (g.Nz != 1) {
    auto vert = [&](long i_) {
      V3<T> mn = load3(m, i_);
      if (is0(mn)) mn = m0;
      const T A = aLUT[symidx(r0, regions ? regions[i_] : 0)];
      const T e = (T(2.0) * A / (cz * cz));
      h.x += e * (mn.x - m0.x);
      h.y += e * (mn.y - m0.y);
      h.z += e * (mn.z - m0.z);
    };
    vert(g.idx(ix, iy, g.lz(iz - 1)));
    vert(g.idx(ix, iy, g.hz(iz + 1)));
  }
  return h;
}

// ---- uniaxial anisotropy, kernels_src/cl/uniaxialanisotropy2.cl:17-35 -------
template <typename T>
__device__ __forceinline__ V3<T> uniaxial_cell(long i, V3<T> m, const Par<T> &Ms, const Par<T> &K1_,
                                               const Par<T> &K2_, const Par<T> *u_) {
  const V3<T> u = vnormalized(v3(u_[0].at(i), u_[1].at(i), u_[2].at(i)));
  const T invMs = inv_msat(Ms, i);
  const T K1 = K1_.at(i) * invMs;
  const T K2 = K2_.at(i) * invMs;
  const T mu = vdot(m, u);
  const T f = T(2.0) * K1 * mu;
  const T q = T(4.0) * K2 * (mu * mu * mu);
  return v3(f * u.x + q * u.x, f * u.y + q * u.y, f * u.z + q * u.z);
}

// ---- cubic anisotropy, kernels_src/cl/cubicanisotropy2.cl:28-62 -------------
template <typename T>
__device__ __forceinline__ V3<T> cubic_cell(long i, V3<T> m, const Par<T> &Ms, const Par<T> &k1_,
                                            const Par<T> &k2_, const Par<T> &k3_, const Par<T> *c1_,
                                            const Par<T> *c2_) {
  const T invMs = inv_msat(Ms, i);
  const T k1 = k1_.at(i) * invMs, k2 = k2_.at(i) * invMs, k3 = k3_.at(i) * invMs;
  const V3<T> u1 = vnormalized(v3(c1_[0].at(i), c1_[1].at(i), c1_[2].at(i)));
  const V3<T> u2 = vnormalized(v3(c2_[0].at(i), c2_[1].at(i), c2_[2].at(i)));
  const V3<T> u3 = vcross(u1, u2);
  const T a = vdot(u1, m), b = vdot(u2, m), c = vdot(u3, m);
  const T a2 = a * a, b2 = b * b, c2 = c * c;
  const T a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
  const T f1 = (b2 + c2) * a, f2 = (a2 + c2) * b, f3 = (a2 + b2) * c;
  const T g1 = (b2 * c2) * a, g2 = (a2 * c2) * b, g3 = (a2 * b2) * c;
  const T h1 = (b4 + c4) * (a * a * a), h2 = (a4 + c4) * (b * b * b), h3 = (a4 + b4) * (c * c * c);
  V3<T> B;
  B.x = T(-2.0) * k1 * (f1 * u1.x + f2 * u2.x + f3 * u3.x) - T(2.0) * k2 * (g1 * u1.x + g2 * u2.x + g3 * u3.x) -
        T(4.0) * k3 * (h1 * u1.x + h2 * u2.x + h3 * u3.x);
  B.y = T(-2.0) * k1 * (f1 * u1.y + f2 * u2.y + f3 * u3.y) - T(2.0) * k2 * (g1 * u1.y + g2 * u2.y + g3 * u3.y) -
        T(4.0) * k3 * (h1 * u1.y + h2 * u2.y + h3 * u3.y);
  B.z = T(-2.0) * k1 * (f1 * u1.z + f2 * u2.z + f3 * u3.z) - T(2.0) * k2 * (g1 * u1.z + g2 * u2.z + g3 * u3.z) -
        T(4.0) * k3 * (h1 * u1.z + h2 * u2.z + h3 * u3.z);
  return B;
}

// ---- Philox-4x32-10 and the thermal normals -------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
// counter (cell, step lo, step hi, stream), key = seed; Box-Muller in FP64
__device__ __forceinline__ void thermal_normals(long cell, uint64_t seed, uint64_t step, uint32_t stream,
                                                double &n0, double &n1, double &n2) {
  uint32_t c[4] = {(uint32_t)cell, (uint32_t)step, (uint32_t)(step >> 32), stream};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double s = 2.3283064365386963e-10;
  const double u0 = ((double)c[0] + 0.5) * s, u1 = ((double)c[1] + 0.5) * s;
  const double u2 = ((double)c[2] + 0.5) * s, u3 = ((double)c[3] + 0.5) * s;
  const double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
  double sn, cs;
  sincospi(2.0 * u1, &sn, &cs);
  n0 = r0 * cs;
  n1 = r0 * sn;
  n2 = r1 * cospi(2.0 * u3);
}
// settemperature2.cl:13-17 with the normals computed in place
template <typename T>
__device__ __forceinline__ V3<T> thermal_cell(long i, uint64_t seed, uint64_t step, uint32_t stream, T k,
                                              const Par<T> &Ms, const Par<T> &temp, const Par<T> &alpha) {
  double n0, n1, n2;
  thermal_normals(i, seed, step, stream, n0, n1, n2);
  const T invMs = inv_msat(Ms, i);
  const T f = sqrt(k * alpha.at(i) * temp.at(i) * invMs);
  return v3((T)n0 * f, (T)n1 * f, (T)n2 * f);
}

// ---- LL torque, kernels_src/cl/lltorque2.cl:11-24 ---------------------------------
template <typename T>
__device__ __forceinline__ V3<T> lltorque_cell(V3<T> m, V3<T> H, T alpha) {
  const V3<T> mxH = vcross(m, H);
  const T gilb = T(-1.0) / (T(1.0) + alpha * alpha);
  const V3<T> mm = vcross(m, mxH);
  return v3(gilb * (mxH.x + alpha * mm.x), gilb * (mxH.y + alpha * mm.y), gilb * (mxH.z + alpha * mm.z));
}

// ---- the fused field + torque pass ----------------------------------------------
template <typename T>
struct FieldArgs {
  Mesh g;
  T wx, wy, wz;          // exchange weights float32(2/c^2) (opencl/exchange.go:17-20)
  T cx, cy, cz;          // cell size (DMI path, opencl/dmi.go:208)
  const T *m[3];
  const T *bd[3];        // demag field or nullptr (zero)
  T *tau[3];             // torque out (nullptr = skip)
  T *beff[3];            // effective field out (nullptr = skip)
  const uint8_t *regions;
  Par<T> Ms;
  const T *aLUT, *dLUT;
  int exch_mode;         // 0 none, 1 exchange, 2 exchange + interfacial DMI
  int openbc;
  int has_uniax;
  Par<T> Ku1, Ku2, u[3];
  int has_cubic;
  Par<T> Kc1, Kc2, Kc3, c1[3], c2[3];
  const T *bextLUT[3];   // region LUT or nullptr
  int has_bext;
  Par<T> bext[3];        // uniform / per-cell part
  int has_therm;
  Par<T> temp;
  T kth;                 // 2 kB / (gamma V dt)
  uint64_t seed, step;
  uint32_t stream;
  Par<T> alpha;
  int precess;
  const T *frozenLUT;    // FrozenSpins per region or nullptr
};

template <typename T> void launch_field_torque(const FieldArgs<T> &a, hipStream_t s);

}  // namespace um
