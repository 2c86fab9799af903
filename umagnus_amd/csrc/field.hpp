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

// ---- exchange + bulk DMI, kernels_src/cl/dmibulk.cl:31-235 -------------------
// A neighbour outside the mesh (or with m = 0) is extrapolated from the bulk
// DMI boundary condition unless OpenBC; z neighbours only in 3D.
template <typename T>
__device__ __forceinline__ V3<T> dmibulk_cell(const Mesh &g, int ix, int iy, int iz, long I, V3<T> m0,
                                              const T *const *m, const uint8_t *regions, const T *aLUT,
                                              const T *dLUT, T cx, T cy, T cz, int openbc) {
  const int r0 = regions ? regions[I] : 0;
  V3<T> h = v3(T(0), T(0), T(0));
  auto nb = [&](long i_, bool inside, int axis, T sgn, T c) {
    V3<T> n = v3(T(0), T(0), T(0));
    if (inside) n = load3(m, i_);
    const bool present = !is0(n);
    const int rn = present ? (regions ? regions[i_] : 0) : r0;
    const T A = aLUT[symidx(r0, rn)];
    const T D = dLUT[symidx(r0, rn)];
    const T D_2A = D / (T(2.0) * A);
    if (!present && openbc) return;
    if (!present) {
      const T f = sgn * c * D_2A;
      if (axis == 0) { n.x = m0.x; n.y = m0.y - (f * m0.z); n.z = m0.z + (f * m0.y); }
      else if (axis == 1) { n.x = m0.x + (f * m0.z); n.y = m0.y; n.z = m0.z - (f * m0.x); }
      else { n.x = m0.x - (f * m0.y); n.y = m0.y + (f * m0.x); n.z = m0.z; }
    }
    const T e = (T(2.0) * A / (c * c));
    h.x += e * (n.x - m0.x);
    h.y += e * (n.y - m0.y);
    h.z += e * (n.z - m0.z);
    const T dc = D / c;
    if (axis == 0) { h.y += dc * (sgn * n.z); h.z -= dc * (sgn * n.y); }
    else if (axis == 1) { h.x -= dc * (sgn * n.z); h.z += dc * (sgn * n.x); }
    else { h.x += dc * (sgn * n.y); h.y -= dc * (sgn * n.x); }
  };
  nb(g.idx(g.lx(ix - 1), iy, iz), (ix - 1 >= 0) || (g.pbc & 1), 0, T(-1), cx);
  nb(g.idx(g.hx(ix + 1), iy, iz), (ix + 1 < g.Nx) || (g.pbc & 1), 0, T(1), cx);
  nb(g.idx(ix, g.ly(iy - 1), iz), (iy - 1 >= 0) || (g.pbc & 2), 1, T(-1), cy);
  nb(g.idx(ix, g.hy(iy + 1), iz), (iy + 1 < g.Ny) || (g.pbc & 2), 1, T(1), cy);
  if (g.Nz != 1) {
    nb(g.idx(ix, iy, g.lz(iz - 1)), (iz - 1 >= 0) || (g.pbc & 4), 2, T(-1), cz);
    nb(g.idx(ix, iy, g.hz(iz + 1)), (iz + 1 < g.Nz) || (g.pbc & 4), 2, T(1), cz);
  }
  return h;
}

// ---- spin-transfer torques -----------------------------------------------------
constexpr double UM_MUB = 9.2740091523E-24;   // Bohr magneton, J/T (constants.clh:7)
constexpr double UM_QE = 1.60217646E-19;      // electron charge, C (:6)
constexpr double UM_GAMMA0 = 1.7595e11;       // rad/Ts (:8)
constexpr double UM_HBAR = 1.05457173E-34;    // J s (:9)

template <typename T>
struct SttPar {
  Par<T> J[3], xi, pol, lambda, epsp, thick, fixp[3];
  T mesh_thick, flp;   // mesh thickness Nz*cz; +1 fixed layer on top, -1 below
};

// Zhang-Li torque (zhangli2.cl:17-57): (u.grad) m from central differences over
// clamped (or periodic) neighbours, in T
template <typename T>
__device__ __forceinline__ V3<T> zhangli_cell(const Mesh &g, int ix, int iy, int iz, long i, V3<T> m0,
                                              const T *const *m, const Par<T> &Ms, const Par<T> &alpha_,
                                              const SttPar<T> &p, T cx, T cy, T cz) {
  const T alpha = alpha_.at(i), xi = p.xi.at(i), pol = p.pol.at(i);
  const T invMs = inv_msat(Ms, i);
  const T b = invMs * T(UM_MUB / (2 * UM_QE * UM_GAMMA0)) / (T(1.0) + xi * xi);
  const V3<T> J = v3(pol * p.J[0].at(i), pol * p.J[1].at(i), pol * p.J[2].at(i));
  V3<T> hs = v3(T(0), T(0), T(0));
  auto add = [&](long a_, long c_, T f) {
    const V3<T> ma = load3(m, a_), mc = load3(m, c_);
    hs.x += f * (ma.x - mc.x);
    hs.y += f * (ma.y - mc.y);
    hs.z += f * (ma.z - mc.z);
  };
  if (J.x != T(0)) add(g.idx(g.hx(ix + 1), iy, iz), g.idx(g.lx(ix - 1), iy, iz), (b / cx) * J.x);
  if (J.y != T(0)) add(g.idx(ix, g.hy(iy + 1), iz), g.idx(ix, g.ly(iy - 1), iz), (b / cy) * J.y);
  if (J.z != T(0)) add(g.idx(ix, iy, g.hz(iz + 1)), g.idx(ix, iy, g.lz(iz - 1)), (b / cz) * J.z);
  const V3<T> mh = vcross(m0, hs), mmh = vcross(m0, mh);
  const T gg = T(-1.0) / (T(1.0) + alpha * alpha);
  const T c1 = T(1.0) + xi * alpha, c2 = xi - alpha;
  return v3(gg * (c1 * mmh.x + c2 * mh.x), gg * (c1 * mmh.y + c2 * mh.y), gg * (c1 * mmh.z + c2 * mh.z));
}

// Slonczewski torque (slonczewski2.cl:22-60); no torque where J or Ms is zero
template <typename T>
__device__ __forceinline__ V3<T> slonczewski_cell(long i, V3<T> m, const Par<T> &Ms_, const Par<T> &alpha_,
                                                  const SttPar<T> &p) {
  const T J = p.J[2].at(i);
  const V3<T> pp = vnormalized(v3(p.fixp[0].at(i), p.fixp[1].at(i), p.fixp[2].at(i)));
  const T Ms = Ms_.at(i);
  const T alpha = alpha_.at(i), pol = p.pol.at(i), lambda = p.lambda.at(i), epsp = p.epsp.at(i);
  T thickness = p.thick.at(i);
  if (thickness == T(0.0)) thickness = p.mesh_thick;
  thickness *= p.flp;
  if (J == T(0.0) || Ms == T(0.0)) return v3(T(0), T(0), T(0));
  const T beta = T(UM_HBAR / UM_QE) * (J / (thickness * Ms));
  const T l2 = lambda * lambda;
  const T eps = pol * l2 / ((l2 + T(1.0)) + (l2 - T(1.0)) * vdot(pp, m));
  const T A = beta * eps, B = beta * epsp;
  const T gilb = T(1.0) / (T(1.0) + alpha * alpha);
  const T f1 = gilb * (A + alpha * B), f2 = gilb * (B - alpha * A);
  const V3<T> pxm = vcross(pp, m), mpm = vcross(m, pxm);
  return v3(f1 * mpm.x + f2 * pxm.x, f1 * mpm.y + f2 * pxm.y, f1 * mpm.z + f2 * pxm.z);
}

// ---- uniaxial anisotropy, kernels_src/cl/uniaxialanisotropy2.cl:17-35 -------
// the field from the unit axis u and K1/Ms, K2/Ms
template <typename T>
__device__ __forceinline__ V3<T> uniaxial_apply(V3<T> m, V3<T> u, T K1, T K2) {
  const T mu = vdot(m, u);
  const T f = T(2.0) * K1 * mu;
  const T q = T(4.0) * K2 * (mu * mu * mu);
  return v3(f * u.x + q * u.x, f * u.y + q * u.y, f * u.z + q * u.z);
}
template <typename T>
__device__ __forceinline__ V3<T> uniaxial_cell(long i, V3<T> m, const Par<T> &Ms, const Par<T> &K1_,
                                               const Par<T> &K2_, const Par<T> *u_) {
  const V3<T> u = vnormalized(v3(u_[0].at(i), u_[1].at(i), u_[2].at(i)));
  const T invMs = inv_msat(Ms, i);
  const T K1 = K1_.at(i) * invMs;
  const T K2 = K2_.at(i) * invMs;
  return uniaxial_apply(m, u, K1, K2);
}

// ---- cubic anisotropy, kernels_src/cl/cubicanisotropy2.cl:28-62 -------------
template <typename T>
__device__ __forceinline__ V3<T> cubic_apply(V3<T> m, V3<T> u1, V3<T> u2, V3<T> u3, T k1, T k2, T k3);
template <typename T>
__device__ __forceinline__ V3<T> cubic_cell(long i, V3<T> m, const Par<T> &Ms, const Par<T> &k1_,
                                            const Par<T> &k2_, const Par<T> &k3_, const Par<T> *c1_,
                                            const Par<T> *c2_) {
  const T invMs = inv_msat(Ms, i);
  const T k1 = k1_.at(i) * invMs, k2 = k2_.at(i) * invMs, k3 = k3_.at(i) * invMs;
  const V3<T> u1 = vnormalized(v3(c1_[0].at(i), c1_[1].at(i), c1_[2].at(i)));
  const V3<T> u2 = vnormalized(v3(c2_[0].at(i), c2_[1].at(i), c2_[2].at(i)));
  const V3<T> u3 = vcross(u1, u2);
  return cubic_apply(m, u1, u2, u3, k1, k2, k3);
}
// the cubic field from the unit axes and K1/Ms, K2/Ms, K3/Ms
template <typename T>
__device__ __forceinline__ V3<T> cubic_apply(V3<T> m, V3<T> u1, V3<T> u2, V3<T> u3, T k1, T k2, T k3) {
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
__device__ __forceinline__ V3<T> lltorque_g(V3<T> m, V3<T> H, T alpha, T gilb) {
  const V3<T> mxH = vcross(m, H);
  const V3<T> mm = vcross(m, mxH);
  return v3(gilb * (mxH.x + alpha * mm.x), gilb * (mxH.y + alpha * mm.y), gilb * (mxH.z + alpha * mm.z));
}
template <typename T> __device__ __forceinline__ T gilbert(T alpha) { return T(-1.0) / (T(1.0) + alpha * alpha); }
template <typename T>
__device__ __forceinline__ V3<T> lltorque_cell(V3<T> m, V3<T> H, T alpha) {
  return lltorque_g(m, H, alpha, gilbert(alpha));
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
  int exch_mode;         // 0 none, 1 exchange, 2 exchange + interfacial DMI, 3 exchange + bulk DMI (dLUT = Dbulk)
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
  const unsigned long long *step_dev;   // when set, the Philox step counter is read here (graph replays)
  uint32_t stream;
  Par<T> alpha;
  int precess;
  const T *frozenLUT;    // FrozenSpins per region or nullptr
  int has_zl, has_sl;    // Zhang-Li / Slonczewski torque on (one-cell kernel only)
  SttPar<T> stt;
  const T *badd[3];      // extra field added last (AddCustomField, engine/customfield.go:68) or nullptr
  const T *tadd[3];      // extra torque added before FreezeSpins (AddCustomTorques, engine/torque.go:115) or nullptr
};

// ---- four consecutive x cells per thread ----------------------------------------
// The same per-cell arithmetic as the one-cell kernels, with the stencil's m
// values loaded as 16-byte vectors (the x neighbours of the inner cells come
// from the thread's own registers) and, without a region map, the exchange /
// DMI lookup tables read once.  Needs Nx % 4 == 0.
template <typename T>
__device__ __forceinline__ void ld4(const T *p, T (&v)[4]) {
  if constexpr (sizeof(T) == 4) {
    const float4 f = *(const float4 *)p;
    v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
  } else {
    const double2 f0 = ((const double2 *)p)[0], f1 = ((const double2 *)p)[1];
    v[0] = f0.x; v[1] = f0.y; v[2] = f1.x; v[3] = f1.y;
  }
}
template <typename T>
__device__ __forceinline__ void st4(T *p, const T (&v)[4]) {
  if constexpr (sizeof(T) == 4) {
    *(float4 *)p = float4{v[0], v[1], v[2], v[3]};
  } else {
    ((double2 *)p)[0] = double2{v[0], v[1]};
    ((double2 *)p)[1] = double2{v[2], v[3]};
  }
}
__device__ __forceinline__ void ld4u8(const uint8_t *p, int (&r)[4]) {
  const uint32_t u = *(const uint32_t *)p;
  r[0] = u & 255; r[1] = (u >> 8) & 255; r[2] = (u >> 16) & 255; r[3] = u >> 24;
}

// Stencil weights of a region-free group, formed once per thread from the one
// uniform A and D with the operations exchange_nb / dmi_nb apply per neighbour.
template <typename T>
struct NbPre {
  T s[3];    // exchange: w_axis * A
  T e[3];    // DMI path: 2 A / c_axis^2
  T dc[2];   // D / c_x, D / c_y
  T fh;      // 0.5 D / A (boundary extrapolation)
};
template <typename T>
__device__ __forceinline__ NbPre<T> nb_pre(T A, T D, T wx, T wy, T wz, T cx, T cy, T cz) {
  NbPre<T> p;
  p.s[0] = wx * A; p.s[1] = wy * A; p.s[2] = wz * A;
  p.e[0] = (T(2.0) * A / (cx * cx)); p.e[1] = (T(2.0) * A / (cy * cy)); p.e[2] = (T(2.0) * A / (cz * cz));
  p.dc[0] = D / cx; p.dc[1] = D / cy;
  p.fh = (T(0.5) * D / A);
  return p;
}

// exchange from gathered neighbours (order x-1, x+1, y-1, y+1, z-1, z+1), exchange.cl:5-87
template <typename T, bool REG, bool PRE = false>
__device__ __forceinline__ V3<T> exchange_nb(V3<T> m0, const V3<T> (&mn)[6], int nnb, int r0, const int (&rn)[6],
                                             const T *aLUT, T A0, T wx, T wy, T wz, const NbPre<T> &p) {
  V3<T> B = v3(T(0), T(0), T(0));
#pragma unroll
  for (int j = 0; j < 6; j++) {
    if (j >= nnb) break;
    V3<T> n = mn[j];
    if (is0(n)) n = m0;
    const T aj = REG ? aLUT[symidx(r0, rn[j])] : A0;
    const T w = j < 2 ? wx : j < 4 ? wy : wz;
    const T s = PRE ? p.s[j >> 1] : w * aj;
    B.x += s * (n.x - m0.x);
    B.y += s * (n.y - m0.y);
    B.z += s * (n.z - m0.z);
  }
  return B;
}

// exchange + interfacial DMI from gathered neighbours, dmi.cl:8-148; an in-plane
// neighbour outside the mesh (inside[j] false) is passed as zero
template <typename T, bool REG, bool PRE = false>
__device__ __forceinline__ V3<T> dmi_nb(V3<T> m0, const V3<T> (&mn)[6], int nnb, int r0, const int (&rn)[6],
                                        const T *aLUT, const T *dLUT, T A0, T D0, T cx, T cy, T cz, int openbc,
                                        const NbPre<T> &p) {
  V3<T> h = v3(T(0), T(0), T(0));
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int axis = j >> 1;
    const T sgn = (j & 1) ? T(1) : T(-1);
    const T c = axis == 0 ? cx : cy;
    V3<T> n = mn[j];
    const int rj = is0(n) ? r0 : rn[j];
    const T A = REG ? aLUT[symidx(r0, rj)] : A0;
    const T D = REG ? dLUT[symidx(r0, rj)] : D0;
    if (!is0(n) || !openbc) {
      if (is0(n)) {
        const T f = sgn * c * (PRE ? p.fh : (T(0.5) * D / A));
        if (axis == 0) {
          n.x = m0.x - (f * m0.z);
          n.y = m0.y;
          n.z = m0.z + (f * m0.x);
        } else {
          n.x = m0.x;
          n.y = m0.y - (f * m0.z);
          n.z = m0.z + (f * m0.y);
        }
      }
      const T e = PRE ? p.e[axis] : (T(2.0) * A / (c * c));
      h.x += e * (n.x - m0.x);
      h.y += e * (n.y - m0.y);
      h.z += e * (n.z - m0.z);
      const T dc = PRE ? p.dc[axis] : D / c;
      if (axis == 0) {
        h.x += dc * (sgn * n.z);
        h.z -= dc * (sgn * n.x);
      } else {
        h.y += dc * (sgn * n.z);
        h.z -= dc * (sgn * n.y);
      }
    }
  }
  if (nnb == 6) {
#pragma unroll
    for (int j = 4; j < 6; j++) {
      V3<T> n = mn[j];
      if (is0(n)) n = m0;
      const T A = REG ? aLUT[symidx(r0, rn[j])] : A0;
      const T e = PRE ? p.e[2] : (T(2.0) * A / (cz * cz));
      h.x += e * (n.x - m0.x);
      h.y += e * (n.y - m0.y);
      h.z += e * (n.z - m0.z);
    }
  }
  return h;
}

// exchange + bulk DMI from gathered neighbours (dmibulk.cl:31-235): the same
// sums as dmibulk_cell; a neighbour outside the mesh (inside false) is passed
// as zero and extrapolated from the boundary condition unless OpenBC
template <typename T, bool REG>
__device__ __forceinline__ V3<T> dmibulk_nb(V3<T> m0, const V3<T> (&mn)[6], int nnb, int r0, const int (&rn)[6],
                                            const T *aLUT, const T *dLUT, T A0, T D0, T cx, T cy, T cz,
                                            int openbc) {
  V3<T> h = v3(T(0), T(0), T(0));
#pragma unroll
  for (int j = 0; j < 6; j++) {
    if (j >= nnb) break;
    const int axis = j >> 1;
    const T sgn = (j & 1) ? T(1) : T(-1);
    const T c = axis == 0 ? cx : axis == 1 ? cy : cz;
    V3<T> n = mn[j];
    const bool present = !is0(n);
    const int rj = present ? rn[j] : r0;
    const T A = REG ? aLUT[symidx(r0, rj)] : A0;
    const T D = REG ? dLUT[symidx(r0, rj)] : D0;
    const T D_2A = D / (T(2.0) * A);
    if (!present && openbc) continue;
    if (!present) {
      const T f = sgn * c * D_2A;
      if (axis == 0) { n.x = m0.x; n.y = m0.y - (f * m0.z); n.z = m0.z + (f * m0.y); }
      else if (axis == 1) { n.x = m0.x + (f * m0.z); n.y = m0.y; n.z = m0.z - (f * m0.x); }
      else { n.x = m0.x - (f * m0.y); n.y = m0.y + (f * m0.x); n.z = m0.z; }
    }
    const T e = (T(2.0) * A / (c * c));
    h.x += e * (n.x - m0.x);
    h.y += e * (n.y - m0.y);
    h.z += e * (n.z - m0.z);
    const T dc = D / c;
    if (axis == 0) { h.y += dc * (sgn * n.z); h.z -= dc * (sgn * n.y); }
    else if (axis == 1) { h.x -= dc * (sgn * n.z); h.z += dc * (sgn * n.x); }
    else { h.x += dc * (sgn * n.y); h.y -= dc * (sgn * n.x); }
  }
  return h;
}

// Inputs of a four-cell group (ix0..ix0+3, iy, iz): its m and region bytes and
// those of the stencil neighbours.  Loading them is split from the arithmetic
// so a caller can issue the loads early and overlap them with other work.
template <typename T>
struct G4In {
  T mc[3][4], mxl[3], mxr[3], myl[3][4], myh[3][4], mzl[3][4], mzh[3][4];
  int rc[4], rxl, rxr, ryl[4], ryh[4], rzl[4], rzh[4];
  T A0, D0;
};

template <typename T, bool REG, int EX>
__device__ __forceinline__ void g4_load(const FieldArgs<T> &a, int ix0, int iy, int iz, G4In<T> &q) {
  const Mesh &g = a.g;
  const long I0 = g.idx(ix0, iy, iz);
#pragma unroll
  for (int c = 0; c < 3; c++) ld4(a.m[c] + I0, q.mc[c]);
#pragma unroll
  for (int w = 0; w < 4; w++) q.rc[w] = q.ryl[w] = q.ryh[w] = q.rzl[w] = q.rzh[w] = 0;
  q.rxl = q.rxr = 0;
  q.A0 = q.D0 = T(0);
  if (REG) ld4u8(a.regions + I0, q.rc);
  if (EX == 0) return;
  const long iL = g.idx(g.lx(ix0 - 1), iy, iz), iR = g.idx(g.hx(ix0 + 4), iy, iz);
  const long iYl = g.idx(ix0, g.ly(iy - 1), iz), iYh = g.idx(ix0, g.hy(iy + 1), iz);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    q.mxl[c] = a.m[c][iL];
    q.mxr[c] = a.m[c][iR];
    ld4(a.m[c] + iYl, q.myl[c]);
    ld4(a.m[c] + iYh, q.myh[c]);
  }
  if (REG) {
    q.rxl = a.regions[iL];
    q.rxr = a.regions[iR];
    ld4u8(a.regions + iYl, q.ryl);
    ld4u8(a.regions + iYh, q.ryh);
  }
  if (g.Nz != 1) {
    const long iZl = g.idx(ix0, iy, g.lz(iz - 1)), iZh = g.idx(ix0, iy, g.hz(iz + 1));
#pragma unroll
    for (int c = 0; c < 3; c++) {
      ld4(a.m[c] + iZl, q.mzl[c]);
      ld4(a.m[c] + iZh, q.mzh[c]);
    }
    if (REG) {
      ld4u8(a.regions + iZl, q.rzl);
      ld4u8(a.regions + iZh, q.rzh);
    }
  }
  if (!REG) {
    q.A0 = a.aLUT[0];
    if (EX >= 2) q.D0 = a.dLUT[0];
  }
}

// Fast path of the exchange stencil for a region-free group of uniform A.  It
// applies when every m the group reads is clearly non-zero (|x|+|y|+|z| >=
// 1e-15, so no dot product m.m can underflow: is0() is false for all of
// them); then no neighbour is vacuum and exchange_nb reduces to its
// per-neighbour sums without the vacuum selects.  The terms are added in the
// same order with the same weights, so the result is the same to the bit.  A
// neighbour outside a non-periodic mesh is loaded clamped (it is m0 itself)
// and adds zero, as in exchange_nb.  (With interfacial DMI the x edges need
// the boundary extrapolation, every wave of a row holds an edge group, and a
// fast path there ran next to the general one: 145 us against 141 without it
// in the fused X inverse kernel at 512x512x32, so it is exchange-only.)
template <typename T>
__device__ __forceinline__ T g4_nzmin(T m, T x, T y, T z) {
  return fmin(m, fabs(x) + fabs(y) + fabs(z));
}
template <typename T>
__device__ __forceinline__ bool g4_fast_ok(const G4In<T> &q, bool z6) {
  T mn = T(1);
#pragma unroll
  for (int w = 0; w < 4; w++) {
    mn = g4_nzmin(mn, q.mc[0][w], q.mc[1][w], q.mc[2][w]);
    mn = g4_nzmin(mn, q.myl[0][w], q.myl[1][w], q.myl[2][w]);
    mn = g4_nzmin(mn, q.myh[0][w], q.myh[1][w], q.myh[2][w]);
    if (z6) {
      mn = g4_nzmin(mn, q.mzl[0][w], q.mzl[1][w], q.mzl[2][w]);
      mn = g4_nzmin(mn, q.mzh[0][w], q.mzh[1][w], q.mzh[2][w]);
    }
  }
  mn = g4_nzmin(mn, q.mxl[0], q.mxl[1], q.mxl[2]);
  mn = g4_nzmin(mn, q.mxr[0], q.mxr[1], q.mxr[2]);
  return mn >= T(1e-15);
}
// exchange field of cell w of a fast group, before the 1/Ms factor:
// exchange_nb's sums term by term
template <typename T>
__device__ __forceinline__ V3<T> g4_exch_fast(const G4In<T> &q, int w, bool z6, const NbPre<T> &p) {
  const V3<T> m0 = v3(q.mc[0][w], q.mc[1][w], q.mc[2][w]);
  V3<T> nb[6];
  nb[0] = w > 0 ? v3(q.mc[0][w - 1], q.mc[1][w - 1], q.mc[2][w - 1]) : v3(q.mxl[0], q.mxl[1], q.mxl[2]);
  nb[1] = w < 3 ? v3(q.mc[0][w + 1], q.mc[1][w + 1], q.mc[2][w + 1]) : v3(q.mxr[0], q.mxr[1], q.mxr[2]);
  nb[2] = v3(q.myl[0][w], q.myl[1][w], q.myl[2][w]);
  nb[3] = v3(q.myh[0][w], q.myh[1][w], q.myh[2][w]);
  nb[4] = z6 ? v3(q.mzl[0][w], q.mzl[1][w], q.mzl[2][w]) : m0;
  nb[5] = z6 ? v3(q.mzh[0][w], q.mzh[1][w], q.mzh[2][w]) : m0;
  V3<T> h = v3(T(0), T(0), T(0));
#pragma unroll
  for (int j = 0; j < 6; j++) {
    if (j >= 4 && !z6) break;
    const T s = p.s[j >> 1];
    h.x += s * (nb[j].x - m0.x);
    h.y += s * (nb[j].y - m0.y);
    h.z += s * (nb[j].z - m0.z);
  }
  return h;
}

// Effective field and torque of the group from its loaded inputs; bdv: demag
// field of the four cells (zero when absent).  Same term order as k_field_torque.
// EX: 0 none, 1 exchange, 2 + interfacial DMI, 3 + bulk DMI; STT: Zhang-Li and
// Slonczewski torques (per the a.has_zl / a.has_sl flags).
// UNI: Ms, alpha, the anisotropy, B_ext and temperature parameters are all
// uniform (no per-cell arrays, checked by the launcher); their per-cell
// quotients (1/Ms, K/Ms, the normalised axes, the Gilbert factor) are then the
// same for the four cells and are formed once.
template <bool UNI, typename T>
__device__ __forceinline__ Par<T> upar(const Par<T> &p) {
  if constexpr (UNI) return Par<T>{nullptr, p.mul};
  else return p;
}
template <typename T>
__host__ bool field_uniform(const FieldArgs<T> &a) {
  const Par<T> *ps[] = {&a.Ms, &a.alpha, &a.Ku1, &a.Ku2, &a.u[0], &a.u[1], &a.u[2], &a.Kc1, &a.Kc2, &a.Kc3,
                        &a.c1[0], &a.c1[1], &a.c1[2], &a.c2[0], &a.c2[1], &a.c2[2], &a.bext[0], &a.bext[1],
                        &a.bext[2], &a.temp};
  for (const Par<T> *p : ps)
    if (p->arr) return false;
  return true;
}

// HOIST: with UNI, form the stencil weights and the anisotropy / torque
// constants once per thread instead of once per cell (fewer instructions,
// more live registers: the choice of the calling kernel).
// TH: the thermal term may be present (a.has_therm).  Its FP64 Box-Muller
// sets the register peak of the whole kernel; a caller that checked
// has_therm == 0 on the host instantiates TH = false and gets the smaller
// register budget.
template <typename T, bool REG, int EX, bool STT = false, bool UNI = false, bool HOIST = UNI, bool TH = true>
__device__ __forceinline__ void g4_compute(const FieldArgs<T> &a, int ix0, int iy, int iz, const G4In<T> &q,
                                           const T (&bdv)[3][4]) {
  const Mesh &g = a.g;
  const Par<T> Ms = upar<UNI>(a.Ms), alpha = upar<UNI>(a.alpha), Ku1 = upar<UNI>(a.Ku1), Ku2 = upar<UNI>(a.Ku2);
  const Par<T> Kc1 = upar<UNI>(a.Kc1), Kc2 = upar<UNI>(a.Kc2), Kc3 = upar<UNI>(a.Kc3), temp = upar<UNI>(a.temp);
  Par<T> u[3], c1[3], c2[3], bext[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    u[c] = upar<UNI>(a.u[c]);
    c1[c] = upar<UNI>(a.c1[c]);
    c2[c] = upar<UNI>(a.c2[c]);
    bext[c] = upar<UNI>(a.bext[c]);
  }
  // quantities that are the same for the four cells, formed once
  constexpr bool H = UNI && HOIST;
  constexpr bool PRE = !REG && H;   // region-free: one A, one D
  NbPre<T> nbp{};
  if constexpr (PRE) {
    if (EX == 1 || EX == 2) nbp = nb_pre(q.A0, q.D0, a.wx, a.wy, a.wz, a.cx, a.cy, a.cz);
  }
  T invMs_u = T(0), K1_u = T(0), K2_u = T(0), k1_u = T(0), k2_u = T(0), k3_u = T(0), alpha_u = T(0), gilb_u = T(0);
  V3<T> u_u{}, c1_u{}, c2_u{}, c3_u{};
  if constexpr (H) {
    invMs_u = inv_msat(Ms, 0);
    if (a.has_uniax) {
      u_u = vnormalized(v3(u[0].mul, u[1].mul, u[2].mul));
      K1_u = Ku1.mul * invMs_u;
      K2_u = Ku2.mul * invMs_u;
    }
    if (a.has_cubic) {
      k1_u = Kc1.mul * invMs_u; k2_u = Kc2.mul * invMs_u; k3_u = Kc3.mul * invMs_u;
      c1_u = vnormalized(v3(c1[0].mul, c1[1].mul, c1[2].mul));
      c2_u = vnormalized(v3(c2[0].mul, c2[1].mul, c2[2].mul));
      c3_u = vcross(c1_u, c2_u);
    }
    alpha_u = alpha.mul;
    gilb_u = gilbert(alpha_u);
  }
  const long I0 = g.idx(ix0, iy, iz);
  const int nnb = g.Nz != 1 ? 6 : 4;
  const bool inxl = (ix0 - 1 >= 0) || (g.pbc & 1), inxr = (ix0 + 4 < g.Nx) || (g.pbc & 1);
  const bool inyl = (iy - 1 >= 0) || (g.pbc & 2), inyh = (iy + 1 < g.Ny) || (g.pbc & 2);
  const bool inzl = (iz - 1 >= 0) || (g.pbc & 4), inzh = (iz + 1 < g.Nz) || (g.pbc & 4);
  // the fast path (exchange only, region-free, uniform parameters, no STT)
  constexpr bool FASTOK = UNI && !REG && !STT && EX == 1;
  bool fast = false;
  NbPre<T> fpre{};
  if constexpr (FASTOK) {
    fast = g4_fast_ok(q, nnb == 6);
    fpre = PRE ? nbp : nb_pre(q.A0, q.D0, a.wx, a.wy, a.wz, a.cx, a.cy, a.cz);
  }
  T ob[3][4], ot[3][4];
#pragma unroll
  for (int w = 0; w < 4; w++) {
    const long I = I0 + w;
    const V3<T> m0 = v3(q.mc[0][w], q.mc[1][w], q.mc[2][w]);
    V3<T> B = v3(bdv[0][w], bdv[1][w], bdv[2][w]);
    if (FASTOK && fast) {
      const V3<T> h = g4_exch_fast<T>(q, w, nnb == 6, fpre);
      const T invMs = H ? invMs_u : inv_msat(Ms, I);
      B.x += h.x * invMs;
      B.y += h.y * invMs;
      B.z += h.z * invMs;
    } else if (EX != 0 && !is0(m0)) {
      V3<T> mn[6];
      int rn[6];
      mn[0] = w > 0 ? v3(q.mc[0][w - 1], q.mc[1][w - 1], q.mc[2][w - 1]) : v3(q.mxl[0], q.mxl[1], q.mxl[2]);
      rn[0] = w > 0 ? q.rc[w - 1] : q.rxl;
      mn[1] = w < 3 ? v3(q.mc[0][w + 1], q.mc[1][w + 1], q.mc[2][w + 1]) : v3(q.mxr[0], q.mxr[1], q.mxr[2]);
      rn[1] = w < 3 ? q.rc[w + 1] : q.rxr;
      mn[2] = v3(q.myl[0][w], q.myl[1][w], q.myl[2][w]);
      rn[2] = q.ryl[w];
      mn[3] = v3(q.myh[0][w], q.myh[1][w], q.myh[2][w]);
      rn[3] = q.ryh[w];
      if (nnb == 6) {
        mn[4] = v3(q.mzl[0][w], q.mzl[1][w], q.mzl[2][w]);
        rn[4] = q.rzl[w];
        mn[5] = v3(q.mzh[0][w], q.mzh[1][w], q.mzh[2][w]);
        rn[5] = q.rzh[w];
      } else {
        mn[4] = mn[5] = m0;
        rn[4] = rn[5] = q.rc[w];
      }
      V3<T> h;
      if (EX == 1) {
        h = exchange_nb<T, REG, PRE>(m0, mn, nnb, q.rc[w], rn, a.aLUT, q.A0, a.wx, a.wy, a.wz, nbp);
      } else if (EX == 3) {
        const V3<T> zero = v3(T(0), T(0), T(0));
        if (w == 0 && !inxl) mn[0] = zero;
        if (w == 3 && !inxr) mn[1] = zero;
        if (!inyl) mn[2] = zero;
        if (!inyh) mn[3] = zero;
        if (!inzl) mn[4] = zero;
        if (!inzh) mn[5] = zero;
        h = dmibulk_nb<T, REG>(m0, mn, nnb, q.rc[w], rn, a.aLUT, a.dLUT, q.A0, q.D0, a.cx, a.cy, a.cz, a.openbc);
      } else {
        const V3<T> zero = v3(T(0), T(0), T(0));
        if (w == 0 && !inxl) mn[0] = zero;
        if (w == 3 && !inxr) mn[1] = zero;
        if (!inyl) mn[2] = zero;
        if (!inyh) mn[3] = zero;
        h = dmi_nb<T, REG, PRE>(m0, mn, nnb, q.rc[w], rn, a.aLUT, a.dLUT, q.A0, q.D0, a.cx, a.cy, a.cz, a.openbc, nbp);
      }
      const T invMs = H ? invMs_u : inv_msat(Ms, I);
      B.x += h.x * invMs;
      B.y += h.y * invMs;
      B.z += h.z * invMs;
    }
    if (a.has_uniax) {
      const V3<T> h = H ? uniaxial_apply(m0, u_u, K1_u, K2_u) : uniaxial_cell(I, m0, Ms, Ku1, Ku2, u);
      B.x += h.x; B.y += h.y; B.z += h.z;
    }
    if (a.has_cubic) {
      const V3<T> h = H ? cubic_apply(m0, c1_u, c2_u, c3_u, k1_u, k2_u, k3_u)
                          : cubic_cell(I, m0, Ms, Kc1, Kc2, Kc3, c1, c2);
      B.x += h.x; B.y += h.y; B.z += h.z;
    }
    if (a.bextLUT[0]) {
      const int r = q.rc[w];
      B.x += a.bextLUT[0][r];
      B.y += a.bextLUT[1][r];
      B.z += a.bextLUT[2][r];
    }
    if (a.has_bext) {
      B.x += bext[0].at(I);
      B.y += bext[1].at(I);
      B.z += bext[2].at(I);
    }
    if (TH && a.has_therm) {
      const uint64_t stp = a.step_dev ? (uint64_t)*a.step_dev : a.step;
      const V3<T> h = thermal_cell(I, a.seed, stp, a.stream, a.kth, Ms, temp, alpha);
      B.x += h.x; B.y += h.y; B.z += h.z;
    }
    if (a.badd[0]) {
      B.x += a.badd[0][I];
      B.y += a.badd[1][I];
      B.z += a.badd[2][I];
    }
    ob[0][w] = B.x; ob[1][w] = B.y; ob[2][w] = B.z;
    V3<T> t;
    if (a.precess) {
      t = H ? lltorque_g(m0, B, alpha_u, gilb_u) : lltorque_cell(m0, B, alpha.at(I));
    } else {
      const V3<T> c = vcross(m0, vcross(m0, B));
      t = v3(-c.x, -c.y, -c.z);
    }
    if constexpr (STT) {
      // AddSTTorque (engine/torque.go:147-196), before FreezeSpins; the
      // Zhang-Li differences use the clamped / wrapped neighbours as zhangli_cell
      if (a.has_zl) {
        const SttPar<T> &p = a.stt;
        const T alpha = a.alpha.at(I), xi = p.xi.at(I), pol = p.pol.at(I);
        const T invMs = inv_msat(a.Ms, I);
        const T b = invMs * T(UM_MUB / (2 * UM_QE * UM_GAMMA0)) / (T(1.0) + xi * xi);
        const V3<T> J = v3(pol * p.J[0].at(I), pol * p.J[1].at(I), pol * p.J[2].at(I));
        V3<T> hs = v3(T(0), T(0), T(0));
        auto add = [&](V3<T> ma, V3<T> mc, T f) {
          hs.x += f * (ma.x - mc.x);
          hs.y += f * (ma.y - mc.y);
          hs.z += f * (ma.z - mc.z);
        };
        const V3<T> xl = w > 0 ? v3(q.mc[0][w - 1], q.mc[1][w - 1], q.mc[2][w - 1]) : v3(q.mxl[0], q.mxl[1], q.mxl[2]);
        const V3<T> xr = w < 3 ? v3(q.mc[0][w + 1], q.mc[1][w + 1], q.mc[2][w + 1]) : v3(q.mxr[0], q.mxr[1], q.mxr[2]);
        if (J.x != T(0)) add(xr, xl, (b / a.cx) * J.x);
        if (J.y != T(0))
          add(v3(q.myh[0][w], q.myh[1][w], q.myh[2][w]), v3(q.myl[0][w], q.myl[1][w], q.myl[2][w]), (b / a.cy) * J.y);
        if (J.z != T(0) && nnb == 6)
          add(v3(q.mzh[0][w], q.mzh[1][w], q.mzh[2][w]), v3(q.mzl[0][w], q.mzl[1][w], q.mzl[2][w]), (b / a.cz) * J.z);
        const V3<T> mh = vcross(m0, hs), mmh = vcross(m0, mh);
        const T gg = T(-1.0) / (T(1.0) + alpha * alpha);
        const T c1 = T(1.0) + xi * alpha, c2 = xi - alpha;
        t.x += gg * (c1 * mmh.x + c2 * mh.x);
        t.y += gg * (c1 * mmh.y + c2 * mh.y);
        t.z += gg * (c1 * mmh.z + c2 * mh.z);
      }
      if (a.has_sl) {
        const V3<T> z = slonczewski_cell(I, m0, a.Ms, a.alpha, a.stt);
        t.x += z.x; t.y += z.y; t.z += z.z;
      }
    }
    if (a.tadd[0]) {
      t.x += a.tadd[0][I];
      t.y += a.tadd[1][I];
      t.z += a.tadd[2][I];
    }
    if (a.frozenLUT && a.frozenLUT[q.rc[w]] != T(0)) t = v3(T(0), T(0), T(0));
    ot[0][w] = t.x; ot[1][w] = t.y; ot[2][w] = t.z;
  }
  if (a.beff[0]) {
#pragma unroll
    for (int c = 0; c < 3; c++) st4(a.beff[c] + I0, ob[c]);
  }
  if (a.tau[0]) {
#pragma unroll
    for (int c = 0; c < 3; c++) st4(a.tau[c] + I0, ot[c]);
  }
}

template <typename T, bool REG, int EX, bool STT = false, bool UNI = false, bool HOIST = UNI, bool TH = true>
__device__ __forceinline__ void field_group4(const FieldArgs<T> &a, int ix0, int iy, int iz, const T (&bdv)[3][4]) {
  G4In<T> q;
  g4_load<T, REG, (STT && EX == 0) ? 1 : EX>(a, ix0, iy, iz, q);   // STT needs the neighbours
  g4_compute<T, REG, EX, STT, UNI, HOIST, TH>(a, ix0, iy, iz, q, bdv);
}

template <typename T> void launch_field_torque(const FieldArgs<T> &a, hipStream_t s);
// ZeroMask (opencl/zeromask.go): dst = 0 where lut[region] != 0
template <typename T> void zeromask(void *dst, const void *lut, const uint8_t *regions, int N, hipStream_t s);

}  // namespace um
