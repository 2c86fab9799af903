// Kernels of the discrete RK4 adjoint (see adjoint.hpp for the linearisation).
// Per step: k_adj_final transposes the final normalisation and turns the
// cotangent on k4 into stage 4's field cotangent w = gB/Ms (the input of both
// transposed field operators).  Per stage the demag plan convolves Ms*w, and
// k_adj_stage gathers the transposed exchange/DMI stencil, the demag and
// anisotropy terms and the direct torque term, transposes the stage
// normalisation, forms the previous stage's torque cotangent and, from it,
// that stage's w.
#include "adjoint.hpp"

namespace um {

constexpr int ADJ_NT = 256;

long adj_blocks(long N) { return (N + ADJ_NT - 1) / ADJ_NT; }

template <typename T>
__device__ __forceinline__ V3<T> vscale(V3<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <typename T>
__device__ __forceinline__ V3<T> vadd(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }

// cotangent on B of the LL torque: -c (l x m + alpha (l x m) x m)
template <typename T>
__device__ __forceinline__ V3<T> adj_gb(V3<T> m, V3<T> l, T alpha) {
  const T gilb = T(-1.0) / (T(1.0) + alpha * alpha);
  const V3<T> lm = vcross(l, m);
  const V3<T> lmm = vcross(lm, m);
  return v3(gilb * (lm.x + alpha * lmm.x), gilb * (lm.y + alpha * lmm.y), gilb * (lm.z + alpha * lmm.z));
}

template <typename T>
__device__ __forceinline__ void st3(T *const *p, long i, V3<T> v) {
  p[0][i] = v.x;
  p[1][i] = v.y;
  p[2][i] = v.z;
}

// h0 = L x0 and h1 = L x1 at cell I: the exchange (+ interfacial DMI) operator
// of the forward field with the mesh boundary rules of exchange.cl / dmi.cl
// (clamped neighbours contribute nothing; without OpenBC an in-plane neighbour
// outside the mesh is the DMI boundary extrapolation of the cell itself).
template <typename T>
__device__ __forceinline__ void lin_pair(const FieldArgs<T> &f, int ix, int iy, int iz, long I,
                                         const T *const *x0p, const T *const *x1p, V3<T> &h0, V3<T> &h1) {
  const Mesh &g = f.g;
  const int r0 = f.regions ? f.regions[I] : 0;
  const V3<T> a0 = load3(x0p, I), a1 = load3(x1p, I);
  const bool dmi = f.exch_mode == 2;
  h0 = v3(T(0), T(0), T(0));
  h1 = h0;
#pragma unroll
  for (int dir = 0; dir < 6; dir++) {
    const int axis = dir >> 1;
    const int sg = (dir & 1) ? 1 : -1;
    if (axis == 2 && g.Nz == 1) continue;
    int jx = ix, jy = iy, jz = iz;
    bool inside;
    if (axis == 0) {
      jx += sg;
      inside = (jx >= 0 && jx < g.Nx) || (g.pbc & 1);
      jx = modp(jx, g.Nx);
    } else if (axis == 1) {
      jy += sg;
      inside = (jy >= 0 && jy < g.Ny) || (g.pbc & 2);
      jy = modp(jy, g.Ny);
    } else {
      jz += sg;
      inside = (jz >= 0 && jz < g.Nz) || (g.pbc & 4);
      jz = modp(jz, g.Nz);
    }
    const T sgn = T(sg);
    const T c = axis == 0 ? f.cx : axis == 1 ? f.cy : f.cz;
    if (inside) {
      const long J = g.idx(jx, jy, jz);
      const int rj = f.regions ? f.regions[J] : 0;
      const T A = f.aLUT[symidx(r0, rj)];
      const V3<T> n0 = load3(x0p, J), n1 = load3(x1p, J);
      const T e = dmi ? (T(2.0) * A / (c * c)) : (axis == 0 ? f.wx : axis == 1 ? f.wy : f.wz) * A;
      h0 = vadd(h0, v3(e * (n0.x - a0.x), e * (n0.y - a0.y), e * (n0.z - a0.z)));
      h1 = vadd(h1, v3(e * (n1.x - a1.x), e * (n1.y - a1.y), e * (n1.z - a1.z)));
      if (dmi && axis < 2) {
        const T dc = f.dLUT[symidx(r0, rj)] / c;
        if (axis == 0) {
          h0.x += dc * (sgn * n0.z); h0.z -= dc * (sgn * n0.x);
          h1.x += dc * (sgn * n1.z); h1.z -= dc * (sgn * n1.x);
        } else {
          h0.y += dc * (sgn * n0.z); h0.z -= dc * (sgn * n0.y);
          h1.y += dc * (sgn * n1.z); h1.z -= dc * (sgn * n1.y);
        }
      }
    } else if (dmi && axis < 2 && !f.openbc) {
      const int sidx = symidx(r0, r0);
      const T A = f.aLUT[sidx], D = f.dLUT[sidx];
      const T fe = sgn * c * (T(0.5) * D / A);
      const T e = (T(2.0) * A / (c * c));
      const T dc = D / c;
      auto ext = [&](V3<T> a, V3<T> &h) {
        V3<T> n = a;
        if (axis == 0) { n.x = a.x - (fe * a.z); n.z = a.z + (fe * a.x); }
        else { n.y = a.y - (fe * a.z); n.z = a.z + (fe * a.y); }
        h.x += e * (n.x - a.x);
        h.y += e * (n.y - a.y);
        h.z += e * (n.z - a.z);
        if (axis == 0) { h.x += dc * (sgn * n.z); h.z -= dc * (sgn * n.x); }
        else { h.y += dc * (sgn * n.z); h.z -= dc * (sgn * n.y); }
      };
      ext(a0, h0);
      ext(a1, h1);
    }
  }
}

// add the block's sum of s[0..2] to its own three partials (fixed order: no
// atomics, so repeated sweeps give identical sums); all threads must call it
__device__ __forceinline__ void block_sum3(double *part, const double (&s)[3]) {
  __shared__ double ws[ADJ_NT / 64][3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double v = s[c];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
#pragma unroll
    for (int wv = 0; wv < ADJ_NT / 64; wv++) t += ws[wv][threadIdx.x];
    part[blockIdx.x * 3 + threadIdx.x] += t;
  }
}

template <typename T>
__device__ __forceinline__ void adj_stage_cell(const AdjArgs<T> &a, long i, double (&sn)[3]) {
  const FieldArgs<T> &f = a.f;
  const int ix = (int)(i % f.g.Nx);
  const int iy = (int)((i / f.g.Nx) % f.g.Ny);
  const int iz = (int)(i / ((long)f.g.Nx * f.g.Ny));
  const V3<T> u = load3(a.u, i), l = load3(a.lk, i);
  V3<T> hu, hw;
  lin_pair(f, ix, iy, iz, i, a.u, (const T *const *)a.w, hu, hw);
  const T ms = f.Ms.at(i);
  const T inv = inv_msat(f.Ms, i);
  const V3<T> Be = vscale(hu, inv);
  // uniaxial anisotropy (uniaxialanisotropy2.cl) and its symmetric Jacobian
  V3<T> Ban = v3(T(0), T(0), T(0)), ax = Ban;
  T kan = T(0);
  if (f.has_uniax) {
    ax = vnormalized(v3(f.u[0].at(i), f.u[1].at(i), f.u[2].at(i)));
    const T K1 = f.Ku1.at(i) * inv, K2 = f.Ku2.at(i) * inv;
    const T mu = vdot(u, ax);
    const T fk = T(2.0) * K1 * mu, q = T(4.0) * K2 * (mu * mu * mu);
    Ban = v3(fk * ax.x + q * ax.x, fk * ax.y + q * ax.y, fk * ax.z + q * ax.z);
    kan = T(2.0) * K1 + T(12.0) * K2 * (mu * mu);
  }
  V3<T> B = vadd(Be, Ban);
  if (a.bd[0]) B = vadd(B, load3(a.bd, i));
  if (f.bextLUT[0]) {
    const int r = f.regions ? f.regions[i] : 0;
    B = vadd(B, v3(f.bextLUT[0][r], f.bextLUT[1][r], f.bextLUT[2][r]));
  }
  if (f.has_bext) B = vadd(B, v3(f.bext[0].at(i), f.bext[1].at(i), f.bext[2].at(i)));
  const T alpha = f.alpha.at(i);
  const T gilb = T(-1.0) / (T(1.0) + alpha * alpha);
  const V3<T> gB = adj_gb(u, l, alpha);
  // direct torque term: -c (B x l + alpha ((m x B) x l + B x (l x m)))
  const V3<T> t1 = vcross(B, l), t2 = vcross(vcross(u, B), l), t3 = vcross(B, vcross(l, u));
  V3<T> gm = v3(gilb * (t1.x + alpha * (t2.x + t3.x)), gilb * (t1.y + alpha * (t2.y + t3.y)),
                gilb * (t1.z + alpha * (t2.z + t3.z)));
  gm = vadd(gm, hw);
  V3<T> Dv = v3(T(0), T(0), T(0));
  if (a.D[0]) {
    Dv = load3(a.D, i);
    gm = vadd(gm, vscale(Dv, a.mu0 * ms));
  }
  if (f.has_uniax) gm = vadd(gm, vscale(ax, kan * vdot(ax, gB)));
  if (a.gms) {
    const double d = (double)(a.mu0 * vdot(u, Dv)) - (double)(vdot(gB, vadd(Be, Ban)) * inv);
    a.gms[i] += d;
  }
  if (a.nrm) {
    const T r = T(1) / a.nrm[i];
    const T ug = vdot(u, gm);
    const V3<T> lv = v3((gm.x - u.x * ug) * r, (gm.y - u.y * ug) * r, (gm.z - u.z * ug) * r);
    st3(a.lacc, i, vadd(load3((const T *const *)a.lacc, i), lv));
    if (a.lknext[0]) {
      const V3<T> l5 = load3(a.lv5, i);
      const V3<T> ln = v3(a.c5 * l5.x + a.cs * lv.x, a.c5 * l5.y + a.cs * lv.y, a.c5 * l5.z + a.cs * lv.z);
      st3(a.lknext, i, ln);
      if (a.wnext[0]) {
        const V3<T> gBn = adj_gb(load3(a.unext, i), ln, alpha);
        st3(a.wnext, i, vscale(gBn, inv));
        sn[0] = gBn.x; sn[1] = gBn.y; sn[2] = gBn.z;
      }
    }
  } else {
    st3(a.lacc, i, vadd(load3((const T *const *)a.lacc, i), gm));
  }
}

template <typename T>
__global__ __launch_bounds__(ADJ_NT) void k_adj_stage(AdjArgs<T> a) {
  const long i = blockIdx.x * (long)ADJ_NT + threadIdx.x;
  double sn[3] = {0.0, 0.0, 0.0};
  if (i < a.N) adj_stage_cell(a, i, sn);
  if (a.gbext && a.wnext[0]) block_sum3(a.gbext, sn);
}

template <typename T>
__global__ __launch_bounds__(ADJ_NT) void k_adj_final(AdjFinalArgs<T> a) {
  const long i = blockIdx.x * (long)ADJ_NT + threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  if (i < a.N) {
    const V3<T> m = load3(a.m1, i), l = load3(a.lm, i);
    const T r = T(1) / a.nrm[i];
    const T ml = vdot(m, l);
    const V3<T> lv = v3((l.x - m.x * ml) * r, (l.y - m.y * ml) * r, (l.z - m.z * ml) * r);
    st3(a.lv5, i, lv);
    st3(a.lacc, i, lv);
    const V3<T> lk = vscale(lv, a.c4);
    st3(a.lk, i, lk);
    const V3<T> gB = adj_gb(load3(a.u4, i), lk, a.alpha.at(i));
    st3(a.w, i, vscale(gB, inv_msat(a.Ms, i)));
    s[0] = gB.x; s[1] = gB.y; s[2] = gB.z;
  }
  if (a.gbext) block_sum3(a.gbext, s);
}

template <typename T>
__global__ void k_adj_cast(T *dst, const double *src, long N) {
  const long i = blockIdx.x * (long)ADJ_NT + threadIdx.x;
  if (i < N) dst[i] = (T)src[i];
}

template <typename T>
void launch_adj_stage(const AdjArgs<T> &a, hipStream_t s) {
  UM_LAUNCH(k_adj_stage<T>, dim3((unsigned)adj_blocks(a.N)), dim3(ADJ_NT), 0, s, a);
  UM_HIP(hipGetLastError());
}
template <typename T>
void launch_adj_final(const AdjFinalArgs<T> &a, hipStream_t s) {
  UM_LAUNCH(k_adj_final<T>, dim3((unsigned)adj_blocks(a.N)), dim3(ADJ_NT), 0, s, a);
  UM_HIP(hipGetLastError());
}
template <typename T>
void launch_adj_cast(T *dst, const double *src, long N, hipStream_t s) {
  UM_LAUNCH(k_adj_cast<T>, dim3((unsigned)adj_blocks(N)), dim3(ADJ_NT), 0, s, dst, src, N);
  UM_HIP(hipGetLastError());
}

#define UM_ADJ_INST(T)                                                        \
  template void launch_adj_stage<T>(const AdjArgs<T> &, hipStream_t);        \
  template void launch_adj_final<T>(const AdjFinalArgs<T> &, hipStream_t);   \
  template void launch_adj_cast<T>(T *, const double *, long, hipStream_t);
UM_ADJ_INST(float)
UM_ADJ_INST(double)

}  // namespace um
