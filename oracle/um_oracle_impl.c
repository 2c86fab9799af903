/*
 * um_oracle_impl.c -- precision-generic part of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (checker for tests/, smoke() and the bench
 * cpu_baseline leg).  Never linked into the product library.
 *
 * Compiled twice: -DUM_REAL=float -DSFX=_f32 and -DUM_REAL=double -DSFX=_f64,
 * with -ffp-contract=off so that every product and sum rounds as written.
 * Each function cites the reference file:line it restates.  The FP64 build
 * is the "all-FP64 oracle": it deliberately does not reproduce the float32
 * roundings that uMagNUS64 inherits from amul(float mul) and pow4 (float32
 * multipliers of its FP64 kernels).
 */
#include "um_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef UM_REAL real;
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define OR(n) CAT(um_or_##n, SFX)

/* ---- helpers: kernels_src/clh/amul.clh, float3.clh, stencil.clh ---------- */
static inline real amul(const real *arr, real mul, int i) { return arr == NULL ? mul : mul * arr[i]; }
static inline real inv_msat(const real *Ms_, real Ms_mul, int i) {
  real ms = amul(Ms_, Ms_mul, i);
  return ms == (real)0 ? (real)0 : (real)1 / ms;
}
typedef struct { real x, y, z; } v3;
static inline v3 mk(real a, real b, real c) { v3 r = {a, b, c}; return r; }
static inline real dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline int is0(v3 m) { return dot3(m, m) == (real)0; }
static inline v3 normalized(v3 a) {
  real l = (real)sqrt((double)dot3(a, a));
  real s = (l != (real)0) ? (real)1 / l : (real)0;
  return mk(s * a.x, s * a.y, s * a.z);
}
#define IDX(ix, iy, iz) ((((long)(iz)) * Ny + (iy)) * Nx + (ix))
static inline int modp(int n, int M) { return ((n % M) + M) % M; }
static inline int hclamp(int i, int N, int p) { return p ? modp(i, N) : (i < N - 1 ? i : N - 1); }
static inline int lclamp(int i, int N, int p) { return p ? modp(i, N) : (i > 0 ? i : 0); }
static inline int symidx(int i, int j) { return (j <= i) ? (i * (i + 1) / 2 + j) : (j * (j + 1) / 2 + i); }

/* ---- exchange, kernels_src/cl/exchange.cl:5-87 -------------------------- */
void OR(addexchange)(real *Bx, real *By, real *Bz, const real *mx, const real *my,
                     const real *mz, const real *Ms_, real Ms_mul, const real *aLUT,
                     const uint8_t *regions, real wx, real wy, real wz, int Nx, int Ny,
                     int Nz, int pbc) {
  const int px = pbc & 1, py = pbc & 2, pz = pbc & 4;
#pragma omp parallel for collapse(2)
  for (int iz = 0; iz < Nz; iz++)
    for (int iy = 0; iy < Ny; iy++)
      for (int ix = 0; ix < Nx; ix++) {
        long I = IDX(ix, iy, iz);
        v3 m0 = mk(mx[I], my[I], mz[I]);
        if (is0(m0)) continue;
        int r0 = regions ? regions[I] : 0;
        v3 B = mk(0, 0, 0);
        long nb[6];
        real w[6];
        int nn = 4;
        nb[0] = IDX(lclamp(ix - 1, Nx, px), iy, iz); w[0] = wx;
        nb[1] = IDX(hclamp(ix + 1, Nx, px), iy, iz); w[1] = wx;
        nb[2] = IDX(ix, lclamp(iy - 1, Ny, py), iz); w[2] = wy;
        nb[3] = IDX(ix, hclamp(iy + 1, Ny, py), iz); w[3] = wy;
        if (Nz != 1) {
          nb[4] = IDX(ix, iy, lclamp(iz - 1, Nz, pz)); w[4] = wz;
          nb[5] = IDX(ix, iy, hclamp(iz + 1, Nz, pz)); w[5] = wz;
          nn = 6;
        }
        for (int k = 0; k < nn; k++) {
          long i_ = nb[k];
          v3 m_ = mk(mx[i_], my[i_], mz[i_]);
          if (is0(m_)) m_ = m0;
          real a = aLUT[symidx(r0, regions ? regions[i_] : 0)];
          real s = w[k] * a;
          B.x += s * (m_.x - m0.x);
          B.y += s * (m_.y - m0.y);
          B.z += s * (m_.z - m0.z);
        }
        real invMs = inv_msat(Ms_, Ms_mul, (int)I);
        Bx[I] += B.x * invMs;
        By[I] += B.y * invMs;
        Bz[I] += B.z * invMs;
      }
}

/* ---- exchange + interfacial DMI, kernels_src/cl/dmi.cl:8-148 ------------ */
void OR(adddmi)(real *Hx, real *Hy, real *Hz, const real *mx, const real *my,
                const real *mz, const real *Ms_, real Ms_mul, const real *aLUT,
                const real *dLUT, const uint8_t *regions, real cx, real cy, real cz,
                int Nx, int Ny, int Nz, int pbc, int openbc) {
  const int px = pbc & 1, py = pbc & 2, pz = pbc & 4;
#define REG(i) (regions ? (int)regions[i] : 0)
#pragma omp parallel for collapse(2)
  for (int iz = 0; iz < Nz; iz++)
    for (int iy = 0; iy < Ny; iy++)
      for (int ix = 0; ix < Nx; ix++) {
        long I = IDX(ix, iy, iz);
        v3 h = mk(0, 0, 0);
        v3 m0 = mk(mx[I], my[I], mz[I]);
        int r0 = REG(I);
        long i_;
        if (is0(m0)) continue;
        /* x, left neighbour (:36-55) */
        {
          v3 m1 = mk(0, 0, 0);
          i_ = IDX(lclamp(ix - 1, Nx, px), iy, iz);
          if ((ix - 1 >= 0) || px) m1 = mk(mx[i_], my[i_], mz[i_]);
          int r1 = is0(m1) ? r0 : REG(i_);
          real A1 = aLUT[symidx(r0, r1)];
          real D1 = dLUT[symidx(r0, r1)];
          if (!is0(m1) || !openbc) {
            if (is0(m1)) {
              m1.x = m0.x - (-cx * ((real)0.5 * D1 / A1) * m0.z);
              m1.y = m0.y;
              m1.z = m0.z + (-cx * ((real)0.5 * D1 / A1) * m0.x);
            }
            real e = ((real)2.0 * A1 / (cx * cx));
            h.x += e * (m1.x - m0.x); h.y += e * (m1.y - m0.y); h.z += e * (m1.z - m0.z);
            h.x += (D1 / cx) * (-m1.z);
            h.z -= (D1 / cx) * (-m1.x);
          }
        }
        /* x, right neighbour (:57-76) */
        {
          v3 m2 = mk(0, 0, 0);
          i_ = IDX(hclamp(ix + 1, Nx, px), iy, iz);
          if ((ix + 1 < Nx) || px) m2 = mk(mx[i_], my[i_], mz[i_]);
          int r2 = is0(m2) ? r0 : REG(i_);
          real A2 = aLUT[symidx(r0, r2)];
          real D2 = dLUT[symidx(r0, r2)];
          if (!is0(m2) || !openbc) {
            if (is0(m2)) {
              m2.x = m0.x - (cx * ((real)0.5 * D2 / A2) * m0.z);
              m2.y = m0.y;
              m2.z = m0.z + (cx * ((real)0.5 * D2 / A2) * m0.x);
            }
            real e = ((real)2.0 * A2 / (cx * cx));
            h.x += e * (m2.x - m0.x); h.y += e * (m2.y - m0.y); h.z += e * (m2.z - m0.z);
            h.x += (D2 / cx) * (m2.z);
            h.z -= (D2 / cx) * (m2.x);
          }
        }
        /* y, back neighbour (:79-98) */
        {
          v3 m1 = mk(0, 0, 0);
          i_ = IDX(ix, lclamp(iy - 1, Ny, py), iz);
          if ((iy - 1 >= 0) || py) m1 = mk(mx[i_], my[i_], mz[i_]);
          int r1 = is0(m1) ? r0 : REG(i_);
          real A1 = aLUT[symidx(r0, r1)];
          real D1 = dLUT[symidx(r0, r1)];
          if (!is0(m1) || !openbc) {
            if (is0(m1)) {
              m1.x = m0.x;
              m1.y = m0.y - (-cy * ((real)0.5 * D1 / A1) * m0.z);
              m1.z = m0.z + (-cy * ((real)0.5 * D1 / A1) * m0.y);
            }
            real e = ((real)2.0 * A1 / (cy * cy));
            h.x += e * (m1.x - m0.x); h.y += e * (m1.y - m0.y); h.z += e * (m1.z - m0.z);
            h.y += (D1 / cy) * (-m1.z);
            h.z -= (D1 / cy) * (-m1.y);
          }
        }
        /* y, front neighbour (:100-119) */
        {
          v3 m2 = mk(0, 0, 0);
          i_ = IDX(ix, hclamp(iy + 1, Ny, py), iz);
          if ((iy + 1 < Ny) || py) m2 = mk(mx[i_], my[i_], mz[i_]);
          int r2 = is0(m2) ? r0 : REG(i_);
          real A2 = aLUT[symidx(r0, r2)];
          real D2 = dLUT[symidx(r0, r2)];
          if (!is0(m2) || !openbc) {
            if (is0(m2)) {
              m2.x = m0.x;
              m2.y = m0.y - (cy * ((real)0.5 * D2 / A2) * m0.z);
              m2.z = m0.z + (cy * ((real)0.5 * D2 / A2) * m0.y);
            }
            real e = ((real)2.0 * A2 / (cy * cy));
            h.x += e * (m2.x - m0.x); h.y += e * (m2.y - m0.y); h.z += e * (m2.z - m0.z);
            h.y += (D2 / cy) * (m2.z);
            h.z -= (D2 / cy) * (m2.y);
          }
        }
        /* z: exchange only (:121-140) */
        if (Nz != 1) {
          i_ = IDX(ix, iy, lclamp(iz - 1, Nz, pz));
          v3 m1 = mk(mx[i_], my[i_], mz[i_]);
          if (is0(m1)) m1 = m0;
          real A1 = aLUT[symidx(r0, REG(i_))];
          real e = ((real)2.0 * A1 / (cz * cz));
          h.x += e * (m1.x - m0.x); h.y += e * (m1.y - m0.y); h.z += e * (m1.z - m0.z);
          i_ = IDX(ix, iy, hclamp(iz + 1, Nz, pz));
          v3 m2 = mk(mx[i_], my[i_], mz[i_]);
          if (is0(m2)) m2 = m0;
          real A2 = aLUT[symidx(r0, REG(i_))];
          e = ((real)2.0 * A2 / (cz * cz));
          h.x += e * (m2.x - m0.x); h.y += e * (m2.y - m0.y); h.z += e * (m2.z - m0.z);
        }
        real invMs = inv_msat(Ms_, Ms_mul, (int)I);
        Hx[I] += h.x * invMs;
        Hy[I] += h.y * invMs;
        Hz[I] += h.z * invMs;
      }
#undef REG
}

/* ---- uniaxial anisotropy, kernels_src/cl/uniaxialanisotropy2.cl:3-36 ---- */
void OR(adduniaxial2)(real *Bx, real *By, real *Bz, const real *mx, const real *my,
                      const real *mz, const real *Ms_, real Ms_mul, const real *K1_,
                      real K1_mul, const real *K2_, real K2_mul, const real *ux_,
                      real ux_mul, const real *uy_, real uy_mul, const real *uz_,
                      real uz_mul, int N) {
#pragma omp parallel for
  for (int i = 0; i < N; i++) {
    v3 u = normalized(mk(amul(ux_, ux_mul, i), amul(uy_, uy_mul, i), amul(uz_, uz_mul, i)));
    real invMs = inv_msat(Ms_, Ms_mul, i);
    real K1 = amul(K1_, K1_mul, i);
    real K2 = amul(K2_, K2_mul, i);
    K1 *= invMs;
    K2 *= invMs;
    v3 m = mk(mx[i], my[i], mz[i]);
    real mu = dot3(m, u);
    real f = (real)2.0 * K1 * mu;
    real g = (real)4.0 * K2 * (mu * mu * mu);
    Bx[i] += f * u.x + g * u.x;
    By[i] += f * u.y + g * u.y;
    Bz[i] += f * u.z + g * u.z;
  }
}

/* ---- cubic anisotropy, kernels_src/cl/cubicanisotropy2.cl:10-63 --------- */
void OR(addcubic2)(real *Bx, real *By, real *Bz, const real *mx, const real *my,
                   const real *mz, const real *Ms_, real Ms_mul, const real *k1_,
                   real k1_mul, const real *k2_, real k2_mul, const real *k3_,
                   real k3_mul, const real *c1x_, real c1x_mul, const real *c1y_,
                   real c1y_mul, const real *c1z_, real c1z_mul, const real *c2x_,
                   real c2x_mul, const real *c2y_, real c2y_mul, const real *c2z_,
                   real c2z_mul, int N) {
#pragma omp parallel for
  for (int i = 0; i < N; i++) {
    real invMs = inv_msat(Ms_, Ms_mul, i);
    real k1 = amul(k1_, k1_mul, i) * invMs;
    real k2 = amul(k2_, k2_mul, i) * invMs;
    real k3 = amul(k3_, k3_mul, i) * invMs;
    v3 u1 = normalized(mk(amul(c1x_, c1x_mul, i), amul(c1y_, c1y_mul, i), amul(c1z_, c1z_mul, i)));
    v3 u2 = normalized(mk(amul(c2x_, c2x_mul, i), amul(c2y_, c2y_mul, i), amul(c2z_, c2z_mul, i)));
    v3 u3 = cross3(u1, u2);
    v3 m = mk(mx[i], my[i], mz[i]);
    real a = dot3(u1, m), b = dot3(u2, m), c = dot3(u3, m);
    real a2 = a * a, b2 = b * b, c2 = c * c;
    real a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
    real f1 = (b2 + c2) * a, f2 = (a2 + c2) * b, f3 = (a2 + b2) * c;
    real g1 = (b2 * c2) * a, g2 = (a2 * c2) * b, g3 = (a2 * b2) * c;
    real h1 = (b4 + c4) * (a * a * a), h2 = (a4 + c4) * (b * b * b), h3 = (a4 + b4) * (c * c * c);
    v3 B;
    B.x = (real)-2.0 * k1 * (f1 * u1.x + f2 * u2.x + f3 * u3.x) - (real)2.0 * k2 * (g1 * u1.x + g2 * u2.x + g3 * u3.x) -
          (real)4.0 * k3 * (h1 * u1.x + h2 * u2.x + h3 * u3.x);
    B.y = (real)-2.0 * k1 * (f1 * u1.y + f2 * u2.y + f3 * u3.y) - (real)2.0 * k2 * (g1 * u1.y + g2 * u2.y + g3 * u3.y) -
          (real)4.0 * k3 * (h1 * u1.y + h2 * u2.y + h3 * u3.y);
    B.z = (real)-2.0 * k1 * (f1 * u1.z + f2 * u2.z + f3 * u3.z) - (real)2.0 * k2 * (g1 * u1.z + g2 * u2.z + g3 * u3.z) -
          (real)4.0 * k3 * (h1 * u1.z + h2 * u2.z + h3 * u3.z);
    Bx[i] += B.x;
    By[i] += B.y;
    Bz[i] += B.z;
  }
}

/* ---- region LUT add, kernels_src/cl/regionaddv.cl:3-18 ------------------ */
void OR(regionaddv)(real *dx, real *dy, real *dz, const real *lx, const real *ly,
                    const real *lz, const uint8_t *regions, int N) {
  for (int i = 0; i < N; i++) {
    int r = regions ? regions[i] : 0;
    dx[i] += lx[r];
    dy[i] += ly[r];
    dz[i] += lz[r];
  }
}

/* ---- zeromask, kernels_src/cl/zeromask.cl:2-13 -------------------------- */
void OR(zeromask)(real *dst, const real *maskLUT, const uint8_t *regions, int N) {
  for (int i = 0; i < N; i++)
    if (maskLUT[regions ? regions[i] : 0] != (real)0) dst[i] = 0;
}

/* ---- exchange + bulk DMI, kernels_src/cl/dmibulk.cl:31-235 -------------- */
/* One neighbour along `axis` (0 x, 1 y, 2 z) on side sgn (-1 lower, +1 upper):
 * missing neighbours are extrapolated from the boundary condition unless
 * OpenBC; the DMI term couples the two components transverse to the axis. */
static void dmibulk_nb(v3 *h, v3 m0, v3 n, int present, int r0, int rn, int axis, real sgn, real c,
                       const real *aLUT, const real *dLUT, int openbc) {
  int r1 = present ? rn : r0;
  real A = aLUT[symidx(r0, r1)];
  real D = dLUT[symidx(r0, r1)];
  real D_2A = D / ((real)2.0 * A);
  if (!present && openbc) return;
  if (!present) {
    real f = sgn * c * D_2A;
    if (axis == 0) { n.x = m0.x; n.y = m0.y - (f * m0.z); n.z = m0.z + (f * m0.y); }
    else if (axis == 1) { n.x = m0.x + (f * m0.z); n.y = m0.y; n.z = m0.z - (f * m0.x); }
    else { n.x = m0.x - (f * m0.y); n.y = m0.y + (f * m0.x); n.z = m0.z; }
  }
  real e = ((real)2.0 * A / (c * c));
  h->x += e * (n.x - m0.x); h->y += e * (n.y - m0.y); h->z += e * (n.z - m0.z);
  real dc = D / c;
  if (axis == 0) { h->y += dc * (sgn * n.z); h->z -= dc * (sgn * n.y); }
  else if (axis == 1) { h->x -= dc * (sgn * n.z); h->z += dc * (sgn * n.x); }
  else { h->x += dc * (sgn * n.y); h->y -= dc * (sgn * n.x); }
}

void OR(adddmibulk)(real *Hx, real *Hy, real *Hz, const real *mx, const real *my,
                    const real *mz, const real *Ms_, real Ms_mul, const real *aLUT,
                    const real *dLUT, const uint8_t *regions, real cx, real cy, real cz,
                    int Nx, int Ny, int Nz, int pbc, int openbc) {
  const int px = pbc & 1, py = pbc & 2, pz = pbc & 4;
#pragma omp parallel for collapse(2)
  for (int iz = 0; iz < Nz; iz++)
    for (int iy = 0; iy < Ny; iy++)
      for (int ix = 0; ix < Nx; ix++) {
        long I = IDX(ix, iy, iz);
        v3 h = mk(0, 0, 0);
        v3 m0 = mk(mx[I], my[I], mz[I]);
        int r0 = regions ? regions[I] : 0;
        if (is0(m0)) continue;
        const int nax = Nz != 1 ? 3 : 2;
        for (int axis = 0; axis < nax; axis++)
          for (int side = 0; side < 2; side++) {
            const int sg = side ? 1 : -1;
            int jx = ix, jy = iy, jz = iz, inside, p;
            int NN = axis == 0 ? Nx : axis == 1 ? Ny : Nz;
            int *jj = axis == 0 ? &jx : axis == 1 ? &jy : &jz;
            p = axis == 0 ? px : axis == 1 ? py : pz;
            inside = (*jj + sg >= 0 && *jj + sg < NN) || p;
            *jj = sg > 0 ? hclamp(*jj + 1, NN, p) : lclamp(*jj - 1, NN, p);
            long J = IDX(jx, jy, jz);
            v3 n = mk(0, 0, 0);
            if (inside) n = mk(mx[J], my[J], mz[J]);
            int rn = regions ? regions[J] : 0;
            real c = axis == 0 ? cx : axis == 1 ? cy : cz;
            dmibulk_nb(&h, m0, n, !is0(n), r0, rn, axis, (real)sg, c, aLUT, dLUT, openbc);
          }
        real invMs = inv_msat(Ms_, Ms_mul, (int)I);
        Hx[I] += h.x * invMs;
        Hy[I] += h.y * invMs;
        Hz[I] += h.z * invMs;
      }
}

/* ---- Zhang-Li torque, kernels_src/cl/zhangli2.cl:1-58 -------------------- */
#define UM_MUB 9.2740091523E-24
#define UM_QE 1.60217646E-19
#define UM_GAMMA0 1.7595e11
#define UM_HBAR 1.05457173E-34
void OR(addzhangli)(real *tx, real *ty, real *tz, const real *mx, const real *my,
                    const real *mz, const real *Ms_, real Ms_mul, const real *jx_, real jx_mul,
                    const real *jy_, real jy_mul, const real *jz_, real jz_mul,
                    const real *alpha_, real alpha_mul, const real *xi_, real xi_mul,
                    const real *pol_, real pol_mul, real cx, real cy, real cz, int Nx, int Ny,
                    int Nz, int pbc) {
  const int px = pbc & 1, py = pbc & 2, pz = pbc & 4;
  const real pre = (real)(UM_MUB / (2 * UM_QE * UM_GAMMA0));
#pragma omp parallel for collapse(2)
  for (int iz = 0; iz < Nz; iz++)
    for (int iy = 0; iy < Ny; iy++)
      for (int ix = 0; ix < Nx; ix++) {
        long i = IDX(ix, iy, iz);
        real alpha = amul(alpha_, alpha_mul, (int)i);
        real xi = amul(xi_, xi_mul, (int)i);
        real pol = amul(pol_, pol_mul, (int)i);
        real invMs = inv_msat(Ms_, Ms_mul, (int)i);
        real b = invMs * pre / ((real)1.0 + xi * xi);
        v3 J = mk(pol * amul(jx_, jx_mul, (int)i), pol * amul(jy_, jy_mul, (int)i),
                  pol * amul(jz_, jz_mul, (int)i));
        v3 hs = mk(0, 0, 0);   /* (u . grad) m, central differences over clamped neighbours */
        if (J.x != (real)0) {
          long a = IDX(hclamp(ix + 1, Nx, px), iy, iz), c = IDX(lclamp(ix - 1, Nx, px), iy, iz);
          real f = (b / cx) * J.x;
          hs.x += f * (mx[a] - mx[c]); hs.y += f * (my[a] - my[c]); hs.z += f * (mz[a] - mz[c]);
        }
        if (J.y != (real)0) {
          long a = IDX(ix, hclamp(iy + 1, Ny, py), iz), c = IDX(ix, lclamp(iy - 1, Ny, py), iz);
          real f = (b / cy) * J.y;
          hs.x += f * (mx[a] - mx[c]); hs.y += f * (my[a] - my[c]); hs.z += f * (mz[a] - mz[c]);
        }
        if (J.z != (real)0) {
          long a = IDX(ix, iy, hclamp(iz + 1, Nz, pz)), c = IDX(ix, iy, lclamp(iz - 1, Nz, pz));
          real f = (b / cz) * J.z;
          hs.x += f * (mx[a] - mx[c]); hs.y += f * (my[a] - my[c]); hs.z += f * (mz[a] - mz[c]);
        }
        v3 m = mk(mx[i], my[i], mz[i]);
        v3 mh = cross3(m, hs), mmh = cross3(m, mh);
        real g = (real)-1.0 / ((real)1.0 + alpha * alpha);
        real c1 = (real)1.0 + xi * alpha, c2 = xi - alpha;
        tx[i] += g * (c1 * mmh.x + c2 * mh.x);
        ty[i] += g * (c1 * mmh.y + c2 * mh.y);
        tz[i] += g * (c1 * mmh.z + c2 * mh.z);
      }
}

/* ---- Slonczewski torque, kernels_src/cl/slonczewski2.cl:4-61 ------------- */
/* A cell with J == 0 or Ms == 0 gets no torque.  (The reference returns from
 * its grid-stride loop there, which also skips the thread's later cells; the
 * per-cell skip is what the loop intends.) */
void OR(addslonczewski)(real *tx, real *ty, real *tz, const real *mx, const real *my,
                        const real *mz, const real *Ms_, real Ms_mul, const real *jz_, real jz_mul,
                        const real *px_, real px_mul, const real *py_, real py_mul,
                        const real *pz_, real pz_mul, const real *alpha_, real alpha_mul,
                        const real *pol_, real pol_mul, const real *lambda_, real lambda_mul,
                        const real *eps_, real eps_mul, const real *thick_, real thick_mul,
                        real meshThickness, real flp, int N) {
#pragma omp parallel for
  for (int i = 0; i < N; i++) {
    v3 m = mk(mx[i], my[i], mz[i]);
    real J = amul(jz_, jz_mul, i);
    v3 p = normalized(mk(amul(px_, px_mul, i), amul(py_, py_mul, i), amul(pz_, pz_mul, i)));
    real Ms = amul(Ms_, Ms_mul, i);
    real alpha = amul(alpha_, alpha_mul, i);
    real pol = amul(pol_, pol_mul, i);
    real lambda = amul(lambda_, lambda_mul, i);
    real epsPrime = amul(eps_, eps_mul, i);
    real thickness = amul(thick_, thick_mul, i);
    if (thickness == (real)0.0) thickness = meshThickness;
    thickness *= flp;
    if (J == (real)0.0 || Ms == (real)0.0) continue;
    real beta = (real)(UM_HBAR / UM_QE) * (J / (thickness * Ms));
    real l2 = lambda * lambda;
    real eps = pol * l2 / ((l2 + (real)1.0) + (l2 - (real)1.0) * dot3(p, m));
    real A = beta * eps, B = beta * epsPrime;
    real gilb = (real)1.0 / ((real)1.0 + alpha * alpha);
    real f1 = gilb * (A + alpha * B), f2 = gilb * (B - alpha * A);
    v3 pxm = cross3(p, m), mpm = cross3(m, pxm);
    tx[i] += f1 * mpm.x + f2 * pxm.x;
    ty[i] += f1 * mpm.y + f2 * pxm.y;
    tz[i] += f1 * mpm.z + f2 * pxm.z;
  }
}

/* ---- LL torque, kernels_src/cl/lltorque2.cl:1-25 ------------------------ */
void OR(lltorque2)(real *tx, real *ty, real *tz, const real *mx, const real *my,
                   const real *mz, const real *hx, const real *hy, const real *hz,
                   const real *alpha_, real alpha_mul, int N) {
#pragma omp parallel for
  for (int i = 0; i < N; i++) {
    v3 m = mk(mx[i], my[i], mz[i]);
    v3 H = mk(hx[i], hy[i], hz[i]);
    real alpha = amul(alpha_, alpha_mul, i);
    v3 mxH = cross3(m, H);
    real gilb = (real)-1.0 / ((real)1.0 + alpha * alpha);
    v3 mmH = cross3(m, mxH);
    tx[i] = gilb * (mxH.x + alpha * mmH.x);
    ty[i] = gilb * (mxH.y + alpha * mmH.y);
    tz[i] = gilb * (mxH.z + alpha * mmH.z);
  }
}

/* ---- LL torque without precession, kernels_src/cl/llnoprecess.cl:1-23 --- */
void OR(llnoprecess)(real *tx, real *ty, real *tz, const real *mx, const real *my,
                     const real *mz, const real *hx, const real *hy, const real *hz, int N) {
  for (int i = 0; i < N; i++) {
    v3 m = mk(mx[i], my[i], mz[i]);
    v3 H = mk(hx[i], hy[i], hz[i]);
    v3 t = cross3(m, cross3(m, H));
    tx[i] = -t.x;
    ty[i] = -t.y;
    tz[i] = -t.z;
  }
}

/* ---- linear combination, kernels_src/cl/madd2.cl .. madd7.cl ------------ */
/* madd3 groups its last two terms (madd3.cl:12); the others sum left to right. */
void OR(madd)(real *dst, int nsrc, const real *const *src, const real *fac, int N) {
#pragma omp parallel for
  for (int i = 0; i < N; i++) {
    real acc;
    if (nsrc == 3) {
      acc = (fac[0] * src[0][i]) + (fac[1] * src[1][i] + fac[2] * src[2][i]);
    } else {
      acc = fac[0] * src[0][i];
      for (int k = 1; k < nsrc; k++) acc = acc + fac[k] * src[k][i];
    }
    dst[i] = acc;
  }
}

/* ---- normalize, kernels_src/cl/normalize2.cl:1-24 ----------------------- */
void OR(normalize2)(real *vx, real *vy, real *vz, const real *vol, int N) {
#pragma omp parallel for
  for (int i = 0; i < N; i++) {
    real v = vol == NULL ? (real)1 : vol[i];
    v3 V = normalized(mk(v * vx[i], v * vy[i], v * vz[i]));
    if (v == (real)0) {
      vx[i] = 0; vy[i] = 0; vz[i] = 0;
    } else {
      vx[i] = V.x; vy[i] = V.y; vz[i] = V.z;
    }
  }
}

/* ---- thermal field scale, kernels_src/cl/temperature2.cl:1-19 ----------- */
void OR(settemperature2)(real *B, const real *noise, real kB2_VgammaDt, const real *Ms_,
                         real Ms_mul, const real *temp_, real temp_mul, const real *alpha_,
                         real alpha_mul, int N) {
  for (int i = 0; i < N; i++) {
    real invMs = inv_msat(Ms_, Ms_mul, i);
    real temp = amul(temp_, temp_mul, i);
    real alpha = amul(alpha_, alpha_mul, i);
    B[i] = noise[i] * (real)sqrt((double)(kB2_VgammaDt * alpha * temp * invMs));
  }
}

/* ---- reductions, opencl/reduce.go:26-251 -------------------------------- */
/* MaxVecNorm: max over cells of dot(v,v) in working precision, then the host
 * takes sqrt in float64 (reduce.go:180-209, reducemaxvecnorm2.cl:22). */
double OR(maxvecnorm)(const real *x, const real *y, const real *z, int N) {
  real mx = 0;
  for (int i = 0; i < N; i++) {
    real d = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
    if (d > mx) mx = d;
  }
  return sqrt((double)mx);
}
/* MaxVecDiff (reduce.go:214-251).  The kernel widens the scalar distance to a
 * vector (reducemaxvecdiff2.cl:24), so the reference returns sqrt(3)*|a-b|;
 * sqrt3_quirk selects that behaviour. */
double OR(maxvecdiff)(const real *x1, const real *y1, const real *z1, const real *x2,
                      const real *y2, const real *z2, int N, int sqrt3_quirk) {
  real mx = 0;
  for (int i = 0; i < N; i++) {
    real dx = x1[i] - x2[i], dy = y1[i] - y2[i], dz = z1[i] - z2[i];
    real d = dx * dx + dy * dy + dz * dz;
    if (sqrt3_quirk) {
      real dist = (real)sqrt((double)d);
      d = dist * dist + dist * dist + dist * dist;
    }
    if (d > mx) mx = d;
  }
  return sqrt((double)mx);
}
/* MaxAbs (reduce.go:105-128): the mathematically intended max|x| (reduce_test.go:75-84);
 * the reference kernel (reducemaxabs.cl) omits fabs, which only matters for negative inputs. */
double OR(maxabs)(const real *x, int N) {
  double mx = 0;
  for (int i = 0; i < N; i++) {
    double a = fabs((double)x[i]);
    if (a > mx) mx = a;
  }
  return mx;
}
/* Sum / Dot (reduce.go:26-102): the intended sums, accumulated in FP64.  The
 * runtime-compiled reducesum.cl / reducedot.cl add the first element of each
 * work-item twice (reducesum.cl:12 and :18); the tests' expectations encode
 * the plain sum, which is what this computes. */
double OR(sum)(const real *x, int N) {
  double s = 0;
  for (int i = 0; i < N; i++) s += (double)x[i];
  return s;
}
double OR(dot)(const real *a, const real *b, int N) {
  double s = 0;
  for (int i = 0; i < N; i++) s += (double)a[i] * (double)b[i];
  return s;
}

/* ---- FFT ----------------------------------------------------------------- */
/* The reference delegates to VkFFT (cl/vkfft.go:218-259): unnormalised forward
 * (exp(-i..)) R2C and backward C2R.  This restatement is a mixed-radix Stockham
 * FFT with twiddles computed in FP64, any length (prime factors handled by a
 * direct DFT of that radix).  Pinned against numpy.fft in tests. */
typedef struct { real re, im; } cx;
static void factorize(int n, int *f, int *nf) {
  *nf = 0;
  const int pref[] = {4, 2, 3, 5, 7};
  for (int k = 0; k < 5; k++) {
    int p = pref[k];
    while (n % p == 0 && n > 1) {
      if (p == 4 && (n % 4) != 0) break;
      f[(*nf)++] = p;
      n /= p;
    }
  }
  for (int p = 11; n > 1; p += 2)
    while (n % p == 0) { f[(*nf)++] = p; n /= p; }
}
static void fft_line(cx *x, cx *y, int n, int sign, const double *twr, const double *twi) {
  /* twr/twi: exp(-2 pi i m / n), m = 0..n-1 (FP64) */
  if (n == 1) return;
  int f[64], nf;
  factorize(n, f, &nf);
  cx *a = x, *b = y;
  int p = 1;
  cx u[64], t[64];
  for (int s = 0; s < nf; s++) {
    int R = f[s];
    int nR = n / R;
    int tstep = n / (p * R);
    if (R == 2 || R == 4) {
      /* same Stockham indexing; radix-2/4 butterflies written out */
      for (int i = 0; i < nR; i++) {
        int k = i % p;
        for (int j = 0; j < R; j++) {
          cx v = a[i + j * nR];
          int m = j * k * tstep;
          double wr = twr[m], wi = sign < 0 ? twi[m] : -twi[m];
          u[j].re = (real)(v.re * (real)wr - v.im * (real)wi);
          u[j].im = (real)(v.re * (real)wi + v.im * (real)wr);
        }
        if (R == 2) {
          t[0].re = u[0].re + u[1].re; t[0].im = u[0].im + u[1].im;
          t[1].re = u[0].re - u[1].re; t[1].im = u[0].im - u[1].im;
        } else {
          /* y1 = (a0 - a2) + sign*i*(a1 - a3) */
          cx s02 = {u[0].re + u[2].re, u[0].im + u[2].im}, d02 = {u[0].re - u[2].re, u[0].im - u[2].im};
          cx s13 = {u[1].re + u[3].re, u[1].im + u[3].im}, d13 = {u[1].re - u[3].re, u[1].im - u[3].im};
          cx r = sign < 0 ? (cx){d13.im, -d13.re} : (cx){-d13.im, d13.re};
          t[0].re = s02.re + s13.re; t[0].im = s02.im + s13.im;
          t[2].re = s02.re - s13.re; t[2].im = s02.im - s13.im;
          t[1].re = d02.re + r.re; t[1].im = d02.im + r.im;
          t[3].re = d02.re - r.re; t[3].im = d02.im - r.im;
        }
        int j0 = (i - k) * R + k;
        for (int j = 0; j < R; j++) b[j0 + j * p] = t[j];
      }
      cx *tmp = a; a = b; b = tmp;
      p *= R;
      continue;
    }
    for (int i = 0; i < nR; i++) {
      int k = i % p;
      for (int j = 0; j < R; j++) {
        cx v = a[i + j * nR];
        int m = (j * k * tstep) % n;
        double wr = twr[m], wi = sign < 0 ? twi[m] : -twi[m];
        u[j].re = (real)(v.re * (real)wr - v.im * (real)wi);
        u[j].im = (real)(v.re * (real)wi + v.im * (real)wr);
      }
      /* direct DFT of radix R */
      for (int q = 0; q < R; q++) {
        real sr = 0, si = 0;
        for (int j = 0; j < R; j++) {
          int m = ((j * q) % R) * (n / R);
          double wr = twr[m], wi = sign < 0 ? twi[m] : -twi[m];
          sr += u[j].re * (real)wr - u[j].im * (real)wi;
          si += u[j].re * (real)wi + u[j].im * (real)wr;
        }
        t[q].re = sr;
        t[q].im = si;
      }
      int j0 = (i - k) * R + k;
      for (int j = 0; j < R; j++) b[j0 + j * p] = t[j];
    }
    cx *tmp = a; a = b; b = tmp;
    p *= R;
  }
  if (a != x) memcpy(x, a, sizeof(cx) * n);
}
static void make_tw(int n, double **twr, double **twi) {
  *twr = (double *)malloc(sizeof(double) * n);
  *twi = (double *)malloc(sizeof(double) * n);
  for (int m = 0; m < n; m++) {
    double a = -2.0 * M_PI * (double)m / (double)n;
    (*twr)[m] = cos(a);
    (*twi)[m] = sin(a);
  }
}
void OR(fft)(real *re_im, int n, int sign) {
  double *twr, *twi;
  make_tw(n, &twr, &twi);
  cx *y = (cx *)malloc(sizeof(cx) * n);
  fft_line((cx *)re_im, y, n, sign, twr, twi);
  free(y); free(twr); free(twi);
}
/* 3D complex FFT over the whole array [z][y][x], x fastest. */
static void fft3d_axis(cx *d, const int P[3], int axis, int sign, int nvalid_lines_z) {
  int n = P[axis];
  if (n == 1) return;
  double *twr, *twi;
  make_tw(n, &twr, &twi);
  long stride = axis == 0 ? 1 : (axis == 1 ? P[0] : (long)P[0] * P[1]);
  int nz = (axis == 2) ? 1 : nvalid_lines_z;
  int na = axis == 0 ? P[1] : P[0];
  int nb = axis == 2 ? P[1] : nz;
#pragma omp parallel
  {
    cx *line = (cx *)malloc(sizeof(cx) * n), *tmp = (cx *)malloc(sizeof(cx) * n);
#pragma omp for collapse(2)
    for (int b = 0; b < nb; b++)
      for (int a = 0; a < na; a++) {
        long base;
        if (axis == 0) base = ((long)b * P[1] + a) * P[0];
        else if (axis == 1) base = (long)b * P[0] * P[1] + a;
        else base = (long)b * P[0] + a;
        for (int k = 0; k < n; k++) line[k] = d[base + k * stride];
        fft_line(line, tmp, n, sign, twr, twi);
        for (int k = 0; k < n; k++) d[base + k * stride] = line[k];
      }
    free(line); free(tmp);
  }
  free(twr); free(twi);
}
void OR(fft3d)(real *re_im, const int P[3], int sign) {
  cx *d = (cx *)re_im;
  fft3d_axis(d, P, 0, sign, P[2]);
  fft3d_axis(d, P, 1, sign, P[2]);
  fft3d_axis(d, P, 2, sign, P[2]);
}

/* ---- demag convolution, opencl/conv_demag.go + kernmulrsymm*.cl ---------- */
typedef struct {
  int N[3], P[3], H;       /* input size, padded size, Px/2+1 */
  int is2d;
  int KS[3];               /* stored kernel size: H x (Py/2+1) x (Pz/2+1) */
  real *kern[6];           /* xx yy zz yz xz xy (NULL when absent in 2D) */
  cx *buf[3];              /* R2C output, full [Pz][Py][H] */
} demag_t;

/* forward 3D R2C of a real padded array: out[Pz][Py][H] */
static void r2c3d(const real *in, cx *out, const int P[3], int H) {
  int Px = P[0];
  double *twr, *twi;
  make_tw(Px, &twr, &twi);
#pragma omp parallel
  {
    cx *line = (cx *)malloc(sizeof(cx) * Px), *tmp = (cx *)malloc(sizeof(cx) * Px);
#pragma omp for
    for (long r = 0; r < (long)P[1] * P[2]; r++) {
      for (int k = 0; k < Px; k++) { line[k].re = in[r * Px + k]; line[k].im = 0; }
      fft_line(line, tmp, Px, -1, twr, twi);
      for (int k = 0; k < H; k++) out[r * H + k] = line[k];
    }
    free(line); free(tmp);
  }
  free(twr); free(twi);
  int Q[3] = {H, P[1], P[2]};
  fft3d_axis(out, Q, 1, -1, P[2]);
  fft3d_axis(out, Q, 2, -1, P[2]);
}
/* backward 3D C2R (unnormalised), out real [Pz][Py][Px]; destroys in */
static void c2r3d(cx *in, real *out, const int P[3], int H) {
  int Q[3] = {H, P[1], P[2]};
  fft3d_axis(in, Q, 2, +1, P[2]);
  fft3d_axis(in, Q, 1, +1, P[2]);
  int Px = P[0];
  double *twr, *twi;
  make_tw(Px, &twr, &twi);
#pragma omp parallel
  {
    cx *line = (cx *)malloc(sizeof(cx) * Px), *tmp = (cx *)malloc(sizeof(cx) * Px);
#pragma omp for
    for (long r = 0; r < (long)P[1] * P[2]; r++) {
      for (int k = 0; k < H; k++) line[k] = in[r * H + k];
      for (int k = H; k < Px; k++) { line[k].re = in[r * H + (Px - k)].re; line[k].im = -in[r * H + (Px - k)].im; }
      fft_line(line, tmp, Px, +1, twr, twi);
      for (int k = 0; k < Px; k++) out[r * Px + k] = line[k].re;
    }
    free(line); free(tmp);
  }
  free(twr); free(twi);
}

/* DemagConvolution.init, opencl/conv_demag.go:131-202 and scaleRealParts,
 * opencl/conv_common.go:27-50: FFT each real-space component, keep the
 * non-redundant H x (Py/2+1) x (Pz/2+1) real parts scaled by 1/(Px Py Pz).
 * The FP32 oracle FFTs the float32-rounded tensor in float32 like the
 * reference; the FP64 oracle keeps everything in FP64. */
void *OR(demag_new_from_kernel)(const int N[3], const int P[3], double *const K[6]) {
  demag_t *d = (demag_t *)calloc(1, sizeof(demag_t));
  memcpy(d->N, N, sizeof(int) * 3);
  memcpy(d->P, P, sizeof(int) * 3);
  d->H = P[0] / 2 + 1;
  d->is2d = (N[2] == 1);
  d->KS[0] = d->H; d->KS[1] = P[1] / 2 + 1; d->KS[2] = P[2] / 2 + 1;
  long np = (long)P[0] * P[1] * P[2];
  long nc = (long)d->H * P[1] * P[2];
  long nk = (long)d->KS[0] * d->KS[1] * d->KS[2];
  for (int c = 0; c < 3; c++) d->buf[c] = (cx *)malloc(sizeof(cx) * nc);
  real *tmp = (real *)malloc(sizeof(real) * np);
  real scale = (real)1 / (real)np;
  double maximg = 0;
  for (int c = 0; c < 6; c++) {
    if (d->is2d && (c == 3 || c == 4)) { d->kern[c] = NULL; continue; }
    for (long i = 0; i < np; i++) tmp[i] = (real)K[c][i];
    r2c3d(tmp, d->buf[0], P, d->H);
    d->kern[c] = (real *)malloc(sizeof(real) * nk);
    for (int iz = 0; iz < d->KS[2]; iz++)
      for (int iy = 0; iy < d->KS[1]; iy++)
        for (int ix = 0; ix < d->KS[0]; ix++) {
          cx v = d->buf[0][((long)iz * P[1] + iy) * d->H + ix];
          d->kern[c][((long)iz * d->KS[1] + iy) * d->KS[0] + ix] = v.re * scale;
          if (fabs((double)v.im) > maximg) maximg = fabs((double)v.im);
        }
  }
  free(tmp);
  /* conv_common.go:43-49: imaginary parts must vanish (FFT_IMAG_TOLERANCE 1e-6) */
  if (maximg * sqrt((double)scale) > 1e-6) {
    fprintf(stderr, "um_oracle: demag kernel imaginary part %g\n", maximg * sqrt((double)scale));
  }
  return d;
}

void *OR(demag_new)(const int N[3], const int pbc[3], const double cs[3], double accuracy,
                    int nthreads) {
  (void)nthreads;
  int P[3];
  um_or_pad_size(N, pbc, P);
  long np = (long)P[0] * P[1] * P[2];
  double *K[6];
  for (int c = 0; c < 6; c++) K[c] = (double *)malloc(sizeof(double) * np);
  if (um_or_demag_kernel_bf(N, pbc, cs, accuracy, K) != 0) {
    for (int c = 0; c < 6; c++) free(K[c]);
    return NULL;
  }
  void *h = OR(demag_new_from_kernel)(N, P, K);
  for (int c = 0; c < 6; c++) free(K[c]);
  return h;
}

void OR(demag_free)(void *h) {
  demag_t *d = (demag_t *)h;
  if (!d) return;
  for (int c = 0; c < 6; c++) free(d->kern[c]);
  for (int c = 0; c < 3; c++) free(d->buf[c]);
  free(d);
}

int OR(demag_kspec)(void *h, int comp, real *out) {
  demag_t *d = (demag_t *)h;
  if (!d->kern[comp]) return -1;
  long nk = (long)d->KS[0] * d->KS[1] * d->KS[2];
  memcpy(out, d->kern[comp], sizeof(real) * nk);
  return 0;
}

/* copypadmul2.cl:3-21 (into a zeroed padded buffer) then R2C */
static void fw_component(demag_t *d, int c, const real *m, const real *vol, const real *Ms_,
                         real Ms_mul, real *pad) {
  const int *N = d->N, *P = d->P;
  long np = (long)P[0] * P[1] * P[2];
  memset(pad, 0, sizeof(real) * np);
  const real MU0 = (real)(4 * M_PI * 1e-7);
  for (int iz = 0; iz < N[2]; iz++)
    for (int iy = 0; iy < N[1]; iy++)
      for (int ix = 0; ix < N[0]; ix++) {
        long sI = ((long)iz * N[1] + iy) * N[0] + ix;
        real Bsat = MU0 * amul(Ms_, Ms_mul, (int)sI);
        real v = amul(vol, (real)1.0, (int)sI);
        pad[((long)iz * P[1] + iy) * P[0] + ix] = Bsat * v * m[sI];
      }
  (void)c;
}
static void bw_component(demag_t *d, cx *spec, real *out, real *pad) {
  const int *N = d->N, *P = d->P;
  c2r3d(spec, pad, P, d->H);
  /* copyunpad.cl:2-13 */
  for (int iz = 0; iz < N[2]; iz++)
    for (int iy = 0; iy < N[1]; iy++)
      for (int ix = 0; ix < N[0]; ix++)
        out[((long)iz * N[1] + iy) * N[0] + ix] = pad[((long)iz * P[1] + iy) * P[0] + ix];
}

/* DemagConvolution.Exec, opencl/conv_demag.go:43-85 */
void OR(demag_exec)(void *h, real *Bx, real *By, real *Bz, const real *mx, const real *my,
                    const real *mz, const real *vol, const real *Ms_, real Ms_mul) {
  demag_t *d = (demag_t *)h;
  const int *P = d->P;
  long np = (long)P[0] * P[1] * P[2];
  real *pad = (real *)malloc(sizeof(real) * np);
  const real *m[3] = {mx, my, mz};
  real *B[3] = {Bx, By, Bz};
  for (int c = 0; c < 3; c++) {
    fw_component(d, c, m[c], vol, Ms_, Ms_mul, pad);
    r2c3d(pad, d->buf[c], P, d->H);
  }
  const int Hh = d->H, Py = P[1], Pz = P[2];
  if (!d->is2d) {
    /* kernmulRSymm3D, kernmulrsymm3d.cl:27-91 */
    const real *Kxx = d->kern[0], *Kyy = d->kern[1], *Kzz = d->kern[2];
    const real *Kyz = d->kern[3], *Kxz = d->kern[4], *Kxy = d->kern[5];
#pragma omp parallel for collapse(2)
    for (int iz = 0; iz < Pz; iz++)
      for (int iy = 0; iy < Py; iy++)
        for (int ix = 0; ix < Hh; ix++) {
          long e = ((long)iz * Py + iy) * Hh + ix;
          cx M0 = d->buf[0][e], M1 = d->buf[1][e], M2 = d->buf[2][e];
          real sYZ = 1, sXZ = 1, sXY = 1;
          int jy = iy, jz = iz;
          if (jy > Py / 2) { jy = Py - jy; sYZ = -sYZ; sXY = -sXY; }
          if (jz > Pz / 2) { jz = Pz - jz; sYZ = -sYZ; sXZ = -sXZ; }
          long I = ((long)jz * (Py / 2 + 1) + jy) * Hh + ix;
          real xx = Kxx[I], yy = Kyy[I], zz = Kzz[I];
          real yz = Kyz[I] * sYZ, xz = Kxz[I] * sXZ, xy = Kxy[I] * sXY;
          d->buf[0][e].re = M0.re * xx + M1.re * xy + M2.re * xz;
          d->buf[0][e].im = M0.im * xx + M1.im * xy + M2.im * xz;
          d->buf[1][e].re = M0.re * xy + M1.re * yy + M2.re * yz;
          d->buf[1][e].im = M0.im * xy + M1.im * yy + M2.im * yz;
          d->buf[2][e].re = M0.re * xz + M1.re * yz + M2.re * zz;
          d->buf[2][e].im = M0.im * xz + M1.im * yz + M2.im * zz;
        }
  } else {
    /* kernmulRSymm2Dz + kernmulRSymm2Dxy, kernmulrsymm2dz.cl:4-29, 2dxy.cl:5-41 */
    const real *Kxx = d->kern[0], *Kyy = d->kern[1], *Kzz = d->kern[2], *Kxy = d->kern[5];
    for (int iy = 0; iy < Py; iy++)
      for (int ix = 0; ix < Hh; ix++) {
        long e = (long)iy * Hh + ix;
        real fxy = 1;
        int jy = iy;
        if (jy > Py / 2) { jy = Py - jy; fxy = -fxy; }
        long I = (long)jy * Hh + ix;
        cx M0 = d->buf[0][e], M1 = d->buf[1][e], M2 = d->buf[2][e];
        real xx = Kxx[I], yy = Kyy[I], xy = fxy * Kxy[I], zz = Kzz[I];
        d->buf[2][e].re = M2.re * zz;
        d->buf[2][e].im = M2.im * zz;
        d->buf[0][e].re = M0.re * xx + M1.re * xy;
        d->buf[0][e].im = M0.im * xx + M1.im * xy;
        d->buf[1][e].re = M0.re * xy + M1.re * yy;
        d->buf[1][e].im = M0.im * xy + M1.im * yy;
      }
  }
  for (int c = 0; c < 3; c++) bw_component(d, d->buf[c], B[c], pad);
  free(pad);
}

/* Brute-force real-space convolution (opencl/conv_selftest.go:64-119): the
 * O(N^2) sum of K(r_i - r_j) M(r_j) with K indexed modulo the padded size.
 * Uses the FP64 real-space tensor directly. */
void OR(demag_brute)(const int N[3], const int P[3], double *const K[6], real *Bx, real *By,
                     real *Bz, const real *mx, const real *my, const real *mz, const real *vol,
                     const real *Ms_, real Ms_mul) {
  const double MU0 = 4 * M_PI * 1e-7;
  long n = (long)N[0] * N[1] * N[2];
  double *M = (double *)malloc(sizeof(double) * 3 * n);
  for (long j = 0; j < n; j++) {
    double s = MU0 * (double)amul(Ms_, Ms_mul, (int)j) * (double)amul(vol, (real)1, (int)j);
    M[3 * j] = s * mx[j]; M[3 * j + 1] = s * my[j]; M[3 * j + 2] = s * mz[j];
  }
#pragma omp parallel for
  for (long i = 0; i < n; i++) {
    int ix = (int)(i % N[0]), iy = (int)((i / N[0]) % N[1]), iz = (int)(i / ((long)N[0] * N[1]));
    double b[3] = {0, 0, 0};
    for (long j = 0; j < n; j++) {
      int jx = (int)(j % N[0]), jy = (int)((j / N[0]) % N[1]), jz = (int)(j / ((long)N[0] * N[1]));
      int dx = ((ix - jx) % P[0] + P[0]) % P[0];
      int dy = ((iy - jy) % P[1] + P[1]) % P[1];
      int dz = ((iz - jz) % P[2] + P[2]) % P[2];
      long k = ((long)dz * P[1] + dy) * P[0] + dx;
      double xx = K[0][k], yy = K[1][k], zz = K[2][k], yz = K[3][k], xz = K[4][k], xy = K[5][k];
      double m0 = M[3 * j], m1 = M[3 * j + 1], m2 = M[3 * j + 2];
      b[0] += xx * m0 + xy * m1 + xz * m2;
      b[1] += xy * m0 + yy * m1 + yz * m2;
      b[2] += xz * m0 + yz * m1 + zz * m2;
    }
    Bx[i] = (real)b[0]; By[i] = (real)b[1]; Bz[i] = (real)b[2];
  }
  free(M);
}

/* ---- thermal normals ------------------------------------------------------ */
/* The reference draws Threefry normals into a staging buffer once per
 * accepted step (engine/temperature.go:62-127, RNG_common.clh:115 Box-Muller).
 * The product replaces it by a counter-based Philox-4x32-10 stream that is
 * recomputed wherever needed: counter = (cell, step lo, step hi, stream),
 * key = seed; words 0,1 -> (nx, ny), words 2,3 -> (nz, unused) by Box-Muller
 * with u = (w + 0.5) * 2^-32.  Computed here in FP64. */
void OR(thermal_normals)(real *nx, real *ny, real *nz, int N, uint64_t seed, uint64_t step,
                         uint32_t stream) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < N; i++) {
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)step, (uint32_t)(step >> 32), stream};
    uint32_t w[4];
    um_or_philox4x32_10(ctr, key, w);
    double u0 = ((double)w[0] + 0.5) * 2.3283064365386963e-10;
    double u1 = ((double)w[1] + 0.5) * 2.3283064365386963e-10;
    double u2 = ((double)w[2] + 0.5) * 2.3283064365386963e-10;
    double u3 = ((double)w[3] + 0.5) * 2.3283064365386963e-10;
    double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
    nx[i] = (real)(r0 * cos(2 * M_PI * u1));
    ny[i] = (real)(r0 * sin(2 * M_PI * u1));
    nz[i] = (real)(r1 * cos(2 * M_PI * u3));
  }
}

/* ========================================================================== */
/* Engine restatement: effective field, torque, steppers, step controller.    */
/* ========================================================================== */
typedef struct { real *arr; double mul; } sparam;
typedef struct {
  int N[3], n;
  double cs[3];
  int pbc[3], pbccode;
  sparam p[UM_P_COUNT][3];           /* scalar params use [0] */
  uint8_t *regions;
  real *aLUT, *dLUT;                 /* symmetric LUTs, UM_OR_SYMLUT entries */
  real *bextLUT[3];                  /* B_ext per region (NULL = none) */
  real *frozenLUT;
  real *nodemagLUT;   /* NoDemagSpins per region (NULL: none set) */
  real *vol;
  int enable_demag, openbc, precess, heun_sqrt3, has_dmi;
  real *dbLUT; int has_dmibulk;      /* bulk DMI (Dbulk) */
  int disable_zl, disable_sl; double flp;   /* STT switches, engine/torque.go:26-29 */
  double demag_acc;
  int newell, newell_arad, newell_zsd; /* EnableNewellDemag and its two settings */
  void *demag;
  uint64_t seed; uint32_t stream;
  real *m, *tmp[16];
  /* solver state, engine/run.go:13-43 */
  int solver;
  double Time, dt, fixdt, maxerr, relerr, minslope, headroom, mindt, maxdt, gammaLL, alarm;
  double lastErr, peakErr, lastTorque;
  long nsteps, nundone, nevals;
  real *k1; int k1_valid;
  int relaxing;                      /* engine/relax.go:22: no thermal field while relaxing */
} sim_t;

#define MU0_D (4 * M_PI * 1e-7)
#define KB_D 1.380650424e-23

void *OR(sim_new)(int Nx, int Ny, int Nz, double cx, double cy, double cz, int pbcx, int pbcy,
                  int pbcz) {
  sim_t *s = (sim_t *)calloc(1, sizeof(sim_t));
  s->N[0] = Nx; s->N[1] = Ny; s->N[2] = Nz;
  s->n = Nx * Ny * Nz;
  s->cs[0] = cx; s->cs[1] = cy; s->cs[2] = cz;
  s->pbc[0] = pbcx; s->pbc[1] = pbcy; s->pbc[2] = pbcz;
  s->pbccode = (pbcx ? 1 : 0) | (pbcy ? 2 : 0) | (pbcz ? 4 : 0);
  s->enable_demag = 1; s->precess = 1; s->heun_sqrt3 = 1;
  s->demag_acc = 6.0;
  s->newell_arad = 32;               /* asymptotic_radius, engine/demag.go:27 */
  s->dbLUT = (real *)calloc(UM_OR_SYMLUT, sizeof(real));
  s->flp = 1.0;                      /* FIXEDLAYER_TOP, engine/torque.go:368-371 */
  s->p[UM_P_POL][0].mul = 1;         /* Pol = 1, Lambda = 1 by default (torque.go:73-74) */
  s->p[UM_P_LAMBDA][0].mul = 1;
  s->aLUT = (real *)calloc(UM_OR_SYMLUT, sizeof(real));
  s->dLUT = (real *)calloc(UM_OR_SYMLUT, sizeof(real));
  s->p[UM_P_ANISU][0].mul = 0; s->p[UM_P_ANISU][1].mul = 0; s->p[UM_P_ANISU][2].mul = 0;
  s->m = (real *)calloc(3 * (size_t)s->n, sizeof(real));
  for (int i = 0; i < 16; i++) s->tmp[i] = (real *)calloc(3 * (size_t)s->n, sizeof(real));
  s->k1 = (real *)calloc(3 * (size_t)s->n, sizeof(real));
  /* solver defaults, engine/run.go:20-25, torque.go:183 */
  s->solver = UM_SOLVER_RK45DP;
  s->dt = 1e-15; s->maxerr = 1e-5; s->relerr = 1e-4; s->minslope = 1e8; s->headroom = 0.8;
  s->gammaLL = 1.7595e11;
  return s;
}

void OR(sim_free)(void *h) {
  sim_t *s = (sim_t *)h;
  if (!s) return;
  for (int i = 0; i < UM_P_COUNT; i++) for (int c = 0; c < 3; c++) free(s->p[i][c].arr);
  free(s->regions); free(s->aLUT); free(s->dLUT); free(s->dbLUT); free(s->frozenLUT); free(s->nodemagLUT); free(s->vol);
  for (int c = 0; c < 3; c++) free(s->bextLUT[c]);
  if (s->demag) OR(demag_free)(s->demag);
  free(s->m); free(s->k1);
  for (int i = 0; i < 16; i++) free(s->tmp[i]);
  free(s);
}

static real *dupn(const real *a, int n) {
  if (!a) return NULL;
  real *r = (real *)malloc(sizeof(real) * n);
  memcpy(r, a, sizeof(real) * n);
  return r;
}
void OR(sim_set_scalar)(void *h, int id, const real *arr, double mul) {
  sim_t *s = (sim_t *)h;
  free(s->p[id][0].arr);
  s->p[id][0].arr = dupn(arr, s->n);
  s->p[id][0].mul = mul;
}
void OR(sim_set_vector)(void *h, int id, const real *ax, const real *ay, const real *az,
                        double mx, double my, double mz) {
  sim_t *s = (sim_t *)h;
  const real *a[3] = {ax, ay, az};
  double m[3] = {mx, my, mz};
  for (int c = 0; c < 3; c++) {
    free(s->p[id][c].arr);
    s->p[id][c].arr = dupn(a[c], s->n);
    s->p[id][c].mul = m[c];
  }
}
void OR(sim_set_regions)(void *h, const uint8_t *regions) {
  sim_t *s = (sim_t *)h;
  free(s->regions);
  s->regions = NULL;
  if (regions) {
    s->regions = (uint8_t *)malloc(s->n);
    memcpy(s->regions, regions, s->n);
  }
}
void OR(sim_set_exchange_luts)(void *h, const real *aLUT, const real *dLUT) {
  sim_t *s = (sim_t *)h;
  memcpy(s->aLUT, aLUT, sizeof(real) * UM_OR_SYMLUT);
  s->has_dmi = 0;
  if (dLUT) {
    memcpy(s->dLUT, dLUT, sizeof(real) * UM_OR_SYMLUT);
    for (int i = 0; i < UM_OR_SYMLUT; i++) if (dLUT[i] != 0) s->has_dmi = 1;
  } else {
    memset(s->dLUT, 0, sizeof(real) * UM_OR_SYMLUT);
  }
}
void OR(sim_set_dmi_bulk)(void *h, const real *lut) {
  sim_t *s = (sim_t *)h;
  s->has_dmibulk = 0;
  if (lut) {
    memcpy(s->dbLUT, lut, sizeof(real) * UM_OR_SYMLUT);
    for (int i = 0; i < UM_OR_SYMLUT; i++) if (lut[i] != 0) s->has_dmibulk = 1;
  } else {
    memset(s->dbLUT, 0, sizeof(real) * UM_OR_SYMLUT);
  }
}
void OR(sim_set_stt_flags)(void *h, int disable_zl, int disable_sl, double flp) {
  sim_t *s = (sim_t *)h;
  s->disable_zl = disable_zl; s->disable_sl = disable_sl; s->flp = flp;
}
void OR(sim_set_bext_lut)(void *h, const real *lx, const real *ly, const real *lz) {
  sim_t *s = (sim_t *)h;
  const real *l[3] = {lx, ly, lz};
  for (int c = 0; c < 3; c++) {
    free(s->bextLUT[c]);
    s->bextLUT[c] = dupn(l[c], UM_OR_NREGION);
  }
}
/* NoDemagSpins (engine/demag.go:22-23): per-region flag, NULL clears */
void OR(sim_set_nodemag_lut)(void *h, const real *lut) {
  sim_t *s = (sim_t *)h;
  free(s->nodemagLUT);
  s->nodemagLUT = NULL;
  if (!lut) return;
  for (int r = 0; r < UM_OR_NREGION; r++)
    if (lut[r] != 0) { s->nodemagLUT = dupn(lut, UM_OR_NREGION); break; }
}
void OR(sim_set_frozen_lut)(void *h, const real *lut) {
  sim_t *s = (sim_t *)h;
  free(s->frozenLUT);
  s->frozenLUT = dupn(lut, UM_OR_NREGION);
}
void OR(sim_set_geometry)(void *h, const real *vol) {
  sim_t *s = (sim_t *)h;
  free(s->vol);
  s->vol = dupn(vol, s->n);
}
void OR(sim_set_flags)(void *h, int enable_demag, int openbc, int precess, int heun_sqrt3) {
  sim_t *s = (sim_t *)h;
  s->enable_demag = enable_demag; s->openbc = openbc; s->precess = precess; s->heun_sqrt3 = heun_sqrt3;
}
void OR(sim_set_demag_accuracy)(void *h, double acc) {
  sim_t *s = (sim_t *)h;
  s->demag_acc = acc;
  if (s->demag) { OR(demag_free)(s->demag); s->demag = NULL; }
}
void OR(sim_set_demag_newell)(void *h, int enable, int asymptotic_radius, int zero_self_demag) {
  sim_t *s = (sim_t *)h;
  s->newell = enable; s->newell_arad = asymptotic_radius; s->newell_zsd = zero_self_demag;
  if (s->demag) { OR(demag_free)(s->demag); s->demag = NULL; }
}
static void *sim_newell_plan(sim_t *s) {
  int P[3];
  um_or_pad_size(s->N, s->pbc, P);
  long np = (long)P[0] * P[1] * P[2];
  double *K[6];
  for (int c = 0; c < 6; c++) K[c] = (double *)malloc(sizeof(double) * np);
  void *h = NULL;
  if (um_or_demag_kernel_newell(s->N, s->pbc, s->cs, s->newell_arad, s->newell_zsd, K) == 0)
    h = OR(demag_new_from_kernel)(s->N, P, K);
  for (int c = 0; c < 6; c++) free(K[c]);
  return h;
}
void OR(sim_set_thermal_seed)(void *h, uint64_t seed, uint32_t stream) {
  sim_t *s = (sim_t *)h;
  s->seed = seed; s->stream = stream;
}
void OR(sim_set_m)(void *h, const real *mx, const real *my, const real *mz) {
  sim_t *s = (sim_t *)h;
  memcpy(s->m, mx, sizeof(real) * s->n);
  memcpy(s->m + s->n, my, sizeof(real) * s->n);
  memcpy(s->m + 2 * s->n, mz, sizeof(real) * s->n);
  /* magnetization.SetArray normalises, engine/magnetization.go:45-51 */
  OR(normalize2)(s->m, s->m + s->n, s->m + 2 * s->n, s->vol, s->n);
  s->k1_valid = 0;
}
void OR(sim_get_m)(void *h, real *mx, real *my, real *mz) {
  sim_t *s = (sim_t *)h;
  memcpy(mx, s->m, sizeof(real) * s->n);
  memcpy(my, s->m + s->n, sizeof(real) * s->n);
  memcpy(mz, s->m + 2 * s->n, sizeof(real) * s->n);
}
static int param_nonzero(sim_t *s, int id, int nc) {
  for (int c = 0; c < nc; c++) {
    if (s->p[id][c].mul == 0) continue;
    if (!s->p[id][c].arr) return 1;
    for (int i = 0; i < s->n; i++) if (s->p[id][c].arr[i] != 0) return 1;
  }
  return 0;
}
#define PA(id, c) s->p[id][c].arr, (real)s->p[id][c].mul
#define PA_ARR(id) ((const real *)s->p[id][0].arr)
#define PA_MUL(id) ((real)s->p[id][0].mul)

/* SetDemagField, engine/demag.go:42-65; with NoDemagSpins set, the masked
 * field of setMaskedDemagField (:67-98): the geometry (or 1) zeroed in the
 * flagged regions is the convolution's vol, and the field is zeroed there. */
static void set_demag(sim_t *s, real *B) {
  int n = s->n;
  if (!s->enable_demag) { memset(B, 0, sizeof(real) * 3 * n); return; }
  if (!s->demag)
    s->demag = s->newell ? sim_newell_plan(s) : OR(demag_new)(s->N, s->pbc, s->cs, s->demag_acc, 0);
  if (!s->nodemagLUT) {
    OR(demag_exec)(s->demag, B, B + n, B + 2 * n, s->m, s->m + n, s->m + 2 * n, s->vol, PA(UM_P_MSAT, 0));
    return;
  }
  real *buf = (real *)malloc(sizeof(real) * n);
  for (int i = 0; i < n; i++) buf[i] = s->vol ? s->vol[i] : (real)1;
  OR(zeromask)(buf, s->nodemagLUT, s->regions, n);
  OR(demag_exec)(s->demag, B, B + n, B + 2 * n, s->m, s->m + n, s->m + 2 * n, buf, PA(UM_P_MSAT, 0));
  free(buf);
  for (int c = 0; c < 3; c++) OR(zeromask)(B + c * n, s->nodemagLUT, s->regions, n);
}

/* SetEffectiveField, engine/effectivefield.go:14-24 */
/* B += B_therm of the current step (temperature.go:55-127, temperature2.cl) */
static void add_thermal(sim_t *s, real *B) {
  int n = s->n;
  real *eta = s->tmp[15];
  OR(thermal_normals)(eta, eta + n, eta + 2 * n, n, s->seed, (uint64_t)s->nsteps, s->stream);
  double V = s->cs[0] * s->cs[1] * s->cs[2];
  real k2 = (real)(2 * KB_D / (s->gammaLL * V * s->dt));
  real *bt = s->tmp[14];
  for (int c = 0; c < 3; c++) {
    OR(settemperature2)(bt + c * n, eta + c * n, k2, PA(UM_P_MSAT, 0), PA(UM_P_TEMP, 0), PA(UM_P_ALPHA, 0), n);
    for (int i = 0; i < n; i++) B[c * n + i] += bt[c * n + i];
  }
}

/* AddExchangeField, engine/exchange.go:49-66: exchange, + interfacial or + bulk DMI */
static void add_exchange_field(sim_t *s, real *B) {
  int n = s->n;
  real *m = s->m;
  if (s->has_dmi) {
    OR(adddmi)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), s->aLUT, s->dLUT, s->regions,
               (real)s->cs[0], (real)s->cs[1], (real)s->cs[2], s->N[0], s->N[1], s->N[2], s->pbccode,
               s->openbc);
  } else if (s->has_dmibulk) {
    OR(adddmibulk)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), s->aLUT, s->dbLUT,
                   s->regions, (real)s->cs[0], (real)s->cs[1], (real)s->cs[2], s->N[0], s->N[1], s->N[2],
                   s->pbccode, s->openbc);
  } else {
    real wx = (real)(2 / (s->cs[0] * s->cs[0])), wy = (real)(2 / (s->cs[1] * s->cs[1])),
         wz = (real)(2 / (s->cs[2] * s->cs[2]));
    OR(addexchange)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), s->aLUT, s->regions,
                    wx, wy, wz, s->N[0], s->N[1], s->N[2], s->pbccode);
  }
}

static void set_beff(sim_t *s, real *B) {
  int n = s->n;
  real *m = s->m;
  set_demag(s, B);
  add_exchange_field(s, B);
  /* AddAnisotropyField, engine/anisotropy.go:156-161 */
  if (param_nonzero(s, UM_P_KU1, 1) || param_nonzero(s, UM_P_KU2, 1))
    OR(adduniaxial2)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), PA(UM_P_KU1, 0),
                     PA(UM_P_KU2, 0), PA(UM_P_ANISU, 0), PA(UM_P_ANISU, 1), PA(UM_P_ANISU, 2), n);
  if (param_nonzero(s, UM_P_KC1, 1) || param_nonzero(s, UM_P_KC2, 1) || param_nonzero(s, UM_P_KC3, 1))
    OR(addcubic2)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), PA(UM_P_KC1, 0),
                  PA(UM_P_KC2, 0), PA(UM_P_KC3, 0), PA(UM_P_ANISC1, 0), PA(UM_P_ANISC1, 1),
                  PA(UM_P_ANISC1, 2), PA(UM_P_ANISC2, 0), PA(UM_P_ANISC2, 1), PA(UM_P_ANISC2, 2), n);
  /* B_ext.AddTo, engine/excitation.go:41-53: region LUT, then uniform/array part */
  if (s->bextLUT[0])
    OR(regionaddv)(B, B + n, B + 2 * n, s->bextLUT[0], s->bextLUT[1], s->bextLUT[2], s->regions, n);
  if (param_nonzero(s, UM_P_BEXT, 3))
    for (int c = 0; c < 3; c++)
      for (int i = 0; i < n; i++) B[c * n + i] += amul(PA(UM_P_BEXT, c), i);
  /* B_therm.AddTo, engine/temperature.go:55-127; skipped while relaxing (effectivefield.go:20) */
  if (param_nonzero(s, UM_P_TEMP, 1) && !s->relaxing) add_thermal(s, B);
}

/* SetTorque, engine/torque.go:124-144, FreezeSpins :348-352 */
static void torque_fn(sim_t *s, real *dst) {
  int n = s->n;
  set_beff(s, dst);
  if (s->precess)
    OR(lltorque2)(dst, dst + n, dst + 2 * n, s->m, s->m + n, s->m + 2 * n, dst, dst + n, dst + 2 * n,
                  PA(UM_P_ALPHA, 0), n);
  else
    OR(llnoprecess)(dst, dst + n, dst + 2 * n, s->m, s->m + n, s->m + 2 * n, dst, dst + n, dst + 2 * n, n);
  /* AddSTTorque, engine/torque.go:147-196 (J != 0) */
  if (param_nonzero(s, UM_P_J, 3)) {
    if (!s->disable_zl)
      OR(addzhangli)(dst, dst + n, dst + 2 * n, s->m, s->m + n, s->m + 2 * n, PA(UM_P_MSAT, 0),
                     PA(UM_P_J, 0), PA(UM_P_J, 1), PA(UM_P_J, 2), PA(UM_P_ALPHA, 0), PA(UM_P_XI, 0),
                     PA(UM_P_POL, 0), (real)s->cs[0], (real)s->cs[1], (real)s->cs[2], s->N[0], s->N[1],
                     s->N[2], s->pbccode);
    if (!s->disable_sl && param_nonzero(s, UM_P_FIXEDLAYER, 3))
      OR(addslonczewski)(dst, dst + n, dst + 2 * n, s->m, s->m + n, s->m + 2 * n, PA(UM_P_MSAT, 0),
                         PA(UM_P_J, 2), PA(UM_P_FIXEDLAYER, 0), PA(UM_P_FIXEDLAYER, 1), PA(UM_P_FIXEDLAYER, 2),
                         PA(UM_P_ALPHA, 0), PA(UM_P_POL, 0), PA(UM_P_LAMBDA, 0), PA(UM_P_EPSPRIME, 0),
                         PA(UM_P_FREELAYER, 0), (real)(s->N[2] * s->cs[2]), (real)s->flp, n);
  }
  if (s->frozenLUT)
    for (int c = 0; c < 3; c++) OR(zeromask)(dst + c * n, s->frozenLUT, s->regions, n);
  s->nevals++;
}

void OR(sim_beff)(void *h, real *bx, real *by, real *bz) {
  sim_t *s = (sim_t *)h;
  real *B = s->tmp[13];
  set_beff(s, B);
  memcpy(bx, B, sizeof(real) * s->n);
  memcpy(by, B + s->n, sizeof(real) * s->n);
  memcpy(bz, B + 2 * s->n, sizeof(real) * s->n);
}
void OR(sim_demag_field)(void *h, real *bx, real *by, real *bz) {
  sim_t *s = (sim_t *)h;
  real *B = s->tmp[13];
  set_demag(s, B);
  memcpy(bx, B, sizeof(real) * s->n);
  memcpy(by, B + s->n, sizeof(real) * s->n);
  memcpy(bz, B + 2 * s->n, sizeof(real) * s->n);
}
void OR(sim_torque)(void *h, real *tx, real *ty, real *tz) {
  sim_t *s = (sim_t *)h;
  real *T = s->tmp[13];
  torque_fn(s, T);
  s->nevals--;
  memcpy(tx, T, sizeof(real) * s->n);
  memcpy(ty, T + s->n, sizeof(real) * s->n);
  memcpy(tz, T + 2 * s->n, sizeof(real) * s->n);
}
void OR(sim_set_solver)(void *h, int solver) {
  sim_t *s = (sim_t *)h;
  s->solver = solver;
  s->k1_valid = 0;
}
void OR(sim_set_solver_params)(void *h, const double *k) {
  sim_t *s = (sim_t *)h;
  s->dt = k[0]; s->fixdt = k[1]; s->maxerr = k[2]; s->relerr = k[3]; s->minslope = k[4];
  s->headroom = k[5]; s->mindt = k[6]; s->maxdt = k[7]; s->gammaLL = k[8];
}
void OR(sim_get_state)(void *h, double *o) {
  sim_t *s = (sim_t *)h;
  o[0] = s->Time; o[1] = s->dt; o[2] = (double)s->nsteps; o[3] = (double)s->nundone;
  o[4] = (double)s->nevals; o[5] = s->lastErr; o[6] = s->peakErr; o[7] = s->lastTorque;
}

static void set_last_err(sim_t *s, double err) {
  s->lastErr = err;
  if (err > s->peakErr) s->peakErr = err;
}
static void adapt(sim_t *s, double corr) {
  double dt = um_or_adapt_dt(s->dt, corr, s->fixdt, s->headroom, s->mindt, s->maxdt, s->Time, s->alarm);
  if (dt < 0) { fprintf(stderr, "um_oracle: time step too small\n"); dt = s->dt; }
  s->dt = dt;
}
static void copy3(sim_t *s, real *dst, const real *src) { memcpy(dst, src, sizeof(real) * 3 * s->n); }
static void madd_vec(sim_t *s, real *dst, int k, const real *const *src, const real *fac) {
  int n = s->n;
  for (int c = 0; c < 3; c++) {
    const real *sc[8];
    for (int j = 0; j < k; j++) sc[j] = src[j] + c * n;
    OR(madd)(dst + c * n, k, sc, fac, n);
  }
}
static double maxvecnorm3(sim_t *s, const real *v) {
  int n = s->n;
  return OR(maxvecnorm)(v, v + n, v + 2 * n, n);
}
static void normalize_m(sim_t *s) {
  int n = s->n;
  OR(normalize2)(s->m, s->m + n, s->m + 2 * n, s->vol, n);
}

/* integralController, engine/stepcontroller.go:35-97 */
static void integral_controller(sim_t *s, const real *Err, const real *delM, real *k1, const real *m0,
                                double t0, double h, int accOrder, int rejOrder, int FSAL) {
  int n = s->n;
  double err = maxvecnorm3(s, Err) * h;
  if (err < s->maxerr || s->dt <= s->mindt || s->fixdt != 0) {
    real *errnorm = s->tmp[10], *ddtnorm = s->tmp[11];
    for (int i = 0; i < n; i++) {
      errnorm[i] = (real)sqrt((double)(Err[i] * Err[i] + Err[n + i] * Err[n + i] + Err[2 * n + i] * Err[2 * n + i]));
      ddtnorm[i] = (real)sqrt((double)(delM[i] * delM[i] + delM[n + i] * delM[n + i] + delM[2 * n + i] * delM[2 * n + i]));
    }
    double maxdm = maxvecnorm3(s, delM);
    int fail = 0;
    double rlerr = 0;
    if (maxdm < s->minslope) {
      fail = 0;
    } else {
      for (int i = 0; i < n; i++) errnorm[i] = (ddtnorm[i] == 0) ? (real)0 : errnorm[i] / ddtnorm[i];
      rlerr = OR(maxabs)(errnorm, n);
      fail = 1;
    }
    if (fail == 0 || s->relerr <= 0.0 || rlerr < s->relerr || s->dt <= s->mindt || s->fixdt != 0) {
      set_last_err(s, err);
      s->lastTorque = maxvecnorm3(s, delM);
      s->nsteps++;
      s->Time = t0 + s->dt;
      if (fail == 0) adapt(s, pow(s->maxerr / err, 1. / accOrder));
      else adapt(s, pow(s->relerr / rlerr, 1. / accOrder));
      if (FSAL && k1 && delM) { copy3(s, k1, delM); s->k1_valid = 1; }
    } else {
      s->Time = t0;
      copy3(s, s->m, m0);
      s->nundone++;
      adapt(s, pow(s->relerr / rlerr, 1. / rejOrder));
    }
  } else {
    s->Time = t0;
    copy3(s, s->m, m0);
    s->nundone++;
    adapt(s, pow(s->maxerr / err, 1. / rejOrder));
  }
}

/* RK45DP.Step, engine/rk45dp.go:12-98 */
static void step_rk45dp(sim_t *s) {
  real *m = s->m;
  if (s->fixdt != 0) s->dt = s->fixdt;
  if (!s->k1_valid) { torque_fn(s, s->k1); s->k1_valid = 1; }
  if (param_nonzero(s, UM_P_TEMP, 1)) torque_fn(s, s->k1);
  double t0 = s->Time;
  real *m0 = s->tmp[0], *k2 = s->tmp[1], *k3 = s->tmp[2], *k4 = s->tmp[3], *k5 = s->tmp[4],
       *k6 = s->tmp[5], *Err = s->tmp[6];
  real *k1 = s->k1;
  copy3(s, m0, m);
  real h = (real)(s->dt * s->gammaLL);
  {
    const real *src[] = {m, k1};
    real fac[] = {1, (real)(1. / 5.) * h};
    s->Time = t0 + (1. / 5.) * s->dt;
    madd_vec(s, m, 2, src, fac); normalize_m(s); torque_fn(s, k2);
  }
  {
    const real *src[] = {m0, k1, k2};
    real fac[] = {1, (real)(3. / 40.) * h, (real)(9. / 40.) * h};
    s->Time = t0 + (3. / 10.) * s->dt;
    madd_vec(s, m, 3, src, fac); normalize_m(s); torque_fn(s, k3);
  }
  {
    const real *src[] = {m0, k1, k2, k3};
    real fac[] = {1, (real)(44. / 45.) * h, (real)(-56. / 15.) * h, (real)(32. / 9.) * h};
    s->Time = t0 + (4. / 5.) * s->dt;
    madd_vec(s, m, 4, src, fac); normalize_m(s); torque_fn(s, k4);
  }
  {
    const real *src[] = {m0, k1, k2, k3, k4};
    real fac[] = {1, (real)(19372. / 6561.) * h, (real)(-25360. / 2187.) * h, (real)(64448. / 6561.) * h,
                  (real)(-212. / 729.) * h};
    s->Time = t0 + (8. / 9.) * s->dt;
    madd_vec(s, m, 5, src, fac); normalize_m(s); torque_fn(s, k5);
  }
  {
    const real *src[] = {m0, k1, k2, k3, k4, k5};
    real fac[] = {1, (real)(9017. / 3168.) * h, (real)(-355. / 33.) * h, (real)(46732. / 5247.) * h,
                  (real)(49. / 176.) * h, (real)(-5103. / 18656.) * h};
    s->Time = t0 + s->dt;
    madd_vec(s, m, 6, src, fac); normalize_m(s); torque_fn(s, k6);
  }
  real *k7 = k2;
  {
    const real *src[] = {m0, k1, k3, k4, k5, k6};
    real fac[] = {1, (real)(35. / 384.) * h, (real)(500. / 1113.) * h, (real)(125. / 192.) * h,
                  (real)(-2187. / 6784.) * h, (real)(11. / 84.) * h};
    s->Time = t0 + s->dt;
    madd_vec(s, m, 6, src, fac); normalize_m(s); torque_fn(s, k7);
  }
  {
    const real *src[] = {k1, k3, k4, k5, k6, k7};
    real fac[] = {(real)(35. / 384. - 5179. / 57600.), (real)(500. / 1113. - 7571. / 16695.),
                  (real)(125. / 192. - 393. / 640.), (real)(-2187. / 6784. - -92097. / 339200.),
                  (real)(11. / 84. - 187. / 2100.), (real)(0. - 1. / 40.)};
    madd_vec(s, Err, 6, src, fac);
  }
  integral_controller(s, Err, k7, k1, m0, t0, (double)h, 5, 6, 1);
}

/* Heun.Step, engine/heun.go:15-59 */
static void step_heun(sim_t *s) {
  int n = s->n;
  real *y = s->m, *dy0 = s->tmp[0], *m0 = s->tmp[1], *dy = s->tmp[2];
  copy3(s, m0, y);
  if (s->fixdt != 0) s->dt = s->fixdt;
  real dt = (real)(s->dt * s->gammaLL);
  torque_fn(s, dy0);
  {
    const real *src[] = {y, dy0};
    real fac[] = {1, dt};
    madd_vec(s, y, 2, src, fac);
  }
  s->Time += s->dt;
  torque_fn(s, dy);
  double err = OR(maxvecdiff)(dy0, dy0 + n, dy0 + 2 * n, dy, dy + n, dy + 2 * n, n, s->heun_sqrt3) * (double)dt;
  if (err < s->maxerr || s->dt <= s->mindt || s->fixdt != 0) {
    const real *src[] = {y, dy, dy0};
    real fac[] = {1, (real)0.5 * dt, (real)-0.5 * dt};
    madd_vec(s, y, 3, src, fac);
    normalize_m(s);
    s->nsteps++;
    adapt(s, pow(s->maxerr / err, 1. / 2.));
    set_last_err(s, err);
    s->lastTorque = maxvecnorm3(s, dy);
  } else {
    s->Time -= s->dt;
    copy3(s, y, m0);
    s->nundone++;
    adapt(s, pow(s->maxerr / err, 1. / 3.));
  }
}

/* RK4.Step, engine/rk4.go:11-69 and simpleController, stepcontroller.go:13-32 */
static void step_rk4(sim_t *s) {
  real *m = s->m;
  if (s->fixdt != 0) s->dt = s->fixdt;
  double t0 = s->Time;
  real *m0 = s->tmp[0], *k1 = s->tmp[1], *k2 = s->tmp[2], *k3 = s->tmp[3], *k4 = s->tmp[4], *Err = s->tmp[5];
  copy3(s, m0, m);
  real h = (real)(s->dt * s->gammaLL);
  torque_fn(s, k1);
  s->Time = t0 + (1. / 2.) * s->dt;
  { const real *src[] = {m, k1}; real fac[] = {1, (real)(1. / 2.) * h}; madd_vec(s, m, 2, src, fac); }
  normalize_m(s); torque_fn(s, k2);
  { const real *src[] = {m0, k2}; real fac[] = {1, (real)(1. / 2.) * h}; madd_vec(s, m, 2, src, fac); }
  normalize_m(s); torque_fn(s, k3);
  s->Time = t0 + s->dt;
  { const real *src[] = {m0, k3}; real fac[] = {1, (real)1. * h}; madd_vec(s, m, 2, src, fac); }
  normalize_m(s); torque_fn(s, k4);
  { const real *src[] = {k1, k4}; real fac[] = {1, -1}; madd_vec(s, Err, 2, src, fac); }
  double err = maxvecnorm3(s, Err) * (double)h;
  int ok;
  if (err < s->maxerr || s->dt <= s->mindt || s->fixdt != 0) {
    s->nsteps++;
    adapt(s, pow(s->maxerr / err, 1. / 4.));
    set_last_err(s, err);
    ok = 1;
  } else {
    s->nundone++;
    adapt(s, pow(s->maxerr / err, 1. / 5.));
    ok = 0;
  }
  if (ok) {
    const real *src[] = {m0, k1, k2, k3, k4};
    real fac[] = {1, (real)(1. / 6.) * h, (real)(1. / 3.) * h, (real)(1. / 3.) * h, (real)(1. / 6.) * h};
    madd_vec(s, m, 5, src, fac);
    normalize_m(s);
    s->lastTorque = maxvecnorm3(s, k4);
  } else {
    s->Time = t0;
    copy3(s, m, m0);
  }
}

/* RK23 (Bogacki-Shampine).Step, engine/rk23.go:21-84, with integralController
 * (accept order 3, reject order 4, FSAL) */
static void step_rk23(sim_t *s) {
  real *m = s->m;
  if (s->fixdt != 0) s->dt = s->fixdt;
  if (!s->k1_valid) { torque_fn(s, s->k1); s->k1_valid = 1; }
  if (param_nonzero(s, UM_P_TEMP, 1)) torque_fn(s, s->k1);   /* FSAL is not used with temperature */
  double t0 = s->Time;
  real *m0 = s->tmp[0], *k2 = s->tmp[1], *k3 = s->tmp[2], *k4 = s->tmp[3];
  real *k1 = s->k1;
  copy3(s, m0, m);
  real h = (real)(s->dt * s->gammaLL);
  s->Time = t0 + (1. / 2.) * s->dt;
  { const real *src[] = {m, k1}; real fac[] = {1, (real)(1. / 2.) * h}; madd_vec(s, m, 2, src, fac); }
  normalize_m(s); torque_fn(s, k2);
  s->Time = t0 + (3. / 4.) * s->dt;
  { const real *src[] = {m0, k2}; real fac[] = {1, (real)(3. / 4.) * h}; madd_vec(s, m, 2, src, fac); }
  normalize_m(s); torque_fn(s, k3);
  {
    const real *src[] = {m0, k1, k2, k3};
    real fac[] = {1, (real)(2. / 9.) * h, (real)(1. / 3.) * h, (real)(4. / 9.) * h};
    madd_vec(s, m, 4, src, fac);
  }
  normalize_m(s);
  s->Time = t0 + s->dt;
  torque_fn(s, k4);
  real *Err = k2;   /* re-use k2 */
  {
    const real *src[] = {k1, k2, k3, k4};
    real fac[] = {(real)(7. / 24. - 2. / 9.), (real)(1. / 4. - 1. / 3.), (real)(1. / 3. - 4. / 9.), (real)(1. / 8.)};
    madd_vec(s, Err, 4, src, fac);
  }
  integral_controller(s, Err, k4, k1, m0, t0, (double)h, 3, 4, 1);
}

void OR(sim_step)(void *h) {
  sim_t *s = (sim_t *)h;
  switch (s->solver) {
  case UM_SOLVER_HEUN: step_heun(s); break;
  case UM_SOLVER_RK23: step_rk23(s); break;
  case UM_SOLVER_RK4: step_rk4(s); break;
  default: step_rk45dp(s); break;
  }
}
/* Steps(n), engine/run.go:197-200 (RunWhile drops the FSAL torque, :207) */
void OR(sim_steps)(void *h, int n) {
  sim_t *s = (sim_t *)h;
  s->k1_valid = 0;
  long stop = s->nsteps + n;
  while (s->nsteps < stop) OR(sim_step)(h);
}
/* Run(seconds), engine/run.go:190-194 */
void OR(sim_run)(void *h, double seconds) {
  sim_t *s = (sim_t *)h;
  s->k1_valid = 0;
  double stop = s->Time + seconds;
  s->alarm = stop;
  while (s->Time < stop) OR(sim_step)(h);
}

/* relaxSteps, engine/relax.go:104-112: n steps, Time restored */
static void relax_steps(sim_t *s, int n) {
  double t0 = s->Time;
  long stop = s->nsteps + n;
  while (s->nsteps < stop) OR(sim_step)(s);
  s->Time = t0;
}
static double selfdot3(sim_t *s, const real *a) {
  double d = 0;
  for (long i = 0; i < 3L * s->n; i++) d += (double)a[i] * (double)a[i];
  return d;
}
/* Relax, engine/relax.go:24-100 */
/* Minimize, engine/minimizer.go:56-160: steepest descent (minimize.cl) with
 * Barzilai-Borwein step lengths alternating on NSteps parity, torque without
 * precession and without thermal noise; stops when the last `samples` max|dm|
 * are all <= stop.  Time does not advance. */
void OR(sim_minimize)(void *h, double stop, int samples) {
  sim_t *s = (sim_t *)h;
  int n = s->n;
  int prev_precess = s->precess;
  s->precess = 0;
  s->relaxing = 1;
  double *ring = (double *)calloc(samples, sizeof(double));
  int count = 0, tail = 0;
  real *k = s->tmp[0], *k0 = s->tmp[1], *m0 = s->tmp[2], *dm = s->tmp[3], *dk = s->tmp[4];
  real hh = (real)1e-4;
  torque_fn(s, k);
  for (;;) {
    double rmax = ring[0];
    for (int i = 1; i < count; i++) if (ring[i] > rmax) rmax = ring[i];
    if (!(count < samples || rmax > stop)) break;
    memcpy(m0, s->m, sizeof(real) * 3 * n);
    /* minimize.cl:11-20 */
    for (int i = 0; i < n; i++) {
      real t[3] = {k[i], k[n + i], k[2 * n + i]};
      real t2 = hh * hh * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      real div = (real)4.0 + t2;
      for (int c = 0; c < 3; c++)
        s->m[c * n + i] = (((real)4.0 - t2) * m0[c * n + i] + (real)4.0 * hh * t[c]) / div;
    }
    memcpy(k0, k, sizeof(real) * 3 * n);
    torque_fn(s, k);
    s->lastTorque = OR(maxvecnorm)(k, k + n, k + 2 * n, n);
    for (int i = 0; i < 3 * n; i++) { dm[i] = s->m[i] - m0[i]; dk[i] = -k[i] + k0[i]; }
    ring[tail] = OR(maxvecnorm)(dm, dm + n, dm + 2 * n, n);
    tail = (tail + 1) % samples;
    if (count < samples) count++;
    rmax = ring[0];
    for (int i = 1; i < count; i++) if (ring[i] > rmax) rmax = ring[i];
    s->lastErr = rmax;
    if (rmax > s->peakErr) s->peakErr = rmax;
    real nom, dv;
    if (s->nsteps % 2 == 0) { nom = (real)OR(dot)(dm, dm, 3 * n); dv = (real)OR(dot)(dm, dk, 3 * n); }
    else { nom = (real)OR(dot)(dm, dk, 3 * n); dv = (real)OR(dot)(dk, dk, 3 * n); }
    hh = dv != (real)0 ? nom / dv : (real)1e-4;
    normalize_m(s);
    s->nsteps++;
  }
  free(ring);
  s->precess = prev_precess;
  s->relaxing = 0;
  s->k1_valid = 0;
}

void OR(sim_relax)(void *h, double torque_threshold) {
  sim_t *s = (sim_t *)h;
  int prev_solver = s->solver, prev_precess = s->precess;
  double prev_maxerr = s->maxerr, prev_fixdt = s->fixdt;
  s->solver = UM_SOLVER_RK23;
  s->k1_valid = 0;
  s->fixdt = 0;
  s->precess = 0;
  s->relaxing = 1;
  const int N = 3;
  relax_steps(s, N);
  double E0 = OR(sim_energy)(h, 9);
  relax_steps(s, N);
  double E1 = OR(sim_energy)(h, 9);
  while (E1 < E0) {
    relax_steps(s, N);
    E0 = E1;
    E1 = OR(sim_energy)(h, 9);
  }
  if (torque_threshold > 0) {
    for (;;) {
      while (maxvecnorm3(s, s->k1) > torque_threshold) relax_steps(s, N);
      s->maxerr /= sqrt(2.0);
      if (s->maxerr < 1e-9) break;
    }
  } else {
    double T0 = 0, T1 = selfdot3(s, s->k1);
    while (s->maxerr > 1e-9) {
      s->maxerr /= sqrt(2.0);
      relax_steps(s, N);
      T0 = T1; T1 = selfdot3(s, s->k1);
      while (T1 < T0) {
        relax_steps(s, N);
        T0 = T1; T1 = selfdot3(s, s->k1);
      }
    }
  }
  s->solver = prev_solver;
  s->maxerr = prev_maxerr;
  s->fixdt = prev_fixdt;
  s->precess = prev_precess;
  s->relaxing = 0;
  s->k1_valid = 0;
}
void OR(sim_average_m)(void *h, double out[3]) {
  sim_t *s = (sim_t *)h;
  /* sAverageMagnet, engine/average.go:28-40: Sum(m_c)/NCell, or with a
   * geometry Dot(m_c, vol)/Sum(vol) */
  double w = 0;
  for (int c = 0; c < 3; c++) out[c] = 0;
  for (int i = 0; i < s->n; i++) {
    double v = s->vol ? (double)s->vol[i] : 1.0;
    w += v;
    for (int c = 0; c < 3; c++) out[c] += (double)s->m[c * s->n + i] * v;
  }
  if (w > 0) for (int c = 0; c < 3; c++) out[c] /= w;
}
/* Energies, engine/energy.go:27-61: term 0 demag (-1/2), 1 exchange (-1/2),
 * 2 Zeeman (-1), 3 anisotropy (via its field: -1/2 of K1 part only when K2=0),
 * 4 total of 0..2 + anisotropy. E = prefactor * V * sum(M_full . B_term). */
double OR(sim_energy)(void *h, int term) {
  sim_t *s = (sim_t *)h;
  int n = s->n;
  double V = s->cs[0] * s->cs[1] * s->cs[2];
  real *B = s->tmp[12], *m = s->m;
  double e = 0;
  memset(B, 0, sizeof(real) * 3 * n);
  if (term == 0) { set_demag(s, B); }
  else if (term == 1) {
    add_exchange_field(s, B);
  } else if (term == 2) {
    if (s->bextLUT[0])
      OR(regionaddv)(B, B + n, B + 2 * n, s->bextLUT[0], s->bextLUT[1], s->bextLUT[2], s->regions, n);
    for (int c = 0; c < 3; c++)
      for (int i = 0; i < n; i++) B[c * n + i] += amul(PA(UM_P_BEXT, c), i);
  } else if (term == 3 || term == 4) {
    /* AddAnisotropyEnergyDensity, engine/anisotropy.go:163-190: K1 and K2 separately */
    if (term == 3)
      OR(adduniaxial2)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), PA(UM_P_KU1, 0),
                       (const real *)NULL, (real)0, PA(UM_P_ANISU, 0), PA(UM_P_ANISU, 1), PA(UM_P_ANISU, 2), n);
    else
      OR(adduniaxial2)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), (const real *)NULL, (real)0,
                       PA(UM_P_KU2, 0), PA(UM_P_ANISU, 0), PA(UM_P_ANISU, 1), PA(UM_P_ANISU, 2), n);
  } else if (term >= 5 && term <= 7) {
    /* cubic Kc1, Kc2, Kc3 one at a time (anisotropy.go:192-207) */
    const real *ka = NULL, *kb = NULL, *kc = NULL;
    real ma = 0, mb = 0, mc = 0;
    if (term == 5) { ka = PA_ARR(UM_P_KC1); ma = PA_MUL(UM_P_KC1); }
    if (term == 6) { kb = PA_ARR(UM_P_KC2); mb = PA_MUL(UM_P_KC2); }
    if (term == 7) { kc = PA_ARR(UM_P_KC3); mc = PA_MUL(UM_P_KC3); }
    OR(addcubic2)(B, B + n, B + 2 * n, m, m + n, m + 2 * n, PA(UM_P_MSAT, 0), ka, ma, kb, mb, kc, mc,
                  PA(UM_P_ANISC1, 0), PA(UM_P_ANISC1, 1), PA(UM_P_ANISC1, 2), PA(UM_P_ANISC2, 0),
                  PA(UM_P_ANISC2, 1), PA(UM_P_ANISC2, 2), n);
  } else if (term == 8) {
    /* GetThermalEnergy, engine/temperature.go:129-135 */
    if (!param_nonzero(s, UM_P_TEMP, 1) || s->relaxing) return 0;
    add_thermal(s, B);
  } else if (term == 9) {
    /* GetTotalEnergy, engine/energy.go:27-34 */
    double tot = 0;
    for (int t = 0; t <= 8; t++) tot += OR(sim_energy)(h, t);
    return tot;
  }
  static const double prefs[9] = {-0.5, -0.5, -1.0, -0.5, -0.25, -0.25, -1. / 6., -1. / 8., -1.0};
  double pref = prefs[term];
  for (int i = 0; i < n; i++) {
    double ms = (double)amul(PA(UM_P_MSAT, 0), i) * (s->vol ? (double)s->vol[i] : 1.0);
    for (int c = 0; c < 3; c++) e += ms * (double)m[c * n + i] * (double)B[c * n + i];
  }
  return pref * V * e;
}
