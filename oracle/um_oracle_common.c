/*
 * um_oracle_common.c -- precision independent part of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (checker for tests/, smoke() and the bench
 * cpu_baseline leg).  Never linked into the product library.
 */
#include "um_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Multipliers and Weyl constants as published. */
void um_or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* padSize, mag/demagkernel.go:404-420 (SMALL_N = 5, :426) */
void um_or_pad_size(const int N[3], const int pbc[3], int P[3]) {
  for (int i = 0; i < 3; i++) {
    if (pbc[i] != 0) { P[i] = N[i]; continue; }
    if (i != 2 || N[i] > 5) P[i] = 2 * N[i];
    else P[i] = 2 * N[i] - 1;
  }
}

/* kernelRanges, mag/demagkernel.go:339-352 */
static void kernel_ranges(const int size[3], const int pbc[3], int r1[3], int r2[3]) {
  for (int c = 0; c < 3; c++) {
    if (pbc[c] == 0) {
      r1[c] = -(size[c] - 1) / 2;
      r2[c] = (size[c] - 1) / 2;
    } else {
      r1[c] = -(size[c] * pbc[c] - 1);
      r2[c] = (size[c] * pbc[c] - 1);
    }
  }
  if (size[2] == 1 && pbc[2] == 0) r2[2] = 0;
}

/* delta, mag/demagkernel.go:362-370 */
static double delta_(int d) {
  if (d < 0) d = -d;
  if (d > 0) d -= 1;
  return (double)d;
}

/* wrap, mag/demagkernel.go:373-381 */
static int wrap_(int number, int max) {
  while (number < 0) number += max;
  while (number >= max) number -= max;
  return number;
}

/* CalcDemagKernel, mag/demagkernel.go:115-336.  The FP64 accumulation order of
 * the integration loops follows the reference exactly; the reference then
 * stores float32(B[d]) (:267), which the FP32 oracle does when it consumes
 * this array.  With PBC the reference adds float32(B) of every periodic image
 * into its float32 array; here the images are summed in FP64 (same order) and
 * rounded once, which differs by at most (images) x 2^-24 relative.
 * Component order of K: xx, yy, zz, yz, xz, xy. */
int um_or_demag_kernel_bf(const int N[3], const int pbc[3], const double cs[3],
                          double accuracy, double *K[6]) {
  int size[3];
  um_or_pad_size(N, pbc, size);
  if (!(size[0] > 0 && size[1] > 0 && size[2] > 0)) return -1;
  if (!(cs[0] > 0 && cs[1] > 0 && cs[2] > 0) || !(accuracy > 0)) return -1;
  const long P = (long)size[0] * size[1] * size[2];
  /* array[s][d] for d >= s, mapped onto K[].  The loop nest is parallel over
   * (s, zw): every task owns the elements of one wrapped z plane and visits
   * the offsets z of the range that wrap onto it in ascending order, so each
   * element sums its periodic images in the reference's loop order (z, y, x
   * ascending, :165-268). */
  double *arr[3][3] = {{K[0], K[5], K[4]}, {NULL, K[1], K[3]}, {NULL, NULL, K[2]}};
  for (int c = 0; c < 6; c++) memset(K[c], 0, sizeof(double) * P);

  int r1[3], r2[3];
  kernel_ranges(size, pbc, r1, r2);
  double L = cs[0];
  if (cs[1] < L) L = cs[1];
  if (cs[2] < L) L = cs[2];

#define AT(a, z, y, x) (a)[((long)(z) * size[1] + (y)) * size[0] + (x)]
  const int nzw = size[2] / 2 + 1;
#pragma omp parallel for schedule(dynamic)
  for (int sz = 0; sz < 3 * nzw; sz++) {
    const int s = sz / nzw, zt = sz % nzw;
    int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
    double R[3], R2[3], pole[3];
    for (int z = r1[2]; z <= r2[2]; z++) {
      int zw = wrap_(z, size[2]);
      if (zw != zt) continue;
      R[2] = (double)z * cs[2];
      for (int y = r1[1]; y <= r2[1]; y++) {
        int yw = wrap_(y, size[1]);
        if (yw > size[1] / 2) continue;
        R[1] = (double)y * cs[1];
        for (int x = r1[0]; x <= r2[0]; x++) {
          int xw = wrap_(x, size[0]);
          if (xw > size[0] / 2) continue;
          R[0] = (double)x * cs[0];
          double dx = delta_(x) * cs[0], dy = delta_(y) * cs[1], dz = delta_(z) * cs[2];
          double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (d == 0) d = L;
          double maxSize = d / accuracy;
          int nv = (int)(fmax(cs[v] / maxSize, 1) + 0.5);
          int nw = (int)(fmax(cs[w] / maxSize, 1) + 0.5);
          int nx = (int)(fmax(cs[0] / maxSize, 1) + 0.5);
          int ny = (int)(fmax(cs[1] / maxSize, 1) + 0.5);
          int nz = (int)(fmax(cs[2] / maxSize, 1) + 0.5);
          nv *= 2;
          nw *= 2;
          /* Go's int is 64-bit: the sub-cell count overflows 32 bits at accuracy ~40 */
          double scale = 1 / (double)((long long)nv * nw * nx * ny * nz);
          double surface = cs[v] * cs[w];
          double charge = surface * scale;
          double pu1 = cs[u] / 2.;
          double pu2 = -pu1;
          double B[3] = {0, 0, 0};
          for (int i = 0; i < nv; i++) {
            double pv = -(cs[v] / 2.) + cs[v] / (double)(2 * nv) + (double)i * (cs[v] / (double)nv);
            pole[v] = pv;
            for (int j = 0; j < nw; j++) {
              double pw = -(cs[w] / 2.) + cs[w] / (double)(2 * nw) + (double)j * (cs[w] / (double)nw);
              pole[w] = pw;
              for (int a = 0; a < nx; a++) {
                double rx = R[0] - cs[0] / 2 + cs[0] / (double)(2 * nx) + (cs[0] / (double)nx) * (double)a;
                for (int b = 0; b < ny; b++) {
                  double ry = R[1] - cs[1] / 2 + cs[1] / (double)(2 * ny) + (cs[1] / (double)ny) * (double)b;
                  for (int g = 0; g < nz; g++) {
                    double rz = R[2] - cs[2] / 2 + cs[2] / (double)(2 * nz) + (cs[2] / (double)nz) * (double)g;
                    pole[u] = pu1;
                    R2[0] = rx - pole[0]; R2[1] = ry - pole[1]; R2[2] = rz - pole[2];
                    double r = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
                    double qr = charge / (4 * M_PI * r * r * r);
                    double bx = R2[0] * qr, by = R2[1] * qr, bz = R2[2] * qr;
                    pole[u] = pu2;
                    R2[0] = rx - pole[0]; R2[1] = ry - pole[1]; R2[2] = rz - pole[2];
                    r = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
                    qr = -charge / (4 * M_PI * r * r * r);
                    B[0] += (bx + R2[0] * qr);
                    B[1] += (by + R2[1] * qr);
                    B[2] += (bz + R2[2] * qr);
                  }
                }
              }
            }
          }
          for (int dd = s; dd < 3; dd++) AT(arr[s][dd], zw, yw, xw) += B[dd];
        }
      }
    }
  }

  double *XX = K[0], *YY = K[1], *ZZ = K[2], *YZ = K[3], *XZ = K[4], *XY = K[5];
  /* mirror in X, :280-293 */
  for (int z = 0; z < size[2]; z++)
    for (int y = 0; y < size[1]; y++)
      for (int x = size[0] / 2 + 1; x < size[0]; x++) {
        int x2 = size[0] - x;
        AT(XX, z, y, x) = AT(XX, z, y, x2);
        AT(XY, z, y, x) = -AT(XY, z, y, x2);
        AT(XZ, z, y, x) = -AT(XZ, z, y, x2);
        AT(YY, z, y, x) = AT(YY, z, y, x2);
        AT(YZ, z, y, x) = AT(YZ, z, y, x2);
        AT(ZZ, z, y, x) = AT(ZZ, z, y, x2);
      }
  /* mirror in Y, :295-309 */
  for (int z = 0; z < size[2]; z++)
    for (int y = size[1] / 2 + 1; y < size[1]; y++) {
      int y2 = size[1] - y;
      for (int x = 0; x < size[0]; x++) {
        AT(XX, z, y, x) = AT(XX, z, y2, x);
        AT(XY, z, y, x) = -AT(XY, z, y2, x);
        AT(XZ, z, y, x) = AT(XZ, z, y2, x);
        AT(YY, z, y, x) = AT(YY, z, y2, x);
        AT(YZ, z, y, x) = -AT(YZ, z, y2, x);
        AT(ZZ, z, y, x) = AT(ZZ, z, y2, x);
      }
    }
  /* mirror in Z, :311-324 */
  for (int z = size[2] / 2 + 1; z < size[2]; z++) {
    int z2 = size[2] - z;
    for (int y = 0; y < size[1]; y++)
      for (int x = 0; x < size[0]; x++) {
        AT(XX, z, y, x) = AT(XX, z2, y, x);
        AT(XY, z, y, x) = AT(XY, z2, y, x);
        AT(XZ, z, y, x) = -AT(XZ, z2, y, x);
        AT(YY, z, y, x) = AT(YY, z2, y, x);
        AT(YZ, z, y, x) = -AT(YZ, z2, y, x);
        AT(ZZ, z, y, x) = AT(ZZ, z2, y, x);
      }
  }
  /* 2D: xz and yz are not used, :327-330 */
  if (size[2] == 1) {
    memset(XZ, 0, sizeof(double) * P);
    memset(YZ, 0, sizeof(double) * P);
  }
#undef AT
  return 0;
}

/* symmidx, engine/exchange.go:181-187 */
int um_or_symmidx(int i, int j) {
  if (j <= i) return i * (i + 1) / 2 + j;
  return j * (j + 1) / 2 + i;
}

/* exchAverage, engine/exchange.go:193-202 */
float um_or_exch_average(float exi, float exj) {
  if (exi * exj >= 0.0f) {
    return 2.0f / (1.0f / exi + 1.0f / exj);
  }
  double a = (double)exi, b = (double)exj;
  double sign = copysign(1.0, a + b);
  double magn = sqrt(sqrt(-a * b) * fabs(a + b) / 2);
  return (float)(sign * magn);
}

/* adaptDt, engine/run.go:150-187 */
double um_or_adapt_dt(double dt, double corr, double fixdt, double headroom,
                      double mindt, double maxdt, double time, double alarm) {
  if (fixdt != 0) return fixdt;
  if (isnan(corr)) corr = 1;
  if (corr == 0) return -1;
  corr *= headroom;
  if (corr > 2) corr = 2;
  if (corr < 1. / 2.) corr = 1. / 2.;
  dt *= corr;
  if (mindt != 0 && dt < mindt) dt = mindt;
  if (maxdt != 0 && dt > maxdt) dt = maxdt;
  if (dt == 0) return -1;
  if (time < alarm && time + dt > alarm) dt = alarm - time;
  if (!(dt > 0)) return -1;
  return dt;
}
