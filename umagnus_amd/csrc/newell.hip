// Newell demag tensor in FP64 on the GPU: mag/newellDemagKernel.go:1108-1324
// (CalcNewellDemagKernel, a port of OOMMF's demagcoef.cc).
//
// Near field (offset |R| <= asymptotic_radius * cbrt(V)): Newell's closed
// forms f (diagonal) and g (off-diagonal) combined by the 27-point second
// difference of CalculateSDA00/01, i.e. the product of the 1D stencil
// (-1, 2, -1) on each axis.  Far field: OOMMF's asymptotic expansion to
// 1/R^9 of Nxx and Nxy, evaluated on sub-cells when the cell is not cubic
// (DemagAsymptoticRefineData, refine ratio 1.5).
//
// One thread per non-negative offset (x, y, z) <= (P-1)/2 computes all six
// components; the other octants follow by mirror symmetry (the same pass as
// the brute-force tensor).  Sums of many cancelling terms use an error-free
// two-sum accumulation, which is at least as accurate as the reference's
// magnitude-sorted compensated sum (accSum, :506-526); results differ from
// it only by rounding.
#include "common.hpp"
#include "sim.hpp"

#include <algorithm>
#include <cmath>

namespace um {

// Running sum with the rounding error of every addition carried separately
// (Knuth's two-sum): the result is the exact sum rounded once, up to n^2 u^2.
struct TwoSumAcc {
  double s = 0.0, c = 0.0;
  __host__ __device__ void add(double x) {
    const double t = s + x;
    const double bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  __host__ __device__ double get() const { return s + c; }
};

// Newell's f (Newell_f, :749-806)
__host__ __device__ inline double newell_f(double x, double y, double z) {
#pragma clang fp contract(off)
  x = fabs(x); y = fabs(y); z = fabs(z);
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double R2 = x2 + y2 + z2;
  if (R2 <= 0.0) return 0.0;
  const double R = sqrt(R2);
  TwoSumAcc a;
  if (z > 0.0) {
    a.add(2.0 * (2.0 * x2 - y2 - z2) * R);
    const double xyz = x * y * z;
    if (xyz > 0.0) a.add(-12.0 * xyz * atan2(y * z, x * R));
    const double sxz = x2 + z2;
    if (y > 0.0 && sxz > 0.0) a.add(3.0 * y * (z2 - x2) * log(((y + R) * (y + R)) / sxz));
    const double sxy = x2 + y2;
    if (sxy > 0.0) a.add(3.0 * z * (y2 - x2) * log(((z + R) * (z + R)) / sxy));
  } else if (x == y) {
    const double K = 2.0 * sqrt(2.0) - 6.0 * log(1.0 + sqrt(2.0));
    a.add(K * x2 * x);
  } else {
    a.add(2.0 * (2.0 * x2 - y2) * R);
    if (y > 0.0 && x > 0.0) a.add(-6.0 * y * x2 * log((y + R) / x));
  }
  return a.get() / 12.0;
}

// Newell's g (Newell_g, :809-866), odd in x and in y
__host__ __device__ inline double newell_g(double x, double y, double z) {
#pragma clang fp contract(off)
  double sign = 1.0;
  if (x < 0.0) sign = -sign;
  if (y < 0.0) sign = -sign;
  x = fabs(x); y = fabs(y); z = fabs(z);
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double R2 = x2 + y2 + z2;
  if (R2 <= 0.0) return 0.0;
  const double R = sqrt(R2);
  TwoSumAcc a;
  a.add(-2.0 * x * y * R);
  if (z > 0.0) {
    a.add(-z * z2 * atan2(x * y, z * R));
    a.add(-3.0 * z * y2 * atan2(x * z, y * R));
    a.add(-3.0 * z * x2 * atan2(y * z, x * R));
    const double sxy = x2 + y2, syz = y2 + z2, sxz = x2 + z2;
    if (sxy > 0.0) a.add(3.0 * x * y * z * log((z + R) * (z + R) / sxy));
    if (syz > 0.0) a.add(0.5 * y * (3.0 * z2 - y2) * log((x + R) * (x + R) / syz));
    if (sxz > 0.0) a.add(0.5 * x * (3.0 * z2 - x2) * log((y + R) * (y + R) / sxz));
  } else {
    if (y > 0.0) a.add(-y * y2 * log((x + R) / y));
    if (x > 0.0) a.add(-x * x2 * log((y + R) / x));
  }
  return sign * a.get() / 6.0;
}

// 27-point second difference: sum over i, j, k in {-1, 0, 1} of
// w_i w_j w_k F(x + i a, y + j b, z + k c) with w = (-1, 2, -1).
template <bool G>
__device__ inline double second_diff27(double x, double y, double z, double a, double b, double c) {
#pragma clang fp contract(off)
  TwoSumAcc acc;
#pragma unroll 1
  for (int k = -1; k <= 1; k++) {
    const double wk = k == 0 ? 2.0 : -1.0;
#pragma unroll 1
    for (int j = -1; j <= 1; j++) {
      const double wjk = (j == 0 ? 2.0 : -1.0) * wk;
#pragma unroll 1
      for (int i = -1; i <= 1; i++) {
        const double w = (i == 0 ? 2.0 : -1.0) * wjk;
        const double px = x + i * a, py = y + j * b, pz = z + k * c;
        acc.add(w * (G ? newell_g(px, py, pz) : newell_f(px, py, pz)));
      }
    }
  }
  return acc.get();
}

// Self-demag factor of an x*y*z prism along x (SelfDemagNx, :869-933)
static double self_demag_nx(double x, double y, double z) {
#pragma clang fp contract(off)
  if (x <= 0.0 || y <= 0.0 || z <= 0.0) return 0.0;
  if (x == y && y == z) return 1.0 / 3.0;
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double R = sqrt(x2 + y2 + z2);
  const double Rxy = sqrt(x2 + y2), Rxz = sqrt(x2 + z2), Ryz = sqrt(y2 + z2);
  // log(1 + p), switching to the explicit ratio where log1p loses accuracy
  auto lg = [](double p, double num, double den) { return p > -0.5 ? log1p(p) : log(num / den); };
  TwoSumAcc a;
  a.add(2.0 * x * y * z *
        ((x / (x + Rxy) + (2 * x2 + y2 + z2) / (R * Rxy + x * Rxz)) / (x + Rxz) +
         (x / (x + Rxz) + (2 * x2 + y2 + z2) / (R * Rxz + x * Rxy)) / (x + Rxy)) /
        ((x + R) * (Rxy + Rxz + R)));
  a.add(-1.0 * x * y * z *
        ((y / (y + Rxy) + (2 * y2 + x2 + z2) / (R * Rxy + y * Ryz)) / (y + Ryz) +
         (y / (y + Ryz) + (2 * y2 + x2 + z2) / (R * Ryz + y * Rxy)) / (y + Rxy)) /
        ((y + R) * (Rxy + Ryz + R)));
  a.add(-1.0 * x * y * z *
        ((z / (z + Rxz) + (2 * z2 + x2 + y2) / (R * Rxz + z * Ryz)) / (z + Ryz) +
         (z / (z + Ryz) + (2 * z2 + x2 + y2) / (R * Ryz + z * Rxz)) / (z + Rxz)) /
        ((z + R) * (Rxz + Ryz + R)));
  a.add(6.0 * atan(y * z / (x * R)));
  const double p4 = -1.0 * y * z * z * (1.0 / (x + Rxz) + y / (Rxy * Rxz + x * R)) / (Rxz * (y + Rxy));
  a.add(3.0 * x * lg(p4, x * (y + R), Rxz * (y + Rxy)) / z);
  const double p5 = -1.0 * y * y * z * (1.0 / (x + Rxy) + z / (Rxy * Rxz + x * R)) / (Rxy * (z + Rxz));
  a.add(3.0 * x * lg(p5, x * (z + R), Rxy * (z + Rxz)) / y);
  const double p6 = -1.0 * x2 * z * (1 / (y + Rxy) + z / (Rxy * Ryz + y * R)) / (Rxy * (z + Ryz));
  a.add(-3.0 * y * lg(p6, y * (z + R), Rxy * (z + Ryz)) / x);
  const double p7 = -1.0 * x2 * y * (1 / (z + Rxz) + y / (Rxz * Ryz + z * R)) / (Rxz * (y + Ryz));
  a.add(-3.0 * z * lg(p7, z * (y + R), Rxz * (y + Ryz)) / x);
  return a.get() / (3.0 * M_PI);
}

// ---- far field --------------------------------------------------------------
// One far-field evaluator: the Nxx-type (kind 0) or Nxy-type (kind 1)
// expansion for a cell with edges rd (after refinement), applied with the
// axes permuted: role axis i is original axis perm[i] (:442-504).
struct NewellFar {
  int kind, cubic;
  int perm[3], cnt[3];
  double rd[3], scale, lead, self_demag;
  double a[6], b[10], c[15];
};

// DemagAsymptoticRefineData.Init (:101-133): the two longer edges are cut
// into pieces at most maxratio times the shortest one.
static void far_refine(NewellFar &F, const double d[3], double maxratio) {
  int s;
  if (d[2] <= d[0] && d[2] <= d[1]) s = 2;
  else if (d[1] <= d[0] && d[1] <= d[2]) s = 1;
  else s = 0;
  for (int i = 0; i < 3; i++) {
    if (i == s) { F.cnt[i] = 1; F.rd[i] = d[i]; continue; }
    const double ratio = std::ceil(d[i] / (maxratio * d[s]));
    F.cnt[i] = (int)ratio;
    F.rd[i] = d[i] / ratio;
  }
  F.scale = 1.0 / ((double)F.cnt[0] * (double)F.cnt[1] * (double)F.cnt[2]);
}

// Polynomial coefficient of a row over the monomials of one degree in
// (dx^2, dy^2, dz^2); the rows are OOMMF's published expansion coefficients.
static double poly4(const double m[6], double dx2, double dy2, double dz2) {
#pragma clang fp contract(off)
  return m[0] * (dx2 * dx2) + m[1] * dx2 * dy2 + m[2] * dx2 * dz2 + m[3] * (dy2 * dy2) + m[4] * dy2 * dz2 +
         m[5] * (dz2 * dz2);
}
static double poly6(const double m[10], double dx2, double dy2, double dz2) {
#pragma clang fp contract(off)
  const double dx4 = dx2 * dx2, dy4 = dy2 * dy2, dz4 = dz2 * dz2;
  return m[0] * (dx4 * dx2) + m[1] * dx4 * dy2 + m[2] * dx4 * dz2 + m[3] * dx2 * dy4 + m[4] * dx2 * dy2 * dz2 +
         m[5] * dx2 * dz4 + m[6] * (dy4 * dy2) + m[7] * dy4 * dz2 + m[8] * dy2 * dz4 + m[9] * (dz4 * dz2);
}

// Oxs_DemagNxxAsymptoticBase coefficients (:137-218)
static void far_init_xx(NewellFar &F) {
#pragma clang fp contract(off)
  const double dx = F.rd[0], dy = F.rd[1], dz = F.rd[2];
  F.self_demag = self_demag_nx(dx, dy, dz);
  const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  const double dx4 = dx2 * dx2, dx6 = dx4 * dx2;
  F.lead = (-1.0 * dx * dy * dz) / (4.0 * M_PI);
  F.cubic = !((dx2 != dy2) || (dx2 != dz2) || (dy2 != dz2));
  const double ai = F.lead / 4.0, bi = F.lead / 16.0, ci = F.lead / 192.0;
  if (F.cubic) {
    for (int i = 0; i < 6; i++) F.a[i] = 0.0;
    static const double mb[10] = {-14, 105, 105, -105, 0, -105, 7, 0, 0, 7};
    for (int i = 0; i < 10; i++) F.b[i] = mb[i] == 0 ? 0.0 : bi * (mb[i] * dx4);
    static const double mc[15] = {32, -448, -448, -150, 7620, -150, 314, -3810, -3810, 314,
                                  -16, 134, 300, 134, -16};
    for (int i = 0; i < 15; i++) F.c[i] = ci * (mc[i] * dx6);
    return;
  }
  static const double ma[6][3] = {{8, -4, -4}, {-24, 27, -3}, {-24, -3, 27}, {3, -4, 1}, {6, -3, -3}, {3, 1, -4}};
  for (int i = 0; i < 6; i++) F.a[i] = ai * (ma[i][0] * dx2 + ma[i][1] * dy2 + ma[i][2] * dz2);
  static const double mb[10][6] = {
      {32, -40, -40, 12, 10, 12},         {-240, 580, 20, -202, -75, 22},  {-240, 20, 580, 22, -75, -202},
      {180, -505, 55, 232, -75, 8},       {360, -450, -450, -180, 900, -180}, {180, 55, -505, 8, -75, 232},
      {-10, 30, -5, -16, 10, -2},         {-30, 55, 20, 8, -75, 22},       {-30, 20, 55, 22, -75, 8},
      {-10, -5, 30, -2, 10, -16}};
  for (int i = 0; i < 10; i++) F.b[i] = bi * poly4(mb[i], dx2, dy2, dz2);
  static const double mc[15][10] = {
      {384, -896, -896, 672, 560, 672, -120, -112, -112, -120},
      {-5376, 22624, 2464, -19488, -7840, 672, 3705, 2198, 938, -345},
      {-5376, 2464, 22624, 672, -7840, -19488, -345, 938, 2198, 3705},
      {10080, -48720, 1680, 49770, -2625, -630, -10440, -1050, 2100, -315},
      {20160, -47040, -47040, -6300, 133350, -6300, 7065, -26670, -26670, 7065},
      {10080, 1680, -48720, -630, -2625, 49770, -315, 2100, -1050, -10440},
      {-3360, 17290, -1610, -19488, 5495, -588, 4848, -3136, 938, -75},
      {-10080, 32970, 14070, -6300, -66675, 12600, -10080, 53340, -26670, 3015},
      {-10080, 14070, 32970, 12600, -66675, -6300, 3015, -26670, 53340, -10080},
      {-3360, -1610, 17290, -588, 5495, -19488, -75, 938, -3136, 4848},
      {105, -560, 70, 672, -280, 42, -192, 224, -112, 15},
      {420, -1610, -350, 672, 2345, -588, 528, -3136, 2198, -345},
      {630, -1470, -1470, -630, 5250, -630, 360, -1050, -1050, 360},
      {420, -350, -1610, -588, 2345, 672, -345, 2198, -3136, 528},
      {105, 70, -560, 42, -280, 672, 15, -112, 224, -192}};
  for (int i = 0; i < 15; i++) F.c[i] = ci * poly6(mc[i], dx2, dy2, dz2);
}

// Oxs_DemagNxyAsymptoticBase coefficients (:296-360)
static void far_init_xy(NewellFar &F) {
#pragma clang fp contract(off)
  const double dx = F.rd[0], dy = F.rd[1], dz = F.rd[2];
  const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  const double dx4 = dx2 * dx2, dx6 = dx4 * dx2;
  F.self_demag = 0.0;
  F.lead = (-dx * dy * dz / (4 * M_PI));
  F.cubic = !(dx2 != dy2 || dx2 != dz2 || dy2 != dz2);
  const double ai = (F.lead * 5.0) / 4.0, bi = (F.lead * 7.0) / 16.0, ci = F.lead / 64.0;
  for (int i = 3; i < 6; i++) F.a[i] = 0.0;
  for (int i = 6; i < 10; i++) F.b[i] = 0.0;
  for (int i = 10; i < 15; i++) F.c[i] = 0.0;
  if (F.cubic) {
    F.a[0] = F.a[1] = F.a[2] = 0.0;
    static const double mb[6] = {-7, 19, 13, -7, 13, -13};
    for (int i = 0; i < 6; i++) F.b[i] = bi * (mb[i] * dx4);
    static const double mc[10] = {48, -142, -582, -142, 2840, -450, 48, -582, -450, 180};
    for (int i = 0; i < 10; i++) F.c[i] = ci * (mc[i] * dx6);
    return;
  }
  static const double ma[3][3] = {{4, -3, -1}, {-3, 4, -1}, {-3, -3, 6}};
  for (int i = 0; i < 3; i++) F.a[i] = ai * (ma[i][0] * dx2 + ma[i][1] * dy2 + ma[i][2] * dz2);
  static const double mb[6][6] = {{16, -30, -10, 10, 5, 2},     {-40, 105, -5, -40, -5, 4},
                                  {-40, -15, 115, 20, -35, -32}, {10, -30, 5, 16, -10, 2},
                                  {20, -15, -35, -40, 115, -32}, {10, 15, -40, 10, -40, 32}};
  for (int i = 0; i < 6; i++) F.b[i] = bi * poly4(mb[i], dx2, dy2, dz2);
  static const double mc[10][10] = {
      {576, -2016, -672, 1680, 840, 336, -315, -210, -126, -45},
      {-3024, 13664, 448, -12670, -2485, 546, 2520, 910, 84, -135},
      {-3024, 1344, 12768, 2730, -10185, -8694, -945, 1680, 2394, 1350},
      {2520, -12670, 910, 13664, -2485, 84, -3024, 448, 546, -135},
      {5040, -9940, -13580, -9940, 49700, -6300, 5040, -13580, -6300, 2700},
      {2520, 2730, -14490, 420, -7875, 17640, -945, 3990, -840, -3600},
      {-315, 1680, -210, -2016, 840, -126, 576, -672, 336, -45},
      {-945, 2730, 1680, 1344, -10185, 2394, -3024, 12768, -8694, 1350},
      {-945, 420, 3990, 2730, -7875, -840, 2520, -14490, 17640, -3600},
      {-315, -630, 2100, -630, 3150, -3360, -315, 2100, -3360, 1440}};
  for (int i = 0; i < 10; i++) F.c[i] = ci * poly6(mc[i], dx2, dy2, dz2);
}

static NewellFar far_make(int kind, int p0, int p1, int p2, const double d[3]) {
  NewellFar F{};
  F.kind = kind;
  F.perm[0] = p0; F.perm[1] = p1; F.perm[2] = p2;
  const double dp[3] = {d[p0], d[p1], d[p2]};
  far_refine(F, dp, 1.5);   // DEMAG_ASYMPTOTIC_DEFAULT_REFINE_RATIO
  if (kind == 0) far_init_xx(F);
  else far_init_xy(F);
  return F;
}

// Expansion at one point (NxxAsymptotic :220-251 / NxyAsymptotic :362-396)
__device__ inline double far_point(const NewellFar &F, double x, double y, double z) {
#pragma clang fp contract(off)
  const double xx = x * x, yy = y * y, zz = z * z;
  const double R2 = xx + yy + zz;
  if (R2 == 0.0) return F.kind == 0 ? F.self_demag : 0.0;
  const double R4 = R2 * R2;
  const double tx2 = xx / R4, ty2 = yy / R4, tz2 = zz / R4;
  const double iR = 1.0 / sqrt(R2);
  const double *a = F.a, *b = F.b, *c = F.c;
  const double tz4 = tz2 * tz2;
  if (F.kind == 0) {
    const double tz6 = tz4 * tz2;
    const double term3 = (2 * tx2 - ty2 - tz2) * F.lead;
    double term5 = 0.0, term7;
    if (F.cubic) {
      const double ty4 = ty2 * ty2;
      term7 = ((b[0] * tx2 + (b[1] * ty2 + b[2] * tz2)) * tx2 + (b[3] * ty4 + b[5] * tz4)) * tx2 +
              b[6] * ty4 * ty2 + b[9] * tz6;
    } else {
      term5 = (a[0] * tx2 + (a[1] * ty2 + a[2] * tz2)) * tx2 + (a[3] * ty2 + a[4] * tz2) * ty2 + a[5] * tz4;
      term7 = ((b[0] * tx2 + (b[1] * ty2 + b[2] * tz2)) * tx2 + ((b[3] * ty2 + b[4] * tz2) * ty2 + b[5] * tz4)) *
                  tx2 +
              ((b[6] * ty2 + b[7] * tz2) * ty2 + b[8] * tz4) * ty2 + b[9] * tz6;
    }
    const double term9 =
        (((c[0] * tx2 + (c[1] * ty2 + c[2] * tz2)) * tx2 + ((c[3] * ty2 + c[4] * tz2) * ty2 + c[5] * tz4)) * tx2 +
         (((c[6] * ty2 + c[7] * tz2) * ty2 + c[8] * tz4) * ty2 + c[9] * tz6)) *
            tx2 +
        (((c[10] * ty2 + c[11] * tz2) * ty2 + c[12] * tz4) * ty2 + c[13] * tz6) * ty2 + c[14] * tz4 * tz4;
    return (term9 + term7 + term5 + term3) * iR;
  }
  const double term3 = 3.0 * F.lead;
  const double term5 = F.cubic ? 0.0 : a[0] * tx2 + a[1] * ty2 + a[2] * tz2;
  const double term7 = (b[0] * tx2 + (b[1] * ty2 + b[2] * tz2)) * tx2 + (b[3] * ty2 + b[4] * tz2) * ty2 + b[5] * tz4;
  const double term9 =
      ((c[0] * tx2 + (c[1] * ty2 + c[2] * tz2)) * tx2 + ((c[3] * ty2 + c[4] * tz2) * ty2 + c[5] * tz4)) * tx2 +
      ((c[6] * ty2 + c[7] * tz2) * ty2 + c[8] * tz4) * ty2 + c[9] * tz4 * tz2;
  const double iR2 = 1.0 / R2;
  const double iR5 = iR2 * iR2 * iR;
  return (term9 + term7 + term5 + term3) * iR5 * x * y;
}

// Refined expansion (:260-292, :404-440): the sum over sub-cell pairs folded
// into offsets weighted by (count - |offset|).
__device__ inline double far_eval(const NewellFar &F, const double R[3]) {
#pragma clang fp contract(off)
  const double x = R[F.perm[0]], y = R[F.perm[1]], z = R[F.perm[2]];
  const int xc = F.cnt[0], yc = F.cnt[1], zc = F.cnt[2];
  double zsum = 0.0;
  for (int k = 1 - zc; k < zc; k++) {
    const double zo = z + (double)k * F.rd[2];
    double ysum = 0.0;
    for (int j = 1 - yc; j < yc; j++) {
      const double yo = y + (double)j * F.rd[1];
      double xsum = (double)xc * far_point(F, x, yo, zo);
      for (int i = 1; i < xc; i++)
        xsum += (double)(xc - i) *
                (far_point(F, x + (double)i * F.rd[0], yo, zo) + far_point(F, x - (double)i * F.rd[0], yo, zo));
      ysum += (double)(yc - abs(j)) * xsum;
    }
    zsum += (double)(zc - abs(k)) * ysum;
  }
  return zsum * F.scale;
}

struct NewellArgs {
  int P[3];
  int r1[3], r2[3];      // image offsets per axis (kernelRanges)
  int H[3];              // wrapped destinations per axis: 0..P/2
  double d[3];
  double scale;          // -1 / (4 pi V): Newell's sums are 4 pi V N, the kernel is -N
  double arad, arad2;    // far field where |R|^2 > arad2 (arad >= 0)
  double self[3];        // kernel values at the origin (xx, yy, zz)
  NewellFar far[6];      // xx yy zz yz xz xy
  double *K[6];
};

// One thread per wrapped destination element; its periodic images (the
// offsets of the kernelRanges box that wrap onto it) are summed in FP64 in the
// reference loop order, z outer, y, x inner.  Each image takes the asymptotic
// expansion beyond the radius and the exact 27-term sums inside it
// (:1161-1246).  The origin element is the self term alone: the reference
// overwrites it after the image sums (:1300-1310).
__global__ void __launch_bounds__(64) k_demag_newell(NewellArgs a) {
#pragma clang fp contract(off)
  const long total = (long)a.H[0] * a.H[1] * a.H[2];
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int xw = (int)(t % a.H[0]);
  const int yw = (int)((t / a.H[0]) % a.H[1]);
  const int zw = (int)(t / ((long)a.H[0] * a.H[1]));
  const long I = ((long)zw * a.P[1] + yw) * a.P[0] + xw;
  if (xw == 0 && yw == 0 && zw == 0) {
    a.K[0][I] = a.self[0]; a.K[1][I] = a.self[1]; a.K[2][I] = a.self[2];
    a.K[3][I] = 0.0; a.K[4][I] = 0.0; a.K[5][I] = 0.0;
    return;
  }
  const double dx = a.d[0], dy = a.d[1], dz = a.d[2];
  double S[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
  for (int z = image_first(zw, a.P[2], a.r1[2]); z <= a.r2[2]; z += a.P[2])
#pragma unroll 1
    for (int y = image_first(yw, a.P[1], a.r1[1]); y <= a.r2[1]; y += a.P[1])
#pragma unroll 1
      for (int x = image_first(xw, a.P[0], a.r1[0]); x <= a.r2[0]; x += a.P[0]) {
        const double R[3] = {(double)x * dx, (double)y * dy, (double)z * dz};
        const bool far = a.arad >= 0.0 && R[0] * R[0] > a.arad2 - R[2] * R[2] - R[1] * R[1];
        if (far) {
#pragma unroll 1
          for (int c = 0; c < 6; c++) S[c] += -1.0 * far_eval(a.far[c], R);
          continue;
        }
        // CalculateSDA00/11/22 and 01/02/12 with their argument permutations
        S[0] += second_diff27<false>(R[0], R[1], R[2], dx, dy, dz) * a.scale;
        S[1] += second_diff27<false>(R[1], R[0], R[2], dy, dx, dz) * a.scale;
        S[2] += second_diff27<false>(R[2], R[1], R[0], dz, dy, dx) * a.scale;
        S[3] += (R[1] == 0.0 || R[2] == 0.0) ? 0.0 : second_diff27<true>(R[1], R[2], R[0], dy, dz, dx) * a.scale;
        S[4] += (R[0] == 0.0 || R[2] == 0.0) ? 0.0 : second_diff27<true>(R[0], R[2], R[1], dx, dz, dy) * a.scale;
        S[5] += (R[0] == 0.0 || R[1] == 0.0) ? 0.0 : second_diff27<true>(R[0], R[1], R[2], dx, dy, dz) * a.scale;
      }
#pragma unroll
  for (int c = 0; c < 6; c++) a.K[c][I] = S[c];
}

// ---- cell normalisation -------------------------------------------------------
static double gcd_floor(double m, double n) {   // Gcd (:937-952)
  m = std::floor(std::fabs(m));
  n = std::floor(std::fabs(n));
  if (m == 0.0 || n == 0.0) return 0.0;
  double t = m - std::floor(m / n) * n;
  while (t > 0.0) {
    m = n;
    n = t;
    t = m - std::floor(m / n) * n;
  }
  return n;
}

// FindRatApprox (:967-1010): nearest-integer continued fraction p/q of x/y;
// true when it meets relerr with q below maxq.
static bool rat_approx(double x, double y, double relerr, double maxq, double *p, double *q) {
#pragma clang fp contract(off)
  x = std::fabs(x);
  y = std::fabs(y);
  const bool swap = x < y;
  if (swap) std::swap(x, y);
  const double x0 = x, y0 = y;
  double p0 = 0.0, p1 = 1.0, q0 = 1.0, q1 = 0.0;
  double m = std::floor(x / y);
  double r = x - m * y;
  double p2 = m * p1 + p0, q2 = m * q1 + q0;
  while (q2 < maxq && std::fabs(x0 * q2 - p2 * y0) > relerr * x0 * q2) {
    x = y;
    y = r;
    m = std::floor(x / y);
    r = x - m * y;
    p0 = p1; p1 = p2; p2 = m * p1 + p0;
    q0 = q1; q1 = q2; q2 = m * q1 + q0;
  }
  *p = swap ? q2 : p2;
  *q = swap ? p2 : q2;
  return std::fabs(x0 * q2 - p2 * y0) <= relerr * x0 * q2;
}

// Cell edges in units where only their ratios matter (:1147-1166)
static void newell_cell(const double cs[3], double d[3]) {
#pragma clang fp contract(off)
  double p1, q1, p2, q2;
  d[0] = cs[0]; d[1] = cs[1]; d[2] = cs[2];
  if (rat_approx(d[0], d[1], 1e-12, 1000, &p1, &q1) && rat_approx(d[2], d[1], 1e-12, 1000, &p2, &q2)) {
    const double g = gcd_floor(q1, q2);
    d[0] = p1 * q2 / g;
    d[1] = q1 * q2 / g;
    d[2] = p2 * q1 / g;
  } else {
    const double mx = std::max(d[0], std::max(d[1], d[2]));
    d[0] /= mx; d[1] /= mx; d[2] /= mx;
  }
}

void demag_kernel_mirror(const int P[3], double *const K[6], hipStream_t s);

void demag_kernel_newell_device(const int N[3], const int pbc[3], const double cs[3], int asymptotic_radius,
                                int zero_self_demag, double *const Kdev[6], int P[3], hipStream_t s) {
  for (int i = 0; i < 3; i++)
    UM_REQUIRE(pbc[i] >= 0 && pbc[i] <= (1 << 16), UM_EINVAL, "PBC repetitions out of range");
  UM_REQUIRE(cs[0] > 0 && cs[1] > 0 && cs[2] > 0, UM_EINVAL, "bad cell size");
  UM_REQUIRE(zero_self_demag >= 0, UM_EINVAL, "zero_self_demag must be >= 0");
  pad_size(N, pbc, P);
  NewellArgs a{};
  kernel_ranges(P, pbc, a.r1, a.r2);
  for (int i = 0; i < 3; i++) {
    a.P[i] = P[i];
    a.H[i] = P[i] / 2 + 1;
    UM_REQUIRE((long)P[i] * pbc[i] < (1L << 30), UM_EINVAL, "PBC image range too large");
    a.K[i] = Kdev[i];
    a.K[i + 3] = Kdev[i + 3];
  }
  newell_cell(cs, a.d);
  const double V = a.d[0] * a.d[1] * a.d[2];
  a.scale = -1.0 / (4.0 * M_PI * V);
  a.arad = (double)asymptotic_radius * std::pow(V, 1.0 / 3.0);
  a.arad2 = a.arad * a.arad;
  // xx yy zz yz xz xy
  a.far[0] = far_make(0, 0, 1, 2, a.d);
  a.far[1] = far_make(0, 1, 0, 2, a.d);
  a.far[2] = far_make(0, 2, 1, 0, a.d);
  a.far[3] = far_make(1, 1, 2, 0, a.d);
  a.far[4] = far_make(1, 0, 2, 1, a.d);
  a.far[5] = far_make(1, 0, 1, 2, a.d);
  // origin (:1300-1310): -N_self, plus 1/3 per axis when the self term is zeroed
  const double third = zero_self_demag != 0 ? 1.0 / 3.0 : 0.0;
  a.self[0] = -self_demag_nx(a.d[0], a.d[1], a.d[2]) + third;
  a.self[1] = -self_demag_nx(a.d[1], a.d[2], a.d[0]) + third;
  a.self[2] = -self_demag_nx(a.d[2], a.d[0], a.d[1]) + third;

  const size_t np = (size_t)P[0] * P[1] * P[2];
  for (int c = 0; c < 6; c++) fill_value<double>(Kdev[c], 0.0, np, s);
  const long total = (long)a.H[0] * a.H[1] * a.H[2];
  UM_REQUIRE(total < (1L << 31), UM_EINVAL, "Newell tensor too large");
  UM_LAUNCH(k_demag_newell, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, s, a);
  UM_HIP(hipGetLastError());
  demag_kernel_mirror(P, Kdev, s);
  if (P[2] == 1) {
    fill_value<double>(Kdev[3], 0.0, np, s);
    fill_value<double>(Kdev[4], 0.0, np, s);
  }
}

}  // namespace um
