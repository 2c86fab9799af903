/*
 * um_oracle_newell.c -- Newell (OOMMF) demag tensor for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (checker for tests/, smoke() and the bench
 * cpu_baseline leg).  Never linked into the product library.
 *
 * Restates mag/newellDemagKernel.go (itself a port of OOMMF's demagcoef.cc):
 *   - Newell's f and g functions (Newell_f :749-806, Newell_g :809-866),
 *   - the 27-point second differences CalculateSDA00/01 (:648-733) written as
 *     the tensor product of the 1D stencil (-1, 2, -1) on each axis,
 *   - the self-demag factor SelfDemagNx (:869-933),
 *   - the far-field expansions of Nxx and Nxy up to 1/R^9 with the cell
 *     refinement of DemagAsymptoticRefineData (:17-490),
 *   - the rational cell-ratio approximation FindRatApprox/Gcd (:943-1010),
 *   - CalcNewellDemagKernel (:1108-1324).
 * The magnitude-sorted compensated sum accSum (:506-526) is followed step by
 * step; equal-magnitude values may be ordered differently from the
 * reference's quicksort (:529-601), which can move the last bit of a sum.
 */
#include "um_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- accSum: sort by magnitude, then a compensated sum from the largest -- */
static double acc_sum(double *v, int n) {
  /* insertion sort, ascending |v| */
  for (int i = 1; i < n; i++) {
    double t = v[i];
    int j = i - 1;
    while (j >= 0 && fabs(v[j]) > fabs(t)) { v[j + 1] = v[j]; j--; }
    v[j + 1] = t;
  }
  double sum = v[n - 1], corr = 0.0;
  for (int i = n - 2; i >= 0; i--) {
    const double x = v[i];
    const double y = corr + x;
    double tmp = y - corr;
    const double u = x - tmp;
    const double t = y + sum;
    tmp = t - sum;
    const double w = y - tmp;
    const double z = u + w;
    sum = t + z;
    tmp = sum - t;
    corr = z - tmp;
  }
  return sum;
}

/* ---- Newell's f and g --------------------------------------------------- */
double um_or_newell_f(double x, double y, double z) {
  x = fabs(x); y = fabs(y); z = fabs(z);
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double R2 = x2 + y2 + z2;
  if (R2 <= 0.0) return 0.0;
  const double R = sqrt(R2);
  double p[4];
  int n = 0;
  if (z > 0.0) {
    p[n++] = 2.0 * (2.0 * x2 - y2 - z2) * R;
    const double xyz = x * y * z;
    if (xyz > 0.0) p[n++] = -12.0 * xyz * atan2(y * z, x * R);
    const double sxz = x2 + z2;
    if (y > 0.0 && sxz > 0.0) p[n++] = 3.0 * y * (z2 - x2) * log(((y + R) * (y + R)) / sxz);
    const double sxy = x2 + y2;
    if (sxy > 0.0) p[n++] = 3.0 * z * (y2 - x2) * log(((z + R) * (z + R)) / sxy);
  } else if (x == y) {
    const double K = 2.0 * sqrt(2.0) - 6.0 * log(1.0 + sqrt(2.0));
    p[n++] = K * x2 * x;
  } else {
    p[n++] = 2.0 * (2.0 * x2 - y2) * R;
    if (y > 0.0 && x > 0.0) p[n++] = -6.0 * y * x2 * log((y + R) / x);
  }
  return acc_sum(p, n) / 12.0;
}

double um_or_newell_g(double x, double y, double z) {
  double sign = 1.0;
  if (x < 0.0) sign = -sign;
  if (y < 0.0) sign = -sign;
  x = fabs(x); y = fabs(y); z = fabs(z);
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double R2 = x2 + y2 + z2;
  if (R2 <= 0.0) return 0.0;
  const double R = sqrt(R2);
  double p[7];
  int n = 0;
  p[n++] = -2.0 * x * y * R;
  if (z > 0.0) {
    p[n++] = -z * z2 * atan2(x * y, z * R);
    p[n++] = -3.0 * z * y2 * atan2(x * z, y * R);
    p[n++] = -3.0 * z * x2 * atan2(y * z, x * R);
    const double sxy = x2 + y2, syz = y2 + z2, sxz = x2 + z2;
    if (sxy > 0.0) p[n++] = 3.0 * x * y * z * log((z + R) * (z + R) / sxy);
    if (syz > 0.0) p[n++] = 0.5 * y * (3.0 * z2 - y2) * log((x + R) * (x + R) / syz);
    if (sxz > 0.0) p[n++] = 0.5 * x * (3.0 * z2 - x2) * log((y + R) * (y + R) / sxz);
  } else {
    if (y > 0.0) p[n++] = -y * y2 * log((x + R) / y);
    if (x > 0.0) p[n++] = -x * x2 * log((y + R) / x);
  }
  return sign * acc_sum(p, n) / 6.0;
}

/* ---- self-demag factor of an x*y*z prism along x (SelfDemagNx) ---------- */
/* log of 1 + p where p = piece; the reference switches to the explicit ratio
 * form for p <= -0.5 to avoid log1p's loss there. */
static double sd_log(double piece, double alt_num, double alt_den) {
  return piece > -0.5 ? log1p(piece) : log(alt_num / alt_den);
}

double um_or_self_demag_nx(double x, double y, double z) {
  if (x <= 0.0 || y <= 0.0 || z <= 0.0) return 0.0;
  if (x == y && y == z) return 1.0 / 3.0;
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double R = sqrt(x2 + y2 + z2);
  const double Rxy = sqrt(x2 + y2), Rxz = sqrt(x2 + z2), Ryz = sqrt(y2 + z2);
  double a[8];
  a[0] = 2.0 * x * y * z *
         ((x / (x + Rxy) + (2 * x2 + y2 + z2) / (R * Rxy + x * Rxz)) / (x + Rxz) +
          (x / (x + Rxz) + (2 * x2 + y2 + z2) / (R * Rxz + x * Rxy)) / (x + Rxy)) /
         ((x + R) * (Rxy + Rxz + R));
  a[1] = -1.0 * x * y * z *
         ((y / (y + Rxy) + (2 * y2 + x2 + z2) / (R * Rxy + y * Ryz)) / (y + Ryz) +
          (y / (y + Ryz) + (2 * y2 + x2 + z2) / (R * Ryz + y * Rxy)) / (y + Rxy)) /
         ((y + R) * (Rxy + Ryz + R));
  a[2] = -1.0 * x * y * z *
         ((z / (z + Rxz) + (2 * z2 + x2 + y2) / (R * Rxz + z * Ryz)) / (z + Ryz) +
          (z / (z + Ryz) + (2 * z2 + x2 + y2) / (R * Ryz + z * Rxz)) / (z + Rxz)) /
         ((z + R) * (Rxz + Ryz + R));
  a[3] = 6.0 * atan(y * z / (x * R));
  const double p4 = -1.0 * y * z * z * (1.0 / (x + Rxz) + y / (Rxy * Rxz + x * R)) / (Rxz * (y + Rxy));
  a[4] = 3.0 * x * sd_log(p4, x * (y + R), Rxz * (y + Rxy)) / z;
  const double p5 = -1.0 * y * y * z * (1.0 / (x + Rxy) + z / (Rxy * Rxz + x * R)) / (Rxy * (z + Rxz));
  a[5] = 3.0 * x * sd_log(p5, x * (z + R), Rxy * (z + Rxz)) / y;
  const double p6 = -1.0 * x2 * z * (1 / (y + Rxy) + z / (Rxy * Ryz + y * R)) / (Rxy * (z + Ryz));
  a[6] = -3.0 * y * sd_log(p6, y * (z + R), Rxy * (z + Ryz)) / x;
  const double p7 = -1.0 * x2 * y * (1 / (z + Rxz) + y / (Rxz * Ryz + z * R)) / (Rxz * (y + Ryz));
  a[7] = -3.0 * z * sd_log(p7, z * (y + R), Rxz * (y + Ryz)) / x;
  return acc_sum(a, 8) / (3.0 * M_PI);
}

/* ---- 27-point second differences --------------------------------------- */
/* sum over i,j,k in {-1,0,1} of w_i w_j w_k F(x + i a, y + j b, z + k c),
 * w = (-1, 2, -1): weights 8 (centre), -4 (faces), 2 (edges), -1 (corners),
 * the term list of CalculateSDA00/01. */
static double second_diff27(double (*F)(double, double, double), double x, double y, double z,
                            double a, double b, double c) {
  static const double w[3] = {-1.0, 2.0, -1.0};
  double v[27];
  int n = 0;
  for (int k = -1; k <= 1; k++)
    for (int j = -1; j <= 1; j++)
      for (int i = -1; i <= 1; i++)
        v[n++] = w[i + 1] * w[j + 1] * w[k + 1] * F(x + i * a, y + j * b, z + k * c);
  return acc_sum(v, 27);
}

/* CalculateSDA00 (:648-687): 4 pi V Nxx of the cell pair at offset (x,y,z) */
double um_or_newell_sda00(double x, double y, double z, double dx, double dy, double dz) {
  if (x == 0.0 && y == 0.0 && z == 0.0) return um_or_self_demag_nx(dx, dy, dz) * (4.0 * M_PI * dx * dy * dz);
  return second_diff27(um_or_newell_f, x, y, z, dx, dy, dz);
}
/* CalculateSDA01 (:697-733) */
double um_or_newell_sda01(double x, double y, double z, double dx, double dy, double dz) {
  if (x == 0.0 || y == 0.0) return 0.0;
  return second_diff27(um_or_newell_g, x, y, z, dx, dy, dz);
}

/* ---- far field: OOMMF asymptotic expansions ------------------------------ */
typedef struct {           /* DemagNabData.Set (:74-94) */
  double x, y, z, tx2, ty2, tz2, R, iR, R2, iR2;
} nab_t;

static void nab_set(nab_t *p, double x, double y, double z) {
  p->x = x; p->y = y; p->z = z;
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  p->R2 = x2 + y2 + z2;
  const double R4 = p->R2 * p->R2;
  p->R = sqrt(p->R2);
  if (p->R2 != 0.0) {
    p->tx2 = x2 / R4; p->ty2 = y2 / R4; p->tz2 = z2 / R4;
    p->iR2 = 1.0 / p->R2; p->iR = 1.0 / p->R;
  } else {
    p->tx2 = p->ty2 = p->tz2 = 0.0;
    p->iR2 = p->R = p->iR = 0.0;
  }
}

typedef struct {           /* DemagAsymptoticRefineData (:101-133) */
  double rd[3], scale;
  int cnt[3];
} refine_t;

/* the two longer edges are cut into pieces at most maxratio times the shortest */
static void refine_init(refine_t *r, double dx, double dy, double dz, double maxratio) {
  const double d[3] = {dx, dy, dz};
  int s;                     /* index of the shortest edge (ties: z, then y, then x) */
  if (dz <= dx && dz <= dy) s = 2;
  else if (dy <= dx && dy <= dz) s = 1;
  else s = 0;
  for (int i = 0; i < 3; i++) {
    if (i == s) { r->cnt[i] = 1; r->rd[i] = d[i]; continue; }
    const double ratio = ceil(d[i] / (maxratio * d[s]));
    r->cnt[i] = (int)ratio;
    r->rd[i] = d[i] / ratio;
  }
  r->scale = 1.0 / ((double)r->cnt[0] * (double)r->cnt[1] * (double)r->cnt[2]);
}

typedef struct {           /* DemagNxxAsymptoticBase (:137-218) */
  int cubic;
  double self_demag, lead;
  double a[6], b[10], c[15];
} nxx_t;

static void nxx_init(nxx_t *s, const refine_t *r) {
  const double dx = r->rd[0], dy = r->rd[1], dz = r->rd[2];
  s->self_demag = um_or_self_demag_nx(dx, dy, dz);
  const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  const double dx4 = dx2 * dx2, dy4 = dy2 * dy2, dz4 = dz2 * dz2;
  const double dx6 = dx4 * dx2, dy6 = dy4 * dy2, dz6 = dz4 * dz2;
  s->lead = (-1.0 * dx * dy * dz) / (4.0 * M_PI);
  s->cubic = !((dx2 != dy2) || (dx2 != dz2) || (dy2 != dz2));
  /* 1/R^5 */
  const double ai = s->lead / 4.0;
  if (!s->cubic) {
    s->a[0] = ai * (8.0 * dx2 - 4.0 * dy2 - 4.0 * dz2);
    s->a[1] = ai * (-24.0 * dx2 + 27.0 * dy2 - 3.0 * dz2);
    s->a[2] = ai * (-24.0 * dx2 - 3.0 * dy2 + 27.0 * dz2);
    s->a[3] = ai * (3.0 * dx2 - 4.0 * dy2 + dz2);
    s->a[4] = ai * (6.0 * dx2 - 3.0 * dy2 - 3.0 * dz2);
    s->a[5] = ai * (3.0 * dx2 + dy2 - 4.0 * dz2);
  } else {
    for (int i = 0; i < 6; i++) s->a[i] = 0.0;
  }
  /* 1/R^7 */
  const double bi = s->lead / 16.0;
  if (s->cubic) {
    const double m[10] = {-14, 105, 105, -105, 0, -105, 7, 0, 0, 7};
    for (int i = 0; i < 10; i++) s->b[i] = m[i] == 0 ? 0.0 : bi * (m[i] * dx4);
  } else {
    /* coefficient rows over (dx4, dx2dy2, dx2dz2, dy4, dy2dz2, dz4) */
    static const double m[10][6] = {
        {32, -40, -40, 12, 10, 12},        {-240, 580, 20, -202, -75, 22},
        {-240, 20, 580, 22, -75, -202},    {180, -505, 55, 232, -75, 8},
        {360, -450, -450, -180, 900, -180}, {180, 55, -505, 8, -75, 232},
        {-10, 30, -5, -16, 10, -2},        {-30, 55, 20, 8, -75, 22},
        {-30, 20, 55, 22, -75, 8},         {-10, -5, 30, -2, 10, -16}};
    for (int i = 0; i < 10; i++)
      s->b[i] = bi * (m[i][0] * dx4 + m[i][1] * dx2 * dy2 + m[i][2] * dx2 * dz2 + m[i][3] * dy4 +
                      m[i][4] * dy2 * dz2 + m[i][5] * dz4);
  }
  /* 1/R^9 */
  const double ci = s->lead / 192.0;
  if (s->cubic) {
    static const double m[15] = {32, -448, -448, -150, 7620, -150, 314, -3810, -3810, 314,
                                 -16, 134, 300, 134, -16};
    for (int i = 0; i < 15; i++) s->c[i] = ci * (m[i] * dx6);
  } else {
    /* rows over (dx6, dx4dy2, dx4dz2, dx2dy4, dx2dy2dz2, dx2dz4, dy6, dy4dz2, dy2dz4, dz6) */
    static const double m[15][10] = {
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
    for (int i = 0; i < 15; i++)
      s->c[i] = ci * (m[i][0] * dx6 + m[i][1] * dx4 * dy2 + m[i][2] * dx4 * dz2 + m[i][3] * dx2 * dy4 +
                      m[i][4] * dx2 * dy2 * dz2 + m[i][5] * dx2 * dz4 + m[i][6] * dy6 +
                      m[i][7] * dy4 * dz2 + m[i][8] * dy2 * dz4 + m[i][9] * dz6);
  }
}

/* Oxs_DemagNxxAsymptoticBase::NxxAsymptotic (:220-251) */
static double nxx_eval(const nxx_t *s, const nab_t *p) {
  if (p->iR2 <= 0.0) return s->self_demag;
  const double tx2 = p->tx2, ty2 = p->ty2, tz2 = p->tz2;
  const double tz4 = tz2 * tz2, tz6 = tz4 * tz2;
  const double term3 = (2 * tx2 - ty2 - tz2) * s->lead;
  double term5 = 0.0, term7;
  const double *a = s->a, *b = s->b, *c = s->c;
  if (s->cubic) {
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
  return (term9 + term7 + term5 + term3) * p->iR;
}

typedef struct {           /* DemagNxyAsymptoticBase (:296-360) */
  int cubic;
  double lead;
  double a[3], b[6], c[10];
} nxy_t;

static void nxy_init(nxy_t *s, const refine_t *r) {
  const double dx = r->rd[0], dy = r->rd[1], dz = r->rd[2];
  const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  const double dx4 = dx2 * dx2, dy4 = dy2 * dy2, dz4 = dz2 * dz2;
  const double dx6 = dx4 * dx2, dy6 = dy4 * dy2, dz6 = dz4 * dz2;
  s->lead = (-dx * dy * dz / (4 * M_PI));
  s->cubic = !(dx2 != dy2 || dx2 != dz2 || dy2 != dz2);
  if (!s->cubic) {
    const double ai = (s->lead * 5.0) / 4.0;
    s->a[0] = ai * (4.0 * dx2 - 3.0 * dy2 - dz2);
    s->a[1] = ai * (-3.0 * dx2 + 4.0 * dy2 - dz2);
    s->a[2] = ai * (-3.0 * dx2 - 3.0 * dy2 + 6.0 * dz2);
  } else {
    s->a[0] = s->a[1] = s->a[2] = 0.0;
  }
  const double bi = (s->lead * 7.0) / 16.0;
  if (s->cubic) {
    static const double m[6] = {-7, 19, 13, -7, 13, -13};
    for (int i = 0; i < 6; i++) s->b[i] = bi * (m[i] * dx4);
  } else {
    static const double m[6][6] = {{16, -30, -10, 10, 5, 2},    {-40, 105, -5, -40, -5, 4},
                                   {-40, -15, 115, 20, -35, -32}, {10, -30, 5, 16, -10, 2},
                                   {20, -15, -35, -40, 115, -32}, {10, 15, -40, 10, -40, 32}};
    for (int i = 0; i < 6; i++)
      s->b[i] = bi * (m[i][0] * dx4 + m[i][1] * dx2 * dy2 + m[i][2] * dx2 * dz2 + m[i][3] * dy4 +
                      m[i][4] * dy2 * dz2 + m[i][5] * dz4);
  }
  const double ci = s->lead / 64.0;
  if (s->cubic) {
    static const double m[10] = {48, -142, -582, -142, 2840, -450, 48, -582, -450, 180};
    for (int i = 0; i < 10; i++) s->c[i] = ci * (m[i] * dx6);
  } else {
    static const double m[10][10] = {
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
    for (int i = 0; i < 10; i++)
      s->c[i] = ci * (m[i][0] * dx6 + m[i][1] * dx4 * dy2 + m[i][2] * dx4 * dz2 + m[i][3] * dx2 * dy4 +
                      m[i][4] * dx2 * dy2 * dz2 + m[i][5] * dx2 * dz4 + m[i][6] * dy6 +
                      m[i][7] * dy4 * dz2 + m[i][8] * dy2 * dz4 + m[i][9] * dz6);
  }
}

/* Oxs_DemagNxyAsymptoticBase::NxyAsymptotic (:362-396) */
static double nxy_eval(const nxy_t *s, const nab_t *p) {
  if (p->R2 <= 0.0) return 0.0;
  const double tx2 = p->tx2, ty2 = p->ty2, tz2 = p->tz2;
  const double *a = s->a, *b = s->b, *c = s->c;
  const double term3 = 3.0 * s->lead;
  const double term5 = s->cubic ? 0.0 : a[0] * tx2 + a[1] * ty2 + a[2] * tz2;
  const double tz4 = tz2 * tz2;
  const double term7 = (b[0] * tx2 + (b[1] * ty2 + b[2] * tz2)) * tx2 + (b[3] * ty2 + b[4] * tz2) * ty2 + b[5] * tz4;
  const double term9 =
      ((c[0] * tx2 + (c[1] * ty2 + c[2] * tz2)) * tx2 + ((c[3] * ty2 + c[4] * tz2) * ty2 + c[5] * tz4)) * tx2 +
      ((c[6] * ty2 + c[7] * tz2) * ty2 + c[8] * tz4) * ty2 + c[9] * tz4 * tz2;
  const double iR5 = p->iR2 * p->iR2 * p->iR;
  return (term9 + term7 + term5 + term3) * iR5 * p->x * p->y;
}

/* Refined evaluation (:260-292, :404-440): the source cell is cut into
 * cnt[0] x cnt[1] x cnt[2] sub-cells; the pair sum over sub-cell offsets is
 * folded into one sum over offsets weighted by (cnt - |offset|). */
typedef struct {
  refine_t r;
  int kind;                /* 0: Nxx-type, 1: Nxy-type */
  nxx_t xx;
  nxy_t xy;
} asym_t;

static double asym_base(const asym_t *A, const nab_t *p) {
  return A->kind == 0 ? nxx_eval(&A->xx, p) : nxy_eval(&A->xy, p);
}

static double asym_eval(const asym_t *A, double x, double y, double z) {
  const int xc = A->r.cnt[0], yc = A->r.cnt[1], zc = A->r.cnt[2];
  nab_t p, q;
  double zsum = 0.0;
  for (int k = 1 - zc; k < zc; k++) {
    const double zo = z + (double)k * A->r.rd[2];
    double ysum = 0.0;
    for (int j = 1 - yc; j < yc; j++) {
      const double yo = y + (double)j * A->r.rd[1];
      nab_set(&p, x, yo, zo);
      double xsum = (double)xc * asym_base(A, &p);
      for (int i = 1; i < xc; i++) {
        nab_set(&p, x + (double)i * A->r.rd[0], yo, zo);
        nab_set(&q, x - (double)i * A->r.rd[0], yo, zo);
        xsum += (double)(xc - i) * (asym_base(A, &p) + asym_base(A, &q));
      }
      ysum += (double)(yc - abs(j)) * xsum;
    }
    zsum += (double)(zc - abs(k)) * ysum;
  }
  return zsum * A->r.scale;
}

/* The six far-field evaluators: Nxx, Nyy, Nzz from the Nxx form and Nxy, Nxz,
 * Nyz from the Nxy form, with the axes permuted (:442-504).  perm[i] names
 * the original axis that plays the role of axis i. */
typedef struct {
  asym_t A;
  int perm[3];
} asym_comp_t;

static void asym_comp_init(asym_comp_t *c, int kind, int p0, int p1, int p2, const double d[3]) {
  c->perm[0] = p0; c->perm[1] = p1; c->perm[2] = p2;
  c->A.kind = kind;
  refine_init(&c->A.r, d[p0], d[p1], d[p2], 1.5);
  if (kind == 0) nxx_init(&c->A.xx, &c->A.r);
  else nxy_init(&c->A.xy, &c->A.r);
}

static double asym_comp_eval(const asym_comp_t *c, const double R[3]) {
  return asym_eval(&c->A, R[c->perm[0]], R[c->perm[1]], R[c->perm[2]]);
}

/* ---- rational approximation of the edge ratios -------------------------- */
static double gcd_(double m, double n) {   /* Gcd (:937-952) */
  m = floor(fabs(m));
  n = floor(fabs(n));
  if (m == 0.0 || n == 0.0) return 0.0;
  double t = m - floor(m / n) * n;
  while (t > 0.0) {
    m = n;
    n = t;
    t = m - floor(m / n) * n;
  }
  return n;
}

/* FindRatApprox (:967-1010): nearest-integer continued fraction of x/y */
static int rat_approx(double x, double y, double relerr, double maxq, double *p, double *q) {
  if (x < 0.0) x = -x;
  if (y < 0.0) y = -y;
  int swap = 0;
  if (x < y) { double t = x; x = y; y = t; swap = 1; }
  const double x0 = x, y0 = y;
  double p0 = 0.0, p1 = 1.0, q0 = 1.0, q1 = 0.0;
  double m = floor(x / y);
  double r = x - m * y;
  double p2 = m * p1 + p0, q2 = m * q1 + q0;
  while (q2 < maxq && fabs(x0 * q2 - p2 * y0) > relerr * x0 * q2) {
    x = y;
    y = r;
    m = floor(x / y);
    r = x - m * y;
    p0 = p1; p1 = p2; p2 = m * p1 + p0;
    q0 = q1; q1 = q2; q2 = m * q1 + q0;
  }
  if (!swap) { *p = p2; *q = q2; }
  else { *p = q2; *q = p2; }
  return fabs(x0 * q2 - p2 * y0) <= relerr * x0 * q2;
}

/* Normalised cell edges (:1147-1166): integers when the ratios are rational
 * with denominators below 1000, else scaled so the longest edge is 1. */
void um_or_newell_cell(const double cs[3], double d[3]) {
  double p1, q1, p2, q2;
  d[0] = cs[0]; d[1] = cs[1]; d[2] = cs[2];
  if (rat_approx(d[0], d[1], 1e-12, 1000, &p1, &q1) && rat_approx(d[2], d[1], 1e-12, 1000, &p2, &q2)) {
    const double g = gcd_(q1, q2);
    d[0] = p1 * q2 / g;
    d[1] = q1 * q2 / g;
    d[2] = p2 * q1 / g;
  } else {
    double mx = d[0];
    if (d[1] > mx) mx = d[1];
    if (d[2] > mx) mx = d[2];
    d[0] /= mx; d[1] /= mx; d[2] /= mx;
  }
}

/* Far-field value of component c (xx yy zz yz xz xy order) at offset R, in the
 * normalised cell units; exposed for the expansion tests. */
double um_or_newell_asymptotic(int comp, const double d[3], double Rx, double Ry, double Rz) {
  asym_comp_t c;
  switch (comp) {
    case 0: asym_comp_init(&c, 0, 0, 1, 2, d); break;  /* xx */
    case 1: asym_comp_init(&c, 0, 1, 0, 2, d); break;  /* yy */
    case 2: asym_comp_init(&c, 0, 2, 1, 0, d); break;  /* zz */
    case 3: asym_comp_init(&c, 1, 1, 2, 0, d); break;  /* yz */
    case 4: asym_comp_init(&c, 1, 0, 2, 1, d); break;  /* xz */
    default: asym_comp_init(&c, 1, 0, 1, 2, d); break; /* xy */
  }
  const double R[3] = {Rx, Ry, Rz};
  return asym_comp_eval(&c, R);
}

/* Exact value of component c (same order) times -1/(4 pi V): the demag
 * kernel element -N at offset R (CalculateSDA* times `scale`, :1170). */
double um_or_newell_exact(int comp, const double d[3], double Rx, double Ry, double Rz) {
  const double scale = -1.0 / (4.0 * M_PI * d[0] * d[1] * d[2]);
  switch (comp) {
    case 0: return um_or_newell_sda00(Rx, Ry, Rz, d[0], d[1], d[2]) * scale;
    case 1: return um_or_newell_sda00(Ry, Rx, Rz, d[1], d[0], d[2]) * scale;
    case 2: return um_or_newell_sda00(Rz, Ry, Rx, d[2], d[1], d[0]) * scale;
    case 3: return um_or_newell_sda01(Ry, Rz, Rx, d[1], d[2], d[0]) * scale;
    case 4: return um_or_newell_sda01(Rx, Rz, Ry, d[0], d[2], d[1]) * scale;
    default: return um_or_newell_sda01(Rx, Ry, Rz, d[0], d[1], d[2]) * scale;
  }
}

/* CalcNewellDemagKernel, mag/newellDemagKernel.go:1108-1324, for open
 * boundaries.  K: xx yy zz yz xz xy, each of size P = padSize(N), FP64 (the
 * reference rounds each element to float32 when it stores it).  In 2D the yz
 * and xz outputs are zero.  Returns 0 on success. */
int um_or_demag_kernel_newell(const int N[3], const int pbc[3], const double cs[3],
                              int asymptotic_radius, int zero_self_demag, double *K[6]) {
  int size[3];
  for (int i = 0; i < 3; i++)
    if (pbc[i] < 0) return -1;
  um_or_pad_size(N, pbc, size);
  if (!(size[0] > 0 && size[1] > 0 && size[2] > 0)) return -1;
  if (!(cs[0] > 0 && cs[1] > 0 && cs[2] > 0) || zero_self_demag < 0) return -1;
  const long P = (long)size[0] * size[1] * size[2];
  for (int c = 0; c < 6; c++) memset(K[c], 0, sizeof(double) * P);
  /* kernelRanges (mag/demagkernel.go:339-352) */
  int r1[3], r2[3];
  for (int c = 0; c < 3; c++) {
    if (pbc[c] == 0) { r1[c] = -(size[c] - 1) / 2; r2[c] = (size[c] - 1) / 2; }
    else { r1[c] = -(size[c] * pbc[c] - 1); r2[c] = size[c] * pbc[c] - 1; }
  }
  if (size[2] == 1 && pbc[2] == 0) r2[2] = 0;

  double d[3];
  um_or_newell_cell(cs, d);
  const double arad = (double)asymptotic_radius * pow(d[0] * d[1] * d[2], 1.0 / 3.0);
  const double arad2 = arad * arad;
  asym_comp_t far[6];
  asym_comp_init(&far[0], 0, 0, 1, 2, d);
  asym_comp_init(&far[1], 0, 1, 0, 2, d);
  asym_comp_init(&far[2], 0, 2, 1, 0, d);
  asym_comp_init(&far[3], 1, 1, 2, 0, d);
  asym_comp_init(&far[4], 1, 0, 2, 1, d);
  asym_comp_init(&far[5], 1, 0, 1, 2, d);

#define AT(c, z, y, x) K[c][((long)(z) * size[1] + (y)) * size[0] + (x)]
  /* The reference runs one goroutine per offset z (:1165); with PBC the
   * goroutines of the z images of one plane add into the same float32
   * elements unsynchronised.  Here one task owns each wrapped plane and sums
   * its images in ascending z (then y, x) in FP64. */
  const int nzw = size[2] / 2 + 1;
#pragma omp parallel for schedule(dynamic)
  for (int zt = 0; zt < nzw; zt++) {
    for (int z = r1[2]; z <= r2[2]; z++) {
      const int zw = (z % size[2] + size[2]) % size[2];
      if (zw != zt) continue;
      double R[3];
      R[2] = (double)z * d[2];
      for (int y = r1[1]; y <= r2[1]; y++) {
        const int yw = (y % size[1] + size[1]) % size[1];
        if (yw > size[1] / 2) continue;
        R[1] = (double)y * d[1];
        for (int x = r1[0]; x <= r2[0]; x++) {
          const int xw = (x % size[0] + size[0]) % size[0];
          if (xw > size[0] / 2) continue;
          R[0] = (double)x * d[0];
          const int farfield = arad >= 0.0 && R[0] * R[0] > arad2 - R[2] * R[2] - R[1] * R[1];
          for (int c = 0; c < 6; c++) {
            const double v = farfield ? -1.0 * asym_comp_eval(&far[c], R)
                                      : um_or_newell_exact(c, d, R[0], R[1], R[2]);
            AT(c, zw, yw, xw) += v;
          }
        }
      }
    }
  }
  /* mirror images (:1253-1298): odd components change sign with their axis */
  for (int z = 0; z < size[2]; z++)
    for (int y = 0; y < size[1]; y++)
      for (int x = 0; x < size[0]; x++) {
        const int mx = x > size[0] / 2, my = y > size[1] / 2, mz = z > size[2] / 2;
        if (!(mx || my || mz)) continue;
        const int x2 = mx ? size[0] - x : x, y2 = my ? size[1] - y : y, z2 = mz ? size[2] - z : z;
        const double sx = mx ? -1.0 : 1.0, sy = my ? -1.0 : 1.0, sz = mz ? -1.0 : 1.0;
        AT(0, z, y, x) = AT(0, z2, y2, x2);
        AT(1, z, y, x) = AT(1, z2, y2, x2);
        AT(2, z, y, x) = AT(2, z2, y2, x2);
        AT(3, z, y, x) = sy * sz * AT(3, z2, y2, x2);
        AT(4, z, y, x) = sx * sz * AT(4, z2, y2, x2);
        AT(5, z, y, x) = sx * sy * AT(5, z2, y2, x2);
      }
  /* self term (:1300-1310); with PBC the images summed into the origin
   * element are discarded with it, as in the reference */
  double third = zero_self_demag != 0 ? 1.0 / 3.0 : 0.0;
  AT(0, 0, 0, 0) = -um_or_self_demag_nx(d[0], d[1], d[2]) + third;
  AT(1, 0, 0, 0) = -um_or_self_demag_nx(d[1], d[2], d[0]) + third;
  AT(2, 0, 0, 0) = -um_or_self_demag_nx(d[2], d[0], d[1]) + third;
  AT(3, 0, 0, 0) = AT(4, 0, 0, 0) = AT(5, 0, 0, 0) = 0.0;
  if (size[2] == 1)
    for (long i = 0; i < P; i++) K[3][i] = K[4][i] = 0.0;
#undef AT
  return 0;
}
