/*
 * um_oracle.h -- CPU oracle for the uMagNUS effective-field + time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm (OpenCL kernels under kernels_src/cl/, Go host code under engine/,
 * opencl/ and mag/), used as the checker by tests/, by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py.  The product (libumhip.so) never
 * links, loads or calls anything from here.
 *
 * Every function is built twice from um_oracle_impl.c: once with real = float
 * (suffix _f32) and once with real = double (suffix _f64).  Functions without
 * a suffix are precision independent.
 *
 * Parity status of each function (what pins it) is listed next to it:
 *   [golden]   reference test expectation (test/NAME.mx3, opencl/reduce_test.go)
 *   [closed]   closed form / invariant / independent library (numpy.fft)
 *   [energy]   finite-difference derivative of an independently written energy
 */
#ifndef UM_ORACLE_H
#define UM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UM_OR_NREGION 256
#define UM_OR_SYMLUT (UM_OR_NREGION * (UM_OR_NREGION + 1) / 2)

/* ---- precision independent ---------------------------------------------- */

/* Philox-4x32-10 block (Salmon et al., SC'11).  [closed: Random123 KAT] */
void um_or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Brute-force demag tensor, mag/demagkernel.go:115-336.  Writes the six
 * real-space components (xx,yy,zz,yz,xz,xy) of size P = padSize(N, pbc) in
 * FP64; with PBC each element sums its periodic images (kernelRanges).  In
 * 2D (P[2]==1) the yz and xz outputs are zero.  Returns 0 on success.
 * [golden: test/demagSmall*.mx3, demagcube.mx3, demagodd.mx3, demag2Dpbc*.mx3,
 *  demagrodpbc.mx3; closed: sum of open-boundary images] */
int um_or_demag_kernel_bf(const int N[3], const int pbc[3], const double cs[3],
                          double accuracy, double *K[6]);
/* padSize, mag/demagkernel.go:404-420 */
void um_or_pad_size(const int N[3], const int pbc[3], int P[3]);

/* Newell demag tensor, mag/newellDemagKernel.go:1108-1324, with periodic
 * images as the brute-force one (the origin keeps the self term alone).
 * Same output layout as um_or_demag_kernel_bf.
 * [closed: prism demag factors (Aharoni), block sums, far-field expansion vs
 *  the exact form; golden: test/standardproblem4_newell-3d.mx3,
 *  standardproblem4b_newell.mx3] */
int um_or_demag_kernel_newell(const int N[3], const int pbc[3], const double cs[3],
                              int asymptotic_radius, int zero_self_demag, double *K[6]);
/* pieces of it, for the tests: Newell's f and g, SelfDemagNx, the 27-point
 * sums CalculateSDA00/01, the normalised cell edges, and one tensor element
 * (xx yy zz yz xz xy = 0..5) from the exact form or the far-field expansion,
 * both as kernel values (-N) for the exact one and N for the expansion. */
double um_or_newell_f(double x, double y, double z);
double um_or_newell_g(double x, double y, double z);
double um_or_self_demag_nx(double x, double y, double z);
double um_or_newell_sda00(double x, double y, double z, double dx, double dy, double dz);
double um_or_newell_sda01(double x, double y, double z, double dx, double dy, double dz);
void um_or_newell_cell(const double cs[3], double d[3]);
double um_or_newell_exact(int comp, const double d[3], double Rx, double Ry, double Rz);
double um_or_newell_asymptotic(int comp, const double d[3], double Rx, double Ry, double Rz);

/* symmidx, engine/exchange.go:181-187 */
int um_or_symmidx(int i, int j);
/* exchAverage, engine/exchange.go:193-202 (float32 arithmetic, as Go does) */
float um_or_exch_average(float exi, float exj);

/* adaptDt, engine/run.go:150-187.  State passed explicitly.  Returns new dt,
 * or a negative value where the reference would abort (corr==0 / dt==0). */
double um_or_adapt_dt(double dt, double corr, double fixdt, double headroom,
                      double mindt, double maxdt, double time, double alarm);

/* ---- precision generic: declared for both suffixes ---------------------- */
#define UM_OR_DECL(SFX, real)                                                            \
  void um_or_addexchange##SFX(real *Bx, real *By, real *Bz, const real *mx,             \
        const real *my, const real *mz, const real *Ms_, real Ms_mul,                    \
        const real *aLUT, const uint8_t *regions, real wx, real wy, real wz,             \
        int Nx, int Ny, int Nz, int pbc);                                                \
  void um_or_adddmi##SFX(real *Hx, real *Hy, real *Hz, const real *mx, const real *my,  \
        const real *mz, const real *Ms_, real Ms_mul, const real *aLUT,                  \
        const real *dLUT, const uint8_t *regions, real cx, real cy, real cz,             \
        int Nx, int Ny, int Nz, int pbc, int openbc);                                    \
  void um_or_adduniaxial2##SFX(real *Bx, real *By, real *Bz, const real *mx,            \
        const real *my, const real *mz, const real *Ms_, real Ms_mul, const real *K1_,   \
        real K1_mul, const real *K2_, real K2_mul, const real *ux_, real ux_mul,         \
        const real *uy_, real uy_mul, const real *uz_, real uz_mul, int N);              \
  void um_or_addcubic2##SFX(real *Bx, real *By, real *Bz, const real *mx,               \
        const real *my, const real *mz, const real *Ms_, real Ms_mul, const real *k1_,   \
        real k1_mul, const real *k2_, real k2_mul, const real *k3_, real k3_mul,         \
        const real *c1x_, real c1x_mul, const real *c1y_, real c1y_mul,                  \
        const real *c1z_, real c1z_mul, const real *c2x_, real c2x_mul,                  \
        const real *c2y_, real c2y_mul, const real *c2z_, real c2z_mul, int N);          \
  void um_or_regionaddv##SFX(real *dx, real *dy, real *dz, const real *lx,              \
        const real *ly, const real *lz, const uint8_t *regions, int N);                  \
  void um_or_zeromask##SFX(real *dst, const real *maskLUT, const uint8_t *regions, int N);\
  void um_or_adddmibulk##SFX(real *Hx, real *Hy, real *Hz, const real *mx, const real *my, \
        const real *mz, const real *Ms_, real Ms_mul, const real *aLUT,                  \
        const real *dLUT, const uint8_t *regions, real cx, real cy, real cz,             \
        int Nx, int Ny, int Nz, int pbc, int openbc);                                    \
  void um_or_addzhangli##SFX(real *tx, real *ty, real *tz, const real *mx,              \
        const real *my, const real *mz, const real *Ms_, real Ms_mul, const real *jx_,   \
        real jx_mul, const real *jy_, real jy_mul, const real *jz_, real jz_mul,         \
        const real *alpha_, real alpha_mul, const real *xi_, real xi_mul,                \
        const real *pol_, real pol_mul, real cx, real cy, real cz, int Nx, int Ny,       \
        int Nz, int pbc);                                                                \
  void um_or_addslonczewski##SFX(real *tx, real *ty, real *tz, const real *mx,          \
        const real *my, const real *mz, const real *Ms_, real Ms_mul, const real *jz_,   \
        real jz_mul, const real *px_, real px_mul, const real *py_, real py_mul,         \
        const real *pz_, real pz_mul, const real *alpha_, real alpha_mul,                \
        const real *pol_, real pol_mul, const real *lambda_, real lambda_mul,            \
        const real *eps_, real eps_mul, const real *thick_, real thick_mul,              \
        real meshThickness, real flp, int N);                                            \
  void um_or_lltorque2##SFX(real *tx, real *ty, real *tz, const real *mx,               \
        const real *my, const real *mz, const real *hx, const real *hy, const real *hz,  \
        const real *alpha_, real alpha_mul, int N);                                      \
  void um_or_llnoprecess##SFX(real *tx, real *ty, real *tz, const real *mx,             \
        const real *my, const real *mz, const real *hx, const real *hy, const real *hz,  \
        int N);                                                                          \
  void um_or_madd##SFX(real *dst, int nsrc, const real *const *src, const real *fac,     \
        int N);                                                                          \
  void um_or_normalize2##SFX(real *vx, real *vy, real *vz, const real *vol, int N);      \
  void um_or_settemperature2##SFX(real *B, const real *noise, real kB2_VgammaDt,         \
        const real *Ms_, real Ms_mul, const real *temp_, real temp_mul,                  \
        const real *alpha_, real alpha_mul, int N);                                      \
  double um_or_maxvecnorm##SFX(const real *x, const real *y, const real *z, int N);      \
  double um_or_maxvecdiff##SFX(const real *x1, const real *y1, const real *z1,           \
        const real *x2, const real *y2, const real *z2, int N, int sqrt3_quirk);         \
  double um_or_maxabs##SFX(const real *x, int N);                                        \
  double um_or_sum##SFX(const real *x, int N);                                           \
  double um_or_dot##SFX(const real *a, const real *b, int N);                            \
  void um_or_fft##SFX(real *re_im, int n, int sign);                                     \
  void um_or_fft3d##SFX(real *re_im, const int P[3], int sign);                          \
  void *um_or_demag_new##SFX(const int N[3], const int pbc[3], const double cs[3],       \
        double accuracy, int nthreads);                                                  \
  void *um_or_demag_new_from_kernel##SFX(const int N[3], const int P[3],                 \
        double *const K[6]);                                                             \
  void um_or_demag_free##SFX(void *h);                                                   \
  void um_or_demag_exec##SFX(void *h, real *Bx, real *By, real *Bz, const real *mx,      \
        const real *my, const real *mz, const real *vol, const real *Ms_, real Ms_mul);  \
  int um_or_demag_kspec##SFX(void *h, int comp, real *out);                              \
  void um_or_demag_brute##SFX(const int N[3], const int P[3], double *const K[6],        \
        real *Bx, real *By, real *Bz, const real *mx, const real *my, const real *mz,    \
        const real *vol, const real *Ms_, real Ms_mul);                                  \
  void um_or_thermal_normals##SFX(real *nx, real *ny, real *nz, int N, uint64_t seed,    \
        uint64_t step, uint32_t stream);

UM_OR_DECL(_f32, float)
UM_OR_DECL(_f64, double)

/* ---- the simulation (engine-level restatement) --------------------------- */
/* Parameter ids used by um_or_sim_set_* (names follow the .mx3 quantities). */
enum {
  UM_P_MSAT = 0, UM_P_KU1, UM_P_KU2, UM_P_KC1, UM_P_KC2, UM_P_KC3, UM_P_ALPHA,
  UM_P_TEMP, UM_P_ANISU, UM_P_ANISC1, UM_P_ANISC2, UM_P_BEXT,
  /* spin-transfer torque (engine/torque.go:11-20): xi, Pol, Lambda, EpsilonPrime,
   * FreeLayerThickness, J (vector, A/m^2), FixedLayer (vector) */
  UM_P_XI, UM_P_POL, UM_P_LAMBDA, UM_P_EPSPRIME, UM_P_FREELAYER, UM_P_J, UM_P_FIXEDLAYER,
  UM_P_COUNT
};
enum { UM_SOLVER_HEUN = 2, UM_SOLVER_RK23 = 3, UM_SOLVER_RK4 = 4, UM_SOLVER_RK45DP = 5 };

#define UM_OR_SIM_DECL(SFX, real)                                                          \
  void *um_or_sim_new##SFX(int Nx, int Ny, int Nz, double cx, double cy, double cz,      \
        int pbcx, int pbcy, int pbcz);                                                     \
  void um_or_sim_free##SFX(void *h);                                                       \
  /* scalar param: arr (per cell, may be NULL) times mul; vectors use 3 arrays */          \
  void um_or_sim_set_scalar##SFX(void *h, int id, const real *arr, double mul);            \
  void um_or_sim_set_vector##SFX(void *h, int id, const real *ax, const real *ay,          \
        const real *az, double mx, double my, double mz);                                  \
  void um_or_sim_set_regions##SFX(void *h, const uint8_t *regions);                        \
  void um_or_sim_set_exchange_luts##SFX(void *h, const real *aLUT, const real *dLUT);      \
  void um_or_sim_set_bext_lut##SFX(void *h, const real *lx, const real *ly,                \
        const real *lz);                                                                   \
  void um_or_sim_set_frozen_lut##SFX(void *h, const real *lut);                            \
  void um_or_sim_set_nodemag_lut##SFX(void *h, const real *lut);                           \
  void um_or_sim_set_geometry##SFX(void *h, const real *vol);                              \
  void um_or_sim_set_flags##SFX(void *h, int enable_demag, int openbc, int precess,        \
        int heun_sqrt3);                                                                   \
  void um_or_sim_set_demag_accuracy##SFX(void *h, double acc);                             \
  void um_or_sim_minimize##SFX(void *h, double stop, int samples);                         \
  /* bulk DMI LUT (Dbulk, engine/exchange.go:49-66); NULL: none */                         \
  void um_or_sim_set_dmi_bulk##SFX(void *h, const real *lut);                               \
  /* DisableZhangLiTorque, DisableSlonczewskiTorque, FixedLayerPosition sign (+1 top) */   \
  void um_or_sim_set_stt_flags##SFX(void *h, int disable_zl, int disable_sl, double flp);   \
  /* EnableNewellDemag (engine/demag.go:26-56): 1 = Newell tensor, 0 = brute force */      \
  void um_or_sim_set_demag_newell##SFX(void *h, int enable, int asymptotic_radius,         \
        int zero_self_demag);                                                              \
  void um_or_sim_set_thermal_seed##SFX(void *h, uint64_t seed, uint32_t stream);           \
  void um_or_sim_set_m##SFX(void *h, const real *mx, const real *my, const real *mz);      \
  void um_or_sim_get_m##SFX(void *h, real *mx, real *my, real *mz);                        \
  void um_or_sim_beff##SFX(void *h, real *bx, real *by, real *bz);                         \
  void um_or_sim_torque##SFX(void *h, real *tx, real *ty, real *tz);                       \
  void um_or_sim_demag_field##SFX(void *h, real *bx, real *by, real *bz);                  \
  void um_or_sim_set_solver##SFX(void *h, int solver);                                     \
  /* solver knobs: dt, fixdt, maxerr, relerr, minslope, headroom, mindt, maxdt, gammaLL */ \
  void um_or_sim_set_solver_params##SFX(void *h, const double *knobs9);                    \
  void um_or_sim_get_state##SFX(void *h, double *out8);                                    \
  void um_or_sim_step##SFX(void *h);                                                       \
  void um_or_sim_steps##SFX(void *h, int n);                                               \
  void um_or_sim_run##SFX(void *h, double seconds);                                        \
  void um_or_sim_relax##SFX(void *h, double torque_threshold);                              \
  void um_or_sim_average_m##SFX(void *h, double out[3]);                                   \
  /* energy terms: 0 demag, 1 exchange(+DMI), 2 Zeeman, 3 uniaxial K1, 4 uniaxial K2,      \
   * 5..7 cubic Kc1..Kc3, 8 thermal, 9 total (engine/energy.go:27-34) */                   \
  double um_or_sim_energy##SFX(void *h, int term);

UM_OR_SIM_DECL(_f32, float)
UM_OR_SIM_DECL(_f64, double)

#ifdef __cplusplus
}
#endif
#endif
