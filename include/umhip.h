/*
 * umhip.h -- C ABI of libumhip.so, the MI355X (gfx950) implementation of the
 * uMagNUS effective-field and time-stepping path.
 *
 * The library replaces the reference's cgo OpenCL binding (the Go files of cl/) and its
 * VkFFT wrapper (cl/vkfft.go, cl/fft_interface.h) underneath the exported API
 * of package opencl / opencl64.  Every entry point below names the reference
 * function it stands in for.  Conventions:
 *
 *  - Every function returns 0 on success and a negative UM_E* code on error;
 *    um_last_error() then describes the failure (thread-local string).
 *  - "prec" selects the arithmetic type: UM_F32 (package opencl) or UM_F64
 *    (package opencl64).  Device arrays are plain device pointers of that type.
 *  - Vector quantities are three separate component arrays (SoA), x fastest:
 *    idx = (iz*Ny + iy)*Nx + ix, as data.Slice (data/slice.go:26-32).
 *  - A parameter in the reference's MSlice convention (opencl/mslice.go:11-45)
 *    is passed as (arr, mul): arr == NULL means "uniform, value mul", else the
 *    per-cell value is mul*arr[i].
 *  - All device work is enqueued in order on one HIP stream per library
 *    (um_set_stream) or per simulation handle, like the reference's single
 *    in-order command queue (opencl/init.go:145).  Host pointers are never
 *    retained past a call (cgo pointer rules).
 */
#ifndef UMHIP_H
#define UMHIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { UM_F32 = 0, UM_F64 = 1 };
enum {
  UM_OK = 0,
  UM_EINVAL = -1,   /* bad argument (size mismatch, NULL where not allowed) */
  UM_EHIP = -2,     /* HIP runtime error */
  UM_ENOMEM = -3,   /* device allocation failed */
  UM_EUNSUP = -4,   /* unsupported size or option */
  UM_ESTATE = -5    /* call out of order (e.g. before um_init) */
};

/* ---- runtime: opencl/init.go, opencl/slice.go, opencl/buffer.go ----------- */
/* Init(gpu int), opencl/init.go:48 */
int um_init(int device);
/* description of the last error on this thread */
const char *um_last_error(void);
/* name and compute-unit count of the current device */
int um_device_info(char *name, int name_len, int *cu_count, size_t *mem_bytes);
/* stream used by the unfused ops from now on; NULL selects the HIP null
 * stream.  Until called, the library uses a stream of its own. */
int um_set_stream(void *hip_stream);
/* Finish() on the command queue */
int um_sync(void);
/* MemAlloc / NewSlice, opencl/slice.go:14-45 */
int um_malloc(void **ptr, size_t bytes);
/* memFree, opencl/slice.go:46 */
int um_free(void *ptr);
/* MemCpyHtoD / MemCpyDtoH / MemCpy, opencl/slice.go:53-155 (blocking) */
int um_memcpy_h2d(void *dst, const void *src, size_t bytes);
int um_memcpy_d2h(void *dst, const void *src, size_t bytes);
int um_memcpy_d2d(void *dst, const void *src, size_t bytes);
/* Zero / Memset, opencl/slice.go:156-194 (kernel fill, safe inside graphs) */
int um_zero(void *dst, size_t bytes);
int um_memset(int prec, void *dst, double value, size_t n);

/* ---- field terms: opencl/exchange.go, dmi.go, anisotropy.go, region.go ---- */
/* AddExchange, opencl/exchange.go:16 (kernels_src/cl/exchange.cl) */
int um_add_exchange(int prec, void *const B[3], const void *const m[3], const void *Ms,
                    double Ms_mul, const void *aLUT, const uint8_t *regions, double wx,
                    double wy, double wz, int Nx, int Ny, int Nz, int pbc);
/* AddDMI, opencl/dmi.go:15 (kernels_src/cl/dmi.cl) */
int um_add_dmi(int prec, void *const B[3], const void *const m[3], const void *Ms,
               double Ms_mul, const void *aLUT, const void *dLUT, const uint8_t *regions,
               double cx, double cy, double cz, int Nx, int Ny, int Nz, int pbc, int openbc);
/* AddUniaxialAnisotropy2, opencl/anisotropy.go:153 */
int um_add_uniaxial2(int prec, void *const B[3], const void *const m[3], const void *Ms,
                     double Ms_mul, const void *K1, double K1_mul, const void *K2,
                     double K2_mul, const void *const u[3], const double u_mul[3], int N);
/* AddCubicAnisotropy2, opencl/anisotropy.go:12 */
int um_add_cubic2(int prec, void *const B[3], const void *const m[3], const void *Ms,
                  double Ms_mul, const void *K1, double K1_mul, const void *K2, double K2_mul,
                  const void *K3, double K3_mul, const void *const c1[3],
                  const double c1_mul[3], const void *const c2[3], const double c2_mul[3],
                  int N);
/* RegionAddV, opencl/region.go:13 */
int um_region_addv(int prec, void *const dst[3], const void *const lut[3],
                   const uint8_t *regions, int N);
/* ZeroMask, opencl/zeromask.go:12 */
int um_zeromask(int prec, void *dst, const void *maskLUT, const uint8_t *regions, int N);
/* LLTorque, opencl/lltorque.go:16 */
int um_lltorque(int prec, void *const t[3], const void *const m[3], const void *const B[3],
                const void *alpha, double alpha_mul, int N);
/* LLNoPrecess, opencl/lltorque.go:97 */
int um_llnoprecess(int prec, void *const t[3], const void *const m[3],
                   const void *const B[3], int N);
/* SetTemperature, opencl/temperature.go:14, with the noise drawn in place from
 * the Philox-4x32-10 stream (seed, step, stream) instead of a staging buffer
 * filled by Generator.Normal (opencl/rng.go:146).  k = 2 kB / (gamma V dt). */
int um_thermal_field(int prec, void *const B[3], uint64_t seed, uint64_t step,
                     uint32_t stream, double k, const void *Ms, double Ms_mul,
                     const void *temp, double temp_mul, const void *alpha, double alpha_mul,
                     int N);

/* ---- vector arithmetic: opencl/madd.go, normalize.go, vecnorm.go ---------- */
/* AddDMIBulk, opencl/dmibulk.go:14 (dmibulk.cl) */
int um_add_dmi_bulk(int prec, void *const B[3], const void *const m[3], const void *Ms, double Ms_mul,
                    const void *aLUT, const void *dLUT, const uint8_t *regions, double cx, double cy, double cz,
                    int Nx, int Ny, int Nz, int pbc, int openbc);
/* AddZhangLiTorque, opencl/zhangli.go:12 (zhangli2.cl): tau += torque */
int um_add_zhangli(int prec, void *const tau[3], const void *const m[3], const void *Ms, double Ms_mul,
                   const void *const J[3], const double J_mul[3], const void *alpha, double alpha_mul,
                   const void *xi, double xi_mul, const void *pol, double pol_mul, double cx, double cy,
                   double cz, int Nx, int Ny, int Nz, int pbc);
/* AddSlonczewskiTorque2, opencl/slonczewski.go:12 (slonczewski2.cl): tau += torque;
 * mesh_thickness = Nz*cz, flp = +1 (fixed layer on top) or -1 */
int um_add_slonczewski(int prec, void *const tau[3], const void *const m[3], const void *Ms, double Ms_mul,
                       const void *Jz, double Jz_mul, const void *const fixedP[3], const double fixedP_mul[3],
                       const void *alpha, double alpha_mul, const void *pol, double pol_mul, const void *lambda,
                       double lambda_mul, const void *epsPrime, double epsPrime_mul, const void *thickness,
                       double thickness_mul, double mesh_thickness, double flp, int N);

/* Madd2..Madd7, opencl/madd.go:111-489: dst = sum fac[k]*src[k], nsrc in 1..7 */
int um_madd(int prec, void *dst, int nsrc, const void *const *src, const double *fac, int N);
/* Normalize, opencl/normalize.go:13 */
int um_normalize(int prec, void *const v[3], const void *vol, int N);
/* VecNorm, opencl/vecnorm.go:12 */
int um_vecnorm(int prec, void *dst, const void *const a[3], int N);
/* Div, opencl/madd.go:60 (0 where b == 0) */
int um_div(int prec, void *dst, const void *a, const void *b, int N);
/* Mul, opencl/madd.go:13 */
int um_mul(int prec, void *dst, const void *a, const void *b, int N);

/* ---- fused evaluation and RK stage entry points ----------------------------
 * What an unchanged engine needs to run the fused kernels through its own
 * call sequence: one pass per torque evaluation (SetTorque without B_demag),
 * one pass per RK stage (MaddN + Normalize) and the step controller's norms
 * with one readback (integralController). */

/* A parameter in the MSlice convention (opencl/mslice.go:11-45): arr == NULL
 * means uniform with value mul, else the per-cell value is mul*arr[i]. */
typedef struct {
  const void *arr;
  double mul;
} um_param;

/* Inputs of um_field_torque.  Arrays are device pointers of the call's
 * precision, vectors are three component arrays (SoA).  Zero-initialise the
 * struct: every NULL pointer / zero parameter switches its term off. */
typedef struct {
  int Nx, Ny, Nz;
  int pbc;                       /* PBC_code (data/mesh.go): bit 0 x, bit 1 y, bit 2 z */
  double cellsize[3];
  const void *m[3];              /* magnetization (M.Buffer()) */
  const void *bdemag[3];         /* B_demag (um_demag_exec output), all NULL: no demag term */
  void *torque[3];               /* out, SetTorque (engine/torque.go:124), all NULL: not written; may alias bdemag */
  void *beff[3];                 /* out, SetEffectiveField (engine/effectivefield.go:14), all NULL: not written */
  const uint8_t *regions;        /* region byte per cell, NULL: every cell in region 0 */
  um_param Msat;
  const void *aex_lut;           /* Aex_red SymmLUT, 256*257/2 reals (engine/exchange.go:153-177); NULL: no exchange */
  const void *dmi_lut;           /* Dind_red SymmLUT, or Dbulk's when dmi_bulk = 1; NULL: no DMI (needs aex_lut) */
  int dmi_bulk;
  int openbc;                    /* OpenBC (engine/exchange.go:28) */
  um_param Ku1, Ku2, anisU[3];   /* AddUniaxialAnisotropy2 (opencl/anisotropy.go:153) */
  um_param Kc1, Kc2, Kc3, anisC1[3], anisC2[3];   /* AddCubicAnisotropy2 (opencl/anisotropy.go:12) */
  const void *bext_lut[3];       /* B_ext region LUT, 256 reals per component (RegionAddV), or NULL */
  um_param Bext[3];              /* B_ext's uniform / per-cell part */
  int thermal;                   /* 1: B_therm from the Philox stream (seed, step, stream), as um_thermal_field */
  uint64_t seed, step;
  uint32_t stream;
  um_param Temp;
  double thermal_k;              /* 2 kB / (gammaLL V dt) (engine/temperature.go:109) */
  const void *badd[3];           /* field added to B_eff last (AddCustomField, engine/customfield.go:68), or NULL */
  um_param alpha;
  int precess;                   /* 1: LLTorque, 0: LLNoPrecess (opencl/lltorque.go:16,97) */
  int zhangli, slonczewski;      /* AddSTTorque terms (engine/torque.go:147-196) */
  um_param J[3], xi, pol, lambda, eps_prime, free_layer_thickness, fixed_layer[3];
  double fixed_layer_top;        /* CurrentSignFromFixedLayerPosition: +1 fixed layer on top, -1 below */
  const void *tadd[3];           /* torque added before FreezeSpins (AddCustomTorques, engine/torque.go:115), or NULL */
  const void *frozen_lut;        /* FrozenSpins per region (256 reals, non-zero: frozen), or NULL */
} um_field_args;

/* SetTorque (engine/torque.go:124-133) after SetDemagField: B_eff = bdemag +
 * exchange/DMI + anisotropy + B_ext + B_therm + badd in the order of
 * SetEffectiveField (engine/effectivefield.go:14-24), then LLTorque or
 * LLNoPrecess, the spin-transfer torques, tadd and FreezeSpins, in one pass
 * per cell on the library stream. */
int um_field_torque(int prec, const um_field_args *a);

/* One RK stage: Madd<nsrc>(dst, src...) then Normalize(dst, vol) in one pass
 * per cell (engine/rk45dp.go:50-87, opencl/madd.go:111-489).  src holds nsrc
 * vectors as 3*nsrc component pointers (src[3*j + c]), nsrc in 1..8; a
 * three-source sum is grouped x0 + (x1 + x2) like madd3.cl.  normalize = 0
 * leaves the sum as it is (Heun's first stage, engine/heun.go:32).  dst may
 * alias a source. */
int um_rk45dp_stage(int prec, void *const dst[3], int nsrc, const void *const *src, const double *fac,
                    const void *vol, int normalize, int N);

/* The norms integralController needs (engine/stepcontroller.go:35-62) from
 * the stage torques in one pass and one readback, without an Err buffer:
 * Err = sum_j e[j]*k_j (rk45dp.go:95; k = k1 k3 k4 k5 k6 k7 for RK45DP,
 * k1..k4 for RK23), delM = k[nk-1].  nk in 1..6; k holds 3*nk component
 * pointers (k[3*j + c]).  out[0] = MaxVecNorm(Err)*h (err), out[1] =
 * MaxAbs(VecNorm(Err)/VecNorm(delM)) with 0 where |delM| = 0 (rlerr),
 * out[2] = MaxVecNorm(delM) (maxdm).  A NaN anywhere in Err gives err = NaN. */
int um_rk45dp_error(int prec, int nk, const void *const *k, const double *e, double h, int N, double out[3]);

/* um_rk45dp_error without any allocation and without a readback; no
 * counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops): the three results (err,
 * rlerr, maxdm, same values as um_rk45dp_error) are written to out_dev, a
 * device array of 3 doubles, on the library stream, and nothing waits for
 * them.  The error kernel's max slots live in a 32-byte device array held by
 * each GPU until the process ends (its size does not depend on N); um_init
 * must have been called.  Calls on one GPU share that array, so calls that
 * overlap on one GPU (other streams, other threads) race. */
int um_rk45dp_error_noalloc(int prec, int nk, const void *const *k, const double *e, double h, int N,
                            double *out_dev);

/* ---- reductions: opencl/reduce.go (results in FP64, one readback) --------- */
int um_reduce_sum(int prec, const void *x, int N, double *out);
int um_reduce_dot(int prec, const void *a, const void *b, int N, double *out);
int um_reduce_maxabs(int prec, const void *x, int N, double *out);
int um_reduce_maxvecnorm(int prec, const void *const v[3], int N, double *out);
/* sqrt3_quirk=1 reproduces the widening in reducemaxvecdiff2.cl:24 */
int um_reduce_maxvecdiff(int prec, const void *const a[3], const void *const b[3], int N,
                         int sqrt3_quirk, double *out);

/* ---- demag: opencl/conv_demag.go, mag/demagkernel.go ---------------------- */
/* NewDemag(inputSize, PBC, DemagKernel(...)), opencl/conv_demag.go:27 with the
 * brute-force tensor of mag/demagkernel.go:115 computed on the GPU in FP64.
 * accuracy = DemagAccuracy (engine/demag.go:49).  pbc: SetPBC repetitions per
 * axis (NULL or 0: open); a periodic axis is not padded (P = N, odd Nx
 * included) and its tensor sums the periodic images of kernelRanges
 * (mag/demagkernel.go:339-352). */
int um_demag_create(int prec, int Nx, int Ny, int Nz, const int pbc[3], const double cs[3],
                    double accuracy, void **plan);
/* NewDemag with the Newell tensor: newellDemagConv (engine/demag.go:134-142)
 * with mag.NewellDemagKernel (mag/newellDemagKernel.go:1006, CalcNewellDemagKernel
 * :1108-1324) computed on the GPU in FP64.  asymptotic_radius: far-field
 * expansion beyond this many cells (reference default 32, negative: never);
 * zero_self_demag != 0 adds 1/3 to the diagonal self terms (default 0).
 * pbc as for um_demag_create; as in the reference the origin element holds
 * the self term alone (the images summed into it are discarded, :1300-1310). */
int um_demag_create_newell(int prec, int Nx, int Ny, int Nz, const int pbc[3], const double cs[3],
                           int asymptotic_radius, int zero_self_demag, void **plan);
/* NewDemag from a caller-supplied FP64 host tensor (xx,yy,zz,yz,xz,xy), each of
 * the padded size P (x fastest) -- e.g. the Newell tensor or an OVF cache. */
int um_demag_create_from_kernel(int prec, int Nx, int Ny, int Nz, const int P[3],
                                const double *const K[6], void **plan);
/* (*DemagConvolution).Exec(B, m, vol, Msat), opencl/conv_demag.go:43 */
int um_demag_exec(void *plan, void *const B[3], const void *const m[3], const void *vol,
                  const void *Ms, double Ms_mul);
/* (*DemagConvolution).Free, opencl/conv_demag.go:204 */
int um_demag_destroy(void *plan);
/* padded size of a plan and the kernel spectrum component c (0..5) copied to
 * host (H x (Py/2+1) x (Pz/2+1) values of the plan's precision) */
int um_demag_info(void *plan, int P[3]);
int um_demag_kspec(void *plan, int c, void *host_out);
/* CalcDemagKernel, mag/demagkernel.go:115-336, on the GPU in FP64; K[c] are
 * host buffers of the padded size P = padSize(N, pbc). */
int um_demag_kernel_bf(int Nx, int Ny, int Nz, const int pbc[3], const double cs[3],
                       double accuracy, double *const K[6]);
/* CalcNewellDemagKernel, mag/newellDemagKernel.go:1108-1324, on the GPU in
 * FP64 into host buffers K[c] of the padded size. */
int um_demag_kernel_newell(int Nx, int Ny, int Nz, const int pbc[3], const double cs[3],
                           int asymptotic_radius, int zero_self_demag, double *const K[6]);

/* ---- autotuner (no counterpart in the reference, whose VkFFT plans are fixed) ----
 * Launch schedules of the demag passes picked per shape by timing: rows per
 * work-group of the X passes, the four-step X inverse, lines of the generic
 * Y pass, columns of the convolution pass, and the pass order of the generic
 * Stockham engine (radix schedule).  mode 0: static tables (default); 1: time
 * every new shape when its plan is created (results kept per precision and
 * size until the next call); 2: apply known results only.  file (or NULL):
 * the known results are read from it now and new ones appended, so later
 * processes reuse them. */
int um_set_autotune(int mode, const char *file);
/* the schedule of a plan: nl_x nl_xi four_step nl_y B_conv sched_x sched_y sched_z */
int um_demag_schedule(void *plan, int out[8]);

/* ---- launch trace (diagnostics; no counterpart in the reference) ----
 * um_trace_kernels(1) clears the trace and records the demangled name of
 * every kernel the library launches from then on (e.g.
 * "void um::k_xfwd<float, 512, 8>(um::XArgs<float>)"), one per launch, until
 * um_trace_kernels(0).  um_traced_kernels copies the names, one per line,
 * NUL-terminated and truncated to out_len; *needed (nullable) receives the
 * full length including the NUL. */
int um_trace_kernels(int on);
int um_traced_kernels(char *out, size_t out_len, size_t *needed);

/* ---- demag tensor cache: mag/demagkernel.go:17-109, newellDemagKernel.go:1006-1106 ----
 * The -cache flag (engine/gofiles.go:17): with a directory set, um_demag_create,
 * um_demag_create_newell and the simulation handle read the tensor from
 * <dir>/uMagNUSkernel_<N>_<pbc>_<cellsize>_<accuracy | zero_self_demag>_<ij>.ovf
 * (the reference's names and OVF2 "binary 4" files, so caches are shared) and
 * write it there after computing it when a file is missing or unreadable.
 * With a cache the float32 tensor of the files is used (also when freshly
 * computed); without one (NULL or "", the default) the FP64 tensor is. */
int um_set_kernel_cache(const char *dir);
/* the file prefix of a plan's cache entry (fmt.Sprint spelling) into out */
int um_kernel_cache_basename(const char *dir, const int N[3], const int pbc[3], const double cs[3], int newell,
                             double accuracy, int zero_self_demag, char *out, int out_len);
/* LoadKernel (mag/demagkernel.go:97): a scalar OVF1/OVF2 field of size P into K */
int um_load_kernel(const char *fname, const int P[3], double *K);
/* SaveKernel (mag/demagkernel.go:102): K of size P as an OVF2 binary 4 scalar field */
int um_save_kernel(const char *fname, const int P[3], const double cs[3], const char *name, const double *K);

/* ---- simulation handle: engine/ effective field, torque, steppers --------- */
/* Parameter ids (the .mx3 quantity names) for um_sim_set_scalar / _vector */
enum {
  UM_P_MSAT = 0, UM_P_KU1, UM_P_KU2, UM_P_KC1, UM_P_KC2, UM_P_KC3, UM_P_ALPHA,
  UM_P_TEMP, UM_P_ANISU, UM_P_ANISC1, UM_P_ANISC2, UM_P_BEXT,
  /* spin-transfer torque, engine/torque.go:11-20: xi, Pol (default 1), Lambda
   * (default 1), EpsilonPrime, FreeLayerThickness (0: mesh thickness), J
   * (vector, A/m^2), FixedLayer (vector) */
  UM_P_XI, UM_P_POL, UM_P_LAMBDA, UM_P_EPSPRIME, UM_P_FREELAYER, UM_P_J, UM_P_FIXEDLAYER,
  UM_P_COUNT
};
/* SetSolver ids, engine/run.go:78-90 */
enum { UM_SOLVER_HEUN = 2, UM_SOLVER_RK23 = 3, UM_SOLVER_RK4 = 4, UM_SOLVER_RK45DP = 5 };

/* A simulation owns m, the parameters, the demag plan and the stepper state on
 * one HIP stream (engine/magnetization.go, parameter.go, run.go). */
int um_sim_create(int prec, int Nx, int Ny, int Nz, double cx, double cy, double cz,
                  int pbcx, int pbcy, int pbcz, void **sim);
int um_sim_destroy(void *sim);
/* Msat, Ku1, ... (engine/parameter.go): host array (per cell, NULL = uniform) x mul */
int um_sim_set_scalar(void *sim, int id, const void *host_arr, double mul);
int um_sim_set_vector(void *sim, int id, const void *ax, const void *ay, const void *az,
                      double mx, double my, double mz);
/* regions (engine/regions.go), host uint8 per cell or NULL (all region 0) */
int um_sim_set_regions(void *sim, const uint8_t *host_regions);
/* lex2 / din2 symmetric LUTs (engine/exchange.go:106-177), 32896 values each;
 * dLUT NULL = no interfacial DMI (AddExchange instead of AddDMI) */
int um_sim_set_exchange_luts(void *sim, const void *aLUT, const void *dLUT);
/* B_ext per region (engine/excitation.go:41-53), 3 x 256 values, NULL clears */
int um_sim_set_bext_lut(void *sim, const void *lx, const void *ly, const void *lz);
/* FrozenSpins per region (engine/torque.go:348-352), 256 values, NULL clears */
int um_sim_set_frozen_lut(void *sim, const void *lut);
/* NoDemagSpins per region (engine/demag.go:22-23, setMaskedDemagField
 * :67-98): flagged cells neither generate nor receive the demag field;
 * 256 values, NULL clears */
int um_sim_set_nodemag_lut(void *sim, const void *lut);
/* Time-dependent parameters (a script's `Ku1 = K * sin(2*pi*f*t)`, re-evaluated
 * per stage by regionwise.update, engine/parameter.go:70-86): fn(t, values,
 * ctx) writes the parameter's ncomp (1 or 3) uniform multipliers at stage time
 * t before every effective-field evaluation; a per-cell array set with
 * um_sim_set_scalar/_vector is scaled by them.  fn = NULL removes it.  Steps
 * with a time function are not graph-replayed. */
typedef void (*um_time_fn)(double t, double *values, void *ctx);
int um_sim_set_time_function(void *sim, int param, int ncomp, um_time_fn fn, void *ctx);
/* Extra per-cell field and torque terms, each a host array per component
 * (n values of the handle's precision) times a multiplier:
 *  kind 0: a B_ext mask term, Excitation.Add(mask, f(t)) (engine/excitation.go:
 *          41-53, one Madd2 per term), counted in the Zeeman energy;
 *  kind 1: a custom field added to B_eff last (AddFieldTerm,
 *          engine/customfield.go:68), not counted in any energy term;
 *  kind 2: a custom torque added before FreezeSpins (AddTorqueTerm,
 *          engine/torque.go:108-121).
 * fn (nullable) gives the multiplier fn(t)[0] at every stage time instead of
 * mul; steps with such a term are not graph-replayed.  Terms enter the fused
 * kernel as one summed field / torque array.  The adjoint tape rejects them.
 * um_sim_clear_field_terms removes the terms of one kind (-1: all). */
int um_sim_add_field_term(void *sim, int kind, const void *ax, const void *ay, const void *az, double mul,
                          um_time_fn fn, void *ctx);
int um_sim_clear_field_terms(void *sim, int kind);
/* geometry (engine/geom.go): per-cell volume fraction, NULL = full */
int um_sim_set_geometry(void *sim, const void *vol);
/* EnableDemag, OpenBC, DoPrecess, and the MaxVecDiff sqrt(3) widening in Heun */
int um_sim_set_flags(void *sim, int enable_demag, int openbc, int precess, int heun_sqrt3);
int um_sim_set_demag_accuracy(void *sim, double accuracy);
/* Dbulk (engine/exchange.go:49-66): the bulk DMI LUT (same layout as the
 * exchange LUTs; NULL: none).  Refused at evaluation time together with a
 * non-zero interfacial Dind, as the reference does. */
int um_sim_set_dmi_bulk(void *sim, const void *lut);
/* DisableZhangLiTorque, DisableSlonczewskiTorque and FixedLayerPosition
 * (+1 FIXEDLAYER_TOP, -1 FIXEDLAYER_BOTTOM), engine/torque.go:26-29, 368-371 */
int um_sim_set_stt_flags(void *sim, int disable_zhangli, int disable_slonczewski, double fixed_layer_sign);
/* ---- discrete adjoint of the fixed-step RK4 path ------------------------------
 * No counterpart in the reference (it has no differentiation); these serve
 * callers that need gradients of a loss of the final magnetization: inverse
 * design and parameter fitting.  Valid for SetSolver(4) with FixDt > 0,
 * precession on, no thermal field, no cubic anisotropy, no frozen spins, no
 * geometry, and Msat > 0 in every cell.
 *
 * um_sim_tape_begin: subsequent um_sim_step/steps/run calls record their
 *   stages (up to max_steps steps).  checkpoint_every = 0 keeps the whole
 *   trajectory in device memory (28 values per cell per step); K > 0 keeps m
 *   every K steps and K steps of stage data, recomputing the rest during the
 *   backward sweep (K ~ sqrt(max_steps) minimises memory).
 * um_sim_backward: given dm = dL/dm at the last recorded step (3 device arrays
 *   of the handle's precision), writes dL/dm at the first recorded step into
 *   dm0 (device, may be NULL), dL/dMsat per cell into dms (device, N values,
 *   may be NULL) and dL/dB_ext of a uniform field into dbext (host, 3 doubles,
 *   may be NULL).  May be called repeatedly with other cotangents.
 * um_sim_set_m is refused while the tape records; any parameter change
 * invalidates the tape (UM_ESTATE). */
int um_sim_tape_begin(void *sim, int max_steps, int checkpoint_every);
int um_sim_tape_end(void *sim);
int um_sim_tape_steps(void *sim, long long *steps);
int um_sim_backward(void *sim, const void *const dm[3], void *const dm0[3], void *dms, double dbext[3]);
/* EnableNewellDemag (engine/demag.go:26-56): enable=1 selects the Newell tensor
 * (with asymptotic_radius and zero_self_demag as in um_demag_create_newell) in
 * place of the brute-force one; demag itself stays switched by enable_demag. */
int um_sim_set_demag_newell(void *sim, int enable, int asymptotic_radius, int zero_self_demag);
/* ThermSeed (engine/temperature.go:138); stream separates ensemble members */
int um_sim_set_thermal_seed(void *sim, uint64_t seed, uint32_t stream);
/* m.SetArray (engine/magnetization.go:45), normalises; and m readback */
int um_sim_set_m(void *sim, const void *mx, const void *my, const void *mz);
int um_sim_get_m(void *sim, void *mx, void *my, void *mz);
/* device pointers of m (resident state, for callers that keep data on the GPU) */
int um_sim_m_device(void *sim, void *ptrs[3]);
/* B_eff, torque and B_demag quantities (effectivefield.go:14, torque.go:124, demag.go:42) */
int um_sim_beff(void *sim, void *bx, void *by, void *bz);
int um_sim_torque(void *sim, void *tx, void *ty, void *tz);
int um_sim_demag_field(void *sim, void *bx, void *by, void *bz);
/* SetSolver, engine/run.go:92 */
int um_sim_set_solver(void *sim, int solver);
/* dt, FixDt, MaxErr, RelErr, MinSlope, Headroom, MinDt, MaxDt, GammaLL */
int um_sim_set_solver_params(void *sim, const double knobs[9]);
/* t, dt, step, NUndone, NEval, LastErr, PeakErr, LastTorque */
int um_sim_get_state(void *sim, double out[8]);
/* one stepper.Step(); Steps(n) (run.go:197); Run(seconds) (run.go:190) */
int um_sim_step(void *sim);
int um_sim_steps(void *sim, int n);
int um_sim_run(void *sim, double seconds);
/* m.Average() (engine/average.go:28) and energy terms (energy.go):
 * 0 demag, 1 exchange(+DMI), 2 Zeeman, 3 uniaxial K1, 4 uniaxial K2,
 * 5..7 cubic Kc1..Kc3 (anisotropy.go:163-207), 8 thermal (temperature.go:129),
 * 9 total (GetTotalEnergy, energy.go:27-34) */
int um_sim_average_m(void *sim, double out[3]);
int um_sim_energy(void *sim, int term, double *out);
/* Relax (engine/relax.go:24-100): RK23 without precession and without the
 * thermal field, stepping while the total energy decreases, then while the
 * torque decreases (torque_threshold <= 0) or until max |torque| is below it,
 * tightening MaxErr; solver, MaxErr, FixDt and precession are restored. */
int um_sim_relax(void *sim, double torque_threshold);
/* Minimize(), engine/minimizer.go:56-160: steepest descent with Barzilai-Borwein
 * steps on the precession-free torque until the last `samples` max|dm| are all
 * below stop_max_dm (MinimizerStop, default 1e-6; MinimizerSamples, default 10).
 * Time does not advance; LastErr reports the sampled max|dm|. */
int um_sim_minimize(void *sim, double stop_max_dm, int samples);
/* the HIP stream the simulation enqueues on */
int um_sim_stream(void *sim, void **hip_stream);
/* timing with HIP events on the simulation stream, read at the per-step
 * synchronisation (profiling on resets the counters): out_ms = total ms in
 * {demag FFT path, fused field+torque kernel, whole steps}, n_evals =
 * evaluations timed */
int um_sim_set_profiling(void *sim, int on);
int um_sim_kernel_times(void *sim, double out_ms[3], int *n_evals);

/* ---- HIP graphs and ensembles (no counterpart in the reference, whose steps
 * are enqueued call by call) ------------------------------------------------ */
/* With FixDt set, um_sim_steps runs one step eagerly, then replays a captured
 * two-step HIP graph (no host synchronisation per step: the step statistics
 * and the thermal Philox step counter stay on the device) and reads the
 * statistics back once.  Results are identical to eager stepping.  The graph
 * is re-captured after any parameter or solver change.  Adaptive steps
 * (FixDt = 0) always run eagerly. */
int um_sim_set_graph(void *sim, int on);
/* Ensemble stepping, the batched form of the reference's job queue
 * (cmd/uMagNUS/queue.go:86-151): n steps of every simulation in sims[], each on
 * its own stream, with the graph replays of the members interleaved so small
 * grids run concurrently.  Members in graph mode with FixDt replay graphs;
 * others step eagerly in turn. */
int um_sims_steps(int nsims, void *const *sims, int n);
/* Ensemble stepping in batched launches, the one-launch form of the same job
 * queue (cmd/uMagNUS/queue.go:86-151): n steps of every simulation in sims[]
 * (each handle once).  After one eager step per member, the launches of two
 * fixed-dt steps of every member in graph mode are recorded and merged: each
 * kernel variant runs once per phase for all members that reach it, every
 * work-group working on its own member's grid and buffers, so each member
 * computes what it computes alone (um_sims_steps, um_sim_steps), bit for bit.
 * The merged sequence is replayed as one graph per two steps; it is rebuilt
 * when a member's parameters, solver, FixDt or buffers change.  Members not in
 * graph mode with FixDt step eagerly in turn.  *launches_per_two_steps (when
 * not NULL) receives the number of launches that replay two steps of the
 * whole ensemble, or 0 when a member needs a kernel without a batched form
 * and the members were replayed one graph each. */
int um_sims_steps_batched(int nsims, void *const *sims, int n, int *launches_per_two_steps);
/* Generic capture of the library stream (um_set_stream; a stream other than
 * the NULL stream, which HIP cannot capture): every um_* call between begin
 * and end is recorded instead of run; launch replays it on the library
 * stream.  Calls that read results back to the host (um_reduce_*, copies to
 * host) fail while capturing. */
int um_graph_begin(void);
int um_graph_end(void **graph);
int um_graph_launch(void *graph);
int um_graph_destroy(void *graph);

#ifdef __cplusplus
}
#endif
#endif
