// Vector arithmetic, RK stage/error kernels and reductions.
#pragma once
#include "common.hpp"

namespace um {

constexpr int RED_BLOCKS = 1024;   // fixed partial count: deterministic sums
constexpr int RED_THREADS = 256;

template <typename T> struct Vec3P { T *c[3]; };
template <typename T> struct CVec3P { const T *c[3]; };

// dst = sum_k fac[k] * src[k] (nsrc <= 8), then optionally normalise the
// three components with vol (normalize2.cl) -- one pass per cell.
template <typename T>
struct StageArgs {
  int nsrc;
  int ncomp;             // 3 (vectors) or 1 (one component array)
  int group3;            // madd3 grouping (madd3.cl:12)
  int normalize;
  const T *src[8][3];
  T fac[8];
  T *dst[3];
  const T *vol;
  T *nrm;                // when normalising: |sum| per cell (before the scaling) or nullptr
  long N;
  // device-controlled steps: fac[j] (j >= 1) is multiplied by the step's
  // h = T(dev_dt[0] * gam) formed on the device, and nothing is written
  // while *skip != 0 (steps after the end of a run)
  const double *dev_dt;
  double gam;
  const int *skip;
};
template <typename T> void launch_stage(const StageArgs<T> &a, hipStream_t s);

// Device-side control of adaptive RK45DP steps: the state the host keeps
// between steps (engine/run.go, stepcontroller.go), on the device.
struct AdaptState {
  double dt, time, last_err, peak_err, last_torque, stop;
  long long nsteps, nundone, nevals;
  int done, accepted, error, pad;
};
struct AdaptPar {
  double maxerr, relerr, minslope, headroom, mindt, maxdt, gam;
  int evals_per_step, h_f32;
};
// integralController + adaptDt of one RK45DP step from the error kernel's
// slots (cleared afterwards), then k1 <- k7 (accepted) or m <- m0 (rejected
// or after the end of the run); the vectors are 3n long
template <typename T>
void launch_adapt(unsigned long long *res, AdaptState *st, const AdaptPar &p, T *m, const T *m0, T *k1,
                  const T *k7, long n3, hipStream_t s);

// Minimizer step (minimize.cl): m0 <- m, m <- ((4 - h^2|k|^2) m + 4 h k) / (4 + h^2|k|^2)
template <typename T> void launch_min_step(T *m, T *m0, const T *k, T h, long N, hipStream_t s);
// Minimizer statistics (engine/minimizer.go:78-95) with dm = m - m0, dk = k0 - k:
// out = {max|dm|^2, dm.dm, dm.dk, dk.dk, max|k|^2} in FP64 (fixed-order partials:
// 5*RED_BLOCKS doubles)
template <typename T>
void launch_min_stats(const T *m, const T *m0, const T *k, const T *k0, long N, double *partial, double *out,
                      hipStream_t s);

// Fused RK error reductions.  Writes into res[0..2] (as FP64 bit patterns,
// max-combined with atomics; res must be zeroed first):
//  RK45DP: Err = sum e_j k_j;  res = {max|Err|^2, max|k7|^2, max |Err|/|k7|}
//  diff:   res[0] = max|a - b|^2   (Heun's MaxVecDiff, RK4's k1 - k4)
//  norm:   res[0] = max|a|^2
template <typename T>
struct ErrArgs {
  int mode;              // 0 rk45dp, 1 diff, 2 norm
  const T *k[6][3];      // rk45dp: k1 k3 k4 k5 k6 k7; diff: a b; norm: a
  T e[6];
  long N;
  unsigned long long *res;
};
template <typename T> void launch_err(const ErrArgs<T> &a, hipStream_t s);

// Step statistics folded on the device (graph replays): gs = {lastErr, peakErr,
// lastTorque, step counter}.  mode 0: RK45DP (err = sqrt(res0) h, torque =
// sqrt(res1)); 1 / 3: err = sqrt(mode * res0) h (Heun's sqrt3 widening = 3);
// 2: torque = sqrt(res0).
void launch_step_stats(unsigned long long *res, double *gs, double h, int mode, int count, hipStream_t s);
void launch_step_count(unsigned long long *counter, hipStream_t s);

// Deterministic FP64 sums: out[0] = sum_i w_i * (a_i . b_i) over ncomp
// components (w nullable = 1, b nullable = 1).  partial: RED_BLOCKS doubles.
template <typename T>
void launch_wdot(const T *const *a, const T *const *b, int ncomp, const T *w, T wmul, long N, double *partial,
                 double *out, hipStream_t s);
// max |x_i| over N (result as FP64 bits, res zeroed by the caller)
template <typename T> void launch_maxabs(const T *x, long N, unsigned long long *res, hipStream_t s);

template <typename T> void launch_vecnorm(T *dst, const T *const a[3], long N, hipStream_t s);
template <typename T> void launch_div(T *dst, const T *a, const T *b, long N, hipStream_t s);
template <typename T> void launch_mul(T *dst, const T *a, const T *b, long N, hipStream_t s);

}  // namespace um
