// Discrete adjoint of the fixed-step RK4 integrator (engine/rk4.go:11-69)
// through the effective field and LL torque: reverse-mode gradients of a loss
// of the final magnetization with respect to the initial m, the per-cell
// saturation magnetisation and a uniform external field.
//
// Linearisation used by the kernels (per cell, alpha and c = 1/(1+alpha^2)):
//   torque    t = -c (m x B + alpha m x (m x B))
//   cotangent on B:  gB = -c (l x m + alpha (l x m) x m)           (l = dL/dt)
//   direct part on m: gm = -c (B x l + alpha ((m x B) x l + B x (l x m)))
//   field     B = K * (mu0 Ms m) + (1/Ms) L m + B_anis(m) + B_ext
// The demag convolution is self-adjoint (K_ij = K_ji, K(-r) = K(r)) and the
// exchange + interfacial-DMI operator L is symmetric (the DMI stencil is odd
// both in the neighbour offset and in the component coupling; its boundary
// extrapolation adds diag(D^2/2A, 0, D^2/2A)), so both reuse the forward
// operators: dL/dm += mu0 Ms (K * gB) + L (gB / Ms).
#pragma once
#include "field.hpp"

namespace um {

template <typename T>
struct AdjArgs {
  FieldArgs<T> f;        // parameters and mesh of the forward evaluation (f.m unused)
  const T *u[3];         // stage input m (normalised)
  const T *lk[3];        // cotangent on the stage torque
  T *w[3];               // gB / Ms of this stage (written by the previous kernel)
  const T *bd[3];        // demag field of the stage (forward tape) or nullptr
  const T *D[3];         // K * gB or nullptr
  const T *nrm;          // |v| of the stage before normalisation (nullptr: stage 1)
  const T *lv5[3];       // cotangent of the step's final pre-normalisation vector
  T *lacc[3];            // cotangent of the step's m0 (accumulated)
  T *lknext[3];          // cotangent on the previous stage's torque (nullptr: stage 1)
  T c5, cs;              // lknext = c5 lv5 + cs lv
  const T *unext[3];     // input m of the previous stage: with lknext it gives
  T *wnext[3];           //   that stage's gB / Ms, written here (nullptr: stage 1)
  T mu0;
  double *gms;           // per-cell dL/dMs accumulator or nullptr
  double *gbext;         // per-block partial sums of gB (3 per block) or nullptr
  long N;
};

// final step normalisation m1 = v5/|v5|: lv5 = (lm - m1 (m1.lm))/|v5|,
// lacc = lv5, lk = c4 lv5 (the cotangent on k4), and stage 4's gB / Ms into w
template <typename T>
struct AdjFinalArgs {
  const T *m1[3], *lm[3], *u4[3];
  const T *nrm;
  T *lv5[3], *lacc[3], *lk[3], *w[3];
  T c4;
  Par<T> Ms, alpha;
  double *gbext;         // per-block partial sums of gB or nullptr
  long N;
};

template <typename T> void launch_adj_stage(const AdjArgs<T> &a, hipStream_t s);
template <typename T> void launch_adj_final(const AdjFinalArgs<T> &a, hipStream_t s);
// blocks used by the gB partial sums (3 doubles each)
long adj_blocks(long N);
// dst[i] = (T) src[i]
template <typename T> void launch_adj_cast(T *dst, const double *src, long N, hipStream_t s);

}  // namespace um
