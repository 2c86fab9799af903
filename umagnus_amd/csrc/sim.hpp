// Simulation handle: engine-level state (m, parameters, demag plan, solver
// state) and the steppers, restating engine/rk45dp.go, heun.go, rk4.go,
// stepcontroller.go and run.go on top of the fused kernels.
#pragma once
#include "common.hpp"

namespace um {

struct DemagBase;
DemagBase *demag_create(int prec, const int N[3], const int P[3], double *const Kdev[6], hipStream_t s);
void demag_kernel_bf_device(const int N[3], const int pbc[3], const double cs[3], double accuracy,
                            double *const Kdev[6], int P[3], hipStream_t s);
// Newell tensor (mag/newellDemagKernel.go:1108-1324) into Kdev[6], FP64
void demag_kernel_newell_device(const int N[3], const int pbc[3], const double cs[3], int asymptotic_radius,
                                int zero_self_demag, double *const Kdev[6], int P[3], hipStream_t s);
void pad_size(const int N[3], const int pbc[3], int P[3]);
// The tensor of a plan (mag.DemagKernel / NewellDemagKernel): from the OVF
// cache when a cache directory is set, else computed on the GPU (kcache.hip)
void demag_tensor(const int N[3], const int pbc[3], const double cs[3], int newell, double accuracy,
                  int asymptotic_radius, int zero_self_demag, double *const Kdev[6], int P[3], hipStream_t s);
void set_kernel_cache(const char *dir);
std::string kernel_cache_basename(const std::string &dir, const int N[3], const int pbc[3], const double cs[3],
                                  int newell, double accuracy, int zero_self_demag);
void ovf_read(const std::string &fname, int size[3], int *ncomp, std::vector<double> &vals);
void ovf_write_scalar(const std::string &fname, const int size[3], const double cs[3], const std::string &name,
                      const double *K);
// kernelRanges (mag/demagkernel.go:339-352): tensor offsets r1..r2 per axis
void kernel_ranges(const int P[3], const int pbc[3], int r1[3], int r2[3]);
// first offset >= r1 that wraps onto w (0 <= w < P, r1 <= 0); the periodic
// images of w are image_first, +P, +2P, ... up to r2
__host__ __device__ __forceinline__ int image_first(int w, int P, int r1) { return w - P * ((w - r1) / P); }
void demag_exec(DemagBase *d, void *const B[3], const void *const m[3], const void *vol, const void *Ms,
                double Ms_mul, hipStream_t s);
// Demag + effective field + torque in one pipeline; fa: FieldArgs of the
// plan's precision (bd unused).  false when the plan has no fused path.
bool demag_exec_field(DemagBase *d, const void *const m[3], const void *vol, const void *Ms, double Ms_mul,
                      const void *fa, hipStream_t s);
void demag_destroy(DemagBase *d);
void demag_kspec(DemagBase *d, int c, void *host);
void demag_info(DemagBase *d, int P[3], int *prec);
// launch schedule of a plan: nl_x nl_xi xi4 nl_y B_conv sched_x sched_y sched_z
constexpr int DEMAG_NKNOB = 8;
void demag_schedule(DemagBase *d, int out[DEMAG_NKNOB]);
// autotuning of new plans: mode 0 off (static tables), 1 time unknown shapes,
// 2 apply known results only; file: per-shape results read at this call and
// appended as shapes are tuned (NULL: none)
void set_autotune(int mode, const char *file);

struct SimBase {
  virtual ~SimBase() {}
  virtual void set_scalar(int id, const void *arr, double mul) = 0;
  virtual void set_vector(int id, const void *const a[3], const double mul[3]) = 0;
  virtual void set_regions(const uint8_t *r) = 0;
  virtual void set_exchange_luts(const void *a, const void *d) = 0;
  virtual void set_bext_lut(const void *const l[3]) = 0;
  virtual void set_frozen_lut(const void *l) = 0;
  virtual void set_nodemag_lut(const void *l) = 0;
  virtual void set_time_function(int id, int nc, um_time_fn fn, void *ctx) = 0;
  virtual void add_field_term(int kind, const void *const a[3], double mul, um_time_fn fn, void *ctx) = 0;
  virtual void clear_field_terms(int kind) = 0;
  virtual void set_geometry(const void *vol) = 0;
  virtual void set_flags(int demag, int openbc, int precess, int heun_sqrt3) = 0;
  virtual void set_demag_accuracy(double a) = 0;
  virtual void set_demag_newell(int enable, int asymptotic_radius, int zero_self_demag) = 0;
  virtual void set_dmi_bulk(const void *lut) = 0;
  virtual void minimize(double stop_max_dm, int samples) = 0;
  virtual void set_stt_flags(int disable_zl, int disable_sl, double flp) = 0;
  // discrete RK4 adjoint: record steps, then a reverse sweep from dL/dm_final
  virtual void tape_begin(int max_steps, int checkpoint_every) = 0;
  virtual void tape_end() = 0;
  virtual long tape_steps() const = 0;
  virtual void backward(const void *const gm[3], void *const gm0[3], void *gms, double *gbext) = 0;
  virtual void set_thermal_seed(uint64_t seed, uint32_t stream) = 0;
  virtual void set_m(const void *const m[3]) = 0;
  virtual void get_m(void *const m[3]) = 0;
  virtual void m_device(void *p[3]) = 0;
  virtual void beff(void *const out[3]) = 0;
  virtual void torque(void *const out[3]) = 0;
  virtual void demag_field(void *const out[3]) = 0;
  virtual void set_solver(int s) = 0;
  virtual void set_solver_params(const double k[9]) = 0;
  virtual void get_state(double o[8]) = 0;
  virtual void step() = 0;
  virtual void steps(int n) = 0;
  virtual void run(double seconds) = 0;
  virtual void average_m(double out[3]) = 0;
  virtual double energy(int term) = 0;
  virtual void relax(double torque_threshold) = 0;
  virtual hipStream_t stream() = 0;
  virtual void set_profiling(int on) = 0;
  virtual void set_graph(int on) = 0;
  virtual long graph_prepare(long stop) = 0;
  // batched ensembles: as graph_prepare without capturing a graph; then the
  // launches of two fixed-dt steps recorded instead of run (batch.hpp),
  // a key that changes whenever such a recording would, and the host side
  // of two replayed steps
  virtual long batch_prepare(long stop) = 0;
  virtual void graph_ensure() = 0;   // capture the two-step graph if it is stale
  virtual void batch_record(std::vector<struct BatchEntry> &out) = 0;
  virtual std::vector<uint64_t> batch_key() const = 0;
  virtual void batch_advance() = 0;
  virtual void graph_replay() = 0;
  virtual void graph_finish(long stop) = 0;
  virtual long step_count() const = 0;
  virtual void kernel_times(double ms[3], int *nevals) = 0;
};

SimBase *sim_create(int prec, const int N[3], const double cs[3], const int pbc[3]);

}  // namespace um
