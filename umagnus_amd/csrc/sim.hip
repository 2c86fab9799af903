// Simulation handle: the engine-level restatement (engine/effectivefield.go,
// torque.go, rk45dp.go, heun.go, rk4.go, stepcontroller.go, run.go) driving
// the fused kernels.  Stage combinations are fused with normalisation and the
// step-control norms with the error estimate, so a step reads back one
// 24-byte result instead of several reductions.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "adjoint.hpp"
#include "field.hpp"
#include "batch.hpp"
#include "sim.hpp"
#include "vec.hpp"

namespace um {

static constexpr double KB = 1.380650424e-23;  // mag/constants.go:9
static constexpr double MU0_S = 4 * M_PI * 1e-7;   // mag/constants.go:8
static constexpr int NREGION = 256;
static constexpr int SYMLUT = NREGION * (NREGION + 1) / 2;

// demag input mask of NoDemagSpins (engine/demag.go:67-83): the geometry (or 1)
// with the flagged regions zeroed
template <typename T>
__global__ void k_nodemag_vol(T *out, const T *vol, const T *lut, const uint8_t *regions, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const T v = vol ? vol[i] : T(1);
    out[i] = lut[regions ? regions[i] : 0] != T(0) ? T(0) : v;
  }
}

static uint64_t next_serial() {
  static std::atomic<uint64_t> c{1};
  return c++;
}

template <typename T>
struct Sim : SimBase {
  int N[3];
  long n;
  double cs[3];
  int pbc[3], pbccode;
  hipStream_t st = nullptr;

  struct PSlot {
    DevBuf arr;
    double mul = 0;
    bool nonzero = false;   // host-side knowledge: any non-zero value
    bool arr_nz = false;    // the per-cell array holds a non-zero value
  };
  PSlot p[UM_P_COUNT][3];
  // time-dependent parameters: a host function of the stage time gives the
  // uniform multiplier(s), re-evaluated for every evaluation as
  // regionwise.update does per stage (engine/parameter.go:70-86)
  struct TimeFn {
    um_time_fn fn = nullptr;
    void *ctx = nullptr;
    int nc = 1;
  };
  TimeFn tfn[UM_P_COUNT];
  bool any_tfn = false;
  // extra field / torque terms (B_ext mask terms, custom fields and torques),
  // summed into xfield / xtorque before every evaluation
  struct ExtraTerm {
    DevBuf arr;   // 3n values
    double mul = 1;
    um_time_fn fn = nullptr;
    void *ctx = nullptr;
    int kind = 0;
  };
  std::vector<ExtraTerm> xterms;
  DevBuf xfield, xtorque;
  bool xterm_tfn = false;
  DevBuf regions, aLUT, dLUT, dbLUT, bextLUT, frozenLUT, vol;
  DevBuf nodemagLUT, ndvol;   // NoDemagSpins per region, and the masked demag input
  bool has_nodemag = false;
  long nd_gen = -1;
  bool has_regions = false, has_dmi = false, has_dmibulk = false, has_bextlut = false, has_frozen = false,
       has_vol = false;
  int disable_zl = 0, disable_sl = 0;   // DisableZhangLiTorque / DisableSlonczewskiTorque (engine/torque.go:26-27)
  double flp = 1.0;                     // FIXEDLAYER_TOP (engine/torque.go:368-371)
  int enable_demag = 1, openbc = 0, precess = 1, heun_sqrt3 = 1;
  int newell = 0, newell_arad = 32, newell_zsd = 0;   // asymptotic_radius default, engine/demag.go:27
  double demag_acc = 6.0;
  DemagBase *demag = nullptr;
  uint64_t seed = 0;
  uint32_t tstream = 0;

  DevBuf bufs[10];          // m, m0, k1..k6, bd, spare
  T *m, *m0, *k[7], *bd, *spare;  // k[0]=k1 .. k[5]=k6
  DevBuf red;               // 4 x u64 reduction slots
  DevBuf partial, dsum;     // deterministic sums
  unsigned long long *hres = nullptr;  // pinned readback
  double *hsum = nullptr;

  // solver state, engine/run.go:13-43
  int solver = UM_SOLVER_RK45DP;
  int relaxing = 0;              // engine/relax.go:22: no thermal field while relaxing
  double Time = 0, dt = 1e-15, fixdt = 0, maxerr = 1e-5, relerr = 1e-4, minslope = 1e8, headroom = 0.8;
  double mindt = 0, maxdt = 0, gammaLL = 1.7595e11, alarm = 0;
  double lastErr = 0, peakErr = 0, lastTorque = 0;
  long nsteps = 0, nundone = 0, nevals = 0;
  bool k1_valid = false;

  // graph replay of fixed-dt steps: two steps are captured once (the pointer
  // swaps of a step undo themselves after two) and replayed; the step
  // statistics and the Philox step counter live on the device meanwhile
  int graph_mode = 0;
  bool capturing = false;
  bool counted = false;     // the captured step's counter increment is queued
  long gen = 0;                  // bumped by every setter: a captured graph is stale
  const uint64_t serial = next_serial();   // distinguishes handles that reuse an address
  DevBuf gstat;                  // {lastErr, peakErr, lastTorque, step counter}
  bool gstat_live = false;       // gstat holds this session's statistics (uploaded by graph_prepare)
  double *hstat = nullptr;       // pinned mirror
  // device-controlled adaptive RK45DP (graph mode, run()): the step state on
  // the device, its pinned mirror, and the captured ADAPT_K-step graph
  static constexpr int ADAPT_K = 8;
  bool devctl = false;           // steps are being captured in device-controlled form
  DevBuf astate;
  AdaptState *hastate = nullptr;
  hipGraphExec_t agexec = nullptr;
  long a_gen = -1;
  const void *a_ptrs[8] = {};
  hipGraphExec_t gexec = nullptr;
  long g_gen = -1;
  int g_solver = -1;
  double g_dt = 0;
  const void *g_ptrs[8] = {};
  long g_evals = 0;

  // profiling: HIP events around the demag and field kernels of every
  // evaluation, recorded without synchronising and read after the step's
  // own readback has drained the stream
  int profiling = 0;
  std::vector<hipEvent_t> evpool;
  int evused = 0;
  hipEvent_t step_ev[2] = {};
  double t_demag = 0, t_field = 0, t_other = 0;
  int prof_evals = 0;

  Sim(const int Nin[3], const double csin[3], const int pbcin[3]) {
    for (int i = 0; i < 3; i++) { N[i] = Nin[i]; cs[i] = csin[i]; pbc[i] = pbcin[i]; }
    UM_REQUIRE(N[0] > 0 && N[1] > 0 && N[2] > 0, UM_EINVAL, "grid size must be positive");
    for (int i = 0; i < 3; i++) {
      UM_REQUIRE(std::isfinite(cs[i]) && cs[i] > 0, UM_EINVAL, "cell size must be positive and finite");
      UM_REQUIRE(pbc[i] >= 0 && pbc[i] <= (1 << 16), UM_EINVAL, "PBC repetitions out of range");
    }
    ensure_init();
    n = (long)N[0] * N[1] * N[2];
    UM_REQUIRE(n < (1L << 31), UM_EUNSUP, "grid too large for 32-bit cell indices");
    pbccode = (pbc[0] ? 1 : 0) | (pbc[1] ? 2 : 0) | (pbc[2] ? 4 : 0);
    UM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const size_t vb = sizeof(T) * 3 * n;
    for (auto &b : bufs) { b.alloc(vb); fill_bytes(b.p, vb, st); }
    m = bufs[0].as<T>(); m0 = bufs[1].as<T>();
    for (int i = 0; i < 6; i++) k[i] = bufs[2 + i].as<T>();
    k[6] = nullptr;
    bd = bufs[8].as<T>(); spare = bufs[9].as<T>();
    red.alloc(4 * sizeof(unsigned long long));
    partial.alloc(RED_BLOCKS * sizeof(double));
    dsum.alloc(sizeof(double));
    UM_HIP(hipHostMalloc((void **)&hres, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    UM_HIP(hipHostMalloc((void **)&hsum, sizeof(double), hipHostMallocDefault));
    aLUT.alloc(sizeof(T) * SYMLUT);
    dLUT.alloc(sizeof(T) * SYMLUT);
    dbLUT.alloc(sizeof(T) * SYMLUT);
    fill_bytes(aLUT.p, aLUT.bytes, st);
    fill_bytes(dLUT.p, dLUT.bytes, st);
    fill_bytes(dbLUT.p, dbLUT.bytes, st);
    // Pol = 1 and Lambda = 1 by default (engine/torque.go:73-74)
    for (int id : {UM_P_POL, UM_P_LAMBDA}) { p[id][0].mul = 1; p[id][0].nonzero = true; }
    for (auto &e : step_ev) UM_HIP(hipEventCreate(&e));
    gstat.alloc(4 * sizeof(double));
    UM_HIP(hipHostMalloc((void **)&hstat, 4 * sizeof(double), hipHostMallocDefault));
    UM_HIP(hipStreamSynchronize(st));
  }
  ~Sim() override {
    if (st) (void)hipStreamSynchronize(st);
    if (demag) demag_destroy(demag);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (agexec) (void)hipGraphExecDestroy(agexec);
    if (hastate) (void)hipHostFree(hastate);
    if (hstat) (void)hipHostFree(hstat);
    if (hres) (void)hipHostFree(hres);
    if (hsum) (void)hipHostFree(hsum);
    for (auto &e : evpool) (void)hipEventDestroy(e);
    for (auto &e : step_ev) if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
  hipStream_t stream() override { return st; }

  // ---- parameters --------------------------------------------------------------
  void upload(DevBuf &b, const void *host, size_t bytes) {
    if (b.bytes != bytes) b.alloc(bytes);
    UM_HIP(hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, st));
    UM_HIP(hipStreamSynchronize(st));
  }
  static bool any_nonzero(const T *a, long cnt) {
    for (long i = 0; i < cnt; i++) if (a[i] != T(0)) return true;
    return false;
  }
  void set_slot(PSlot &s, const void *arr, double mul) {
    s.mul = mul;
    if (arr) {
      upload(s.arr, arr, sizeof(T) * n);
      s.arr_nz = any_nonzero((const T *)arr, n);
      s.nonzero = (mul != 0) && s.arr_nz;
    } else {
      s.arr.release();
      s.arr_nz = false;
      s.nonzero = (mul != 0);
    }
  }
  void set_time_function(int id, int nc, um_time_fn fn, void *ctx) override {
    UM_REQUIRE(id >= 0 && id < UM_P_COUNT && (nc == 1 || nc == 3), UM_EINVAL, "time function: parameter id / size");
    gen++;
    tfn[id] = TimeFn{fn, ctx, nc};
    any_tfn = false;
    for (auto &f : tfn) any_tfn = any_tfn || f.fn != nullptr;
    if (id == UM_P_MSAT) invalidate_fsal();
  }
  void apply_time_functions() {
    if (!any_tfn) return;
    for (int id = 0; id < UM_P_COUNT; id++) {
      const TimeFn &f = tfn[id];
      if (!f.fn) continue;
      double v[3] = {0, 0, 0};
      f.fn(Time, v, f.ctx);
      for (int c = 0; c < f.nc; c++) {
        PSlot &s = p[id][c];
        s.mul = v[c];
        s.nonzero = v[c] != 0 && (!s.arr.p || s.arr_nz);
      }
    }
  }
  void add_field_term(int kind, const void *const a[3], double mul, um_time_fn fn, void *ctx) override {
    UM_REQUIRE(kind >= 0 && kind <= 2, UM_EINVAL, "field term kind: 0 B_ext mask, 1 custom field, 2 custom torque");
    UM_REQUIRE(a && a[0] && a[1] && a[2], UM_EINVAL, "field term: NULL component");
    UM_REQUIRE(xterms.size() < 64, UM_EUNSUP, "at most 64 extra field / torque terms");
    gen++;
    ExtraTerm t;
    t.arr.alloc(sizeof(T) * 3 * n);
    for (int c = 0; c < 3; c++)
      UM_HIP(hipMemcpyAsync(t.arr.template as<T>() + c * n, a[c], sizeof(T) * n, hipMemcpyHostToDevice, st));
    UM_HIP(hipStreamSynchronize(st));
    t.mul = mul; t.fn = fn; t.ctx = ctx; t.kind = kind;
    xterms.push_back(std::move(t));
    xterm_tfn = false;
    for (auto &x : xterms) xterm_tfn = xterm_tfn || x.fn != nullptr;
  }
  void clear_field_terms(int kind) override {
    UM_REQUIRE(kind >= -1 && kind <= 2, UM_EINVAL, "field term kind: -1 all, 0, 1 or 2");
    gen++;
    UM_HIP(hipStreamSynchronize(st));
    std::vector<ExtraTerm> keep;
    for (auto &x : xterms) if (kind >= 0 && x.kind != kind) keep.push_back(std::move(x));
    xterms = std::move(keep);
    xterm_tfn = false;
    for (auto &x : xterms) xterm_tfn = xterm_tfn || x.fn != nullptr;
  }
  // dst = sum of the terms whose kind bit is set in kinds, each times its
  // multiplier at the current stage time; false (dst untouched) when none
  bool combine_terms(DevBuf &dst, unsigned kinds) {
    std::vector<const ExtraTerm *> sel;
    for (auto &x : xterms) if (kinds & (1u << x.kind)) sel.push_back(&x);
    if (sel.empty()) return false;
    if (!dst.p) dst.alloc(sizeof(T) * 3 * n);
    for (size_t j0 = 0; j0 < sel.size(); j0 += 7) {
      StageArgs<T> a{};
      a.ncomp = 3;
      int ns = 0;
      if (j0 > 0) {   // chain: dst + the next seven terms
        for (int c = 0; c < 3; c++) a.src[ns][c] = dst.as<T>() + c * n;
        a.fac[ns++] = T(1);
      }
      for (size_t j = j0; j < sel.size() && j < j0 + 7; j++) {
        double mul = sel[j]->mul;
        if (sel[j]->fn) {
          double v[3] = {0, 0, 0};
          sel[j]->fn(Time, v, sel[j]->ctx);
          mul = v[0];
        }
        for (int c = 0; c < 3; c++) a.src[ns][c] = sel[j]->arr.template as<T>() + c * n;
        a.fac[ns++] = (T)mul;
      }
      a.nsrc = ns;
      for (int c = 0; c < 3; c++) a.dst[c] = dst.as<T>() + c * n;
      a.N = n;
      launch_stage(a, st);
    }
    return true;
  }
  void set_scalar(int id, const void *arr, double mul) override {
    gen++;
    UM_REQUIRE(id >= 0 && id < UM_P_COUNT, UM_EINVAL, "parameter id");
    set_slot(p[id][0], arr, mul);
    if (id == UM_P_MSAT) invalidate_fsal();
  }
  void set_vector(int id, const void *const a[3], const double mul[3]) override {
    gen++;
    UM_REQUIRE(id >= 0 && id < UM_P_COUNT, UM_EINVAL, "parameter id");
    for (int c = 0; c < 3; c++) set_slot(p[id][c], a[c], mul[c]);
  }
  Par<T> P(int id, int c = 0) const { return Par<T>{(const T *)p[id][c].arr.p, (T)p[id][c].mul}; }
  bool nz(int id, int nc = 1) const {
    for (int c = 0; c < nc; c++) if (p[id][c].nonzero) return true;
    return false;
  }
  void set_regions(const uint8_t *r) override {
    gen++;
    if (r) { upload(regions, r, (size_t)n); has_regions = true; }
    else { regions.release(); has_regions = false; }
  }
  void set_exchange_luts(const void *a, const void *d) override {
    gen++;
    upload(aLUT, a, sizeof(T) * SYMLUT);
    has_dmi = false;
    if (d) {
      upload(dLUT, d, sizeof(T) * SYMLUT);
      has_dmi = any_nonzero((const T *)d, SYMLUT);
    }
  }
  void set_dmi_bulk(const void *l) override {
    gen++;
    has_dmibulk = false;
    if (l) {
      upload(dbLUT, l, sizeof(T) * SYMLUT);
      has_dmibulk = any_nonzero((const T *)l, SYMLUT);
    }
  }
  void set_stt_flags(int zl, int sl, double f) override {
    gen++;
    disable_zl = zl; disable_sl = sl; flp = f;
  }
  void set_bext_lut(const void *const l[3]) override {
    gen++;
    if (!l[0]) { bextLUT.release(); has_bextlut = false; return; }
    std::vector<T> h(3 * NREGION);
    for (int c = 0; c < 3; c++) memcpy(h.data() + c * NREGION, l[c], sizeof(T) * NREGION);
    upload(bextLUT, h.data(), sizeof(T) * 3 * NREGION);
    has_bextlut = any_nonzero(h.data(), 3 * NREGION);
  }
  void set_frozen_lut(const void *l) override {
    gen++;
    if (!l) { frozenLUT.release(); has_frozen = false; return; }
    upload(frozenLUT, l, sizeof(T) * NREGION);
    has_frozen = any_nonzero((const T *)l, NREGION);
  }
  // NoDemagSpins (engine/demag.go:22-23)
  void set_nodemag_lut(const void *l) override {
    gen++;
    if (!l) { nodemagLUT.release(); ndvol.release(); has_nodemag = false; return; }
    upload(nodemagLUT, l, sizeof(T) * NREGION);
    has_nodemag = any_nonzero((const T *)l, NREGION);
  }
  void set_geometry(const void *v) override {
    gen++;
    if (v) { upload(vol, v, sizeof(T) * n); has_vol = true; }
    else { vol.release(); has_vol = false; }
  }
  void set_flags(int d, int ob, int pr, int hs) override {
    gen++;
    enable_demag = d; openbc = ob; precess = pr; heun_sqrt3 = hs;
  }
  void set_demag_accuracy(double a) override {
    gen++;
    if (a != demag_acc && demag) { demag_destroy(demag); demag = nullptr; }
    demag_acc = a;
  }
  // EnableNewellDemag with asymptotic_radius and zero_self_demag (engine/demag.go:26-28)
  void set_demag_newell(int enable, int arad, int zsd) override {
    gen++;
    enable = enable != 0;
    if ((enable != newell || arad != newell_arad || zsd != newell_zsd) && demag) {
      UM_HIP(hipStreamSynchronize(st));
      demag_destroy(demag);
      demag = nullptr;
    }
    newell = enable; newell_arad = arad; newell_zsd = zsd;
  }
  void set_thermal_seed(uint64_t s, uint32_t strm) override { gen++; seed = s; tstream = strm; }
  void invalidate_fsal() { k1_valid = false; }

  // ---- m --------------------------------------------------------------------------
  void set_m(const void *const mh[3]) override {
    UM_REQUIRE(!tape_on, UM_ESTATE, "m cannot be set while the adjoint tape records (end it first)");
    for (int c = 0; c < 3; c++)
      UM_HIP(hipMemcpyAsync(m + c * n, mh[c], sizeof(T) * n, hipMemcpyHostToDevice, st));
    normalize(m);
    UM_HIP(hipStreamSynchronize(st));
    k1_valid = false;
  }
  void get_m(void *const mh[3]) override {
    for (int c = 0; c < 3; c++)
      UM_HIP(hipMemcpyAsync(mh[c], m + c * n, sizeof(T) * n, hipMemcpyDeviceToHost, st));
    UM_HIP(hipStreamSynchronize(st));
  }
  void m_device(void *ptrs[3]) override {
    for (int c = 0; c < 3; c++) ptrs[c] = m + c * n;
  }
  void normalize(T *v) {
    StageArgs<T> a{};
    a.nsrc = 1;
    a.ncomp = 3;
    for (int c = 0; c < 3; c++) { a.src[0][c] = v + c * n; a.dst[c] = v + c * n; }
    a.fac[0] = T(1);
    a.normalize = 1;
    a.vol = (const T *)vol.p;
    a.N = n;
    launch_stage(a, st);
  }

  // ---- effective field and torque ------------------------------------------------------
  void ensure_demag() {
    if (demag || !enable_demag) return;
    int P[3];
    pad_size(N, pbc, P);
    const size_t np = (size_t)P[0] * P[1] * P[2];
    DevBuf K(sizeof(double) * 6 * np);
    double *Kd[6];
    for (int c = 0; c < 6; c++) Kd[c] = K.as<double>() + c * np;
    int P2[3];
    demag_tensor(N, pbc, cs, newell, demag_acc, newell_arad, newell_zsd, Kd, P2, st);
    demag = demag_create(sizeof(T) == 4 ? UM_F32 : UM_F64, N, P2, Kd, st);
    UM_HIP(hipStreamSynchronize(st));
  }
  FieldArgs<T> field_args(const T *mm) {
    apply_time_functions();
    FieldArgs<T> a{};
    a.g = Mesh{N[0], N[1], N[2], pbccode};
    a.wx = (T)(2 / (cs[0] * cs[0]));
    a.wy = (T)(2 / (cs[1] * cs[1]));
    a.wz = (T)(2 / (cs[2] * cs[2]));
    a.cx = (T)cs[0]; a.cy = (T)cs[1]; a.cz = (T)cs[2];
    for (int c = 0; c < 3; c++) a.m[c] = mm + c * n;
    a.regions = has_regions ? (const uint8_t *)regions.p : nullptr;
    a.Ms = P(UM_P_MSAT);
    a.aLUT = aLUT.as<T>();
    // AddExchangeField, engine/exchange.go:49-66
    UM_REQUIRE(!(has_dmi && has_dmibulk), UM_EINVAL, "Cannot have induced and interfacial DMI at the same time");
    a.dLUT = has_dmibulk ? dbLUT.as<T>() : dLUT.as<T>();
    a.exch_mode = has_dmi ? 2 : has_dmibulk ? 3 : 1;
    a.openbc = openbc;
    a.has_uniax = nz(UM_P_KU1) || nz(UM_P_KU2);
    a.Ku1 = P(UM_P_KU1); a.Ku2 = P(UM_P_KU2);
    a.has_cubic = nz(UM_P_KC1) || nz(UM_P_KC2) || nz(UM_P_KC3);
    a.Kc1 = P(UM_P_KC1); a.Kc2 = P(UM_P_KC2); a.Kc3 = P(UM_P_KC3);
    for (int c = 0; c < 3; c++) {
      a.u[c] = P(UM_P_ANISU, c);
      a.c1[c] = P(UM_P_ANISC1, c);
      a.c2[c] = P(UM_P_ANISC2, c);
      a.bext[c] = P(UM_P_BEXT, c);
      a.bextLUT[c] = has_bextlut ? bextLUT.as<T>() + c * NREGION : nullptr;
    }
    a.has_bext = nz(UM_P_BEXT, 3);
    a.has_therm = nz(UM_P_TEMP) && !relaxing;
    a.temp = P(UM_P_TEMP);
    a.kth = (T)(2 * KB / (gammaLL * cs[0] * cs[1] * cs[2] * dt));
    a.seed = seed;
    a.step = (uint64_t)nsteps;
    a.step_dev = capturing ? (const unsigned long long *)(gstat.as<double>() + 3) : nullptr;
    a.stream = tstream;
    a.alpha = P(UM_P_ALPHA);
    a.precess = precess;
    a.frozenLUT = has_frozen ? frozenLUT.as<T>() : nullptr;
    // AddSTTorque, engine/torque.go:147-196
    const bool has_j = nz(UM_P_J, 3);
    UM_REQUIRE(!has_j || nz(UM_P_POL), UM_EINVAL, "spin polarization should not be 0");
    a.has_zl = has_j && !disable_zl;
    a.has_sl = has_j && !disable_sl && nz(UM_P_FIXEDLAYER, 3);
    for (int c = 0; c < 3; c++) {
      a.stt.J[c] = P(UM_P_J, c);
      a.stt.fixp[c] = P(UM_P_FIXEDLAYER, c);
    }
    a.stt.xi = P(UM_P_XI);
    a.stt.pol = P(UM_P_POL);
    a.stt.lambda = P(UM_P_LAMBDA);
    a.stt.epsp = P(UM_P_EPSPRIME);
    a.stt.thick = P(UM_P_FREELAYER);
    a.stt.mesh_thick = (T)(N[2] * cs[2]);
    a.stt.flp = (T)flp;
    if (!xterms.empty()) {
      if (combine_terms(xfield, 1u | 2u))
        for (int c = 0; c < 3; c++) a.badd[c] = xfield.as<T>() + c * n;
      if (combine_terms(xtorque, 4u))
        for (int c = 0; c < 3; c++) a.tadd[c] = xtorque.as<T>() + c * n;
    }
    return a;
  }
  // B_demag into bd (SetDemagField, engine/demag.go:42-65); with NoDemagSpins
  // the flagged cells neither generate nor receive the field (:67-98)
  void compute_demag(const T *mm) {
    if (!enable_demag) return;
    ensure_demag();
    void *B[3] = {bd, bd + n, bd + 2 * n};
    const void *mv[3] = {mm, mm + n, mm + 2 * n};
    const void *v = vol.p;
    if (has_nodemag) {
      if (nd_gen != gen || !ndvol.p) {
        if (!ndvol.p) ndvol.alloc(sizeof(T) * n);
        UM_LAUNCH(k_nodemag_vol<T>, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, st, ndvol.as<T>(),
                           (const T *)vol.p, nodemagLUT.as<T>(), has_regions ? (const uint8_t *)regions.p : nullptr,
                           (long)n);
        UM_HIP(hipGetLastError());
        nd_gen = gen;
      }
      v = ndvol.p;
    }
    demag_exec(demag, B, mv, v, p[UM_P_MSAT][0].arr.p, p[UM_P_MSAT][0].mul, st);
    if (has_nodemag)
      for (int c = 0; c < 3; c++)
        zeromask<T>(bd + c * n, nodemagLUT.p, has_regions ? (const uint8_t *)regions.p : nullptr, n, st);
  }
  // torqueFn (engine/run.go:127-130): SetTorque into dst
  hipEvent_t next_event() {
    if (evused == (int)evpool.size()) {
      hipEvent_t e;
      UM_HIP(hipEventCreate(&e));
      evpool.push_back(e);
    }
    return evpool[evused++];
  }
  // read the events of the evaluations enqueued so far (stream must be drained)
  void prof_collect() {
    if (!profiling) return;
    for (int i = 0; i + 2 < evused; i += 3) {
      float d1 = 0, d2 = 0;
      UM_HIP(hipEventElapsedTime(&d1, evpool[i], evpool[i + 1]));
      UM_HIP(hipEventElapsedTime(&d2, evpool[i + 1], evpool[i + 2]));
      t_demag += d1;
      t_field += d2;
      prof_evals++;
    }
    evused = 0;
  }
  void torque_fn(T *dst) {
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    if (profiling) {
      e0 = next_event();
      e1 = next_event();
      e2 = next_event();
      UM_HIP(hipEventRecord(e0, st));
    }
    FieldArgs<T> a = field_args(m);
    for (int c = 0; c < 3; c++) {
      a.bd[c] = enable_demag ? bd + c * n : nullptr;
      a.tau[c] = dst + c * n;
      a.beff[c] = nullptr;
    }
    bool fused = false;
    if (enable_demag) {
      // one pipeline: B_demag stays on chip (the field share is timed with the demag)
      ensure_demag();
      const void *mv[3] = {m, m + n, m + 2 * n};
      if (!has_nodemag)
        fused = demag_exec_field(demag, mv, vol.p, p[UM_P_MSAT][0].arr.p, p[UM_P_MSAT][0].mul, &a, st);
      if (!fused) compute_demag(m);
    }
    if (profiling) UM_HIP(hipEventRecord(e1, st));
    if (!fused) launch_field_torque(a, st);
    if (profiling) UM_HIP(hipEventRecord(e2, st));
    nevals++;
  }
  void beff(void *const out[3]) override {
    compute_demag(m);
    FieldArgs<T> a = field_args(m);
    for (int c = 0; c < 3; c++) {
      a.bd[c] = enable_demag ? bd + c * n : nullptr;
      a.tau[c] = nullptr;
      a.beff[c] = spare + c * n;
    }
    launch_field_torque(a, st);
    copy_out(spare, out);
  }
  void torque(void *const out[3]) override {
    torque_fn(spare);
    nevals--;
    copy_out(spare, out);
    prof_collect();
  }
  void demag_field(void *const out[3]) override {
    if (enable_demag) {
      compute_demag(m);
      copy_out(bd, out);
    } else {
      fill_bytes(spare, sizeof(T) * 3 * n, st);
      copy_out(spare, out);
    }
  }
  void copy_out(const T *src, void *const out[3]) {
    for (int c = 0; c < 3; c++)
      UM_HIP(hipMemcpyAsync(out[c], src + c * n, sizeof(T) * n, hipMemcpyDeviceToHost, st));
    UM_HIP(hipStreamSynchronize(st));
  }

  // ---- solver ------------------------------------------------------------------------
  void set_solver(int s) override {
    gen++;
    UM_REQUIRE(s == UM_SOLVER_HEUN || s == UM_SOLVER_RK23 || s == UM_SOLVER_RK4 || s == UM_SOLVER_RK45DP, UM_EUNSUP,
               "solver " + std::to_string(s) + " not available (Heun=2, RK23=3, RK4=4, RK45DP=5)");
    solver = s;
    k1_valid = false;
  }
  void set_solver_params(const double kk[9]) override {
    gen++;
    dt = kk[0]; fixdt = kk[1]; maxerr = kk[2]; relerr = kk[3]; minslope = kk[4];
    headroom = kk[5]; mindt = kk[6]; maxdt = kk[7];
    if (kk[8] != gammaLL) k1_valid = false;
    gammaLL = kk[8];
  }
  void get_state(double o[8]) override {
    o[0] = Time; o[1] = dt; o[2] = (double)nsteps; o[3] = (double)nundone; o[4] = (double)nevals;
    o[5] = lastErr; o[6] = peakErr; o[7] = lastTorque;
  }
  void set_last_err(double err) {
    lastErr = err;
    if (err > peakErr) peakErr = err;
  }
  // adaptDt, engine/run.go:150-187
  void adapt(double corr) {
    if (fixdt != 0) { dt = fixdt; return; }
    if (std::isnan(corr)) corr = 1;
    UM_REQUIRE(corr != 0, UM_ESTATE, "Time step too small, check if parameters are sensible");
    corr *= headroom;
    if (corr > 2) corr = 2;
    if (corr < 0.5) corr = 0.5;
    dt *= corr;
    if (mindt != 0 && dt < mindt) dt = mindt;
    if (maxdt != 0 && dt > maxdt) dt = maxdt;
    UM_REQUIRE(dt != 0, UM_ESTATE, "time step too small");
    if (Time < alarm && Time + dt > alarm) dt = alarm - Time;
    UM_REQUIRE(dt > 0, UM_ESTATE, "Time step too small");
  }
  // m_out = normalize(sum fac_j src_j)
  void stage(T *dst, int ns, const T *const *src, const T *fac, bool norm, bool group3 = false,
             T *nrm = nullptr) {
    StageArgs<T> a{};
    a.nrm = nrm;
    a.nsrc = ns;
    a.ncomp = 3;
    a.group3 = group3;
    a.normalize = norm;
    for (int j = 0; j < ns; j++) {
      for (int c = 0; c < 3; c++) a.src[j][c] = src[j] + c * n;
      a.fac[j] = fac[j];
    }
    for (int c = 0; c < 3; c++) a.dst[c] = dst + c * n;
    a.vol = (const T *)vol.p;
    a.N = n;
    if (devctl) {
      AdaptState *d = astate.as<AdaptState>();
      a.dev_dt = &d->dt;
      a.gam = gammaLL;
      a.skip = &d->done;
    }
    launch_stage(a, st);
  }
  // device statistics of a captured step (count: its last statistics launch)
  void step_stats(double h, int mode, bool count) {
    launch_step_stats((unsigned long long *)red.p, gstat.as<double>(), h, mode, count ? 1 : 0, st);
    if (count) counted = true;
  }
  // fused norms -> host (one readback).  A captured step finds the slots
  // cleared: graph_prepare clears them before the replays and every
  // k_step_stats after reading them.
  void err_reduce(int mode, const T *const *src, int nsrc, const T *e, double out[3]) {
    if (!capturing) fill_bytes(red.p, red.bytes, st);
    ErrArgs<T> a{};
    a.mode = mode;
    for (int j = 0; j < nsrc; j++)
      for (int c = 0; c < 3; c++) a.k[j][c] = src[j] + c * n;
    if (e) for (int j = 0; j < 6; j++) a.e[j] = e[j];
    a.N = n;
    a.res = (unsigned long long *)red.p;
    launch_err(a, st);
    if (capturing) {   // the statistics are folded on the device (step_stats); decisions are fixed by FixDt
      out[0] = out[1] = out[2] = 0;
      return;
    }
    UM_HIP(hipMemcpyAsync(hres, red.p, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    UM_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < 3; i++) memcpy(&out[i], &hres[i], 8);
  }

  // RK45DP.Step, engine/rk45dp.go:12-98, with integralController (stepcontroller.go:35-97)
  void step_rk45dp() {
    if (fixdt != 0) dt = fixdt;
    T *k1 = k[0];
    if (!k1_valid) { torque_fn(k1); k1_valid = true; }
    if (nz(UM_P_TEMP)) torque_fn(k1);
    const double t0 = Time;
    std::swap(m, m0);   // m0 <- current magnetization; m is overwritten by every stage
    // device-controlled: the stage kernels multiply T(c) by the step's h
    const T h = devctl ? T(1) : (T)(dt * gammaLL);
    T *k2 = k[1], *k3 = k[2], *k4 = k[3], *k5 = k[4], *k6 = k[5];
    {
      const T *s[] = {m0, k1};
      const T f[] = {T(1), T(1. / 5.) * h};
      Time = t0 + (1. / 5.) * dt;
      stage(m, 2, s, f, true);
      torque_fn(k2);
    }
    {
      const T *s[] = {m0, k1, k2};
      const T f[] = {T(1), T(3. / 40.) * h, T(9. / 40.) * h};
      Time = t0 + (3. / 10.) * dt;
      stage(m, 3, s, f, true, true);
      torque_fn(k3);
    }
    {
      const T *s[] = {m0, k1, k2, k3};
      const T f[] = {T(1), T(44. / 45.) * h, T(-56. / 15.) * h, T(32. / 9.) * h};
      Time = t0 + (4. / 5.) * dt;
      stage(m, 4, s, f, true);
      torque_fn(k4);
    }
    {
      const T *s[] = {m0, k1, k2, k3, k4};
      const T f[] = {T(1), T(19372. / 6561.) * h, T(-25360. / 2187.) * h, T(64448. / 6561.) * h,
                     T(-212. / 729.) * h};
      Time = t0 + (8. / 9.) * dt;
      stage(m, 5, s, f, true);
      torque_fn(k5);
    }
    {
      const T *s[] = {m0, k1, k2, k3, k4, k5};
      const T f[] = {T(1), T(9017. / 3168.) * h, T(-355. / 33.) * h, T(46732. / 5247.) * h,
                     T(49. / 176.) * h, T(-5103. / 18656.) * h};
      Time = t0 + dt;
      stage(m, 6, s, f, true);
      torque_fn(k6);
    }
    T *k7 = k2;
    {
      const T *s[] = {m0, k1, k3, k4, k5, k6};
      const T f[] = {T(1), T(35. / 384.) * h, T(500. / 1113.) * h, T(125. / 192.) * h,
                     T(-2187. / 6784.) * h, T(11. / 84.) * h};
      Time = t0 + dt;
      stage(m, 6, s, f, true);
      torque_fn(k7);
    }
    const T e[6] = {T(35. / 384. - 5179. / 57600.), T(500. / 1113. - 7571. / 16695.),
                    T(125. / 192. - 393. / 640.), T(-2187. / 6784. - -92097. / 339200.),
                    T(11. / 84. - 187. / 2100.), T(0. - 1. / 40.)};
    const T *src[6] = {k1, k3, k4, k5, k6, k7};
    double r[3];
    err_reduce(0, src, 6, e, r);
    if (devctl) {   // accept / reject, the next dt and the FSAL copy on the device
      AdaptPar ap{maxerr, relerr, minslope, headroom, mindt, maxdt, gammaLL, 6, sizeof(T) == 4 ? 1 : 0};
      launch_adapt<T>((unsigned long long *)red.p, astate.as<AdaptState>(), ap, m, m0, k1, k7, 3 * n, st);
      counted = true;
      return;
    }
    const double err = std::sqrt(r[0]) * (double)h;
    const double maxdm = std::sqrt(r[1]);
    if (err < maxerr || dt <= mindt || fixdt != 0) {
      int fail = 0;
      double rlerr = 0;
      if (maxdm < minslope) fail = 0;
      else { rlerr = r[2]; fail = 1; }
      if (fail == 0 || relerr <= 0.0 || rlerr < relerr || dt <= mindt || fixdt != 0) {
        if (capturing) step_stats((double)h, 0, true);
        else {
          set_last_err(err);
          lastTorque = maxdm;
        }
        nsteps++;
        Time = t0 + dt;
        if (fail == 0) adapt(std::pow(maxerr / err, 1. / 5));
        else adapt(std::pow(relerr / rlerr, 1. / 5));
        std::swap(k[0], k[1]);   // FSAL: k1 <- k7
        k1_valid = true;
      } else {
        Time = t0;
        std::swap(m, m0);
        nundone++;
        adapt(std::pow(relerr / rlerr, 1. / 6));
      }
    } else {
      Time = t0;
      std::swap(m, m0);
      nundone++;
      adapt(std::pow(maxerr / err, 1. / 6));
    }
  }

  // RK23 (Bogacki-Shampine).Step, engine/rk23.go:21-84, integralController
  // (accept order 3, reject order 4, FSAL k1 <- k4)
  void step_rk23() {
    if (fixdt != 0) dt = fixdt;
    T *k1 = k[0];
    if (!k1_valid) { torque_fn(k1); k1_valid = true; }
    if (nz(UM_P_TEMP)) torque_fn(k1);   // FSAL is not used with temperature
    const double t0 = Time;
    std::swap(m, m0);
    const T h = (T)(dt * gammaLL);
    T *k2 = k[1], *k3 = k[2], *k4 = k[3];
    {
      const T *s[] = {m0, k1};
      const T f[] = {T(1), T(1. / 2.) * h};
      Time = t0 + (1. / 2.) * dt;
      stage(m, 2, s, f, true);
      torque_fn(k2);
    }
    {
      const T *s[] = {m0, k2};
      const T f[] = {T(1), T(3. / 4.) * h};
      Time = t0 + (3. / 4.) * dt;
      stage(m, 2, s, f, true);
      torque_fn(k3);
    }
    {
      const T *s[] = {m0, k1, k2, k3};
      const T f[] = {T(1), T(2. / 9.) * h, T(1. / 3.) * h, T(4. / 9.) * h};
      stage(m, 4, s, f, true);
    }
    Time = t0 + dt;
    torque_fn(k4);
    // Err = sum e_j k_j over k1..k4 (the two trailing zero terms leave the sum
    // unchanged); the controller's delM is the last source, k4
    const T e[6] = {T(7. / 24. - 2. / 9.), T(1. / 4. - 1. / 3.), T(1. / 3. - 4. / 9.), T(1. / 8.), T(0), T(0)};
    const T *src[6] = {k1, k2, k3, k4, k4, k4};
    double r[3];
    err_reduce(0, src, 6, e, r);
    const double err = std::sqrt(r[0]) * (double)h;
    const double maxdm = std::sqrt(r[1]);
    if (err < maxerr || dt <= mindt || fixdt != 0) {
      int fail = 0;
      double rlerr = 0;
      if (maxdm < minslope) fail = 0;
      else { rlerr = r[2]; fail = 1; }
      if (fail == 0 || relerr <= 0.0 || rlerr < relerr || dt <= mindt || fixdt != 0) {
        if (capturing) step_stats((double)h, 0, true);
        else {
          set_last_err(err);
          lastTorque = maxdm;
        }
        nsteps++;
        Time = t0 + dt;
        if (fail == 0) adapt(std::pow(maxerr / err, 1. / 3));
        else adapt(std::pow(relerr / rlerr, 1. / 3));
        std::swap(k[0], k[3]);   // FSAL: k1 <- k4
        k1_valid = true;
      } else {
        Time = t0;
        std::swap(m, m0);
        nundone++;
        adapt(std::pow(relerr / rlerr, 1. / 4));
      }
    } else {
      Time = t0;
      std::swap(m, m0);
      nundone++;
      adapt(std::pow(maxerr / err, 1. / 4));
    }
  }

  // Heun.Step, engine/heun.go:15-59
  void step_heun() {
    T *dy0 = k[0], *dy = k[1];
    if (fixdt != 0) dt = fixdt;
    const T h = (T)(dt * gammaLL);
    UM_REQUIRE(h > 0, UM_ESTATE, "dt must be positive");
    torque_fn(dy0);
    std::swap(m, m0);
    {
      const T *s[] = {m0, dy0};
      const T f[] = {T(1), h};
      stage(m, 2, s, f, false);   // stage 1 is not normalised (heun.go:32)
    }
    Time += dt;
    torque_fn(dy);
    double r[3];
    const T *src[2] = {dy0, dy};
    err_reduce(1, src, 2, nullptr, r);
    double err = std::sqrt(heun_sqrt3 ? 3.0 * r[0] : r[0]) * (double)h;
    if (capturing) step_stats((double)h, heun_sqrt3 ? 3 : 1, false);
    if (err < maxerr || dt <= mindt || fixdt != 0) {
      const T *s[] = {m, dy, dy0};
      const T f[] = {T(1), T(0.5) * h, T(-0.5) * h};
      stage(m, 3, s, f, true, true);
      nsteps++;
      adapt(std::pow(maxerr / err, 1. / 2.));
      if (!capturing) set_last_err(err);
      const T *sn[1] = {dy};
      double r2[3];
      err_reduce(2, sn, 1, nullptr, r2);
      if (capturing) step_stats(0.0, 2, true);
      else lastTorque = std::sqrt(r2[0]);
    } else {
      Time -= dt;
      std::swap(m, m0);
      nundone++;
      adapt(std::pow(maxerr / err, 1. / 3.));
    }
  }

  // RK4.Step, engine/rk4.go:11-69 with simpleController (stepcontroller.go:13-32)
  void step_rk4() {
    if (fixdt != 0) dt = fixdt;
    const double t0 = Time;
    T *k1 = k[0], *k2 = k[1], *k3 = k[2], *k4 = k[3];
    const T h = (T)(dt * gammaLL);
    torque_fn(k1);
    std::swap(m, m0);
    Time = t0 + (1. / 2.) * dt;
    { const T *s[] = {m0, k1}; const T f[] = {T(1), T(1. / 2.) * h}; stage(m, 2, s, f, true); }
    torque_fn(k2);
    { const T *s[] = {m0, k2}; const T f[] = {T(1), T(1. / 2.) * h}; stage(m, 2, s, f, true); }
    torque_fn(k3);
    Time = t0 + dt;
    { const T *s[] = {m0, k3}; const T f[] = {T(1), T(1.) * h}; stage(m, 2, s, f, true); }
    torque_fn(k4);
    double r[3];
    const T *src[2] = {k1, k4};
    err_reduce(1, src, 2, nullptr, r);
    const double err = std::sqrt(r[0]) * (double)h;
    if (capturing) step_stats((double)h, 1, false);
    bool ok;
    if (err < maxerr || dt <= mindt || fixdt != 0) {
      nsteps++;
      adapt(std::pow(maxerr / err, 1. / 4.));
      if (!capturing) set_last_err(err);
      ok = true;
    } else {
      nundone++;
      adapt(std::pow(maxerr / err, 1. / 5.));
      ok = false;
    }
    if (ok) {
      const T *s[] = {m0, k1, k2, k3, k4};
      const T f[] = {T(1), T(1. / 6.) * h, T(1. / 3.) * h, T(1. / 3.) * h, T(1. / 6.) * h};
      stage(m, 5, s, f, true);
      const T *sn[1] = {k4};
      double r2[3];
      err_reduce(2, sn, 1, nullptr, r2);
      if (capturing) step_stats(0.0, 2, true);
      else lastTorque = std::sqrt(r2[0]);
    } else {
      Time = t0;
      std::swap(m, m0);
    }
  }

  // ---- discrete adjoint of fixed-step RK4 ------------------------------------------------
  // While the tape records, every step stores its stage inputs, the norms of
  // the stage sums before normalisation and the stage demag fields in a
  // segment tape of K steps; m at every K-th step is kept as a checkpoint.
  // The reverse sweep walks the segments backwards, recomputing a segment from
  // its checkpoint when the tape no longer holds it.  K = the step capacity
  // (default) keeps the whole trajectory in HBM and recomputes nothing;
  // K ~ sqrt(steps) bounds the memory to ~2 sqrt(steps) states.
  bool tape_on = false;
  long tape_K = 0, tape_cap = 0, tape_n = 0, seg_start = -1, tape_gen = -1;
  DevBuf tape_ckpt, tape_seg;
  // segment tape: M[0..K] (3n each), then per step 25n: U[3] (3n), NRM[4] (n), BD[4] (3n)
  T *tM(long k) { return tape_seg.as<T>() + (size_t)k * 3 * n; }
  T *tStep(long k) { return tape_seg.as<T>() + (size_t)(tape_K + 1) * 3 * n + (size_t)k * 25 * n; }
  T *tU(long k, int s) { return tStep(k) + (size_t)s * 3 * n; }
  T *tN(long k, int s) { return tStep(k) + 9 * n + (size_t)s * n; }
  T *tBD(long k, int s) { return tStep(k) + 13 * n + (size_t)s * 3 * n; }

  // torqueFn at an explicit input, the demag field into bd_out (kept for the adjoint)
  void torque_at(const T *u, T *bd_out, T *dst) {
    FieldArgs<T> a = field_args(u);
    if (enable_demag) {
      ensure_demag();
      void *B[3] = {bd_out, bd_out + n, bd_out + 2 * n};
      const void *mv[3] = {u, u + n, u + 2 * n};
      demag_exec(demag, B, mv, vol.p, p[UM_P_MSAT][0].arr.p, p[UM_P_MSAT][0].mul, st);
    }
    for (int c = 0; c < 3; c++) {
      a.bd[c] = enable_demag ? bd_out + c * n : nullptr;
      a.tau[c] = dst + c * n;
      a.beff[c] = nullptr;
    }
    launch_field_torque(a, st);
    nevals++;
  }
  // RK4.Step (engine/rk4.go:11-69) with FixDt on tape slot k: M[k] -> M[k+1]
  void step_rk4_taped(long kk) {
    const T h = (T)(fixdt * gammaLL);
    T *k1 = k[0], *k2 = k[1], *k3 = k[2], *k4 = k[3];
    const T *u1 = tM(kk);
    torque_at(u1, tBD(kk, 0), k1);
    { const T *s_[] = {u1, k1}; const T f[] = {T(1), T(1. / 2.) * h}; stage(tU(kk, 0), 2, s_, f, true, false, tN(kk, 0)); }
    torque_at(tU(kk, 0), tBD(kk, 1), k2);
    { const T *s_[] = {u1, k2}; const T f[] = {T(1), T(1. / 2.) * h}; stage(tU(kk, 1), 2, s_, f, true, false, tN(kk, 1)); }
    torque_at(tU(kk, 1), tBD(kk, 2), k3);
    { const T *s_[] = {u1, k3}; const T f[] = {T(1), T(1.) * h}; stage(tU(kk, 2), 2, s_, f, true, false, tN(kk, 2)); }
    torque_at(tU(kk, 2), tBD(kk, 3), k4);
    const T *s_[] = {u1, k1, k2, k3, k4};
    const T f[] = {T(1), T(1. / 6.) * h, T(1. / 3.) * h, T(1. / 3.) * h, T(1. / 6.) * h};
    stage(tM(kk + 1), 5, s_, f, true, false, tN(kk, 3));
  }
  void copy3(T *dst, const T *src) {
    UM_HIP(hipMemcpyAsync(dst, src, sizeof(T) * 3 * n, hipMemcpyDeviceToDevice, st));
  }
  void step_tape() {
    UM_REQUIRE(tape_n < tape_cap, UM_ESTATE, "adjoint tape full (raise max_steps)");
    UM_REQUIRE(gen == tape_gen, UM_ESTATE, "parameters changed while the adjoint tape records");
    const long kk = tape_n % tape_K;
    if (kk == 0) {
      copy3(tape_ckpt.as<T>() + (size_t)(tape_n / tape_K) * 3 * n, m);
      copy3(tM(0), m);
      seg_start = tape_n;
    }
    dt = fixdt;
    step_rk4_taped(kk);
    copy3(m, tM(kk + 1));
    Time += dt;
    nsteps++;
    tape_n++;
  }
  void tape_begin(int cap, int every) override {
    UM_REQUIRE(cap > 0 && every >= 0, UM_EINVAL, "tape: max_steps > 0, checkpoint_every >= 0");
    UM_REQUIRE(solver == UM_SOLVER_RK4 && fixdt > 0, UM_EUNSUP, "the adjoint needs the RK4 solver with FixDt > 0");
    UM_REQUIRE(!nz(UM_P_TEMP), UM_EUNSUP, "no adjoint with a thermal field");
    UM_REQUIRE(!(nz(UM_P_KC1) || nz(UM_P_KC2) || nz(UM_P_KC3)), UM_EUNSUP, "no adjoint of cubic anisotropy");
    UM_REQUIRE(precess && !has_frozen && !has_vol && !relaxing, UM_EUNSUP,
               "the adjoint needs precession on, no frozen spins and no geometry");
    UM_REQUIRE(!has_dmibulk && !nz(UM_P_J, 3), UM_EUNSUP, "no adjoint of bulk DMI or spin-transfer torques");
    UM_REQUIRE(!has_nodemag, UM_EUNSUP, "no adjoint with NoDemagSpins");
    UM_REQUIRE(xterms.empty(), UM_EUNSUP, "no adjoint with extra field or torque terms");
    tape_end();
    tape_K = every > 0 ? std::min<long>(every, cap) : cap;
    const long nck = (cap + tape_K - 1) / tape_K;
    tape_ckpt.alloc(sizeof(T) * 3 * n * nck);
    tape_seg.alloc(sizeof(T) * ((size_t)(tape_K + 1) * 3 * n + (size_t)tape_K * 25 * n));
    tape_cap = cap;
    tape_n = 0;
    seg_start = -1;
    tape_on = true;
    gen++;
    tape_gen = gen;
  }
  void tape_end() override {
    if (tape_seg.p || tape_ckpt.p) UM_HIP(hipStreamSynchronize(st));
    tape_on = false;
    tape_n = tape_cap = 0;
    seg_start = -1;
    tape_ckpt.release();
    tape_seg.release();
  }
  long tape_steps() const override { return tape_n; }
  // one step of the reverse sweep (slot kk): lm -> cotangent of the step's m0
  // stage input m of stage s (1..4) of tape slot kk
  const T *stage_u(long kk, int s) { return s == 1 ? tM(kk) : tU(kk, s - 2); }
  void adjoint_step(long kk, T *&lm, T *&lacc, T *lv5, T *&lk, T *&lk2, T *&w, T *&w2, T *D, double *gms,
                    double *part) {
    const T h = (T)(fixdt * gammaLL);
    AdjFinalArgs<T> fa{};
    for (int c = 0; c < 3; c++) {
      fa.m1[c] = tM(kk + 1) + c * n;
      fa.lm[c] = lm + c * n;
      fa.u4[c] = stage_u(kk, 4) + c * n;
      fa.lv5[c] = lv5 + c * n;
      fa.lacc[c] = lacc + c * n;
      fa.lk[c] = lk + c * n;
      fa.w[c] = w + c * n;
    }
    fa.nrm = tN(kk, 3);
    fa.c4 = T(1. / 6.) * h;
    fa.Ms = P(UM_P_MSAT);
    fa.alpha = P(UM_P_ALPHA);
    fa.gbext = part;
    fa.N = n;
    launch_adj_final(fa, st);
    for (int sidx = 4; sidx >= 1; sidx--) {
      const T *u = stage_u(kk, sidx);
      AdjArgs<T> a{};
      a.f = field_args(u);
      for (int c = 0; c < 3; c++) {
        a.u[c] = u + c * n;
        a.lk[c] = lk + c * n;
        a.w[c] = w + c * n;
        a.bd[c] = enable_demag ? tBD(kk, sidx - 1) + c * n : nullptr;
        a.D[c] = nullptr;
        a.lv5[c] = lv5 + c * n;
        a.lacc[c] = lacc + c * n;
        a.lknext[c] = sidx == 1 ? nullptr : lk2 + c * n;
        a.unext[c] = sidx == 1 ? nullptr : stage_u(kk, sidx - 1) + c * n;
        a.wnext[c] = sidx == 1 ? nullptr : w2 + c * n;
      }
      a.nrm = sidx == 1 ? nullptr : tN(kk, sidx - 2);
      a.c5 = sidx == 2 ? T(1. / 6.) * h : T(1. / 3.) * h;
      a.cs = sidx == 4 ? T(1.) * h : T(1. / 2.) * h;
      a.mu0 = (T)MU0_S;
      a.gms = gms;
      a.gbext = part;
      a.N = n;
      if (enable_demag) {
        // D = K * (Ms w) = K * gB: the forward plan with the mu0 Ms input scaling undone
        void *B[3] = {D, D + n, D + 2 * n};
        const void *mv[3] = {w, w + n, w + 2 * n};
        demag_exec(demag, B, mv, nullptr, p[UM_P_MSAT][0].arr.p, p[UM_P_MSAT][0].mul / MU0_S, st);
        for (int c = 0; c < 3; c++) a.D[c] = D + c * n;
      }
      launch_adj_stage(a, st);
      std::swap(lk, lk2);
      std::swap(w, w2);
    }
    std::swap(lm, lacc);
  }
  void backward(const void *const gm[3], void *const gm0[3], void *gms_out, double *gbext_out) override {
    UM_REQUIRE(tape_on && tape_n > 0, UM_ESTATE, "backward: no recorded steps (tape_begin, then steps)");
    UM_REQUIRE(gen == tape_gen, UM_ESTATE, "parameters changed since the adjoint tape was started");
    for (int c = 0; c < 3; c++) UM_REQUIRE(gm[c], UM_EINVAL, "backward: NULL cotangent");
    if (enable_demag) ensure_demag();
    const size_t vb = sizeof(T) * 3 * n;
    DevBuf b_lm(vb), b_lacc(vb), b_lv5(vb), b_lk(vb), b_lk2(vb), b_w(vb), b_w2(vb), b_D;
    if (enable_demag) b_D.alloc(vb);
    DevBuf b_gms, b_part;
    const long nb = adj_blocks(n);
    if (gms_out) { b_gms.alloc(sizeof(double) * n); fill_bytes(b_gms.p, b_gms.bytes, st); }
    if (gbext_out) { b_part.alloc(sizeof(double) * 3 * nb); fill_bytes(b_part.p, b_part.bytes, st); }
    T *lm = b_lm.as<T>(), *lacc = b_lacc.as<T>(), *lk = b_lk.as<T>(), *lk2 = b_lk2.as<T>();
    T *w = b_w.as<T>(), *w2 = b_w2.as<T>();
    for (int c = 0; c < 3; c++)
      UM_HIP(hipMemcpyAsync(lm + c * n, gm[c], sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    const long K = tape_K, nseg = (tape_n + K - 1) / K;
    for (long sg = nseg - 1; sg >= 0; sg--) {
      const long s0 = sg * K, s1 = std::min(s0 + K, tape_n);
      if (seg_start != s0) {
        copy3(tM(0), tape_ckpt.as<T>() + (size_t)sg * 3 * n);
        for (long t = s0; t < s1; t++) step_rk4_taped(t - s0);
        seg_start = s0;
      }
      for (long t = s1 - 1; t >= s0; t--)
        adjoint_step(t - s0, lm, lacc, b_lv5.as<T>(), lk, lk2, w, w2, b_D.as<T>(), b_gms.as<double>(),
                     b_part.as<double>());
    }
    if (gm0)
      for (int c = 0; c < 3; c++)
        if (gm0[c]) UM_HIP(hipMemcpyAsync(gm0[c], lm + c * n, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    if (gms_out) launch_adj_cast<T>((T *)gms_out, b_gms.as<double>(), n, st);
    if (gbext_out) {
      std::vector<double> h(3 * nb);
      UM_HIP(hipMemcpyAsync(h.data(), b_part.p, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, st));
      UM_HIP(hipStreamSynchronize(st));
      for (int c = 0; c < 3; c++) {
        double t = 0;
        for (long b = 0; b < nb; b++) t += h[3 * b + c];
        gbext_out[c] = t;
      }
    }
    UM_HIP(hipStreamSynchronize(st));
  }

  void step() override {
    if (tape_on) {
      UM_REQUIRE(solver == UM_SOLVER_RK4 && fixdt > 0, UM_ESTATE, "the adjoint tape needs RK4 with FixDt");
      step_tape();
      return;
    }
    if (profiling) UM_HIP(hipEventRecord(step_ev[0], st));
    counted = false;
    switch (solver) {
      case UM_SOLVER_HEUN: step_heun(); break;
      case UM_SOLVER_RK23: step_rk23(); break;
      case UM_SOLVER_RK4: step_rk4(); break;
      default: step_rk45dp(); break;
    }
    // a captured step counts itself in its last k_step_stats
    if (capturing && !counted) launch_step_count((unsigned long long *)(gstat.as<double>() + 3), st);
    if (profiling) {
      UM_HIP(hipEventRecord(step_ev[1], st));
      UM_HIP(hipEventSynchronize(step_ev[1]));
      float d = 0;
      UM_HIP(hipEventElapsedTime(&d, step_ev[0], step_ev[1]));
      t_other += d;  // whole step; the stage/reduction share is this minus the evaluations
      prof_collect();
    }
  }
  // Steps(n), engine/run.go:197-200; RunWhile drops the FSAL torque (:207)
  void steps(int cnt) override {
    const long stop = nsteps + cnt;
    long reps = graph_prepare(stop);
    while (reps-- > 0) graph_replay();
    graph_finish(stop);
  }

  // ---- graph replay (FixDt only) ---------------------------------------------------------
  void set_graph(int on) override { graph_mode = on; }
  long step_count() const override { return nsteps; }
  bool graph_usable() const {
    return graph_mode && fixdt != 0 && !profiling && !tape_on && !any_tfn && !xterm_tfn;
  }
  // one eager step (FSAL setup), then capture two steps if the cached graph is
  // stale; returns how many two-step replays fit before `stop`
  // Every Steps/RunWhile starts without the FSAL torque (engine/run.go:207,
  // stepper.Free()), so the ensemble entry point and Steps agree
  long graph_prepare(long stop) override { return prepare_replays(stop, true); }
  void graph_ensure() override {
    const void *ptrs[8] = {m, m0, k[0], k[1], k[2], k[3], k[4], k[5]};
    bool fresh = gexec && g_gen == gen && g_solver == solver && g_dt == fixdt;
    for (int i = 0; i < 8 && fresh; i++) fresh = (g_ptrs[i] == ptrs[i]);
    if (!fresh) capture2();
  }
  long batch_prepare(long stop) override { return prepare_replays(stop, false); }
  long prepare_replays(long stop, bool capture) {
    k1_valid = false;
    if (!graph_usable() || nsteps >= stop) return 0;
    step();
    if (nsteps + 2 > stop) return 0;
    const void *ptrs[8] = {m, m0, k[0], k[1], k[2], k[3], k[4], k[5]};
    bool fresh = gexec && g_gen == gen && g_solver == solver && g_dt == fixdt;
    for (int i = 0; i < 8 && fresh; i++) fresh = (g_ptrs[i] == ptrs[i]);
    if (capture && !fresh) capture2();
    fill_bytes(red.p, red.bytes, st);   // the replays' reductions start from cleared slots
    hstat[0] = lastErr; hstat[1] = peakErr; hstat[2] = lastTorque;
    const unsigned long long ns = (unsigned long long)nsteps;
    std::memcpy(&hstat[3], &ns, 8);
    UM_HIP(hipMemcpyAsync(gstat.p, hstat, 4 * sizeof(double), hipMemcpyHostToDevice, st));
    gstat_live = true;
    return (stop - nsteps) / 2;
  }
  void capture2() {
    if (gexec) { UM_HIP(hipGraphExecDestroy(gexec)); gexec = nullptr; }
    ensure_demag();
    UM_HIP(hipStreamSynchronize(st));
    const double t0 = Time;
    const long s0 = nsteps, e0 = nevals, u0 = nundone;
    hipGraph_t g = nullptr;
    UM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    capturing = true;
    try {
      step();
      step();
    } catch (...) {
      capturing = false;
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(st, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      throw;
    }
    capturing = false;
    UM_HIP(hipStreamEndCapture(st, &g));
    hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    UM_HIP(e);
    g_evals = nevals - e0;
    Time = t0; nsteps = s0; nevals = e0; nundone = u0;   // capture executed nothing
    g_gen = gen; g_solver = solver; g_dt = fixdt;
    const void *ptrs[8] = {m, m0, k[0], k[1], k[2], k[3], k[4], k[5]};
    for (int i = 0; i < 8; i++) g_ptrs[i] = ptrs[i];
  }
  void graph_replay() override {
    UM_HIP(hipGraphLaunch(gexec, st));
    for (int i = 0; i < 2; i++) Time = Time + dt;
    nsteps += 2;
    nevals += g_evals;
  }
  // the launches of two captured-style steps, recorded (nothing runs)
  long b_evals = 0;
  void batch_record(std::vector<BatchEntry> &out) override {
    UM_REQUIRE(graph_usable(), UM_EUNSUP, "batched steps need graph mode and FixDt");
    ensure_demag();
    const double t0 = Time;
    const long s0 = nsteps, e0 = nevals, u0 = nundone;
    BatchRec rec;
    rec.cur = &out;
    g_rec = &rec;
    capturing = true;
    try {
      step();
      step();
    } catch (...) {
      capturing = false;
      g_rec = nullptr;
      Time = t0; nsteps = s0; nevals = e0; nundone = u0;
      throw;
    }
    capturing = false;
    g_rec = nullptr;
    b_evals = nevals - e0;
    Time = t0; nsteps = s0; nevals = e0; nundone = u0;   // nothing ran
  }
  std::vector<uint64_t> batch_key() const override {
    std::vector<uint64_t> k{serial, (uint64_t)gen, (uint64_t)solver, 0, (uint64_t)(uintptr_t)st};
    std::memcpy(&k[3], &fixdt, 8);
    const void *ptrs[8] = {m, m0, this->k[0], this->k[1], this->k[2], this->k[3], this->k[4], this->k[5]};
    for (const void *q : ptrs) k.push_back((uint64_t)(uintptr_t)q);
    return k;
  }
  void batch_advance() override {
    for (int i = 0; i < 2; i++) Time = Time + dt;
    nsteps += 2;
    nevals += b_evals;
  }
  // device statistics back to the host, then the remaining steps eagerly
  void graph_finish(long stop) override {
    if (gstat_live) {   // only statistics this session's replays folded on the device
      UM_HIP(hipMemcpyAsync(hstat, gstat.p, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
      UM_HIP(hipStreamSynchronize(st));
      lastErr = hstat[0]; peakErr = hstat[1]; lastTorque = hstat[2];
      gstat_live = false;
    }
    while (nsteps < stop) step();
    UM_HIP(hipStreamSynchronize(st));
  }
  // Run(seconds), engine/run.go:190-194
  void run(double seconds) override {
    k1_valid = false;
    const double stop = Time + seconds;
    alarm = stop;
    if (devctl_usable() && Time < stop) {
      run_device(stop);
      return;
    }
    while (Time < stop) step();
    UM_HIP(hipStreamSynchronize(st));
  }
  // adaptive RK45DP in graph mode: accept / reject and the next dt decided on
  // the device, ADAPT_K steps per graph replay, one readback per replay
  bool devctl_usable() const {
    return graph_mode && solver == UM_SOLVER_RK45DP && fixdt == 0 && !profiling && !tape_on && !any_tfn &&
           !xterm_tfn && !nz(UM_P_TEMP);
  }
  void run_device(double stop) {
    if (!k1_valid) { torque_fn(k[0]); k1_valid = true; }   // RunWhile starts without the FSAL torque
    if (!hastate) {
      astate.alloc(sizeof(AdaptState));
      UM_HIP(hipHostMalloc((void **)&hastate, sizeof(AdaptState), hipHostMallocDefault));
    }
    AdaptState h0{};
    h0.dt = dt; h0.time = Time; h0.last_err = lastErr; h0.peak_err = peakErr; h0.last_torque = lastTorque;
    h0.stop = stop; h0.nsteps = nsteps; h0.nundone = nundone; h0.nevals = nevals;
    *hastate = h0;
    fill_bytes(red.p, red.bytes, st);
    UM_HIP(hipMemcpyAsync(astate.p, hastate, sizeof(AdaptState), hipMemcpyHostToDevice, st));
    const void *ptrs[8] = {m, m0, k[0], k[1], k[2], k[3], k[4], k[5]};
    bool fresh = agexec && a_gen == gen;
    for (int i = 0; i < 8 && fresh; i++) fresh = (a_ptrs[i] == ptrs[i]);
    if (!fresh) capture_adaptive();
    do {
      UM_HIP(hipGraphLaunch(agexec, st));
      UM_HIP(hipMemcpyAsync(hastate, astate.p, sizeof(AdaptState), hipMemcpyDeviceToHost, st));
      UM_HIP(hipStreamSynchronize(st));
    } while (!hastate->done);
    dt = hastate->dt; Time = hastate->time; lastErr = hastate->last_err; peakErr = hastate->peak_err;
    lastTorque = hastate->last_torque; nsteps = (long)hastate->nsteps; nundone = (long)hastate->nundone;
    nevals = (long)hastate->nevals;
    UM_REQUIRE(!hastate->error, UM_ESTATE, "Time step too small, check if parameters are sensible");
  }
  void capture_adaptive() {
    if (agexec) { UM_HIP(hipGraphExecDestroy(agexec)); agexec = nullptr; }
    ensure_demag();
    UM_HIP(hipStreamSynchronize(st));
    const double t0 = Time, dt0 = dt;
    const long s0 = nsteps, e0 = nevals, u0 = nundone;
    hipGraph_t g = nullptr;
    UM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    capturing = true;
    devctl = true;
    try {
      for (int i = 0; i < ADAPT_K; i++) step();
    } catch (...) {
      capturing = false;
      devctl = false;
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(st, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      Time = t0; dt = dt0; nsteps = s0; nevals = e0; nundone = u0;
      throw;
    }
    capturing = false;
    devctl = false;
    UM_HIP(hipStreamEndCapture(st, &g));
    hipError_t e = hipGraphInstantiate(&agexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    UM_HIP(e);
    Time = t0; dt = dt0; nsteps = s0; nevals = e0; nundone = u0;   // capture executed nothing
    a_gen = gen;
    const void *ptrs[8] = {m, m0, k[0], k[1], k[2], k[3], k[4], k[5]};
    for (int i = 0; i < 8; i++) a_ptrs[i] = ptrs[i];
  }

  // ---- outputs -----------------------------------------------------------------------
  double wdot(const T *const *a, const T *const *b, int nc, const T *w, T wmul) {
    launch_wdot<T>(a, b, nc, w, wmul, n, partial.as<double>(), dsum.as<double>(), st);
    UM_HIP(hipMemcpyAsync(hsum, dsum.p, sizeof(double), hipMemcpyDeviceToHost, st));
    UM_HIP(hipStreamSynchronize(st));
    return *hsum;
  }
  // sAverageMagnet, engine/average.go:28-40
  void average_m(double out[3]) override {
    const T *w = (const T *)vol.p;
    double denom = (double)n;
    if (w) {
      const T *a1[1] = {w};
      denom = wdot(a1, nullptr, 1, nullptr, T(1));
    }
    for (int c = 0; c < 3; c++) {
      const T *a1[1] = {m + c * n};
      out[c] = wdot(a1, nullptr, 1, w, T(1)) / denom;
    }
  }
  // E = pref * V * sum(Ms vol m . B_term), engine/energy.go, demag.go:146, exchange.go:79,
  // zeeman.go:15, anisotropy.go:163-216 (K1 -1/2, K2 -1/4, Kc1 -1/4, Kc2 -1/6, Kc3 -1/8),
  // temperature.go:129 (thermal, zero while relaxing); term 9 = GetTotalEnergy
  double energy(int term) override {
    UM_REQUIRE(term >= 0 && term <= 9, UM_EINVAL, "energy term 0..9");
    if (term == 9) {
      double tot = 0;
      for (int t = 0; t <= 8; t++) tot += energy(t);
      return tot;
    }
    if (term == 8 && (!nz(UM_P_TEMP) || relaxing)) return 0.0;
    static const double pref[9] = {-0.5, -0.5, -1.0, -0.5, -0.25, -0.25, -1. / 6., -1. / 8., -1.0};
    fill_bytes(spare, sizeof(T) * 3 * n, st);
    const T *B = spare;
    if (term == 0) {
      if (enable_demag) { compute_demag(m); B = bd; }
    } else {
      FieldArgs<T> a = field_args(m);
      // only the requested term: switch every other one off
      if (term != 1) a.exch_mode = 0;
      if (term != 2) {
        a.has_bext = 0;
        for (int c = 0; c < 3; c++) a.bextLUT[c] = nullptr;
      }
      // the Zeeman energy holds B_ext's mask terms; custom fields and torques have none
      for (int c = 0; c < 3; c++) a.badd[c] = a.tadd[c] = nullptr;
      if (term == 2 && combine_terms(xfield, 1u))
        for (int c = 0; c < 3; c++) a.badd[c] = xfield.as<T>() + c * n;
      a.has_uniax = (term == 3 || term == 4) ? a.has_uniax : 0;
      if (term == 3) a.Ku2 = Par<T>{nullptr, T(0)};
      if (term == 4) a.Ku1 = Par<T>{nullptr, T(0)};
      a.has_cubic = (term >= 5 && term <= 7) ? a.has_cubic : 0;
      if (term == 5) { a.Kc2 = Par<T>{nullptr, T(0)}; a.Kc3 = Par<T>{nullptr, T(0)}; }
      if (term == 6) { a.Kc1 = Par<T>{nullptr, T(0)}; a.Kc3 = Par<T>{nullptr, T(0)}; }
      if (term == 7) { a.Kc1 = Par<T>{nullptr, T(0)}; a.Kc2 = Par<T>{nullptr, T(0)}; }
      if (term != 8) a.has_therm = 0;
      for (int c = 0; c < 3; c++) { a.bd[c] = nullptr; a.tau[c] = nullptr; a.beff[c] = spare + c * n; }
      launch_field_torque(a, st);
    }
    // weight = Ms * vol per cell
    DevBuf w(sizeof(T) * n);
    const T *msa = (const T *)p[UM_P_MSAT][0].arr.p;
    if (msa && has_vol) launch_mul<T>(w.as<T>(), msa, (const T *)vol.p, n, st);
    else if (msa) UM_HIP(hipMemcpyAsync(w.p, msa, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    else if (has_vol) UM_HIP(hipMemcpyAsync(w.p, vol.p, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    else fill_value<T>(w.as<T>(), T(1), n, st);
    const T *a3[3] = {m, m + n, m + 2 * n};
    const T *b3[3] = {B, B + n, B + 2 * n};
    const double sum = wdot(a3, b3, 3, w.as<T>(), (T)p[UM_P_MSAT][0].mul);
    const double V = cs[0] * cs[1] * cs[2];
    return pref[term] * V * sum;
  }

  // relaxSteps, engine/relax.go:104-112: n steps, Time restored
  void relax_steps(int cnt) {
    const double t0 = Time;
    const long stop = nsteps + cnt;
    while (nsteps < stop) step();
    Time = t0;
  }
  double k1_maxnorm() {
    const T *src[1] = {k[0]};
    double r[3];
    err_reduce(2, src, 1, nullptr, r);
    return std::sqrt(r[0]);
  }
  double k1_dot() {
    const T *a3[3] = {k[0], k[0] + n, k[0] + 2 * n};
    return wdot(a3, a3, 3, nullptr, T(1));
  }
  // Relax, engine/relax.go:24-100
  void relax(double torque_threshold) override {
    gen++;
    const int prev_solver = solver, prev_precess = precess;
    const double prev_maxerr = maxerr, prev_fixdt = fixdt;
    solver = UM_SOLVER_RK23;
    k1_valid = false;
    fixdt = 0;
    precess = 0;
    relaxing = 1;
    try {
      const int NR = 3;   // energy every 3 steps
      relax_steps(NR);
      double E0 = energy(9);
      relax_steps(NR);
      double E1 = energy(9);
      while (E1 < E0) {
        relax_steps(NR);
        E0 = E1;
        E1 = energy(9);
      }
      if (torque_threshold > 0) {
        for (;;) {
          while (k1_maxnorm() > torque_threshold) relax_steps(NR);
          maxerr /= std::sqrt(2.0);
          if (maxerr < 1e-9) break;
        }
      } else {
        double T0 = 0, T1 = k1_dot();
        while (maxerr > 1e-9) {
          maxerr /= std::sqrt(2.0);
          relax_steps(NR);
          T0 = T1;
          T1 = k1_dot();
          while (T1 < T0) {
            relax_steps(NR);
            T0 = T1;
            T1 = k1_dot();
          }
        }
      }
    } catch (...) {
      solver = prev_solver; maxerr = prev_maxerr; fixdt = prev_fixdt; precess = prev_precess;
      relaxing = 0; k1_valid = false; gen++;
      throw;
    }
    solver = prev_solver;
    maxerr = prev_maxerr;
    fixdt = prev_fixdt;
    precess = prev_precess;
    relaxing = 0;
    k1_valid = false;
    gen++;
    UM_HIP(hipStreamSynchronize(st));
  }

  // Minimize, engine/minimizer.go:56-160 (Exl et al., JAP 115, 17D118): steepest
  // descent on the precession-free torque with Barzilai-Borwein step lengths,
  // until the last `samples` max|dm| are all below `stop`.  Time does not advance.
  void minimize(double stop, int samples) override {
    UM_REQUIRE(samples > 0, UM_EINVAL, "MinimizerSamples must be positive");
    gen++;
    const int prev_precess = precess;
    precess = 0;
    relaxing = 1;
    auto restore = [&] {
      precess = prev_precess;
      relaxing = 0;
      k1_valid = false;
      gen++;
    };
    try {
      std::vector<double> ring(samples, 0.0);
      int count = 0, tail = 0;
      auto ring_max = [&] {
        double mx = ring[0];
        for (int i = 1; i < count; i++) mx = std::max(mx, ring[i]);
        return mx;
      };
      DevBuf part(sizeof(double) * 5 * RED_BLOCKS), res(sizeof(double) * 5);
      double hres5[5];
      T h = T(1e-4);
      T *ka = k[0], *kb = k[1];
      torque_fn(ka);
      while (count < samples || ring_max() > stop) {
        launch_min_step<T>(m, m0, ka, h, n, st);
        torque_fn(kb);
        launch_min_stats<T>(m, m0, kb, ka, n, part.as<double>(), res.as<double>(), st);
        UM_HIP(hipMemcpyAsync(hres5, res.p, sizeof(hres5), hipMemcpyDeviceToHost, st));
        UM_HIP(hipStreamSynchronize(st));
        lastTorque = std::sqrt(hres5[4]);
        ring[tail] = std::sqrt(hres5[0]);
        tail = (tail + 1) % samples;
        count = std::min(count + 1, samples);
        set_last_err(ring_max());
        T nom, div;
        if (nsteps % 2 == 0) { nom = (T)hres5[1]; div = (T)hres5[2]; }
        else { nom = (T)hres5[2]; div = (T)hres5[3]; }
        h = (div != T(0)) ? nom / div : T(1e-4);
        normalize(m);
        nsteps++;
        std::swap(ka, kb);
      }
      if (ka != k[0]) std::swap(k[0], k[1]);
    } catch (...) {
      restore();
      throw;
    }
    restore();
  }

  void set_profiling(int on) override {
    UM_HIP(hipStreamSynchronize(st));
    profiling = on;
    evused = 0;
    t_demag = t_field = t_other = 0;
    prof_evals = 0;
  }
  void kernel_times(double ms[3], int *ne) override {
    ms[0] = t_demag; ms[1] = t_field; ms[2] = t_other;
    *ne = prof_evals;
  }
};

SimBase *sim_create(int prec, const int N[3], const double cs[3], const int pbc[3]) {
  if (prec == UM_F32) return new Sim<float>(N, cs, pbc);
  UM_REQUIRE(prec == UM_F64, UM_EINVAL, "prec must be UM_F32 or UM_F64");
  return new Sim<double>(N, cs, pbc);
}

}  // namespace um
