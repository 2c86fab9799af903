// C ABI of libumhip.so (declared in include/umhip.h).
#include <cxxabi.h>
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "batch.hpp"
#include "field.hpp"
#include "sim.hpp"
#include "vec.hpp"

namespace um {

static thread_local std::string g_err;
void set_last_error(const std::string &msg) { g_err = msg; }

// launch trace: the host stub of a __global__ function carries the kernel's
// mangled name, so dladdr + the C++ demangler name the launched variant
bool g_trace_on = false;
static std::mutex g_trace_mu;
static std::vector<std::string> g_trace;
void trace_kernel(const void *stub) {
  Dl_info info{};
  std::string name = "?";
  if (dladdr(stub, &info) && info.dli_sname) {
    int st = 0;
    char *d = abi::__cxa_demangle(info.dli_sname, nullptr, nullptr, &st);
    name = (st == 0 && d) ? d : info.dli_sname;
    std::free(d);
  }
  std::lock_guard<std::mutex> lk(g_trace_mu);
  g_trace.push_back(name);
}

thread_local BatchRec *g_rec = nullptr;
void batch_refuse(const void *stub) {
  Dl_info info{};
  std::string name = "?";
  if (dladdr(stub, &info) && info.dli_sname) {
    int st = 0;
    char *d = abi::__cxa_demangle(info.dli_sname, nullptr, nullptr, &st);
    name = (st == 0 && d) ? d : info.dli_sname;
    std::free(d);
  }
  throw Error(UM_EUNSUP, "kernel cannot run in a batched launch: " + name);
}

static std::mutex g_mu;
static bool g_inited = false;
static int g_device = 0;
static hipStream_t g_own_stream = nullptr;
static hipStream_t g_stream = nullptr;

// Device scratch of the allocation-free controller norms
// (um_rk45dp_error_noalloc): the error kernel's max slots.  Every GPU the
// library runs on holds its own copy until the process ends, at its own
// address; host code uses that address (g_scratch_dev), never the array's
// name, which on the host is only a stand-in for the symbol.
constexpr int NOALLOC_SLOTS = 4;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned long long scratch[NOALLOC_SLOTS];
static unsigned long long *g_scratch_dev[MAX_DEVICES];

void ensure_init() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_inited) return;
  int n = 0;
  UM_HIP(hipGetDeviceCount(&n));
  UM_REQUIRE(n > 0, UM_EHIP, "no HIP device");
  UM_REQUIRE(g_device >= 0 && g_device < n && g_device < MAX_DEVICES, UM_EINVAL, "um_init: no such device");
  UM_HIP(hipSetDevice(g_device));
  UM_HIP(hipStreamCreateWithFlags(&g_own_stream, hipStreamNonBlocking));
  // with the device current: this GPU's copy of the scratch array
  void *sp = nullptr;
  UM_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  UM_REQUIRE(sp != nullptr, UM_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
  g_scratch_dev[g_device] = (unsigned long long *)sp;
  g_stream = g_own_stream;
  g_inited = true;
}
hipStream_t lib_stream() {
  ensure_init();
  return g_stream;
}

// host launchers defined in field.hip
template <typename T>
void add_exchange(void *const B[3], const void *const m[3], Par<T> Ms, const T *aLUT, const uint8_t *regions, T wx,
                  T wy, T wz, Mesh g, hipStream_t s);
template <typename T>
void add_dmi(void *const B[3], const void *const m[3], Par<T> Ms, const T *aLUT, const T *dLUT,
             const uint8_t *regions, T cx, T cy, T cz, Mesh g, int openbc, hipStream_t s);
template <typename T>
void add_dmi_bulk(void *const B[3], const void *const m[3], Par<T> Ms, const T *aLUT, const T *dLUT,
                  const uint8_t *regions, T cx, T cy, T cz, Mesh g, int openbc, hipStream_t s);
template <typename T>
void add_stt(void *const t[3], const void *const m[3], Par<T> Ms, Par<T> alpha, const SttPar<T> &p, T cx, T cy,
             T cz, Mesh g, int zl, int sl, hipStream_t s);
template <typename T>
void add_uniaxial2(void *const B[3], const void *const m[3], Par<T> Ms, Par<T> K1, Par<T> K2, const Par<T> u[3],
                   int N, hipStream_t s);
template <typename T>
void add_cubic2(void *const B[3], const void *const m[3], Par<T> Ms, Par<T> K1, Par<T> K2, Par<T> K3,
                const Par<T> c1[3], const Par<T> c2[3], int N, hipStream_t s);
template <typename T>
void region_addv(void *const d[3], const void *const l[3], const uint8_t *regions, int N, hipStream_t s);
template <typename T>
void zeromask(void *dst, const void *lut, const uint8_t *regions, int N, hipStream_t s);
template <typename T>
void lltorque(void *const t[3], const void *const m[3], const void *const B[3], Par<T> alpha, int N,
              hipStream_t s);
template <typename T>
void llnoprecess(void *const t[3], const void *const m[3], const void *const B[3], int N, hipStream_t s);
template <typename T>
void thermal_field(void *const B[3], uint64_t seed, uint64_t step, uint32_t stream, T k, Par<T> Ms, Par<T> temp,
                   Par<T> alpha, int N, hipStream_t s);

static void check_prec(int prec) { UM_REQUIRE(prec == UM_F32 || prec == UM_F64, UM_EINVAL, "prec must be UM_F32/UM_F64"); }
static void check3(const void *const p[3], const char *what) {
  UM_REQUIRE(p && p[0] && p[1] && p[2], UM_EINVAL, std::string(what) + ": NULL component");
}

// single-use scratch for reductions on the library stream
struct RedScratch {
  DevBuf red, partial, sum;
  unsigned long long *h = nullptr;
  double *hs = nullptr;
  RedScratch() {
    red.alloc(4 * sizeof(unsigned long long));
    partial.alloc(RED_BLOCKS * sizeof(double));
    sum.alloc(sizeof(double));
    UM_HIP(hipHostMalloc((void **)&h, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    UM_HIP(hipHostMalloc((void **)&hs, sizeof(double), hipHostMallocDefault));
  }
  ~RedScratch() {
    if (h) (void)hipHostFree(h);
    if (hs) (void)hipHostFree(hs);
  }
};
static RedScratch &red_scratch() {
  static RedScratch *r = new RedScratch();
  return *r;
}

static double read_max_slot(hipStream_t s) {
  RedScratch &r = red_scratch();
  UM_HIP(hipMemcpyAsync(r.h, r.red.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  UM_HIP(hipStreamSynchronize(s));
  double v;
  memcpy(&v, &r.h[0], 8);
  return v;
}
template <typename T>
static double err_max(int mode, const void *const a[3], const void *const b[3], long N) {
  hipStream_t s = lib_stream();
  RedScratch &r = red_scratch();
  fill_bytes(r.red.p, r.red.bytes, s);
  ErrArgs<T> e{};
  e.mode = mode;
  for (int c = 0; c < 3; c++) {
    e.k[0][c] = (const T *)a[c];
    if (b) e.k[1][c] = (const T *)b[c];
  }
  e.N = N;
  e.res = (unsigned long long *)r.red.p;
  launch_err(e, s);
  return read_max_slot(s);
}
template <typename T>
static double sum_wdot(const void *a, const void *b, long N) {
  hipStream_t s = lib_stream();
  RedScratch &r = red_scratch();
  const T *aa[1] = {(const T *)a};
  const T *bb[1] = {(const T *)b};
  launch_wdot<T>(aa, b ? bb : nullptr, 1, nullptr, T(1), N, r.partial.as<double>(), r.sum.as<double>(), s);
  UM_HIP(hipMemcpyAsync(r.hs, r.sum.p, sizeof(double), hipMemcpyDeviceToHost, s));
  UM_HIP(hipStreamSynchronize(s));
  return *r.hs;
}

}  // namespace um

using namespace um;

extern "C" {

int um_init(int device) {
  return guard([&] {
    {
      std::lock_guard<std::mutex> lk(g_mu);
      UM_REQUIRE(!g_inited || device == g_device, UM_ESTATE, "um_init: already initialised on another device");
      g_device = device;
    }
    ensure_init();
  });
}
const char *um_last_error(void) { return g_err.c_str(); }
int um_device_info(char *name, int name_len, int *cu_count, size_t *mem_bytes) {
  return guard([&] {
    ensure_init();
    hipDeviceProp_t p;
    UM_HIP(hipGetDeviceProperties(&p, g_device));
    if (name && name_len > 0) {
      std::string s = std::string(p.name) + " (" + p.gcnArchName + ")";
      strncpy(name, s.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (mem_bytes) *mem_bytes = p.totalGlobalMem;
  });
}
int um_set_stream(void *s) {
  return guard([&] {
    ensure_init();
    g_stream = (hipStream_t)s;  // NULL selects the HIP null stream
  });
}
int um_sync(void) { return guard([&] { UM_HIP(hipStreamSynchronize(lib_stream())); }); }
int um_malloc(void **ptr, size_t bytes) {
  return guard([&] {
    ensure_init();
    UM_REQUIRE(ptr, UM_EINVAL, "um_malloc: NULL out pointer");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) throw Error(UM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  });
}
int um_free(void *ptr) { return guard([&] { if (ptr) UM_HIP(hipFree(ptr)); }); }
int um_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  return guard([&] {
    UM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, lib_stream()));
    UM_HIP(hipStreamSynchronize(lib_stream()));
  });
}
int um_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  return guard([&] {
    UM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, lib_stream()));
    UM_HIP(hipStreamSynchronize(lib_stream()));
  });
}
int um_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  return guard([&] { UM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, lib_stream())); });
}
int um_zero(void *dst, size_t bytes) { return guard([&] { fill_bytes(dst, bytes, lib_stream()); }); }
int um_memset(int prec, void *dst, double value, size_t n) {
  return guard([&] {
    check_prec(prec);
    if (prec == UM_F32) fill_value<float>((float *)dst, (float)value, n, lib_stream());
    else fill_value<double>((double *)dst, value, n, lib_stream());
  });
}

#define DISPATCH(prec, CALL_F, CALL_D) \
  do {                                 \
    check_prec(prec);                  \
    if ((prec) == UM_F32) { CALL_F; }  \
    else { CALL_D; }                   \
  } while (0)

int um_add_exchange(int prec, void *const B[3], const void *const m[3], const void *Ms, double Ms_mul,
                    const void *aLUT, const uint8_t *regions, double wx, double wy, double wz, int Nx, int Ny,
                    int Nz, int pbc) {
  return guard([&] {
    check3(B, "B"); check3(m, "m");
    UM_REQUIRE(aLUT && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_add_exchange: bad arguments");
    Mesh g{Nx, Ny, Nz, pbc};
    DISPATCH(prec,
             add_exchange<float>(B, m, par<float>(Ms, Ms_mul), (const float *)aLUT, regions, (float)wx, (float)wy,
                                 (float)wz, g, lib_stream()),
             add_exchange<double>(B, m, par<double>(Ms, Ms_mul), (const double *)aLUT, regions, wx, wy, wz, g,
                                  lib_stream()));
  });
}
int um_add_dmi(int prec, void *const B[3], const void *const m[3], const void *Ms, double Ms_mul, const void *aLUT,
               const void *dLUT, const uint8_t *regions, double cx, double cy, double cz, int Nx, int Ny, int Nz,
               int pbc, int openbc) {
  return guard([&] {
    check3(B, "B"); check3(m, "m");
    UM_REQUIRE(aLUT && dLUT && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_add_dmi: bad arguments");
    Mesh g{Nx, Ny, Nz, pbc};
    DISPATCH(prec,
             add_dmi<float>(B, m, par<float>(Ms, Ms_mul), (const float *)aLUT, (const float *)dLUT, regions,
                            (float)cx, (float)cy, (float)cz, g, openbc, lib_stream()),
             add_dmi<double>(B, m, par<double>(Ms, Ms_mul), (const double *)aLUT, (const double *)dLUT, regions,
                             cx, cy, cz, g, openbc, lib_stream()));
  });
}
int um_add_dmi_bulk(int prec, void *const B[3], const void *const m[3], const void *Ms, double Ms_mul,
                    const void *aLUT, const void *dLUT, const uint8_t *regions, double cx, double cy, double cz,
                    int Nx, int Ny, int Nz, int pbc, int openbc) {
  return guard([&] {
    check3(B, "B"); check3(m, "m");
    UM_REQUIRE(aLUT && dLUT && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_add_dmi_bulk: bad arguments");
    Mesh g{Nx, Ny, Nz, pbc};
    DISPATCH(prec,
             add_dmi_bulk<float>(B, m, par<float>(Ms, Ms_mul), (const float *)aLUT, (const float *)dLUT, regions,
                                 (float)cx, (float)cy, (float)cz, g, openbc, lib_stream()),
             add_dmi_bulk<double>(B, m, par<double>(Ms, Ms_mul), (const double *)aLUT, (const double *)dLUT,
                                  regions, cx, cy, cz, g, openbc, lib_stream()));
  });
}
int um_add_zhangli(int prec, void *const tau[3], const void *const m[3], const void *Ms, double Ms_mul,
                   const void *const J[3], const double J_mul[3], const void *alpha, double alpha_mul,
                   const void *xi, double xi_mul, const void *pol, double pol_mul, double cx, double cy,
                   double cz, int Nx, int Ny, int Nz, int pbc) {
  return guard([&] {
    check3(tau, "tau"); check3(m, "m");
    UM_REQUIRE(J && J_mul && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_add_zhangli: bad arguments");
    Mesh g{Nx, Ny, Nz, pbc};
    auto go = [&](auto zero) {
      using T = decltype(zero);
      SttPar<T> p{};
      for (int c = 0; c < 3; c++) p.J[c] = par<T>(J[c], J_mul[c]);
      p.xi = par<T>(xi, xi_mul);
      p.pol = par<T>(pol, pol_mul);
      add_stt<T>(tau, m, par<T>(Ms, Ms_mul), par<T>(alpha, alpha_mul), p, (T)cx, (T)cy, (T)cz, g, 1, 0,
                 lib_stream());
    };
    DISPATCH(prec, go(0.0f), go(0.0));
  });
}
int um_add_slonczewski(int prec, void *const tau[3], const void *const m[3], const void *Ms, double Ms_mul,
                       const void *Jz, double Jz_mul, const void *const fixedP[3], const double fixedP_mul[3],
                       const void *alpha, double alpha_mul, const void *pol, double pol_mul, const void *lambda,
                       double lambda_mul, const void *epsPrime, double epsPrime_mul, const void *thickness,
                       double thickness_mul, double mesh_thickness, double flp, int N) {
  return guard([&] {
    check3(tau, "tau"); check3(m, "m");
    UM_REQUIRE(fixedP && fixedP_mul && N > 0, UM_EINVAL, "um_add_slonczewski: bad arguments");
    Mesh g{N, 1, 1, 0};
    auto go = [&](auto zero) {
      using T = decltype(zero);
      SttPar<T> p{};
      p.J[2] = par<T>(Jz, Jz_mul);
      for (int c = 0; c < 3; c++) p.fixp[c] = par<T>(fixedP[c], fixedP_mul[c]);
      p.pol = par<T>(pol, pol_mul);
      p.lambda = par<T>(lambda, lambda_mul);
      p.epsp = par<T>(epsPrime, epsPrime_mul);
      p.thick = par<T>(thickness, thickness_mul);
      p.mesh_thick = (T)mesh_thickness;
      p.flp = (T)flp;
      add_stt<T>(tau, m, par<T>(Ms, Ms_mul), par<T>(alpha, alpha_mul), p, T(1), T(1), T(1), g, 0, 1,
                 lib_stream());
    };
    DISPATCH(prec, go(0.0f), go(0.0));
  });
}
int um_add_uniaxial2(int prec, void *const B[3], const void *const m[3], const void *Ms, double Ms_mul,
                     const void *K1, double K1_mul, const void *K2, double K2_mul, const void *const u[3],
                     const double u_mul[3], int N) {
  return guard([&] {
    check3(B, "B"); check3(m, "m");
    UM_REQUIRE(u && u_mul && N >= 0, UM_EINVAL, "um_add_uniaxial2: bad arguments");
    auto go = [&](auto zero) {
      using T = decltype(zero);
      Par<T> up[3];
      for (int c = 0; c < 3; c++) up[c] = par<T>(u[c], u_mul[c]);
      add_uniaxial2<T>(B, m, par<T>(Ms, Ms_mul), par<T>(K1, K1_mul), par<T>(K2, K2_mul), up, N, lib_stream());
    };
    DISPATCH(prec, go(0.0f), go(0.0));
  });
}
int um_add_cubic2(int prec, void *const B[3], const void *const m[3], const void *Ms, double Ms_mul, const void *K1,
                  double K1_mul, const void *K2, double K2_mul, const void *K3, double K3_mul,
                  const void *const c1[3], const double c1_mul[3], const void *const c2[3], const double c2_mul[3],
                  int N) {
  return guard([&] {
    check3(B, "B"); check3(m, "m");
    UM_REQUIRE(c1 && c2 && c1_mul && c2_mul && N >= 0, UM_EINVAL, "um_add_cubic2: bad arguments");
    auto go = [&](auto zero) {
      using T = decltype(zero);
      Par<T> a[3], b[3];
      for (int c = 0; c < 3; c++) { a[c] = par<T>(c1[c], c1_mul[c]); b[c] = par<T>(c2[c], c2_mul[c]); }
      add_cubic2<T>(B, m, par<T>(Ms, Ms_mul), par<T>(K1, K1_mul), par<T>(K2, K2_mul), par<T>(K3, K3_mul), a, b, N,
                    lib_stream());
    };
    DISPATCH(prec, go(0.0f), go(0.0));
  });
}
int um_region_addv(int prec, void *const dst[3], const void *const lut[3], const uint8_t *regions, int N) {
  return guard([&] {
    check3(dst, "dst"); check3(lut, "lut");
    DISPATCH(prec, region_addv<float>(dst, lut, regions, N, lib_stream()),
             region_addv<double>(dst, lut, regions, N, lib_stream()));
  });
}
int um_zeromask(int prec, void *dst, const void *maskLUT, const uint8_t *regions, int N) {
  return guard([&] {
    UM_REQUIRE(dst && maskLUT, UM_EINVAL, "um_zeromask: NULL argument");
    DISPATCH(prec, zeromask<float>(dst, maskLUT, regions, N, lib_stream()),
             zeromask<double>(dst, maskLUT, regions, N, lib_stream()));
  });
}
int um_lltorque(int prec, void *const t[3], const void *const m[3], const void *const B[3], const void *alpha,
                double alpha_mul, int N) {
  return guard([&] {
    check3(t, "t"); check3(m, "m"); check3(B, "B");
    DISPATCH(prec, lltorque<float>(t, m, B, par<float>(alpha, alpha_mul), N, lib_stream()),
             lltorque<double>(t, m, B, par<double>(alpha, alpha_mul), N, lib_stream()));
  });
}
int um_llnoprecess(int prec, void *const t[3], const void *const m[3], const void *const B[3], int N) {
  return guard([&] {
    check3(t, "t"); check3(m, "m"); check3(B, "B");
    DISPATCH(prec, llnoprecess<float>(t, m, B, N, lib_stream()), llnoprecess<double>(t, m, B, N, lib_stream()));
  });
}
int um_thermal_field(int prec, void *const B[3], uint64_t seed, uint64_t step, uint32_t stream, double k,
                     const void *Ms, double Ms_mul, const void *temp, double temp_mul, const void *alpha,
                     double alpha_mul, int N) {
  return guard([&] {
    check3(B, "B");
    DISPATCH(prec,
             thermal_field<float>(B, seed, step, stream, (float)k, par<float>(Ms, Ms_mul), par<float>(temp, temp_mul),
                                  par<float>(alpha, alpha_mul), N, lib_stream()),
             thermal_field<double>(B, seed, step, stream, k, par<double>(Ms, Ms_mul), par<double>(temp, temp_mul),
                                   par<double>(alpha, alpha_mul), N, lib_stream()));
  });
}

int um_madd(int prec, void *dst, int nsrc, const void *const *src, const double *fac, int N) {
  return guard([&] {
    UM_REQUIRE(dst && src && fac && nsrc >= 1 && nsrc <= 8, UM_EINVAL, "um_madd: bad arguments");
    for (int j = 0; j < nsrc; j++) UM_REQUIRE(src[j], UM_EINVAL, "um_madd: NULL source");
    auto go = [&](auto zero) {
      using T = decltype(zero);
      StageArgs<T> a{};
      a.nsrc = nsrc;
      a.ncomp = 1;
      a.group3 = 1;
      a.normalize = 0;
      for (int j = 0; j < nsrc; j++) {
        a.src[j][0] = (const T *)src[j];
        a.fac[j] = (T)fac[j];
      }
      a.dst[0] = (T *)dst;
      a.vol = nullptr;
      a.N = N;
      launch_stage(a, lib_stream());
    };
    DISPATCH(prec, go(0.0f), go(0.0));
  });
}
int um_normalize(int prec, void *const v[3], const void *vol, int N) {
  return guard([&] {
    check3(v, "v");
    auto go = [&](auto zero) {
      using T = decltype(zero);
      StageArgs<T> a{};
      a.nsrc = 1;
      a.ncomp = 3;
      a.normalize = 1;
      for (int c = 0; c < 3; c++) { a.src[0][c] = (const T *)v[c]; a.dst[c] = (T *)v[c]; }
      a.fac[0] = T(1);
      a.vol = (const T *)vol;
      a.N = N;
      launch_stage(a, lib_stream());
    };
    DISPATCH(prec, go(0.0f), go(0.0));
  });
}
int um_vecnorm(int prec, void *dst, const void *const a[3], int N) {
  return guard([&] {
    check3(a, "a");
    DISPATCH(prec, launch_vecnorm<float>((float *)dst, (const float *const *)a, N, lib_stream()),
             launch_vecnorm<double>((double *)dst, (const double *const *)a, N, lib_stream()));
  });
}
int um_div(int prec, void *dst, const void *a, const void *b, int N) {
  return guard([&] {
    DISPATCH(prec, launch_div<float>((float *)dst, (const float *)a, (const float *)b, N, lib_stream()),
             launch_div<double>((double *)dst, (const double *)a, (const double *)b, N, lib_stream()));
  });
}
int um_mul(int prec, void *dst, const void *a, const void *b, int N) {
  return guard([&] {
    DISPATCH(prec, launch_mul<float>((float *)dst, (const float *)a, (const float *)b, N, lib_stream()),
             launch_mul<double>((double *)dst, (const double *)a, (const double *)b, N, lib_stream()));
  });
}

// ---- fused evaluation, RK stage and step-controller norms -------------------------------
extern "C++" {
template <typename T>
static Par<T> upar(const um_param &p) { return Par<T>{(const T *)p.arr, (T)p.mul}; }
static bool unz(const um_param &p) { return p.arr != nullptr || p.mul != 0; }
static bool all_or_none(const void *const p[3]) { return (p[0] && p[1] && p[2]) || (!p[0] && !p[1] && !p[2]); }

template <typename T>
static FieldArgs<T> field_args_from(const um_field_args &u) {
  FieldArgs<T> a{};
  a.g = Mesh{u.Nx, u.Ny, u.Nz, u.pbc & 7};
  const double *cs = u.cellsize;
  // float32(2/c^2) in the FP32 build (opencl/exchange.go:17-20), cell sizes as T (dmi.go)
  a.wx = (T)(2 / (cs[0] * cs[0]));
  a.wy = (T)(2 / (cs[1] * cs[1]));
  a.wz = (T)(2 / (cs[2] * cs[2]));
  a.cx = (T)cs[0]; a.cy = (T)cs[1]; a.cz = (T)cs[2];
  for (int c = 0; c < 3; c++) {
    a.m[c] = (const T *)u.m[c];
    a.bd[c] = (const T *)u.bdemag[c];
    a.tau[c] = (T *)u.torque[c];
    a.beff[c] = (T *)u.beff[c];
    a.u[c] = upar<T>(u.anisU[c]);
    a.c1[c] = upar<T>(u.anisC1[c]);
    a.c2[c] = upar<T>(u.anisC2[c]);
    a.bext[c] = upar<T>(u.Bext[c]);
    a.bextLUT[c] = (const T *)u.bext_lut[c];
    a.badd[c] = (const T *)u.badd[c];
    a.tadd[c] = (const T *)u.tadd[c];
    a.stt.J[c] = upar<T>(u.J[c]);
    a.stt.fixp[c] = upar<T>(u.fixed_layer[c]);
  }
  a.regions = u.regions;
  a.Ms = upar<T>(u.Msat);
  a.aLUT = (const T *)u.aex_lut;
  a.dLUT = (const T *)u.dmi_lut;
  a.exch_mode = !u.aex_lut ? 0 : !u.dmi_lut ? 1 : u.dmi_bulk ? 3 : 2;
  a.openbc = u.openbc;
  a.has_uniax = unz(u.Ku1) || unz(u.Ku2);
  a.Ku1 = upar<T>(u.Ku1); a.Ku2 = upar<T>(u.Ku2);
  a.has_cubic = unz(u.Kc1) || unz(u.Kc2) || unz(u.Kc3);
  a.Kc1 = upar<T>(u.Kc1); a.Kc2 = upar<T>(u.Kc2); a.Kc3 = upar<T>(u.Kc3);
  a.has_bext = unz(u.Bext[0]) || unz(u.Bext[1]) || unz(u.Bext[2]);
  a.has_therm = u.thermal != 0;
  a.temp = upar<T>(u.Temp);
  a.kth = (T)u.thermal_k;
  a.seed = u.seed;
  a.step = u.step;
  a.step_dev = nullptr;
  a.stream = u.stream;
  a.alpha = upar<T>(u.alpha);
  a.precess = u.precess;
  a.frozenLUT = (const T *)u.frozen_lut;
  const bool has_j = unz(u.J[0]) || unz(u.J[1]) || unz(u.J[2]);
  a.has_zl = has_j && u.zhangli;
  a.has_sl = has_j && u.slonczewski && (unz(u.fixed_layer[0]) || unz(u.fixed_layer[1]) || unz(u.fixed_layer[2]));
  a.stt.xi = upar<T>(u.xi);
  a.stt.pol = upar<T>(u.pol);
  a.stt.lambda = upar<T>(u.lambda);
  a.stt.epsp = upar<T>(u.eps_prime);
  a.stt.thick = upar<T>(u.free_layer_thickness);
  a.stt.mesh_thick = (T)(u.Nz * cs[2]);
  a.stt.flp = (T)u.fixed_layer_top;
  return a;
}
}  // extern "C++"

int um_field_torque(int prec, const um_field_args *u) {
  return guard([&] {
    UM_REQUIRE(u, UM_EINVAL, "um_field_torque: NULL arguments");
    UM_REQUIRE(u->Nx > 0 && u->Ny > 0 && u->Nz > 0 && (long)u->Nx * u->Ny * u->Nz < (1L << 31), UM_EINVAL,
               "um_field_torque: grid size");
    for (int c = 0; c < 3; c++)
      UM_REQUIRE(std::isfinite(u->cellsize[c]) && u->cellsize[c] > 0, UM_EINVAL, "um_field_torque: cell size");
    check3(u->m, "um_field_torque: m");
    UM_REQUIRE(all_or_none(u->bdemag) && all_or_none(u->torque) && all_or_none(u->beff) && all_or_none(u->badd) &&
                   all_or_none(u->tadd) && all_or_none(u->bext_lut),
               UM_EINVAL, "um_field_torque: a vector must have all three components or none");
    UM_REQUIRE(u->torque[0] || u->beff[0], UM_EINVAL, "um_field_torque: neither torque nor beff requested");
    UM_REQUIRE(!u->dmi_lut || u->aex_lut, UM_EINVAL, "um_field_torque: DMI needs the exchange LUT");
    const bool has_j = unz(u->J[0]) || unz(u->J[1]) || unz(u->J[2]);
    UM_REQUIRE(!has_j || !(u->zhangli || u->slonczewski) || unz(u->pol), UM_EINVAL,
               "spin polarization should not be 0");
    DISPATCH(prec, launch_field_torque(field_args_from<float>(*u), lib_stream()),
             launch_field_torque(field_args_from<double>(*u), lib_stream()));
  });
}

int um_rk45dp_stage(int prec, void *const dst[3], int nsrc, const void *const *src, const double *fac,
                    const void *vol, int normalize, int N) {
  return guard([&] {
    check3((const void *const *)dst, "um_rk45dp_stage: dst");
    UM_REQUIRE(src && fac && nsrc >= 1 && nsrc <= 8 && N >= 0, UM_EINVAL, "um_rk45dp_stage: bad arguments");
    for (int j = 0; j < 3 * nsrc; j++) UM_REQUIRE(src[j], UM_EINVAL, "um_rk45dp_stage: NULL source component");
    auto go = [&](auto zero) {
      using T = decltype(zero);
      StageArgs<T> a{};
      a.nsrc = nsrc;
      a.ncomp = 3;
      a.group3 = 1;
      a.normalize = normalize != 0;
      for (int j = 0; j < nsrc; j++) {
        for (int c = 0; c < 3; c++) a.src[j][c] = (const T *)src[3 * j + c];
        a.fac[j] = (T)fac[j];
      }
      for (int c = 0; c < 3; c++) a.dst[c] = (T *)dst[c];
      a.vol = (const T *)vol;
      a.N = N;
      launch_stage(a, lib_stream());
    };
    DISPATCH(prec, go(0.0f), go(0.0));
  });
}


// Err = sum e_j k_j with delM = k[nk-1], as the fused six-source error kernel:
// shorter lists repeat delM with zero coefficients
extern "C++" {
template <typename T>
static void rk_err_args(ErrArgs<T> &a, int nk, const void *const *k, const double *e, long N, unsigned long long *res) {
  a.mode = 0;
  for (int j = 0; j < 6; j++) {
    const int s = j < nk - 1 ? j : (j == 5 ? nk - 1 : -1);
    for (int c = 0; c < 3; c++) a.k[j][c] = (const T *)k[3 * (s < 0 ? nk - 1 : s) + c];
    a.e[j] = s < 0 ? T(0) : (T)e[s];
  }
  a.N = N;
  a.res = res;
}
static void check_rk_err(int nk, const void *const *k, const double *e, double h, int N) {
  UM_REQUIRE(k && e && nk >= 1 && nk <= 6 && N >= 0 && std::isfinite(h), UM_EINVAL, "um_rk45dp_error: bad arguments");
  for (int j = 0; j < 3 * nk; j++) UM_REQUIRE(k[j], UM_EINVAL, "um_rk45dp_error: NULL stage component");
}
}  // extern "C++"

int um_rk45dp_error(int prec, int nk, const void *const *k, const double *e, double h, int N, double out[3]) {
  return guard([&] {
    check_rk_err(nk, k, e, h, N);
    UM_REQUIRE(out, UM_EINVAL, "um_rk45dp_error: NULL out");
    hipStream_t s = lib_stream();
    RedScratch &r = red_scratch();
    fill_bytes(r.red.p, r.red.bytes, s);
    auto go = [&](auto zero) {
      using T = decltype(zero);
      ErrArgs<T> a{};
      rk_err_args<T>(a, nk, k, e, N, (unsigned long long *)r.red.p);
      launch_err(a, s);
    };
    DISPATCH(prec, go(0.0f), go(0.0));
    UM_HIP(hipMemcpyAsync(r.h, r.red.p, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    UM_HIP(hipStreamSynchronize(s));
    double v[3];
    for (int i = 0; i < 3; i++) memcpy(&v[i], &r.h[i], 8);
    out[0] = std::sqrt(v[0]) * h;
    out[1] = v[2];
    out[2] = std::sqrt(v[1]);
  });
}

extern "C++" {
// (err, rlerr, maxdm) from the max slots, as um_rk45dp_error does on the host
__global__ void k_rk_err_final(const unsigned long long *res, double h, double *out) {
  if (threadIdx.x == 0) {
    out[0] = __builtin_sqrt(__longlong_as_double((long long)res[0])) * h;
    out[1] = __longlong_as_double((long long)res[2]);
    out[2] = __builtin_sqrt(__longlong_as_double((long long)res[1]));
  }
}
}  // extern "C++"

int um_rk45dp_error_noalloc(int prec, int nk, const void *const *k, const double *e, double h, int N,
                            double *out_dev) {
  return guard([&] {
    check_rk_err(nk, k, e, h, N);
    UM_REQUIRE(out_dev, UM_EINVAL, "um_rk45dp_error_noalloc: NULL out_dev");
    bool inited;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      inited = g_inited;
    }
    // the scratch address is looked up by um_init; nothing is created here
    UM_REQUIRE(inited, UM_ESTATE, "um_rk45dp_error_noalloc: call um_init first");
    unsigned long long *slots = g_scratch_dev[g_device];
    UM_REQUIRE(slots, UM_ESTATE, "um_rk45dp_error_noalloc: no scratch address for this device");
    hipStream_t s = g_stream;
    fill_bytes(slots, NOALLOC_SLOTS * sizeof(unsigned long long), s);   // a kernel, not a memset node
    auto go = [&](auto zero) {
      using T = decltype(zero);
      ErrArgs<T> a{};
      rk_err_args<T>(a, nk, k, e, N, slots);
      launch_err(a, s);
    };
    DISPATCH(prec, go(0.0f), go(0.0));
    UM_LAUNCH(k_rk_err_final, dim3(1), dim3(64), 0, s, (const unsigned long long *)slots, h, out_dev);
    UM_HIP(hipGetLastError());
  });
}

// ---- reductions ---------------------------------------------------------------------
int um_reduce_sum(int prec, const void *x, int N, double *out) {
  return guard([&] {
    UM_REQUIRE(x && out, UM_EINVAL, "um_reduce_sum: NULL argument");
    DISPATCH(prec, *out = sum_wdot<float>(x, nullptr, N), *out = sum_wdot<double>(x, nullptr, N));
  });
}
int um_reduce_dot(int prec, const void *a, const void *b, int N, double *out) {
  return guard([&] {
    UM_REQUIRE(a && b && out, UM_EINVAL, "um_reduce_dot: NULL argument");
    DISPATCH(prec, *out = sum_wdot<float>(a, b, N), *out = sum_wdot<double>(a, b, N));
  });
}
int um_reduce_maxabs(int prec, const void *x, int N, double *out) {
  return guard([&] {
    UM_REQUIRE(x && out, UM_EINVAL, "um_reduce_maxabs: NULL argument");
    hipStream_t s = lib_stream();
    RedScratch &r = red_scratch();
    fill_bytes(r.red.p, r.red.bytes, s);
    DISPATCH(prec, launch_maxabs<float>((const float *)x, N, (unsigned long long *)r.red.p, s),
             launch_maxabs<double>((const double *)x, N, (unsigned long long *)r.red.p, s));
    *out = read_max_slot(s);
  });
}
int um_reduce_maxvecnorm(int prec, const void *const v[3], int N, double *out) {
  return guard([&] {
    check3(v, "v");
    UM_REQUIRE(out, UM_EINVAL, "NULL out");
    DISPATCH(prec, *out = std::sqrt(err_max<float>(2, v, nullptr, N)),
             *out = std::sqrt(err_max<double>(2, v, nullptr, N)));
  });
}
int um_reduce_maxvecdiff(int prec, const void *const a[3], const void *const b[3], int N, int sqrt3_quirk,
                         double *out) {
  return guard([&] {
    check3(a, "a"); check3(b, "b");
    UM_REQUIRE(out, UM_EINVAL, "NULL out");
    double m2 = 0;
    DISPATCH(prec, m2 = err_max<float>(1, a, b, N), m2 = err_max<double>(1, a, b, N));
    *out = std::sqrt(sqrt3_quirk ? 3.0 * m2 : m2);
  });
}

// ---- demag ----------------------------------------------------------------------------
int um_demag_create(int prec, int Nx, int Ny, int Nz, const int pbc[3], const double cs[3], double accuracy,
                    void **plan) {
  return guard([&] {
    check_prec(prec);
    UM_REQUIRE(plan && cs && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_demag_create: bad arguments");
    const int zero[3] = {0, 0, 0};
    const int *pb = pbc ? pbc : zero;
    const int N[3] = {Nx, Ny, Nz};
    int P[3];
    pad_size(N, pb, P);
    hipStream_t s = lib_stream();
    const size_t np = (size_t)P[0] * P[1] * P[2];
    DevBuf K(sizeof(double) * 6 * np);
    double *Kd[6];
    for (int c = 0; c < 6; c++) Kd[c] = K.as<double>() + c * np;
    demag_tensor(N, pb, cs, 0, accuracy, 0, 0, Kd, P, s);
    *plan = demag_create(prec, N, P, Kd, s);
  });
}
int um_demag_create_newell(int prec, int Nx, int Ny, int Nz, const int pbc[3], const double cs[3],
                           int asymptotic_radius, int zero_self_demag, void **plan) {
  return guard([&] {
    check_prec(prec);
    UM_REQUIRE(plan && cs && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_demag_create_newell: bad arguments");
    const int zero[3] = {0, 0, 0};
    const int *pb = pbc ? pbc : zero;
    const int N[3] = {Nx, Ny, Nz};
    int P[3];
    pad_size(N, pb, P);
    hipStream_t s = lib_stream();
    const size_t np = (size_t)P[0] * P[1] * P[2];
    DevBuf K(sizeof(double) * 6 * np);
    double *Kd[6];
    for (int c = 0; c < 6; c++) Kd[c] = K.as<double>() + c * np;
    demag_tensor(N, pb, cs, 1, 0.0, asymptotic_radius, zero_self_demag, Kd, P, s);
    *plan = demag_create(prec, N, P, Kd, s);
  });
}
int um_set_autotune(int mode, const char *file) {
  return guard([&] {
    UM_REQUIRE(mode >= 0 && mode <= 2, UM_EINVAL, "um_set_autotune: mode must be 0, 1 or 2");
    set_autotune(mode, file);
  });
}
int um_demag_schedule(void *plan, int out[8]) {
  return guard([&] {
    UM_REQUIRE(plan && out, UM_EINVAL, "um_demag_schedule: bad arguments");
    demag_schedule((DemagBase *)plan, out);
  });
}
int um_trace_kernels(int on) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(g_trace_mu);
    g_trace.clear();
    g_trace_on = on != 0;
  });
}
int um_traced_kernels(char *out, size_t out_len, size_t *needed) {
  return guard([&] {
    std::string all;
    {
      std::lock_guard<std::mutex> lk(g_trace_mu);
      for (const auto &n : g_trace) { all += n; all += '\n'; }
    }
    if (needed) *needed = all.size() + 1;
    if (out && out_len > 0) {
      const size_t k = std::min(out_len - 1, all.size());
      memcpy(out, all.data(), k);
      out[k] = 0;
    }
  });
}
int um_set_kernel_cache(const char *dir) { return guard([&] { set_kernel_cache(dir); }); }
int um_kernel_cache_basename(const char *dir, const int N[3], const int pbc[3], const double cs[3], int newell,
                             double accuracy, int zero_self_demag, char *out, int out_len) {
  return guard([&] {
    UM_REQUIRE(dir && N && pbc && cs && out && out_len > 0, UM_EINVAL, "um_kernel_cache_basename: bad arguments");
    const std::string b = kernel_cache_basename(dir, N, pbc, cs, newell, accuracy, zero_self_demag);
    UM_REQUIRE((int)b.size() < out_len, UM_EINVAL, "um_kernel_cache_basename: buffer too small");
    std::memcpy(out, b.c_str(), b.size() + 1);
  });
}
int um_load_kernel(const char *fname, const int P[3], double *K) {
  return guard([&] {
    UM_REQUIRE(fname && P && K, UM_EINVAL, "um_load_kernel: bad arguments");
    int sz[3], nc;
    std::vector<double> v;
    ovf_read(fname, sz, &nc, v);
    UM_REQUIRE(nc == 1 && sz[0] == P[0] && sz[1] == P[1] && sz[2] == P[2], UM_EINVAL,
               "um_load_kernel: not a scalar field of the requested size");
    std::memcpy(K, v.data(), sizeof(double) * v.size());
  });
}
int um_save_kernel(const char *fname, const int P[3], const double cs[3], const char *name, const double *K) {
  return guard([&] {
    UM_REQUIRE(fname && P && cs && name && K && P[0] > 0 && P[1] > 0 && P[2] > 0, UM_EINVAL,
               "um_save_kernel: bad arguments");
    ovf_write_scalar(fname, P, cs, name, K);
  });
}
int um_demag_kernel_newell(int Nx, int Ny, int Nz, const int pbc[3], const double cs[3], int asymptotic_radius,
                           int zero_self_demag, double *const K[6]) {
  return guard([&] {
    UM_REQUIRE(cs && K && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_demag_kernel_newell: bad arguments");
    const int zero[3] = {0, 0, 0};
    const int *pb = pbc ? pbc : zero;
    const int N[3] = {Nx, Ny, Nz};
    int P[3];
    pad_size(N, pb, P);
    hipStream_t s = lib_stream();
    const size_t np = (size_t)P[0] * P[1] * P[2];
    DevBuf Kb(sizeof(double) * 6 * np);
    double *Kd[6];
    for (int c = 0; c < 6; c++) Kd[c] = Kb.as<double>() + c * np;
    demag_kernel_newell_device(N, pb, cs, asymptotic_radius, zero_self_demag, Kd, P, s);
    for (int c = 0; c < 6; c++) UM_HIP(hipMemcpyAsync(K[c], Kd[c], sizeof(double) * np, hipMemcpyDeviceToHost, s));
    UM_HIP(hipStreamSynchronize(s));
  });
}
int um_demag_create_from_kernel(int prec, int Nx, int Ny, int Nz, const int P[3], const double *const K[6],
                                void **plan) {
  return guard([&] {
    check_prec(prec);
    UM_REQUIRE(plan && P && K && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_demag_create_from_kernel: bad arguments");
    UM_REQUIRE(P[0] >= Nx && P[1] >= Ny && P[2] >= Nz, UM_EINVAL, "padded size smaller than the grid");
    hipStream_t s = lib_stream();
    const size_t np = (size_t)P[0] * P[1] * P[2];
    DevBuf Kb(sizeof(double) * 6 * np);
    double *Kd[6];
    for (int c = 0; c < 6; c++) {
      Kd[c] = Kb.as<double>() + c * np;
      UM_REQUIRE(K[c], UM_EINVAL, "NULL tensor component");
      UM_HIP(hipMemcpyAsync(Kd[c], K[c], sizeof(double) * np, hipMemcpyHostToDevice, s));
    }
    const int N[3] = {Nx, Ny, Nz};
    *plan = demag_create(prec, N, P, Kd, s);
  });
}
int um_demag_exec(void *plan, void *const B[3], const void *const m[3], const void *vol, const void *Ms,
                  double Ms_mul) {
  return guard([&] {
    UM_REQUIRE(plan, UM_EINVAL, "NULL plan");
    check3(B, "B"); check3(m, "m");
    demag_exec((DemagBase *)plan, B, m, vol, Ms, Ms_mul, lib_stream());
  });
}
int um_demag_destroy(void *plan) {
  return guard([&] {
    if (plan) {
      UM_HIP(hipStreamSynchronize(lib_stream()));
      demag_destroy((DemagBase *)plan);
    }
  });
}
int um_demag_info(void *plan, int P[3]) {
  return guard([&] {
    UM_REQUIRE(plan && P, UM_EINVAL, "NULL argument");
    int prec;
    demag_info((DemagBase *)plan, P, &prec);
  });
}
int um_demag_kspec(void *plan, int c, void *host_out) {
  return guard([&] {
    UM_REQUIRE(plan && host_out, UM_EINVAL, "NULL argument");
    UM_HIP(hipStreamSynchronize(lib_stream()));
    demag_kspec((DemagBase *)plan, c, host_out);
  });
}
int um_demag_kernel_bf(int Nx, int Ny, int Nz, const int pbc[3], const double cs[3], double accuracy,
                       double *const K[6]) {
  return guard([&] {
    UM_REQUIRE(cs && K && Nx > 0 && Ny > 0 && Nz > 0, UM_EINVAL, "um_demag_kernel_bf: bad arguments");
    const int zero[3] = {0, 0, 0};
    const int *pb = pbc ? pbc : zero;
    const int N[3] = {Nx, Ny, Nz};
    int P[3];
    pad_size(N, pb, P);
    hipStream_t s = lib_stream();
    const size_t np = (size_t)P[0] * P[1] * P[2];
    DevBuf Kb(sizeof(double) * 6 * np);
    double *Kd[6];
    for (int c = 0; c < 6; c++) Kd[c] = Kb.as<double>() + c * np;
    demag_kernel_bf_device(N, pb, cs, accuracy, Kd, P, s);
    for (int c = 0; c < 6; c++) UM_HIP(hipMemcpyAsync(K[c], Kd[c], sizeof(double) * np, hipMemcpyDeviceToHost, s));
    UM_HIP(hipStreamSynchronize(s));
  });
}

// ---- simulation handle ---------------------------------------------------------------
#define SIM(h) ((SimBase *)(h))
#define SIMCHK(h) UM_REQUIRE((h) != nullptr, UM_EINVAL, "NULL simulation handle")

int um_sim_create(int prec, int Nx, int Ny, int Nz, double cx, double cy, double cz, int pbcx, int pbcy, int pbcz,
                  void **sim) {
  return guard([&] {
    check_prec(prec);
    UM_REQUIRE(sim, UM_EINVAL, "NULL out pointer");
    const int N[3] = {Nx, Ny, Nz};
    const double cs[3] = {cx, cy, cz};
    const int pbc[3] = {pbcx, pbcy, pbcz};
    *sim = sim_create(prec, N, cs, pbc);
  });
}
int um_sim_destroy(void *sim) { return guard([&] { delete SIM(sim); }); }
int um_sim_set_scalar(void *sim, int id, const void *arr, double mul) {
  return guard([&] { SIMCHK(sim); SIM(sim)->set_scalar(id, arr, mul); });
}
int um_sim_set_vector(void *sim, int id, const void *ax, const void *ay, const void *az, double mx, double my,
                      double mz) {
  return guard([&] {
    SIMCHK(sim);
    const void *a[3] = {ax, ay, az};
    const double m[3] = {mx, my, mz};
    SIM(sim)->set_vector(id, a, m);
  });
}
int um_sim_set_regions(void *sim, const uint8_t *r) { return guard([&] { SIMCHK(sim); SIM(sim)->set_regions(r); }); }
int um_sim_set_exchange_luts(void *sim, const void *a, const void *d) {
  return guard([&] {
    SIMCHK(sim);
    UM_REQUIRE(a, UM_EINVAL, "NULL exchange LUT");
    SIM(sim)->set_exchange_luts(a, d);
  });
}
int um_sim_set_bext_lut(void *sim, const void *lx, const void *ly, const void *lz) {
  return guard([&] {
    SIMCHK(sim);
    const void *l[3] = {lx, ly, lz};
    SIM(sim)->set_bext_lut(l);
  });
}
int um_sim_set_frozen_lut(void *sim, const void *lut) { return guard([&] { SIMCHK(sim); SIM(sim)->set_frozen_lut(lut); }); }
int um_sim_set_time_function(void *sim, int param, int ncomp, um_time_fn fn, void *ctx) {
  return guard([&] { SIMCHK(sim); SIM(sim)->set_time_function(param, ncomp, fn, ctx); });
}
int um_sim_add_field_term(void *sim, int kind, const void *ax, const void *ay, const void *az, double mul,
                          um_time_fn fn, void *ctx) {
  return guard([&] {
    SIMCHK(sim);
    const void *a[3] = {ax, ay, az};
    SIM(sim)->add_field_term(kind, a, mul, fn, ctx);
  });
}
int um_sim_clear_field_terms(void *sim, int kind) {
  return guard([&] { SIMCHK(sim); SIM(sim)->clear_field_terms(kind); });
}
int um_sim_set_nodemag_lut(void *sim, const void *lut) {
  return guard([&] { SIMCHK(sim); SIM(sim)->set_nodemag_lut(lut); });
}
int um_sim_set_geometry(void *sim, const void *vol) { return guard([&] { SIMCHK(sim); SIM(sim)->set_geometry(vol); }); }
int um_sim_set_flags(void *sim, int d, int ob, int pr, int hs) {
  return guard([&] { SIMCHK(sim); SIM(sim)->set_flags(d, ob, pr, hs); });
}
int um_sim_set_demag_accuracy(void *sim, double a) { return guard([&] { SIMCHK(sim); SIM(sim)->set_demag_accuracy(a); }); }
int um_sim_tape_begin(void *sim, int max_steps, int checkpoint_every) {
  return guard([&] { SIMCHK(sim); SIM(sim)->tape_begin(max_steps, checkpoint_every); });
}
int um_sim_tape_end(void *sim) { return guard([&] { SIMCHK(sim); SIM(sim)->tape_end(); }); }
int um_sim_tape_steps(void *sim, long long *steps) {
  return guard([&] {
    SIMCHK(sim);
    UM_REQUIRE(steps, UM_EINVAL, "NULL steps");
    *steps = SIM(sim)->tape_steps();
  });
}
int um_sim_backward(void *sim, const void *const dm[3], void *const dm0[3], void *dms, double dbext[3]) {
  return guard([&] {
    SIMCHK(sim);
    UM_REQUIRE(dm, UM_EINVAL, "NULL cotangent");
    SIM(sim)->backward(dm, dm0, dms, dbext);
  });
}
int um_sim_minimize(void *sim, double stop_max_dm, int samples) {
  return guard([&] { SIMCHK(sim); SIM(sim)->minimize(stop_max_dm, samples); });
}
int um_sim_set_dmi_bulk(void *sim, const void *lut) {
  return guard([&] { SIMCHK(sim); SIM(sim)->set_dmi_bulk(lut); });
}
int um_sim_set_stt_flags(void *sim, int disable_zhangli, int disable_slonczewski, double fixed_layer_sign) {
  return guard([&] {
    SIMCHK(sim);
    UM_REQUIRE(fixed_layer_sign == 1.0 || fixed_layer_sign == -1.0, UM_EINVAL, "fixed layer sign must be +1 or -1");
    SIM(sim)->set_stt_flags(disable_zhangli, disable_slonczewski, fixed_layer_sign);
  });
}
int um_sim_set_demag_newell(void *sim, int enable, int asymptotic_radius, int zero_self_demag) {
  return guard([&] {
    SIMCHK(sim);
    UM_REQUIRE(zero_self_demag >= 0, UM_EINVAL, "zero_self_demag must be >= 0");
    SIM(sim)->set_demag_newell(enable, asymptotic_radius, zero_self_demag);
  });
}
int um_sim_set_thermal_seed(void *sim, uint64_t seed, uint32_t stream) {
  return guard([&] { SIMCHK(sim); SIM(sim)->set_thermal_seed(seed, stream); });
}
int um_sim_set_m(void *sim, const void *mx, const void *my, const void *mz) {
  return guard([&] {
    SIMCHK(sim);
    UM_REQUIRE(mx && my && mz, UM_EINVAL, "NULL m component");
    const void *m[3] = {mx, my, mz};
    SIM(sim)->set_m(m);
  });
}
int um_sim_get_m(void *sim, void *mx, void *my, void *mz) {
  return guard([&] {
    SIMCHK(sim);
    UM_REQUIRE(mx && my && mz, UM_EINVAL, "NULL m component");
    void *m[3] = {mx, my, mz};
    SIM(sim)->get_m(m);
  });
}
int um_sim_m_device(void *sim, void *ptrs[3]) { return guard([&] { SIMCHK(sim); SIM(sim)->m_device(ptrs); }); }
int um_sim_beff(void *sim, void *bx, void *by, void *bz) {
  return guard([&] { SIMCHK(sim); void *o[3] = {bx, by, bz}; SIM(sim)->beff(o); });
}
int um_sim_torque(void *sim, void *tx, void *ty, void *tz) {
  return guard([&] { SIMCHK(sim); void *o[3] = {tx, ty, tz}; SIM(sim)->torque(o); });
}
int um_sim_demag_field(void *sim, void *bx, void *by, void *bz) {
  return guard([&] { SIMCHK(sim); void *o[3] = {bx, by, bz}; SIM(sim)->demag_field(o); });
}
int um_sim_set_solver(void *sim, int s) { return guard([&] { SIMCHK(sim); SIM(sim)->set_solver(s); }); }
int um_sim_set_solver_params(void *sim, const double k[9]) {
  return guard([&] { SIMCHK(sim); UM_REQUIRE(k, UM_EINVAL, "NULL knobs"); SIM(sim)->set_solver_params(k); });
}
int um_sim_get_state(void *sim, double out[8]) { return guard([&] { SIMCHK(sim); SIM(sim)->get_state(out); }); }
int um_sim_step(void *sim) { return guard([&] { SIMCHK(sim); SIM(sim)->step(); }); }
int um_sim_steps(void *sim, int n) { return guard([&] { SIMCHK(sim); SIM(sim)->steps(n); }); }
int um_sim_run(void *sim, double seconds) { return guard([&] { SIMCHK(sim); SIM(sim)->run(seconds); }); }
int um_sim_average_m(void *sim, double out[3]) { return guard([&] { SIMCHK(sim); SIM(sim)->average_m(out); }); }
int um_sim_energy(void *sim, int term, double *out) {
  return guard([&] { SIMCHK(sim); UM_REQUIRE(out, UM_EINVAL, "NULL out"); *out = SIM(sim)->energy(term); });
}
int um_sim_stream(void *sim, void **s) { return guard([&] { SIMCHK(sim); *s = (void *)SIM(sim)->stream(); }); }
int um_sim_set_profiling(void *sim, int on) { return guard([&] { SIMCHK(sim); SIM(sim)->set_profiling(on); }); }
int um_sim_kernel_times(void *sim, double out_ms[3], int *n_evals) {
  return guard([&] { SIMCHK(sim); SIM(sim)->kernel_times(out_ms, n_evals); });
}

int um_sim_relax(void *sim, double torque_threshold) {
  return guard([&] { SIMCHK(sim); SIM(sim)->relax(torque_threshold); });
}

int um_sim_set_graph(void *sim, int on) { return guard([&] { SIMCHK(sim); SIM(sim)->set_graph(on); }); }

int um_sims_steps(int nsims, void *const *sims, int n) {
  return guard([&] {
    UM_REQUIRE(nsims >= 0 && n >= 0 && (nsims == 0 || sims), UM_EINVAL, "um_sims_steps: bad arguments");
    for (int i = 0; i < nsims; i++) SIMCHK(sims[i]);
    std::vector<long> stop(nsims), reps(nsims);
    long most = 0;
    for (int i = 0; i < nsims; i++) {
      stop[i] = SIM(sims[i])->step_count() + n;
      reps[i] = SIM(sims[i])->graph_prepare(stop[i]);
      most = std::max(most, reps[i]);
    }
    // round-robin replays: every member's stream has work queued at once
    for (long r = 0; r < most; r++)
      for (int i = 0; i < nsims; i++)
        if (r < reps[i]) SIM(sims[i])->graph_replay();
    for (int i = 0; i < nsims; i++) SIM(sims[i])->graph_finish(stop[i]);
  });
}

}  // extern "C"

// ---- batched ensembles -----------------------------------------------------------------
// The recorded two-step launch lists of the members, merged into batched
// launches: phase by phase, the kernel family that most members have next runs
// for all of them (one k_batch launch per distinct kernel variant), each
// member's launches keeping their order.  The tables (arguments, work-group
// prefix, member grids) live in one device buffer; the launch sequence is
// captured once into a graph on the plan's stream.
namespace um {
struct BatchGroupLaunch {
  void (*launch)(const void *, const uint32_t *, const uint3 *, int, unsigned, dim3, size_t, hipStream_t);
  size_t a_off, s_off, g_off;
  int nm;
  unsigned total;
  dim3 block;
  size_t lds;
};
struct BatchPlan {
  std::vector<const void *> sims;
  std::vector<std::vector<uint64_t>> keys;
  std::vector<BatchGroupLaunch> launches;
  DevBuf tab;
  hipStream_t st = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t done = nullptr;
  std::vector<hipEvent_t> ready;
  int launches_per_replay() const { return (int)launches.size(); }
  ~BatchPlan() {
    if (st) (void)hipStreamSynchronize(st);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (hipEvent_t e : ready) (void)hipEventDestroy(e);
    if (done) (void)hipEventDestroy(done);
    if (st) (void)hipStreamDestroy(st);
  }
};
static std::mutex g_bmu;
static BatchPlan *g_bplan = nullptr;
static long g_batch_launches = 0;   // batched launches per replay of the last plan

static std::string body_family(const char *mangled) {
  // the body's class name without its template arguments
  std::string f(mangled);
  const size_t i = f.find('I');
  return i == std::string::npos ? f : f.substr(0, i);
}

static BatchPlan *batch_plan_build(const std::vector<SimBase *> &m) {
  const int nm = (int)m.size();
  std::vector<std::vector<BatchEntry>> rec(nm);
  for (int i = 0; i < nm; i++) m[i]->batch_record(rec[i]);
  std::unique_ptr<BatchPlan> P(new BatchPlan());
  for (SimBase *x : m) {
    P->sims.push_back((const void *)x);
    P->keys.push_back(x->batch_key());
  }
  // merge the lists
  std::map<std::string, int> famid;
  std::vector<std::vector<int>> fam(nm);
  for (int i = 0; i < nm; i++)
    for (const BatchEntry &e : rec[i]) {
      auto it = famid.emplace(body_family(e.fam), (int)famid.size()).first;
      fam[i].push_back(it->second);
    }
  std::vector<size_t> head(nm, 0);
  std::vector<unsigned char> host;
  auto put = [&](const void *p, size_t bytes) {
    const size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + bytes);
    std::memcpy(host.data() + off, p, bytes);
    return off;
  };
  for (;;) {
    std::vector<int> count(famid.size(), 0);
    bool any = false;
    for (int i = 0; i < nm; i++)
      if (head[i] < rec[i].size()) { count[fam[i][head[i]]]++; any = true; }
    if (!any) break;
    const int f = (int)(std::max_element(count.begin(), count.end()) - count.begin());
    // members whose next launch is of family f, grouped by kernel variant
    std::vector<std::pair<const BatchEntry *, std::vector<int>>> groups;
    for (int i = 0; i < nm; i++) {
      if (head[i] >= rec[i].size() || fam[i][head[i]] != f) continue;
      const BatchEntry &e = rec[i][head[i]++];
      bool placed = false;
      for (auto &g : groups) {
        const BatchEntry &r = *g.first;
        if (r.key == e.key && r.lds == e.lds && r.block.x == e.block.x && r.block.y == e.block.y &&
            r.block.z == e.block.z) {
          g.second.push_back(i);
          placed = true;
          break;
        }
      }
      if (!placed) groups.push_back({&e, std::vector<int>{i}});
    }
    for (auto &g : groups) {
      const BatchEntry &r = *g.first;
      const int gn = (int)g.second.size();
      std::vector<unsigned char> args((size_t)gn * r.asize);
      std::vector<uint32_t> start(gn);
      std::vector<uint3> grid(gn);
      unsigned long long total = 0;
      for (int j = 0; j < gn; j++) {
        const int i = g.second[j];
        const BatchEntry &e = rec[i][head[i] - 1];
        std::memcpy(args.data() + (size_t)j * r.asize, e.args.data(), r.asize);
        start[j] = (uint32_t)total;
        grid[j] = make_uint3(e.grid.x, e.grid.y, e.grid.z);
        total += (unsigned long long)e.grid.x * e.grid.y * e.grid.z;
      }
      UM_REQUIRE(total > 0 && total < (1ull << 31), UM_EUNSUP, "batched launch too large");
      BatchGroupLaunch L;
      L.launch = r.launch;
      L.a_off = put(args.data(), args.size());
      L.s_off = put(start.data(), start.size() * sizeof(uint32_t));
      L.g_off = put(grid.data(), grid.size() * sizeof(uint3));
      L.nm = gn;
      L.total = (unsigned)total;
      L.block = r.block;
      L.lds = r.lds;
      P->launches.push_back(L);
    }
  }
  P->tab.alloc(std::max<size_t>(host.size(), 256));
  UM_HIP(hipMemcpy(P->tab.p, host.data(), host.size(), hipMemcpyHostToDevice));
  UM_HIP(hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking));
  UM_HIP(hipEventCreateWithFlags(&P->done, hipEventDisableTiming));
  P->ready.resize(nm);
  for (int i = 0; i < nm; i++) UM_HIP(hipEventCreateWithFlags(&P->ready[i], hipEventDisableTiming));
  // the replay: every batched launch in order, captured once
  const unsigned char *base = (const unsigned char *)P->tab.p;
  hipGraph_t g = nullptr;
  UM_HIP(hipStreamBeginCapture(P->st, hipStreamCaptureModeRelaxed));
  try {
    for (const BatchGroupLaunch &L : P->launches)
      L.launch(base + L.a_off, (const uint32_t *)(base + L.s_off), (const uint3 *)(base + L.g_off), L.nm, L.total,
               L.block, L.lds, P->st);
  } catch (...) {
    hipGraph_t junk = nullptr;
    (void)hipStreamEndCapture(P->st, &junk);
    if (junk) (void)hipGraphDestroy(junk);
    throw;
  }
  UM_HIP(hipStreamEndCapture(P->st, &g));
  hipError_t e = hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  UM_HIP(e);
  return P.release();
}

static bool batch_plan_fresh(const BatchPlan *P, const std::vector<SimBase *> &m) {
  if (!P || P->sims.size() != m.size()) return false;
  for (size_t i = 0; i < m.size(); i++)
    if (P->sims[i] != (const void *)m[i] || P->keys[i] != m[i]->batch_key()) return false;
  return true;
}
}  // namespace um

extern "C" {

int um_sims_steps_batched(int nsims, void *const *sims, int n, int *launches_per_two_steps) {
  return guard([&] {
    UM_REQUIRE(nsims >= 0 && n >= 0 && (nsims == 0 || sims), UM_EINVAL, "um_sims_steps_batched: bad arguments");
    for (int i = 0; i < nsims; i++) SIMCHK(sims[i]);
    for (int i = 0; i < nsims; i++)
      for (int j = 0; j < i; j++) UM_REQUIRE(sims[i] != sims[j], UM_EINVAL, "um_sims_steps_batched: a handle twice");
    std::vector<long> stop(nsims), reps(nsims);
    for (int i = 0; i < nsims; i++) {
      stop[i] = SIM(sims[i])->step_count() + n;
      reps[i] = SIM(sims[i])->batch_prepare(stop[i]);
    }
    // members whose every remaining pair of steps can be replayed together
    long R = 0;
    for (int i = 0; i < nsims; i++) R = std::max(R, reps[i]);
    std::vector<SimBase *> mem;
    for (int i = 0; i < nsims; i++)
      if (reps[i] > 0) {
        UM_REQUIRE(reps[i] == R, UM_ESTATE, "batched members out of step");
        mem.push_back(SIM(sims[i]));
      }
    std::lock_guard<std::mutex> lk(g_bmu);
    if (!mem.empty()) {
      bool batched = true;
      if (!batch_plan_fresh(g_bplan, mem)) {
        delete g_bplan;
        g_bplan = nullptr;
        try {
          g_bplan = batch_plan_build(mem);
        } catch (const Error &e) {
          if (e.code != UM_EUNSUP) throw;
          batched = false;   // a kernel without a batched form: graph replays per member
        }
      }
      if (batched) {
        BatchPlan *P = g_bplan;
        g_batch_launches = P->launches_per_replay();
        for (size_t i = 0; i < mem.size(); i++) {
          UM_HIP(hipEventRecord(P->ready[i], mem[i]->stream()));
          UM_HIP(hipStreamWaitEvent(P->st, P->ready[i], 0));
        }
        for (long r = 0; r < R; r++) {
          UM_HIP(hipGraphLaunch(P->gexec, P->st));
          for (SimBase *x : mem) x->batch_advance();
        }
        UM_HIP(hipEventRecord(P->done, P->st));
        for (SimBase *x : mem) UM_HIP(hipStreamWaitEvent(x->stream(), P->done, 0));
      } else {
        g_batch_launches = 0;
        for (SimBase *x : mem) x->graph_ensure();
        for (long r = 0; r < R; r++)
          for (SimBase *x : mem) x->graph_replay();
      }
    }
    if (launches_per_two_steps) *launches_per_two_steps = (int)g_batch_launches;
    for (int i = 0; i < nsims; i++) SIM(sims[i])->graph_finish(stop[i]);
  });
}

int um_graph_begin(void) {
  return guard([&] {
    ensure_init();
    UM_HIP(hipStreamBeginCapture(lib_stream(), hipStreamCaptureModeRelaxed));
  });
}
int um_graph_end(void **graph) {
  return guard([&] {
    UM_REQUIRE(graph, UM_EINVAL, "graph out-pointer is NULL");
    hipGraph_t g = nullptr;
    UM_HIP(hipStreamEndCapture(lib_stream(), &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    UM_HIP(e);
    *graph = (void *)ex;
  });
}
int um_graph_launch(void *graph) {
  return guard([&] {
    UM_REQUIRE(graph, UM_EINVAL, "graph is NULL");
    UM_HIP(hipGraphLaunch((hipGraphExec_t)graph, lib_stream()));
  });
}
int um_graph_destroy(void *graph) {
  return guard([&] {
    if (graph) UM_HIP(hipGraphExecDestroy((hipGraphExec_t)graph));
  });
}

}  // extern "C"
