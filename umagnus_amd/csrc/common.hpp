// Internal declarations shared by the libumhip sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/umhip.h"

namespace um {

// ---- errors ----------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string &msg);

#define UM_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw ::um::Error(UM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define UM_REQUIRE(cond, code, msg)                           \
  do {                                                        \
    if (!(cond)) throw ::um::Error((code), std::string(msg)); \
  } while (0)

// Run f, translating exceptions into a status code + um_last_error().
template <typename F>
int guard(F &&f) {
  try {
    f();
    return UM_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return UM_EHIP;
  }
}

hipStream_t lib_stream();   // stream of the unfused ops
void ensure_init();

// ---- precision traits ------------------------------------------------------
template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };
template <typename T> using C2 = typename cplx<T>::type;

// Parameter in the MSlice convention: arr==nullptr -> uniform value mul.
template <typename T>
struct Par {
  const T *arr;
  T mul;
  __device__ __forceinline__ T at(long i) const { return arr ? mul * arr[i] : mul; }
  bool zero_host() const { return arr == nullptr && mul == T(0); }
};
template <typename T>
inline Par<T> par(const void *arr, double mul) { return Par<T>{(const T *)arr, (T)mul}; }

// ---- launch trace ----------------------------------------------------------
// While tracing is on (um_trace_kernels), every launch through UM_LAUNCH
// records the demangled name of the kernel it starts, so a caller can check
// which compiled variant a plan or a simulation selects.
extern bool g_trace_on;
void trace_kernel(const void *host_stub);
// While a batch recording is installed on the calling thread (batch.hpp), a
// launch of a kernel that cannot be batched refuses the recording.
struct BatchRec;
extern thread_local BatchRec *g_rec;
[[noreturn]] void batch_refuse(const void *host_stub);
#define UM_LAUNCH(kern, ...)                                                \
  do {                                                                      \
    if (::um::g_rec) ::um::batch_refuse((const void *)(kern));             \
    if (::um::g_trace_on) ::um::trace_kernel((const void *)(kern));         \
    hipLaunchKernelGGL(kern, __VA_ARGS__);                                  \
  } while (0)

// ---- launch helpers --------------------------------------------------------
inline unsigned blocks_for(long n, int threads, long cap = 1L << 20) {
  long b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// ---- device buffer (RAII) --------------------------------------------------
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    release();
    p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0;
    return *this;
  }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(UM_ENOMEM, "hipMalloc(" + std::to_string(b) + "): " + hipGetErrorString(e));
    }
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  template <typename T> T *as() const { return (T *)p; }
};

// ---- kernels launched from host code (defined in the .hip files) ----------
void fill_bytes(void *dst, size_t bytes, hipStream_t s);
template <typename T> void fill_value(T *dst, T v, size_t n, hipStream_t s);

}  // namespace um
