// Batched launches: one launch runs the same kernel body for many
// simulations (ensemble members) at once.
//
// A batchable kernel is a body struct K with
//   using Args = <argument struct>;  static constexpr int NT, MINB;
//   static __device__ void run(const Args &a, uint3 blockIdx, uint3 gridDim);
// whose normal __global__ entry passes the hardware block coordinates.  The
// batched entry k_batch<K> covers the work-groups of every member in one
// grid: work-group b finds its member m (start[m] <= b < start[m+1], binary
// search over a prefix table), turns b - start[m] back into the member's own
// block coordinates and runs the body on that member's arguments, so each
// member computes exactly what its own launch would.
//
// Recording: while a BatchRec is installed on the calling thread (g_rec),
// launches of batchable kernels (UM_LAUNCH_B) append an entry to the current
// member's list instead of running; any other launch (UM_LAUNCH) throws, so a
// sequence with a kernel that cannot be batched is never half recorded.
#pragma once
#include <cstring>
#include <type_traits>
#include <typeinfo>
#include <vector>

#include "common.hpp"

namespace um {

__device__ __forceinline__ uint3 hw_bid() { return make_uint3(blockIdx.x, blockIdx.y, blockIdx.z); }
__device__ __forceinline__ uint3 hw_grid() { return make_uint3(gridDim.x, gridDim.y, gridDim.z); }

template <typename K>
__global__ __launch_bounds__(K::NT, K::MINB) void k_batch(const typename K::Args *args, const uint32_t *start,
                                                          const uint3 *grid, int nm) {
  const uint32_t b = blockIdx.x;
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const uint3 g = grid[lo];
  const uint32_t r = b - start[lo];
  K::run(args[lo], make_uint3(r % g.x, (r / g.x) % g.y, r / (g.x * g.y)), g);
}

// one recorded launch of one member
struct BatchEntry {
  const void *key;      // the batched kernel k_batch<K>
  const char *fam;      // the body's name (template arguments included)
  dim3 block, grid;
  size_t lds;
  size_t asize;         // sizeof(K::Args)
  std::vector<unsigned char> args;
  // launches k_batch<K> over `total` work-groups with the device tables
  void (*launch)(const void *args, const uint32_t *start, const uint3 *grid, int nm, unsigned total, dim3 block,
                 size_t lds, hipStream_t s);
};

template <typename K>
void batch_launch(const void *args, const uint32_t *start, const uint3 *grid, int nm, unsigned total, dim3 block,
                  size_t lds, hipStream_t s) {
  static size_t allowed = 65536;   // dynamic LDS this batched kernel may take
  if (lds > allowed) {
    UM_HIP(hipFuncSetAttribute((const void *)k_batch<K>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    allowed = lds;
  }
  if (g_trace_on) trace_kernel((const void *)k_batch<K>);
  hipLaunchKernelGGL(k_batch<K>, dim3(total), block, lds, s, (const typename K::Args *)args, start, grid, nm);
  UM_HIP(hipGetLastError());
}

struct BatchRec {
  std::vector<BatchEntry> *cur = nullptr;   // the member being recorded
  template <typename K>
  void add(dim3 grid, dim3 block, size_t lds, const typename K::Args &a) {
    static_assert(std::is_trivially_copyable<typename K::Args>::value, "batched arguments are copied as bytes");
    UM_REQUIRE(cur, UM_ESTATE, "batch recording without a member");
    UM_REQUIRE((unsigned long long)grid.x * grid.y * grid.z < (1ull << 31), UM_EUNSUP, "batched grid too large");
    BatchEntry e;
    e.key = (const void *)k_batch<K>;
    e.fam = typeid(K).name();
    e.block = block;
    e.grid = grid;
    e.lds = lds;
    e.asize = sizeof(typename K::Args);
    e.args.resize(e.asize);
    std::memcpy(e.args.data(), &a, e.asize);
    e.launch = &batch_launch<K>;
    cur->push_back(std::move(e));
  }
};

// launch of a batchable kernel `kern` (the normal entry of body K): recorded
// while g_rec is set, else launched
template <typename K>
inline void launch_b(void (*kern)(typename K::Args), dim3 grid, dim3 block, size_t lds, hipStream_t s,
                     const typename K::Args &a) {
  if (g_rec) {
    g_rec->template add<K>(grid, block, lds, a);
    return;
  }
  if (g_trace_on) trace_kernel((const void *)kern);
  typename K::Args copy = a;
  void *args[] = {(void *)&copy};
  UM_HIP(hipLaunchKernel((const void *)kern, grid, block, args, lds, s));
}
#define UM_LAUNCH_B(K, kern, grid, block, lds, s, a) ::um::launch_b<K>(kern, grid, block, lds, s, a)

}  // namespace um
