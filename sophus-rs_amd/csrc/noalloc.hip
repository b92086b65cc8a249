// Allocation-free direct block-LDLt solve: sgpu_block_ldlt_solve with every device
// buffer carved from one file-scope device array instead of hipMalloc, for callers that
// cannot allocate device memory.
//
// Each GPU the library runs on holds its own copy of `scratch` (kScratchBytes) from module
// load to process end.  In host code `scratch` names a host-side stand-in, not device
// memory: its device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and only that address is used afterwards.  Every call on one GPU
// reuses the same array and finds the previous call's values there; the solve zeroes what
// must start at zero (the fronts and the step scalars) before it runs.  Calls on one GPU are
// serialised within the process; calls that overlap on one GPU from several processes, or
// with other users of the array, race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "block_solve.hpp"
#include "problem.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(SGPU_NOALLOC_SCRATCH_BYTES);
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_dev_mu[kMaxDevices];
std::mutex g_init_mu;
void* g_scratch[kMaxDevices] = {};

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Device address of this device's `scratch`, looked up on the first call on the device.
void* scratch_address(int dev) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    SGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw sgpu::Error(SGPU_ERR_HIP, "hipGetSymbolAddress returned no address for scratch");
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

// Allocator that only adds up the aligned sizes (no device access).
struct ScratchSize {
  size_t used = 0;
  template <typename T>
  T* alloc(size_t n) {
    used = align_up(used) + std::max<size_t>(n, 1) * sizeof(T);
    return nullptr;
  }
  template <typename T>
  T* up(const T*, size_t n) {
    return alloc<T>(n);
  }
};

// Allocator that hands out consecutive aligned slices of the device array at `base`.
struct ScratchCarver {
  char* base;
  size_t used = 0;
  template <typename T>
  T* alloc(size_t n) {
    used = align_up(used);
    T* p = reinterpret_cast<T*>(base + used);
    used += std::max<size_t>(n, 1) * sizeof(T);
    if (used > kScratchBytes) throw sgpu::Error(SGPU_ERR_OUT_OF_MEMORY, "scratch overrun");
    return p;
  }
  template <typename T>
  T* up(const T* host, size_t n) {
    T* d = alloc<T>(n);
    if (n) SGPU_HIP(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }
};

}  // namespace

using namespace sgpu;

extern "C" sgpu_status sgpu_block_ldlt_solve_noalloc(int64_t num_blocks, const int64_t* col_ptr,
                                                     const int64_t* row_idx, const double* values,
                                                     const double* rhs, double* x) {
  return guarded([&] {
    Symbolic S;
    std::vector<double> blocks;
    parse_lower_blocks(num_blocks, col_ptr, row_idx, values, rhs, x, S, blocks);
    analyze_rcs(S);
    ScratchSize need;
    (void)bind_block_solve(S, blocks, rhs, need);
    if (need.used > kScratchBytes)
      throw Error(SGPU_ERR_OUT_OF_MEMORY,
                  "the allocation-free block LDLt solve needs " + std::to_string(need.used) +
                      " bytes of device scratch for this system; it holds " +
                      std::to_string(kScratchBytes) + " bytes per GPU");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw Error(SGPU_ERR_NO_DEVICE, "no HIP device");
    int dev = 0;
    SGPU_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(SGPU_ERR_UNSUPPORTED, "device ordinal too large");
    std::lock_guard<std::mutex> lk(g_dev_mu[dev]);
    ScratchCarver carve{static_cast<char*>(scratch_address(dev))};
    BlockSolveBufs b = bind_block_solve(S, blocks, rhs, carve);
    run_block_solve(S, b, x);
  });
}
