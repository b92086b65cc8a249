// Internal helpers shared by the host translation units of libsophus_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sophus_gpu.h"

namespace sgpu {

// Thread-local last-error message behind sgpu_last_error().
void clear_error();
sgpu_status set_error(sgpu_status status, const std::string& message);

// Raised inside the library and turned into a status at the C boundary.
struct Error : std::runtime_error {
  sgpu_status status;
  Error(sgpu_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    const sgpu_status s = e == hipErrorOutOfMemory ? SGPU_ERR_OUT_OF_MEMORY
                          : e == hipErrorNoDevice  ? SGPU_ERR_NO_DEVICE
                                                   : SGPU_ERR_HIP;
    throw Error(s, std::string(what) + ": " + hipGetErrorString(e));
  }
}

#define SGPU_HIP(call) ::sgpu::hip_check((call), #call)

// Run `body`, mapping exceptions to a status + last-error message.
template <typename F>
sgpu_status guarded(F&& body) {
  clear_error();
  try {
    body();
    return SGPU_OK;
  } catch (const Error& e) {
    return set_error(e.status, e.what());
  } catch (const std::bad_alloc&) {
    return set_error(SGPU_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return set_error(SGPU_ERR_INVALID_ARGUMENT, e.what());
  }
}

// Device allocations owned by one object, freed together.
struct DeviceArena {
  std::vector<void*> ptrs;
  template <typename T>
  T* alloc(size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    SGPU_HIP(hipMalloc(&p, n * sizeof(T)));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  template <typename T>
  T* upload(const std::vector<T>& v, hipStream_t s) {
    T* p = alloc<T>(v.size());
    (void)s;  // one-time upload: blocking copy, so temporaries may die right after
    if (!v.empty()) SGPU_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  ~DeviceArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

}  // namespace sgpu
