// Device-side symbolic pipeline: the term sort, the S pattern and the per-slot pair lists of
// the Schur BA path built on the GPU (the host analysis in problem.cpp computes the same
// arrays; both are checked against each other element by element).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "api_internal.hpp"
#include "problem.hpp"

namespace sgpu {

// Stable LSD radix sort of (key, value) pairs of 32-bit unsigned keys by their low key_bits
// bits (8-bit digits; equal keys keep their input order).  keys/vals are sorted in place
// through the two scratch arrays of n elements each.
void radix_sort_pairs(hipStream_t s, DeviceArena& arena, uint32_t* keys, uint32_t* vals,
                      uint32_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int key_bits);

// Exclusive prefix sum of n int64 counts (in place), on the device.
void exclusive_scan_i64(hipStream_t s, DeviceArena& arena, int64_t* data, int64_t n);

// The term sort of a problem built with build_host_problem_sorted(.., false), on the device
// (stable radix passes over the sort key's arguments, the last argument first), applied to
// P: the HostProblem build_host_problem returns.
void sort_terms_on_device(HostProblem& P, hipStream_t s);

// analyze() with the S pattern and the pair lists built on the device.  The pairs stay on
// the device (*d_pairs, 2 per pair, allocated from `arena`); S.pairs is left empty.
Symbolic analyze_device(const HostProblem& P, hipStream_t s, DeviceArena& arena,
                        int32_t** d_pairs);

}  // namespace sgpu
