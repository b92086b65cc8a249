// Device-side symbolic pipeline of the Schur BA path (the term sort, the S pattern and the
// per-slot pair lists), with a stable LSD radix sort and an exclusive scan written for it.
//
// The arrays are the ones the host analysis builds (problem.cpp analyze): the same term order
// (CostFn::sort, cost_fn.rs:354-393: stable by the marginalized argument, then the free
// ones, then the conditioned one), the same pattern (block_sparse_matrix_pattern.rs:230-363)
// and the same pair order within each S slot (point order, then the two observations' order).
#include "symbolic.hpp"

#include <algorithm>
#include <chrono>
#include <numeric>

namespace sgpu {
namespace {

constexpr int kThreads = 256;
constexpr int kRadixBits = 8;
constexpr int kDigits = 1 << kRadixBits;
constexpr int kItems = 8;                    // elements per thread per radix tile
constexpr int kRadixTile = kThreads * kItems;     // elements per radix workgroup
constexpr int kScanItems = 4;
constexpr int kScanTile = kThreads * kScanItems;

int blocks_of(int64_t n, int64_t per) { return int((n + per - 1) / per); }

// ---- radix sort ------------------------------------------------------------------------------
// Per tile, the count of each digit: hist[digit * ntiles + tile] (digit-major, so the
// exclusive scan of hist gives each tile's first output position per digit, tiles in order).
__global__ void __launch_bounds__(kThreads)
    k_rs_hist(const uint32_t* __restrict__ keys, int64_t n, int shift, int64_t ntiles,
              int64_t* __restrict__ hist) {
  __shared__ int cnt[kDigits];
  for (int d = threadIdx.x; d < kDigits; d += blockDim.x) cnt[d] = 0;
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kRadixTile;
  for (int k = 0; k < kItems; ++k) {
    const int64_t e = base + k * kThreads + threadIdx.x;
    if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & (kDigits - 1)], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kDigits; d += blockDim.x) hist[int64_t(d) * ntiles + blockIdx.x] = cnt[d];
}

// Stable scatter: the tile's elements go out in their input order within each digit.  The
// tile is processed as kItems rows of 256 elements (row k holds elements k * 256 + tid); an
// element's rank among the equal digits of its row is found with ballots over the digit's
// bits (wavefront), then the earlier wavefronts' counts, then the running per-digit count
// of the tile's earlier rows.
__global__ void __launch_bounds__(kThreads)
    k_rs_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n,
                 int shift, int64_t ntiles, const int64_t* __restrict__ offsets,
                 uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
  __shared__ int64_t run[kDigits];
  __shared__ int wcnt[kThreads / 64][kDigits];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int d = tid; d < kDigits; d += blockDim.x) {
    run[d] = offsets[int64_t(d) * ntiles + blockIdx.x];
    for (int w = 0; w < kThreads / 64; ++w) wcnt[w][d] = 0;
  }
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kRadixTile;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int k = 0; k < kItems; ++k) {
    const int64_t e = base + k * kThreads + tid;
    const bool valid = e < n;
    const uint32_t key = valid ? keys[e] : 0u;
    const uint32_t val = valid ? vals[e] : 0u;
    const int d = int((key >> shift) & (kDigits - 1));
    unsigned long long eq = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const unsigned long long m = __ballot((d >> b) & 1);
      eq &= ((d >> b) & 1) ? m : ~m;
    }
    const int rank = __popcll(eq & lt);
    // the lowest lane of each digit group writes the group's size
    if (valid && (eq & lt) == 0ull) wcnt[wid][d] = __popcll(eq);
    __syncthreads();
    if (valid) {
      int64_t pos = run[d] + rank;
      for (int w = 0; w < wid; ++w) pos += wcnt[w][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    for (int dd = tid; dd < kDigits; dd += blockDim.x) {
      int add = 0;
      for (int w = 0; w < kThreads / 64; ++w) {
        add += wcnt[w][dd];
        wcnt[w][dd] = 0;
      }
      run[dd] += add;
    }
    __syncthreads();
  }
}

// ---- exclusive scan (int64) ------------------------------------------------------------------
// One workgroup per kScanTile elements: the exclusive prefix within the tile, and the tile's
// total into sums[tile].
__global__ void __launch_bounds__(kThreads)
    k_scan_tiles(int64_t* __restrict__ data, int64_t n, int64_t* __restrict__ sums) {
  __shared__ int64_t part[kThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  int64_t v[kScanItems], t = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? data[base + k] : 0;
    t += v[k];
  }
  part[threadIdx.x] = t;
  __syncthreads();
  // inclusive Hillis-Steele scan of the thread totals
  for (int off = 1; off < kThreads; off <<= 1) {
    const int64_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kThreads - 1) sums[blockIdx.x] = part[kThreads - 1];
}

__global__ void k_scan_add(int64_t* __restrict__ data, int64_t n, const int64_t* __restrict__ sums) {
  const int64_t e = int64_t(blockIdx.x) * kScanTile + threadIdx.x;
  const int64_t add = sums[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k)
    if (e + k * kThreads < n) data[e + k * kThreads] += add;
}

__global__ void k_iota(uint32_t* __restrict__ v, int64_t n) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) v[e] = uint32_t(e);
}

// keys[i] = src[order[i]]
__global__ void k_gather_u32(const int32_t* __restrict__ src, const uint32_t* __restrict__ order,
                             uint32_t* __restrict__ keys, int64_t n) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) keys[e] = uint32_t(src[order[e]]);
}

// ---- S pattern ----------------------------------------------------------------------------
// One thread per active point: a bit for every pair of its observations' free cameras (row =
// the larger block, column = the smaller) in the nb x nb bitmap (W words per row).
__global__ void k_pattern_bits(int32_t npt, const int32_t* __restrict__ pt_ptr,
                               const int32_t* __restrict__ pt_block, const int32_t* __restrict__ obs_pose,
                               const int32_t* __restrict__ pose_block, int32_t W,
                               unsigned long long* __restrict__ bits) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= npt || pt_block[b] < 0) return;
  for (int i = pt_ptr[b]; i < pt_ptr[b + 1]; ++i) {
    const int ci = pose_block[obs_pose[i]];
    if (ci < 0) continue;
    for (int j = pt_ptr[b]; j < i; ++j) {
      const int cj = pose_block[obs_pose[j]];
      if (cj < 0) continue;
      const int r = max(ci, cj), c = min(ci, cj);
      atomicOr(&bits[int64_t(r) * W + (c >> 6)], 1ull << (c & 63));
    }
  }
}

// ---- pair lists ------------------------------------------------------------------------------
// The S slot of (r, c), c <= r, from the final bitmap and each row word's first slot.
__device__ __forceinline__ uint32_t slot_of(const unsigned long long* bits, const int32_t* wpre,
                                            int32_t W, int r, int c) {
  const int64_t w = int64_t(r) * W + (c >> 6);
  const unsigned long long below = (c & 63) ? (bits[w] & ((1ull << (c & 63)) - 1ull)) : 0ull;
  return uint32_t(wpre[w] + __popcll(below));
}

// Pairs (i, j), i != j, of one active point's observations with free cameras c_j <= c_i:
// count per point (pass 0) or write them in the host's loop order from the point's offset
// (pass 1): slot key, position, and the two observations' camera-major positions.
__global__ void k_pairs(int pass, int32_t npt, const int32_t* __restrict__ pt_ptr,
                        const int32_t* __restrict__ pt_block, const int32_t* __restrict__ obs_pose,
                        const int32_t* __restrict__ pose_block, const int32_t* __restrict__ cam_pos,
                        const unsigned long long* __restrict__ bits, const int32_t* __restrict__ wpre,
                        int32_t W, int64_t* __restrict__ count_or_off, uint32_t* __restrict__ key,
                        uint32_t* __restrict__ val, int32_t* __restrict__ pi, int32_t* __restrict__ pj) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= npt) return;
  if (pt_block[b] < 0) {
    if (pass == 0) count_or_off[b] = 0;
    return;
  }
  int64_t at = pass == 0 ? 0 : count_or_off[b];
  for (int i = pt_ptr[b]; i < pt_ptr[b + 1]; ++i) {
    const int ci = pose_block[obs_pose[i]];
    if (ci < 0) continue;
    for (int j = pt_ptr[b]; j < pt_ptr[b + 1]; ++j) {
      const int cj = pose_block[obs_pose[j]];
      if (cj < 0 || cj > ci || j == i) continue;
      if (pass == 1) {
        key[at] = slot_of(bits, wpre, W, ci, cj);
        val[at] = uint32_t(at);
        pi[at] = cam_pos[i];
        pj[at] = cam_pos[j];
      }
      ++at;
    }
  }
  if (pass == 0) count_or_off[b] = at;
}

// Sorted pair k: its two positions, and its slot counted into per-slot counts.
__global__ void k_pairs_gather(int64_t n, const uint32_t* __restrict__ key,
                               const uint32_t* __restrict__ order, const int32_t* __restrict__ pi,
                               const int32_t* __restrict__ pj, int32_t* __restrict__ pairs,
                               int64_t* __restrict__ slot_count) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t src = order[k];
  pairs[2 * k] = pi[src];
  pairs[2 * k + 1] = pj[src];
  atomicAdd(reinterpret_cast<unsigned long long*>(&slot_count[key[k]]), 1ull);
}

template <typename T>
T* upload_vec(DeviceArena& a, const std::vector<T>& v) {
  T* p = a.alloc<T>(std::max<size_t>(v.size(), 1));
  if (!v.empty()) SGPU_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

int bit_width(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

}  // namespace

void exclusive_scan_i64(hipStream_t s, DeviceArena& arena, int64_t* data, int64_t n) {
  if (n <= 0) return;
  const int tiles = blocks_of(n, kScanTile);
  int64_t* sums = arena.alloc<int64_t>(size_t(tiles));
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kThreads), 0, s, data, n, sums);
  if (tiles > 1) {
    exclusive_scan_i64(s, arena, sums, tiles);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kThreads), 0, s, data, n, sums);
  }
}

void radix_sort_pairs(hipStream_t s, DeviceArena& arena, uint32_t* keys, uint32_t* vals,
                      uint32_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int key_bits) {
  if (n <= 1 || key_bits <= 0) return;
  const int64_t ntiles = blocks_of(n, kRadixTile);
  int64_t* hist = arena.alloc<int64_t>(size_t(kDigits * ntiles));
  uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
  for (int shift = 0; shift < key_bits; shift += kRadixBits) {
    hipLaunchKernelGGL(k_rs_hist, dim3(unsigned(ntiles)), dim3(kThreads), 0, s, ki, n, shift, ntiles, hist);
    exclusive_scan_i64(s, arena, hist, int64_t(kDigits) * ntiles);
    hipLaunchKernelGGL(k_rs_scatter, dim3(unsigned(ntiles)), dim3(kThreads), 0, s, ki, vi, n, shift,
                       ntiles, hist, ko, vo);
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  if (ki != keys) {  // an odd number of passes: the result is in the scratch arrays
    SGPU_HIP(hipMemcpyAsync(keys, ki, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
    SGPU_HIP(hipMemcpyAsync(vals, vi, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
  }
}

void sort_terms_on_device(HostProblem& P, hipStream_t s) {
  if (P.terms_sorted) return;
  const int64_t n = int64_t(P.obs_point.size());
  if (n <= 1) {
    P.terms_sorted = true;
    return;
  }
  DeviceArena tmp;
  int32_t* d_point = upload_vec(tmp, P.obs_point);
  int32_t* d_pose = upload_vec(tmp, P.obs_pose);
  int32_t* d_intr = upload_vec(tmp, P.obs_intr);
  uint32_t* key = tmp.alloc<uint32_t>(size_t(n));
  uint32_t* val = tmp.alloc<uint32_t>(size_t(n));
  uint32_t* key2 = tmp.alloc<uint32_t>(size_t(n));
  uint32_t* val2 = tmp.alloc<uint32_t>(size_t(n));
  const int g = blocks_of(n, kThreads);
  hipLaunchKernelGGL(k_iota, dim3(g), dim3(kThreads), 0, s, val, n);
  // LSD over the sort key (point, k1, k2): the conditioned argument first, the point last
  const bool intr_free = P.intr_kind == SGPU_VAR_FREE;
  const int32_t* by_key[3] = {intr_free ? d_pose : d_intr, intr_free ? d_intr : d_pose, d_point};
  const std::vector<int32_t>* host_key[3] = {intr_free ? &P.obs_pose : &P.obs_intr,
                                             intr_free ? &P.obs_intr : &P.obs_pose, &P.obs_point};
  for (int k = 0; k < 3; ++k) {
    const int32_t mx = host_key[k]->empty() ? 0 : *std::max_element(host_key[k]->begin(), host_key[k]->end());
    const int bits = bit_width(uint64_t(std::max(mx, 0)));
    if (bits == 0) continue;  // one value: nothing to order
    hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(kThreads), 0, s, by_key[k], val, key, n);
    radix_sort_pairs(s, tmp, key, val, key2, val2, n, bits);
  }
  std::vector<int32_t> ord(static_cast<size_t>(n));
  SGPU_HIP(hipMemcpyAsync(ord.data(), val, size_t(n) * 4, hipMemcpyDeviceToHost, s));
  SGPU_HIP(hipStreamSynchronize(s));
  apply_term_order(P, ord);
}

Symbolic analyze_device(const HostProblem& P, hipStream_t s, DeviceArena& arena,
                        int32_t** d_pairs) {
  Symbolic S;
  analyze_points(P, S);
  const Family& poses = P.fams[P.f_pose];
  const Family& points = P.fams[P.f_point];
  DeviceArena tmp;
  const int32_t nb = S.nb, W = (S.nb + 63) / 64;
  // S pattern: the points' bits on the device, the other entries on the host
  int32_t* d_pt_ptr = upload_vec(tmp, S.pt_ptr);
  int32_t* d_pt_block = upload_vec(tmp, points.block);
  int32_t* d_obs_pose = upload_vec(tmp, P.obs_pose);
  int32_t* d_pose_block = upload_vec(tmp, poses.block);
  const size_t nwords = size_t(nb) * size_t(W);
  unsigned long long* d_bits = tmp.alloc<unsigned long long>(std::max<size_t>(nwords, 1));
  SGPU_HIP(hipMemsetAsync(d_bits, 0, nwords * 8, s));
  if (S.npt > 0)
    hipLaunchKernelGGL(k_pattern_bits, dim3(blocks_of(S.npt, kThreads)), dim3(kThreads), 0, s,
                       S.npt, d_pt_ptr, d_pt_block, d_obs_pose, d_pose_block, W, d_bits);
  std::vector<unsigned long long> bits(nwords);
  SGPU_HIP(hipMemcpyAsync(bits.data(), d_bits, nwords * 8, hipMemcpyDeviceToHost, s));
  SGPU_HIP(hipStreamSynchronize(s));
  for (const auto& rc : pattern_extras(P, S))
    bits[size_t(rc.first) * W + (rc.second >> 6)] |= 1ull << (rc.second & 63);
  S.s_ptr.assign(nb + 1, 0);
  std::vector<int32_t> wpre(nwords);
  for (int32_t r = 0; r < nb; ++r) {
    int32_t at = S.s_ptr[r];
    for (int32_t w = 0; w < W; ++w) {
      wpre[size_t(r) * W + w] = at;
      unsigned long long m = bits[size_t(r) * W + w];
      while (m) {
        const int c = __builtin_ctzll(m);
        S.s_col.push_back(w * 64 + c);
        m &= m - 1;
        ++at;
      }
    }
    S.s_ptr[r + 1] = at;
  }
  const int64_t nslot = int64_t(S.s_col.size());
  if (nslot > int64_t(UINT32_MAX)) throw Error(SGPU_ERR_UNSUPPORTED, "too many S blocks");
  SGPU_HIP(hipMemcpyAsync(d_bits, bits.data(), nwords * 8, hipMemcpyHostToDevice, s));
  int32_t* d_wpre = upload_vec(tmp, wpre);
  // pairs: count per point, offsets, fill in point order, stable sort by slot
  std::vector<int32_t> cam_pos(size_t(S.nobs), -1);
  for (int32_t q = 0; q < int32_t(S.cam_obs.size()); ++q) cam_pos[S.cam_obs[q]] = q;
  int32_t* d_cam_pos = upload_vec(tmp, cam_pos);
  int64_t* d_off = tmp.alloc<int64_t>(size_t(S.npt) + 1);
  const int gp = std::max(1, blocks_of(S.npt, kThreads));
  int64_t total = 0;
  if (S.npt > 0) {
    hipLaunchKernelGGL(k_pairs, dim3(gp), dim3(kThreads), 0, s, 0, S.npt, d_pt_ptr, d_pt_block,
                       d_obs_pose, d_pose_block, d_cam_pos, d_bits, d_wpre, W, d_off, nullptr,
                       nullptr, nullptr, nullptr);
    SGPU_HIP(hipMemsetAsync(d_off + S.npt, 0, 8, s));
    exclusive_scan_i64(s, tmp, d_off, int64_t(S.npt) + 1);
    SGPU_HIP(hipMemcpyAsync(&total, d_off + S.npt, 8, hipMemcpyDeviceToHost, s));
    SGPU_HIP(hipStreamSynchronize(s));
  }
  if (total > int64_t(INT32_MAX) / 2) throw Error(SGPU_ERR_UNSUPPORTED, "too many Schur pairs");
  *d_pairs = arena.alloc<int32_t>(size_t(std::max<int64_t>(2 * total, 1)));
  int64_t* d_slot_count = tmp.alloc<int64_t>(size_t(nslot) + 1);
  SGPU_HIP(hipMemsetAsync(d_slot_count, 0, (size_t(nslot) + 1) * 8, s));
  if (total > 0) {
    uint32_t* key = tmp.alloc<uint32_t>(size_t(total));
    uint32_t* val = tmp.alloc<uint32_t>(size_t(total));
    uint32_t* key2 = tmp.alloc<uint32_t>(size_t(total));
    uint32_t* val2 = tmp.alloc<uint32_t>(size_t(total));
    int32_t* pi = tmp.alloc<int32_t>(size_t(total));
    int32_t* pj = tmp.alloc<int32_t>(size_t(total));
    hipLaunchKernelGGL(k_pairs, dim3(gp), dim3(kThreads), 0, s, 1, S.npt, d_pt_ptr, d_pt_block,
                       d_obs_pose, d_pose_block, d_cam_pos, d_bits, d_wpre, W, d_off, key, val,
                       pi, pj);
    radix_sort_pairs(s, tmp, key, val, key2, val2, total, bit_width(uint64_t(nslot)));
    hipLaunchKernelGGL(k_pairs_gather, dim3(blocks_of(total, kThreads)), dim3(kThreads), 0, s, total,
                       key, val, pi, pj, *d_pairs, d_slot_count);
  }
  exclusive_scan_i64(s, tmp, d_slot_count, nslot + 1);
  S.pair_ptr.resize(size_t(nslot) + 1);
  SGPU_HIP(hipMemcpyAsync(S.pair_ptr.data(), d_slot_count, (size_t(nslot) + 1) * 8,
                          hipMemcpyDeviceToHost, s));
  SGPU_HIP(hipStreamSynchronize(s));
  S.num_schur_pairs = total + S.cam_ptr[S.nb];
  analyze_chunks(S);
  analyze_finish(P, S);
  return S;
}

}  // namespace sgpu

namespace {

template <typename T>
int64_t count_diff(const std::vector<T>& a, const std::vector<T>& b) {
  if (a.size() != b.size()) return int64_t(std::max(a.size(), b.size()));
  int64_t d = 0;
  for (size_t i = 0; i < a.size(); ++i) d += !(a[i] == b[i]);
  return d;
}

}  // namespace

extern "C" sgpu_status sgpu_symbolic_check(const sgpu_var_family* families, int64_t num_families,
                                           const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                           int32_t device, int64_t out[3], double ms[2]) {
  using namespace sgpu;
  return guarded([&] {
    if (!out) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(SGPU_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) throw Error(SGPU_ERR_NO_DEVICE, "device ordinal out of range");
    SGPU_HIP(hipSetDevice(device));
    hipStream_t s = nullptr;
    SGPU_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    const HostProblem Ph = build_host_problem(families, num_families, cost_fns, num_cost_fns);
    const Symbolic Sh = analyze(Ph);
    const auto t1 = clk::now();
    DeviceArena arena;
    int32_t* d_pairs = nullptr;
    HostProblem Pd = build_host_problem_sorted(families, num_families, cost_fns, num_cost_fns,
                                               nullptr, 0, false);
    sort_terms_on_device(Pd, s);
    Symbolic Sd = analyze_device(Pd, s, arena, &d_pairs);
    const auto t2 = clk::now();
    Sd.pairs.resize(Sh.pairs.size());
    if (!Sh.pairs.empty() && Sd.pair_ptr.size() == Sh.pair_ptr.size() &&
        Sd.pair_ptr.back() * 2 == int64_t(Sh.pairs.size()))
      SGPU_HIP(hipMemcpy(Sd.pairs.data(), d_pairs, Sh.pairs.size() * 4, hipMemcpyDeviceToHost));
    int64_t d = 0;
    d += count_diff(Ph.obs_pose, Pd.obs_pose) + count_diff(Ph.obs_point, Pd.obs_point) +
         count_diff(Ph.obs_intr, Pd.obs_intr) + count_diff(Ph.obs_uv, Pd.obs_uv);
    d += count_diff(Sh.pt_ptr, Sd.pt_ptr) + count_diff(Sh.pt_nu, Sd.pt_nu) + count_diff(Sh.cam_nu, Sd.cam_nu);
    d += count_diff(Sh.cam_ptr, Sd.cam_ptr) + count_diff(Sh.cam_obs, Sd.cam_obs);
    d += count_diff(Sh.s_ptr, Sd.s_ptr) + count_diff(Sh.s_col, Sd.s_col);
    d += count_diff(Sh.pair_ptr, Sd.pair_ptr) + count_diff(Sh.pairs, Sd.pairs);
    d += count_diff(Sh.own_chunks, Sd.own_chunks) + count_diff(Sh.cam_chunk_ptr, Sd.cam_chunk_ptr);
    d += count_diff(Sh.intr_slot, Sd.intr_slot) + (Sh.intr_diag_slot != Sd.intr_diag_slot);
    d += count_diff(Sh.perm, Sd.perm) + count_diff(Sh.etree_parent, Sd.etree_parent);
    d += count_diff(Sh.s_off, Sd.s_off) + count_diff(Sh.s_ld, Sd.s_ld);
    d += count_diff(Sh.sn_c0, Sd.sn_c0) + count_diff(Sh.sn_rows, Sd.sn_rows) +
         count_diff(Sh.launch_tiles, Sd.launch_tiles);
    d += (Sh.num_schur_pairs != Sd.num_schur_pairs) + (Sh.nnz_l != Sd.nnz_l) +
         (Sh.n_dense_tail != Sd.n_dense_tail);
    out[0] = d;
    out[1] = Sh.num_schur_pairs;
    out[2] = int64_t(Sh.s_col.size());
    if (ms) {
      ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
      ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
  });
}
