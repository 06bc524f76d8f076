// Device utilities: initial ranking state, list validation, exclusive scan,
// stable LSD radix sort of (key, value) pairs, order-preserving pack/unpack.
#include <algorithm>

#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;

// list_ranking.cpp:60-80
__global__ void __launch_bounds__(kThreads)
initial_state_kernel(const uint64_t* __restrict__ succ, uint64_t n, uint64_t* __restrict__ root,
                     int64_t* __restrict__ rank) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = succ[i];
  root[i] = s;
  rank[i] = (s == i) ? 0 : 1;
}

// is_list (successor_utils.cpp:31-63): no element may have two predecessors.
__global__ void __launch_bounds__(kThreads)
indegree_kernel(const uint64_t* __restrict__ succ, uint64_t n, uint32_t* __restrict__ indeg,
                uint32_t* err) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = succ[i];
  if (s >= n) {
    raise_err(err, kErrRange);
    return;
  }
  if (s == i) return;
  if (atomicAdd(&indeg[s], 1u) != 0) raise_err(err, kErrNotList);
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ scan ---
constexpr unsigned kScanItems = 16;
constexpr unsigned kScanTile = kThreads * kScanItems;

__device__ uint64_t block_exclusive_scan(uint64_t x, uint64_t* s_warp, uint64_t* total) {
  const unsigned lane = __lane_id();
  const unsigned wave = threadIdx.x / 64;
  uint64_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(incl, d);
    if (lane >= static_cast<unsigned>(d)) incl += y;
  }
  if (lane == 63) s_warp[wave] = incl;
  __syncthreads();
  uint64_t off = 0, tot = 0;
#pragma unroll
  for (unsigned w = 0; w < kThreads / 64; w++) {
    if (w < wave) off += s_warp[w];
    tot += s_warp[w];
  }
  *total = tot;
  __syncthreads();
  return off + incl - x;
}

__global__ void __launch_bounds__(kThreads)
scan_tile_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n,
                 uint64_t* __restrict__ tile_sums) {
  __shared__ uint64_t s_data[kScanTile];
  __shared__ uint64_t s_warp[kThreads / 64];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  for (unsigned r = 0; r < kScanItems; r++) {
    uint64_t i = base + r * kThreads + threadIdx.x;
    s_data[r * kThreads + threadIdx.x] = i < n ? in[i] : 0;
  }
  __syncthreads();
  uint64_t acc = 0;
  uint64_t vals[kScanItems];
#pragma unroll
  for (unsigned j = 0; j < kScanItems; j++) {
    vals[j] = s_data[threadIdx.x * kScanItems + j];
    acc += vals[j];
  }
  uint64_t total = 0;
  uint64_t off = block_exclusive_scan(acc, s_warp, &total);
#pragma unroll
  for (unsigned j = 0; j < kScanItems; j++) {
    s_data[threadIdx.x * kScanItems + j] = off;
    off += vals[j];
  }
  __syncthreads();
  for (unsigned r = 0; r < kScanItems; r++) {
    uint64_t i = base + r * kThreads + threadIdx.x;
    if (i < n) out[i] = s_data[r * kThreads + threadIdx.x];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kThreads)
scan_add_kernel(uint64_t* __restrict__ out, uint64_t n, const uint64_t* __restrict__ tile_off) {
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  const uint64_t add = tile_off[blockIdx.x];
  for (unsigned r = 0; r < kScanItems; r++) {
    uint64_t i = base + r * kThreads + threadIdx.x;
    if (i < n) out[i] += add;
  }
}

// ------------------------------------------------------------ radix sort ---
// LSD radix sort, up to 8 bits per pass.  Each pass: a per-tile histogram, a
// scan of the (digit-major) histogram table, then a scatter that first sorts
// the tile by digit in LDS (stable) and writes each digit's run contiguously,
// so global writes are coalesced runs instead of scattered words.
constexpr unsigned kMaxRadixBits = 8;
constexpr unsigned kMaxBins = 1u << kMaxRadixBits;
constexpr unsigned kSortItems = 16;
constexpr unsigned kSortTile = kThreads * kSortItems;
constexpr unsigned kWaves = kThreads / 64;

template <typename K>
__global__ void __launch_bounds__(kThreads)
radix_hist_kernel(const K* __restrict__ keys, uint64_t n, unsigned shift, unsigned bits,
                  uint64_t nblocks, uint64_t* __restrict__ counts) {
  __shared__ uint32_t s_hist[kMaxBins];
  const unsigned bins = 1u << bits;
  for (unsigned d = threadIdx.x; d < bins; d += kThreads) s_hist[d] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  for (unsigned r = 0; r < kSortItems; r++) {
    uint64_t i = base + r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&s_hist[static_cast<unsigned>(keys[i] >> shift) & (bins - 1)], 1u);
  }
  __syncthreads();
  for (unsigned d = threadIdx.x; d < bins; d += kThreads)
    counts[static_cast<uint64_t>(d) * nblocks + blockIdx.x] = s_hist[d];
}

template <typename K, typename V>
__global__ void __launch_bounds__(kThreads)
radix_scatter_kernel(const K* __restrict__ keys, const V* __restrict__ vals, uint64_t n,
                     unsigned shift, unsigned bits, uint64_t nblocks,
                     const uint64_t* __restrict__ offsets, K* __restrict__ keys_out,
                     V* __restrict__ vals_out) {
  __shared__ uint32_t s_run[kMaxBins];      // next local slot of each digit
  __shared__ uint32_t s_start[kMaxBins];    // local start of each digit's run
  __shared__ uint32_t s_wcnt[kWaves][kMaxBins];
  __shared__ K s_keys[kSortTile];
  __shared__ V s_vals[kSortTile];
  __shared__ uint32_t s_scan[kThreads / 64];
  const unsigned bins = 1u << bits;
  const unsigned wave = threadIdx.x / 64;
  const uint64_t lt = lanemask_lt();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  const unsigned count = static_cast<unsigned>(n - base < kSortTile ? n - base : kSortTile);

  // local histogram -> local run starts (exclusive scan over digits)
  for (unsigned d = threadIdx.x; d < kMaxBins; d += kThreads) {
    s_run[d] = 0;
    for (unsigned w = 0; w < kWaves; w++) s_wcnt[w][d] = 0;
  }
  __syncthreads();
  for (unsigned r = 0; r < kSortItems; r++) {
    const unsigned li = r * kThreads + threadIdx.x;
    if (li < count) atomicAdd(&s_run[static_cast<unsigned>(keys[base + li] >> shift) & (bins - 1)], 1u);
  }
  __syncthreads();
  {
    // kThreads == kMaxBins: thread t owns digit t
    const unsigned c = s_run[threadIdx.x];
    unsigned incl = c;
    const unsigned lane = __lane_id();
    for (int d = 1; d < 64; d <<= 1) {
      unsigned y = __shfl_up(incl, d);
      if (lane >= static_cast<unsigned>(d)) incl += y;
    }
    if (lane == 63) s_scan[wave] = incl;
    __syncthreads();
    unsigned off = 0;
    for (unsigned w = 0; w < wave; w++) off += s_scan[w];
    s_start[threadIdx.x] = off + incl - c;
    s_run[threadIdx.x] = off + incl - c;
  }
  __syncthreads();

  // stable local ranking in (round, wave, lane) order; place into LDS
  for (unsigned r = 0; r < kSortItems; r++) {
    const unsigned li = r * kThreads + threadIdx.x;
    const bool valid = li < count;
    K key = 0;
    V val = 0;
    unsigned d = 0;
    if (valid) {
      key = keys[base + li];
      val = vals[base + li];
      d = static_cast<unsigned>(key >> shift) & (bins - 1);
    }
    uint64_t peers = __ballot(valid);
    for (unsigned b = 0; b < bits; b++) {
      const uint64_t bb = __ballot(valid && ((d >> b) & 1u));
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const unsigned rank_in_wave = __popcll(peers & lt);
    if (valid && rank_in_wave == 0) s_wcnt[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      unsigned pos = s_run[d] + rank_in_wave;
      for (unsigned w = 0; w < wave; w++) pos += s_wcnt[w][d];
      s_keys[pos] = key;
      s_vals[pos] = val;
    }
    __syncthreads();
    {
      const unsigned dd = threadIdx.x;
      uint32_t t = 0;
      for (unsigned w = 0; w < kWaves; w++) {
        t += s_wcnt[w][dd];
        s_wcnt[w][dd] = 0;
      }
      s_run[dd] += t;
    }
    __syncthreads();
  }

  // coalesced write-out: consecutive slots of a digit go to consecutive addresses
  for (unsigned li = threadIdx.x; li < count; li += kThreads) {
    const K key = s_keys[li];
    const unsigned d = static_cast<unsigned>(key >> shift) & (bins - 1);
    const uint64_t dst = offsets[static_cast<uint64_t>(d) * nblocks + blockIdx.x] + (li - s_start[d]);
    keys_out[dst] = key;
    vals_out[dst] = s_vals[li];
  }
}

template <typename K, typename V>
void radix_sort_pairs(Context* ctx, K* keys, V* vals, uint64_t n, unsigned key_bits,
                      hipStream_t st) {
  if (n <= 1) return;
  Phase ph(ctx, "radix_sort", st);
  Scratch scratch(ctx);
  K* k2 = scratch.get<K>(n);
  V* v2 = scratch.get<V>(n);
  const uint64_t nblocks = (n + kSortTile - 1) / kSortTile;
  uint64_t* counts = scratch.get<uint64_t>(nblocks * kMaxBins);
  uint64_t* offs = scratch.get<uint64_t>(nblocks * kMaxBins);
  K* src_k = keys;
  V* src_v = vals;
  K* dst_k = k2;
  V* dst_v = v2;
  if (key_bits == 0) key_bits = 1;
  const unsigned passes = (key_bits + kMaxRadixBits - 1) / kMaxRadixBits;
  const unsigned bits = (key_bits + passes - 1) / passes;  // even split, <= 8
  for (unsigned p = 0; p < passes; p++) {
    const unsigned shift = p * bits;
    const unsigned pbits = std::min(bits, key_bits - shift);
    const uint64_t nbins = uint64_t{1} << pbits;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(radix_hist_kernel<K>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(kThreads), 0, st, src_k, n, shift, pbits, nblocks, counts);
    exclusive_scan_u64(ctx, counts, offs, nblocks * nbins, st);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(radix_scatter_kernel<K, V>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(kThreads), 0, st, src_k, src_v, n, shift, pbits, nblocks, offs, dst_k, dst_v);
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
  }
  if (src_k != keys) {
    KC_HIP(hipMemcpyAsync(keys, src_k, n * sizeof(K), hipMemcpyDeviceToDevice, st));
    KC_HIP(hipMemcpyAsync(vals, src_v, n * sizeof(V), hipMemcpyDeviceToDevice, st));
  }
}

// ------------------------------------------------------------- pack ---------
// pack.hpp:23-80 on one device: packed id of active[k] is k.
__global__ void __launch_bounds__(kThreads)
pack_kernel(const uint64_t* __restrict__ succ, const int64_t* __restrict__ rank, uint64_t n,
            const uint64_t* __restrict__ active, uint64_t na, uint64_t* __restrict__ psucc,
            int64_t* __restrict__ prank, uint32_t* err) {
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= na) return;
  const uint64_t a = active[k];
  if (a >= n || (k > 0 && active[k - 1] >= a)) {
    raise_err(err, kErrRange);
    return;
  }
  const uint64_t s = succ[a];
  uint64_t lo = 0, hi = na;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (active[mid] < s) lo = mid + 1; else hi = mid;
  }
  if (lo >= na || active[lo] != s) {
    raise_err(err, kErrRange);
    return;
  }
  psucc[k] = lo;
  prank[k] = rank[a];
}

// pack.hpp:81-175 (write_unpacked)
__global__ void __launch_bounds__(kThreads)
unpack_kernel(const uint64_t* __restrict__ psucc, const int64_t* __restrict__ prank,
              const uint64_t* __restrict__ active, uint64_t na, uint64_t n,
              uint64_t* __restrict__ succ, int64_t* __restrict__ rank, uint32_t* err) {
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= na) return;
  const uint64_t a = active[k];
  const uint64_t ps = psucc[k];
  if (a >= n || ps >= na) {
    raise_err(err, kErrRange);
    return;
  }
  succ[a] = active[ps];
  rank[a] = prank[k];
}

}  // namespace

uint64_t exclusive_scan_u64(Context* ctx, const uint64_t* in, uint64_t* out, uint64_t n,
                            hipStream_t st) {
  if (n == 0) return 0;
  Scratch scratch(ctx);
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  uint64_t* sums = scratch.get<uint64_t>(tiles);
  hipLaunchKernelGGL(scan_tile_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0, st, in,
                     out, n, sums);
  if (tiles > 1) {
    uint64_t* soff = scratch.get<uint64_t>(tiles);
    exclusive_scan_u64(ctx, sums, soff, tiles, st);
    hipLaunchKernelGGL(scan_add_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0, st,
                       out, n, soff);
  }
  return 0;
}

void radix_sort_pairs_u32(Context* ctx, uint32_t* keys, uint32_t* vals, uint64_t n,
                          unsigned key_bits, hipStream_t st) {
  radix_sort_pairs<uint32_t, uint32_t>(ctx, keys, vals, n, key_bits, st);
}

void radix_sort_pairs_u64(Context* ctx, uint64_t* keys, uint64_t* vals, uint64_t n,
                          unsigned key_bits, hipStream_t st) {
  radix_sort_pairs<uint64_t, uint64_t>(ctx, keys, vals, n, key_bits, st);
}

void validate_list(Context* ctx, const uint64_t* succ, uint64_t n, hipStream_t st) {
  Scratch scratch(ctx);
  uint32_t* indeg = scratch.get<uint32_t>(n);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, indeg, n, 0u);
  hipLaunchKernelGGL(indegree_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, succ, n,
                     indeg, ctx->d_err);
  check_err(ctx, st);
}

}  // namespace kc

using kc::Context;

extern "C" {

kc_status kc_set_initial_ranking_state(void* vctx, const uint64_t* succ, uint64_t n,
                                       uint64_t* root_out, int64_t* rank_out, void* stream);
kc_status kc_pack(void* vctx, const uint64_t* succ, const int64_t* rank, uint64_t n,
                  const uint64_t* active, uint64_t nactive, uint64_t* psucc_out,
                  int64_t* prank_out, void* stream);
kc_status kc_unpack(void* vctx, const uint64_t* psucc, const int64_t* prank,
                    const uint64_t* active, uint64_t nactive, uint64_t* succ_inout,
                    int64_t* rank_inout, uint64_t n, void* stream);
}

namespace kc {
// Launch helpers used by api_util.cpp.
void initial_state_device(const uint64_t* succ, uint64_t n, uint64_t* root, int64_t* rank,
                          hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(initial_state_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, succ,
                     n, root, rank);
}
void pack_device(Context* ctx, const uint64_t* succ, const int64_t* rank, uint64_t n,
                 const uint64_t* active, uint64_t na, uint64_t* psucc, int64_t* prank,
                 hipStream_t st) {
  if (na == 0) return;
  hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(na, kThreads)), dim3(kThreads), 0, st, succ, rank, n,
                     active, na, psucc, prank, ctx->d_err);
  check_err(ctx, st);
}
void unpack_device(Context* ctx, const uint64_t* psucc, const int64_t* prank,
                   const uint64_t* active, uint64_t na, uint64_t n, uint64_t* succ, int64_t* rank,
                   hipStream_t st) {
  if (na == 0) return;
  hipLaunchKernelGGL(unpack_kernel, dim3(blocks_for(na, kThreads)), dim3(kThreads), 0, st, psucc,
                     prank, active, na, n, succ, rank, ctx->d_err);
  check_err(ctx, st);
}
}  // namespace kc
