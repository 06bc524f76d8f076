// Seeded synthetic inputs, generated on the device one element per thread.
// The arithmetic (splitmix64, a 4-round Feistel permutation with cycle
// walking, weight and tree-parent hashes) is the same as the CPU generators
// used by the tests, so both produce identical arrays.
#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;
constexpr int kFeistelRounds = 4;
constexpr uint64_t kLocalBlock = 10;

struct Perm {
  uint64_t n;
  unsigned h;
  uint64_t mask;
  uint64_t keys[kFeistelRounds];
};

Perm make_perm(uint64_t n, uint64_t seed) {
  Perm p{};
  p.n = n;
  unsigned b = 2;
  while (b < 64 && (uint64_t{1} << b) < n) b++;
  p.h = (b + 1) / 2;
  p.mask = (uint64_t{1} << p.h) - 1;
  for (int k = 0; k < kFeistelRounds; k++) p.keys[k] = splitmix64(seed + static_cast<uint64_t>(k));
  return p;
}

__device__ __forceinline__ uint64_t perm_fwd(const Perm& p, uint64_t x) {
  do {
    uint64_t l = x >> p.h, r = x & p.mask;
#pragma unroll
    for (int k = 0; k < kFeistelRounds; k++) {
      uint64_t f = splitmix64(r ^ p.keys[k]) & p.mask;
      uint64_t t = l ^ f;
      l = r;
      r = t;
    }
    x = (l << p.h) | r;
  } while (x >= p.n);
  return x;
}

__device__ __forceinline__ uint64_t perm_inv(const Perm& p, uint64_t y) {
  do {
    uint64_t l = y >> p.h, r = y & p.mask;
#pragma unroll
    for (int k = kFeistelRounds - 1; k >= 0; k--) {
      uint64_t f = splitmix64(l ^ p.keys[k]) & p.mask;
      uint64_t t = r ^ f;
      r = l;
      l = t;
    }
    y = (l << p.h) | r;
  } while (y >= p.n);
  return y;
}

__device__ __forceinline__ int64_t gen_weight(uint64_t v, uint64_t wseed) {
  return 1 + static_cast<int64_t>(splitmix64(wseed ^ (v * 0xD1B54A32D192ED03ull)) % 100);
}

__global__ void __launch_bounds__(kThreads)
gen_list_kernel(Perm p, uint64_t wseed, int weighted, uint64_t* __restrict__ succ,
                int64_t* __restrict__ w) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (v >= p.n) return;
  uint64_t i = perm_inv(p, v);
  bool tail = (i + 1 == p.n);
  succ[v] = tail ? v : perm_fwd(p, i + 1);
  if (w) w[v] = tail ? 0 : (weighted ? gen_weight(v, wseed) : 1);
}

// Blocks of 10 consecutive ids, visited in permuted block order.
__global__ void __launch_bounds__(kThreads)
gen_list_local_kernel(uint64_t n, Perm pb, uint64_t wseed, int weighted,
                      uint64_t* __restrict__ succ, int64_t* __restrict__ w) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (v >= n) return;
  uint64_t b = v / kLocalBlock;
  uint64_t bend = (b + 1) * kLocalBlock < n ? (b + 1) * kLocalBlock : n;
  bool tail = false;
  uint64_t s;
  if (v + 1 < bend) {
    s = v + 1;
  } else {
    uint64_t q = perm_inv(pb, b);
    if (q + 1 == pb.n) {
      s = v;
      tail = true;
    } else {
      s = perm_fwd(pb, q + 1) * kLocalBlock;
    }
  }
  succ[v] = s;
  if (w) w[v] = tail ? 0 : (weighted ? gen_weight(v, wseed) : 1);
}

__global__ void __launch_bounds__(kThreads)
gen_tree_kernel(Perm p, uint64_t pseed, uint64_t* __restrict__ parent) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (v >= p.n) return;
  uint64_t i = perm_inv(p, v);
  if (i == 0) {
    parent[v] = v;
    return;
  }
  uint64_t gp = splitmix64(pseed ^ (i * 0x9E3779B97F4A7C15ull)) % i;
  parent[v] = perm_fwd(p, gp);
}

// Segmented lists: one block-stride loop per segment (a workgroup per segment
// at a time), local u32 ids.  Same per-segment arithmetic as gen_list /
// gen_list_local with the segment's seed.
__global__ void __launch_bounds__(kThreads)
gen_segmented_kernel(const uint64_t* __restrict__ seg_off, uint64_t nseg,
                     const uint64_t* __restrict__ seeds, const uint8_t* __restrict__ local_flags,
                     int weighted, uint32_t* __restrict__ succ, int64_t* __restrict__ w) {
  for (uint64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const uint64_t base = seg_off[s];
    const uint64_t n = seg_off[s + 1] - base;
    if (n == 0) continue;
    const uint64_t seed = seeds[s];
    const uint64_t wseed = splitmix64(seed ^ 0x5EEDull);
    const bool local = local_flags && local_flags[s];
    Perm p{};
    {
      const uint64_t dom = local ? (n + kLocalBlock - 1) / kLocalBlock : n;
      unsigned b = 2;
      while (b < 64 && (uint64_t{1} << b) < dom) b++;
      p.n = dom;
      p.h = (b + 1) / 2;
      p.mask = (uint64_t{1} << p.h) - 1;
#pragma unroll
      for (int k = 0; k < kFeistelRounds; k++) p.keys[k] = splitmix64(seed + static_cast<uint64_t>(k));
    }
    for (uint64_t v = threadIdx.x; v < n; v += kThreads) {
      uint64_t sv;
      bool tail = false;
      if (local) {
        uint64_t b = v / kLocalBlock;
        uint64_t bend = (b + 1) * kLocalBlock < n ? (b + 1) * kLocalBlock : n;
        if (v + 1 < bend) {
          sv = v + 1;
        } else {
          uint64_t q = perm_inv(p, b);
          if (q + 1 == p.n) {
            sv = v;
            tail = true;
          } else {
            sv = perm_fwd(p, q + 1) * kLocalBlock;
          }
        }
      } else {
        uint64_t i = perm_inv(p, v);
        tail = (i + 1 == n);
        sv = tail ? v : perm_fwd(p, i + 1);
      }
      succ[base + v] = static_cast<uint32_t>(sv);
      if (w) w[base + v] = tail ? 0 : (weighted ? gen_weight(v, wseed) : 1);
    }
  }
}

}  // namespace

void gen_list_device(uint64_t n, uint64_t seed, int weighted, int local, uint64_t* succ,
                     int64_t* w, hipStream_t st) {
  if (n == 0) return;
  const uint64_t wseed = splitmix64(seed ^ 0x5EEDull);
  if (local) {
    Perm pb = make_perm((n + kLocalBlock - 1) / kLocalBlock, seed);
    hipLaunchKernelGGL(gen_list_local_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st,
                       n, pb, wseed, weighted, succ, w);
  } else {
    Perm p = make_perm(n, seed);
    hipLaunchKernelGGL(gen_list_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, p,
                       wseed, weighted, succ, w);
  }
  KC_HIP(hipGetLastError());
}

void gen_tree_device(uint64_t n, uint64_t seed, uint64_t* parent, hipStream_t st) {
  if (n == 0) return;
  Perm p = make_perm(n, seed);
  const uint64_t pseed = splitmix64(seed ^ 0x7EEull);
  hipLaunchKernelGGL(gen_tree_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, p, pseed,
                     parent);
  KC_HIP(hipGetLastError());
}

void gen_lists_segmented_device(const uint64_t* seg_off, uint64_t nseg, uint64_t total,
                                const uint64_t* seeds, const uint8_t* local_flags, int weighted,
                                uint32_t* succ, int64_t* w, hipStream_t st) {
  if (nseg == 0 || total == 0) return;
  const unsigned grid = static_cast<unsigned>(nseg < 65536 ? nseg : 65536);
  hipLaunchKernelGGL(gen_segmented_kernel, dim3(grid), dim3(kThreads), 0, st, seg_off, nseg, seeds,
                     local_flags, weighted, succ, w);
  KC_HIP(hipGetLastError());
}

}  // namespace kc
