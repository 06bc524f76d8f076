// Random-gather ceiling probe: dst[i] = src[idx[i]], idx a seeded permutation.
#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {
constexpr unsigned kThreads = 256;

__global__ void __launch_bounds__(kThreads)
perm_idx_kernel(uint64_t* __restrict__ idx, uint64_t n, uint64_t seed) {
  // multiplicative permutation of [0, n) for n a power of two: odd multiplier + xor-shift free
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint64_t mask = n - 1;
  uint64_t x = (i * (splitmix64(seed) | 1)) & mask;
  x = (x ^ (x >> 7)) & mask;               // invertible on [0, 2^k)
  x = (x * 0x9E3779B97F4A7C15ull + (seed & mask)) & mask;
  idx[i] = x;
}

__global__ void __launch_bounds__(kThreads)
table_idx_kernel(uint64_t* __restrict__ idx, uint64_t n, uint64_t m, uint64_t seed) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) idx[i] = splitmix64(seed ^ (i * 0x9E3779B97F4A7C15ull)) % m;
}

__global__ void __launch_bounds__(kThreads)
fill_kernel(uint64_t* __restrict__ a, uint64_t n) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) a[i] = i * 3 + 1;
}

__global__ void __launch_bounds__(kThreads)
gather_kernel(const uint64_t* __restrict__ src, const uint64_t* __restrict__ idx,
              uint64_t* __restrict__ dst, uint64_t n) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void __launch_bounds__(kThreads)
gather_nt_kernel(const uint64_t* __restrict__ src, const uint64_t* __restrict__ idx,
                 uint64_t* __restrict__ dst, uint64_t n) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) dst[i] = __builtin_nontemporal_load(&src[idx[i]]);
}
// 16-byte records on one random cycle: rec[idx[k]] = {idx[k + 1], 1}
__global__ void __launch_bounds__(kThreads)
cycle_init_kernel(Rec* __restrict__ rec, const uint64_t* __restrict__ idx, uint64_t n) {
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (k < n) rec[idx[k]] = Rec{idx[(k + 1) & (n - 1)], 1};
}

// independent random record accesses at idx[i]: 0 read, 1 write, 2 read + claim write
__global__ void __launch_bounds__(kThreads)
record_access_kernel(Rec* __restrict__ rec, const uint64_t* __restrict__ idx, int64_t* __restrict__ out,
                     uint64_t n, int pattern) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint64_t p = idx[i];
  if (pattern == 0) {
    out[i] = rec[p].w;
  } else if (pattern == 1) {
    rec[p] = Rec{idx[(i + 1) & (n - 1)], 1};
  } else {
    const Rec v = rec[p];
    rec[p] = Rec{v.s | kRootFlag, v.w};
  }
}

// dependent walks: walker j follows the cycle for L steps from idx[j * L]
// (the walkers cover the cycle exactly once); claim != 0 writes a 16-byte
// claim into every visited record, as the ruler chase does
__global__ void __launch_bounds__(kThreads)
record_walk_kernel(Rec* __restrict__ rec, const uint64_t* __restrict__ idx, int64_t* __restrict__ out,
                   uint64_t walkers, uint64_t L, int claim) {
  uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= walkers) return;
  uint64_t cur = idx[j * L];
  int64_t d = 0;
  for (uint64_t t = 0; t < L; ++t) {
    const Rec v = rec[cur];
    if (claim) rec[cur] = Rec{v.s | kRootFlag, v.w};
    d += v.w;
    cur = v.s & kIdMask;
  }
  out[j] = d;
}

// The same cycle and walks over 8-byte records {u32 successor, i32 weight}
// (the narrow level-0 form); the claim sets the record's root bit.
__global__ void __launch_bounds__(kThreads)
cycle_init8_kernel(Rec8* __restrict__ rec, const uint64_t* __restrict__ idx, uint64_t n) {
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (k < n) rec[idx[k]] = Rec8{static_cast<uint32_t>(idx[(k + 1) & (n - 1)]), 1};
}

__global__ void __launch_bounds__(kThreads)
record_walk8_kernel(Rec8* __restrict__ rec, const uint64_t* __restrict__ idx, int64_t* __restrict__ out,
                    uint64_t walkers, uint64_t L, int claim) {
  uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= walkers) return;
  uint64_t cur = idx[j * L];
  int64_t d = 0;
  for (uint64_t t = 0; t < L; ++t) {
    const Rec8 v = rec[cur];
    if (claim) rec[cur] = Rec8{v.s | kRoot8, v.w};
    d += v.w;
    cur = v.s & (kAux8 - 1);
  }
  out[j] = d;
}
}  // namespace
}  // namespace kc

// mode 0: plain loads from ordinary device memory; 1: non-temporal loads;
// 2: plain loads from an uncached allocation (kept until the process ends);
// 3: plain loads from a fine-grained allocation (kept until the process ends).
extern "C" kc_status kc_probe_random_gather_mode(void* vctx, uint64_t n, uint32_t reps, int mode,
                                                 double* gbps_out, double* ms_out) {
  auto* ctx = static_cast<kc::Context*>(vctx);
  if (!ctx || n == 0 || (n & (n - 1)) != 0 || reps == 0 || mode < 0 || mode > 3) return KC_EINVAL;
  try {
    KC_HIP(hipSetDevice(ctx->device));
    kc::Scratch scratch(ctx);
    uint64_t* src = nullptr;
    if (mode == 2 || mode == 3) {
      // special allocations are never freed (an uncached buffer that a kernel has read must
      // stay allocated for the life of the process)
      static uint64_t* keep[4] = {nullptr, nullptr, nullptr, nullptr};
      static uint64_t keep_n[4] = {0, 0, 0, 0};
      if (keep_n[mode] < n) {
        void* p = nullptr;
        KC_HIP(hipExtMallocWithFlags(&p, n * sizeof(uint64_t),
                                     mode == 2 ? hipDeviceMallocUncached : hipDeviceMallocFinegrained));
        keep[mode] = static_cast<uint64_t*>(p);
        keep_n[mode] = n;
      }
      src = keep[mode];
    } else {
      src = scratch.get<uint64_t>(n);
    }
    uint64_t* idx = scratch.get<uint64_t>(n);
    uint64_t* dst = scratch.get<uint64_t>(n);
    hipStream_t st = ctx->stream;
    const unsigned blocks = kc::blocks_for(n, 256);
    hipLaunchKernelGGL(kc::perm_idx_kernel, dim3(blocks), dim3(256), 0, st, idx, n, 12345ull);
    hipLaunchKernelGGL(kc::fill_kernel, dim3(blocks), dim3(256), 0, st, src, n);
    hipLaunchKernelGGL(kc::gather_kernel, dim3(blocks), dim3(256), 0, st, src, idx, dst, n);
    hipEvent_t a, b;
    KC_HIP(hipEventCreate(&a));
    KC_HIP(hipEventCreate(&b));
    KC_HIP(hipEventRecord(a, st));
    for (uint32_t r = 0; r < reps; r++) {
      if (mode == 1)
        hipLaunchKernelGGL(kc::gather_nt_kernel, dim3(blocks), dim3(256), 0, st, src, idx, dst, n);
      else
        hipLaunchKernelGGL(kc::gather_kernel, dim3(blocks), dim3(256), 0, st, src, idx, dst, n);
    }
    KC_HIP(hipEventRecord(b, st));
    KC_HIP(hipEventSynchronize(b));
    float ms = 0;
    KC_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const double per = ms / reps;
    if (ms_out) *ms_out = per;
    if (gbps_out) *gbps_out = 24.0 * static_cast<double>(n) / (per * 1e-3) / 1e9;
    return KC_OK;
  } catch (const kc::KcError& e) {
    ctx->last_error = e.what;
    return e.status;
  } catch (const std::bad_alloc&) {
    return KC_ENOMEM;
  }
}

extern "C" kc_status kc_probe_random_gather(void* vctx, uint64_t n, uint32_t reps, double* gbps_out,
                                            double* ms_out) {
  return kc_probe_random_gather_mode(vctx, n, reps, 0, gbps_out, ms_out);
}

// Gather from a table of m u64 (random indices in [0, m)), n gathers: the rate
// as a function of the table size shows which level of the memory hierarchy
// serves a random lookup.
extern "C" kc_status kc_probe_table_gather(void* vctx, uint64_t n, uint64_t m, uint32_t reps,
                                           double* gathers_per_s, double* ms_out) {
  auto* ctx = static_cast<kc::Context*>(vctx);
  if (!ctx || n == 0 || m == 0 || m > n || reps == 0) return KC_EINVAL;
  try {
    KC_HIP(hipSetDevice(ctx->device));
    kc::Scratch scratch(ctx);
    uint64_t* src = scratch.get<uint64_t>(m);
    uint64_t* idx = scratch.get<uint64_t>(n);
    uint64_t* dst = scratch.get<uint64_t>(n);
    hipStream_t st = ctx->stream;
    const unsigned blocks = kc::blocks_for(n, 256);
    hipLaunchKernelGGL(kc::table_idx_kernel, dim3(blocks), dim3(256), 0, st, idx, n, m, 777ull);
    hipLaunchKernelGGL(kc::fill_kernel, dim3(kc::blocks_for(m, 256)), dim3(256), 0, st, src, m);
    hipLaunchKernelGGL(kc::gather_kernel, dim3(blocks), dim3(256), 0, st, src, idx, dst, n);
    hipEvent_t a, b;
    KC_HIP(hipEventCreate(&a));
    KC_HIP(hipEventCreate(&b));
    KC_HIP(hipEventRecord(a, st));
    for (uint32_t r = 0; r < reps; r++)
      hipLaunchKernelGGL(kc::gather_kernel, dim3(blocks), dim3(256), 0, st, src, idx, dst, n);
    KC_HIP(hipEventRecord(b, st));
    KC_HIP(hipEventSynchronize(b));
    float ms = 0;
    KC_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const double per = ms / reps;
    if (ms_out) *ms_out = per;
    if (gathers_per_s) *gathers_per_s = static_cast<double>(n) / (per * 1e-3);
    return KC_OK;
  } catch (const kc::KcError& e) {
    ctx->last_error = e.what;
    return e.status;
  } catch (const std::bad_alloc&) {
    return KC_ENOMEM;
  }
}

// Access patterns of 16-byte records over a random cycle of n records (n a
// power of two): 0 independent reads, 1 independent writes, 2 independent
// read + 16-byte write of the same record, 3 dependent walks (n/256 walkers of
// 256 steps) reading only, 4 the same walks writing a claim into every record;
// 5 and 6 the walks of 3 and 4 over 8-byte records (n at most 2^30).
// Reports records per second and the kernel time.
extern "C" kc_status kc_probe_record_access(void* vctx, uint64_t n, int pattern, uint32_t reps,
                                            double* records_per_s, double* ms_out) {
  auto* ctx = static_cast<kc::Context*>(vctx);
  constexpr uint64_t kWalk = 256;
  if (!ctx || n < kWalk || (n & (n - 1)) != 0 || reps == 0 || pattern < 0 || pattern > 6 ||
      (pattern >= 5 && n > kc::kNarrowMaxIds))
    return KC_EINVAL;
  try {
    KC_HIP(hipSetDevice(ctx->device));
    kc::Scratch scratch(ctx);
    kc::Rec* rec = scratch.get<kc::Rec>(n);
    uint64_t* idx = scratch.get<uint64_t>(n);
    int64_t* out = scratch.get<int64_t>(n);
    hipStream_t st = ctx->stream;
    const unsigned blocks = kc::blocks_for(n, 256);
    hipLaunchKernelGGL(kc::perm_idx_kernel, dim3(blocks), dim3(256), 0, st, idx, n, 4242ull);
    kc::Rec8* rec8 = reinterpret_cast<kc::Rec8*>(rec);  // patterns 5, 6: the first half of the buffer
    if (pattern >= 5)
      hipLaunchKernelGGL(kc::cycle_init8_kernel, dim3(blocks), dim3(256), 0, st, rec8, idx, n);
    else
      hipLaunchKernelGGL(kc::cycle_init_kernel, dim3(blocks), dim3(256), 0, st, rec, idx, n);
    const uint64_t walkers = n / kWalk;
    auto launch = [&] {
      if (pattern >= 5)
        hipLaunchKernelGGL(kc::record_walk8_kernel, dim3(kc::blocks_for(walkers, 256)), dim3(256), 0, st,
                           rec8, idx, out, walkers, kWalk, pattern == 6 ? 1 : 0);
      else if (pattern <= 2)
        hipLaunchKernelGGL(kc::record_access_kernel, dim3(blocks), dim3(256), 0, st, rec, idx, out, n,
                           pattern);
      else
        hipLaunchKernelGGL(kc::record_walk_kernel, dim3(kc::blocks_for(walkers, 256)), dim3(256), 0, st,
                           rec, idx, out, walkers, kWalk, pattern == 4 ? 1 : 0);
    };
    launch();
    hipEvent_t a, b;
    KC_HIP(hipEventCreate(&a));
    KC_HIP(hipEventCreate(&b));
    KC_HIP(hipEventRecord(a, st));
    for (uint32_t r = 0; r < reps; r++) launch();
    KC_HIP(hipEventRecord(b, st));
    KC_HIP(hipEventSynchronize(b));
    float ms = 0;
    KC_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const double per = ms / reps;
    if (ms_out) *ms_out = per;
    if (records_per_s) *records_per_s = static_cast<double>(n) / (per * 1e-3);
    return KC_OK;
  } catch (const kc::KcError& e) {
    ctx->last_error = e.what;
    return e.status;
  } catch (const std::bad_alloc&) {
    return KC_ENOMEM;
  }
}
