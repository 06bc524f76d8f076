// List and forest utilities on the device: the reference's successor_utils
// (is_list, LeafInfo, reverse_list, roots; lib/kascade/successor_utils.cpp)
// and the benchmark statistics (compute_stats, bench/detail/stats.cpp:24-162).
#include <climits>

#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;
constexpr uint64_t kNone = ~uint64_t{0};

unsigned grid_for(const Context* ctx, uint64_t n) {
  const uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 8;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(cap, blocks_for(n, kThreads))));
}

__device__ __forceinline__ void flag_store(unsigned long long* flag) {
  __hip_atomic_store(flag, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t wave_sum(int64_t x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ int64_t wave_max(int64_t x) {
  for (int o = 32; o > 0; o >>= 1) {
    const int64_t y = __shfl_xor(x, o);
    x = y > x ? y : x;
  }
  return x;
}
__device__ __forceinline__ uint64_t wave_sum_u(uint64_t x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ uint64_t wave_max_u(uint64_t x) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t y = __shfl_xor(x, o);
    x = y > x ? y : x;
  }
  return x;
}

// c[0..7] = 0, then c[1] = INT64_MIN when max_slot is set (a running maximum)
__global__ void counters_init_kernel(uint64_t* c, int max_slot) {
  if (threadIdx.x < 8) c[threadIdx.x] = (max_slot && threadIdx.x == 1) ? static_cast<uint64_t>(LLONG_MIN) : 0;
}

__global__ void fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads)
    p[i] = v;
}

// Has-predecessor bitmap (LeafInfo, successor_utils.cpp:72-123): one atomic OR
// per non-root element.  A bit that was already set is a second predecessor
// (is_list, :31-63).
__global__ void __launch_bounds__(kThreads)
pred_bits_kernel(const uint64_t* __restrict__ succ, uint64_t n, uint32_t* __restrict__ bits,
                 uint32_t* err, unsigned long long* dup) {
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t s = succ[v];
    if (s >= n) {
      raise_err(err, kErrRange);
      continue;
    }
    if (s == v) continue;
    const uint32_t bit = 1u << (s & 31);
    if (atomicOr(&bits[s >> 5], bit) & bit) flag_store(dup);
  }
}

// is_leaf[v] = no predecessor (LeafInfo::is_leaf: a root without predecessor
// counts as a leaf); the count is reduced per wave.
__global__ void __launch_bounds__(kThreads)
leaf_kernel(const uint32_t* __restrict__ bits, uint64_t n, uint8_t* __restrict__ is_leaf,
            unsigned long long* count) {
  uint64_t c = 0;
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const bool leaf = !((bits[v >> 5] >> (v & 31)) & 1u);
    if (is_leaf) is_leaf[v] = leaf ? 1 : 0;
    c += leaf;
  }
  c = wave_sum_u(c);
  if (__lane_id() == 0 && c) atomicAdd(count, static_cast<unsigned long long>(c));
}

// reverse_list (successor_utils.cpp:159-232): pred[succ[u]] = u and
// dist_to_pred[succ[u]] = dist_to_succ[u]; elements without predecessor keep
// (self, 0).  The exchange returns the previous predecessor, so a second one
// is detected.
__global__ void __launch_bounds__(kThreads)
self_pred_kernel(uint64_t* __restrict__ pred, int64_t* __restrict__ dpred, uint64_t n) {
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kThreads) {
    pred[v] = v;
    dpred[v] = 0;
  }
}

__global__ void __launch_bounds__(kThreads)
reverse_kernel(const uint64_t* __restrict__ succ, const int64_t* __restrict__ w, uint64_t n,
               uint64_t* __restrict__ pred, int64_t* __restrict__ dpred, uint32_t* err) {
  for (uint64_t u = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; u < n;
       u += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t s = succ[u];
    if (s >= n) {
      raise_err(err, kErrRange);
      continue;
    }
    if (s == u) continue;
    const unsigned long long old =
        atomicExch(reinterpret_cast<unsigned long long*>(&pred[s]), static_cast<unsigned long long>(u));
    if (old != s) raise_err(err, kErrNotList);
    dpred[s] = w ? w[u] : 1;
  }
}

// Order-preserving compaction of the ids i < n with pred(i), in two passes over
// tiles of kThreads elements: per-tile counts, an exclusive scan of the counts,
// then each tile writes its ids at its offset.
struct RootPred {
  const uint64_t* a;
  __device__ bool operator()(uint64_t i) const { return a[i] == i; }
};
struct NontrivialPred {
  const uint64_t* size;
  __device__ bool operator()(uint64_t i) const { return size[i] > 1; }
};

template <class P>
__global__ void __launch_bounds__(kThreads)
tile_count_kernel(P pred, uint64_t n, uint64_t* __restrict__ counts) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const int c = __syncthreads_count(i < n && pred(i));
  if (threadIdx.x == 0) counts[blockIdx.x] = static_cast<uint64_t>(c);
}

// Writes out[k] = i for the k-th selected id; pos[i] = k for selected ids and
// kNone otherwise when pos is given.
template <class P>
__global__ void __launch_bounds__(kThreads)
tile_write_kernel(P pred, uint64_t n, const uint64_t* __restrict__ offs, uint64_t* __restrict__ out,
                  uint64_t cap, uint64_t* __restrict__ pos) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const bool sel = i < n && pred(i);
  const uint64_t m = __ballot(sel);
  const unsigned wave = threadIdx.x / 64;
  if (__lane_id() == 0) s_wave[wave] = static_cast<uint32_t>(__popcll(m));
  __syncthreads();
  uint64_t k = offs[blockIdx.x];
  for (unsigned q = 0; q < wave; q++) k += s_wave[q];
  k += __popcll(m & lanemask_lt());
  if (sel && out && k < cap) out[k] = i;
  if (pos && i < n) pos[i] = sel ? k : kNone;
}

// ---- statistics ------------------------------------------------------------
// Basic statistics (compute_basic_stats, stats.cpp:115-145): number of roots,
// maximum rank, rank sum.
__global__ void __launch_bounds__(kThreads)
basic_stats_kernel(const uint64_t* __restrict__ root, const int64_t* __restrict__ rank, uint64_t n,
                   unsigned long long* roots, long long* max_rank, unsigned long long* rank_sum) {
  uint64_t c = 0, s = 0;
  int64_t mx = LLONG_MIN;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads) {
    c += root[i] == i;
    const int64_t r = rank[i];
    s += static_cast<uint64_t>(r);
    mx = r > mx ? r : mx;
  }
  c = wave_sum_u(c);
  s = wave_sum_u(s);
  mx = wave_max(mx);
  if (__lane_id() == 0) {
    if (c) atomicAdd(roots, static_cast<unsigned long long>(c));
    atomicAdd(rank_sum, static_cast<unsigned long long>(s));
    atomicMax(max_rank, static_cast<long long>(mx));
  }
}

__global__ void __launch_bounds__(kThreads)
tree_init_kernel(uint64_t* __restrict__ size, int64_t* __restrict__ sum, int64_t* __restrict__ mx,
                 uint64_t m) {
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; t < m;
       t += static_cast<uint64_t>(gridDim.x) * kThreads) {
    size[t] = 0;
    sum[t] = 0;
    mx[t] = LLONG_MIN;
  }
}

// Per-tree size, rank sum and maximum rank (compute_per_tree_stats,
// stats.cpp:24-113), keyed by the tree index of each element's root.  Lanes of
// a wave that share a tree are combined first, so one atomic per tree and
// wave reaches memory (a single large tree does not serialise on one address).
__global__ void __launch_bounds__(kThreads)
tree_accumulate_kernel(const uint64_t* __restrict__ root, const int64_t* __restrict__ rank, uint64_t n,
                       const uint64_t* __restrict__ tid, uint64_t* __restrict__ size,
                       int64_t* __restrict__ sum, int64_t* __restrict__ mx, uint32_t* err) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads;
  const uint64_t base0 = static_cast<uint64_t>(blockIdx.x) * kThreads;
  // every lane of a wave runs the same number of iterations (wave-uniform loop)
  for (uint64_t base = base0; base < n; base += stride) {
    const uint64_t i = base + threadIdx.x;
    uint64_t t = kNone;
    int64_t r = 0;
    if (i < n) {
      const uint64_t ro = root[i];
      if (ro >= n) {
        raise_err(err, kErrRange);
      } else {
        t = tid[ro];
        if (t == kNone) raise_err(err, kErrNotList);  // root value that is not a root
        r = rank[i];
      }
    }
    uint64_t left = __ballot(t != kNone);
    while (left) {
      const int leader = __ffsll(static_cast<long long>(left)) - 1;
      const uint64_t lt = __shfl(t, leader);
      const bool mine = t == lt && ((left >> __lane_id()) & 1);
      const uint64_t m = __ballot(mine);
      const int64_t s = wave_sum(mine ? r : 0);
      const int64_t x = wave_max(mine ? r : LLONG_MIN);
      if (__lane_id() == static_cast<unsigned>(leader)) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&size[lt]), static_cast<unsigned long long>(__popcll(m)));
        atomicAdd(reinterpret_cast<unsigned long long*>(&sum[lt]), static_cast<unsigned long long>(s));
        atomicMax(reinterpret_cast<long long*>(&mx[lt]), static_cast<long long>(x));
      }
      left &= ~m;
    }
  }
}

// Aggregates over trees (stats.cpp:60-72): nontrivial trees (size > 1), their
// size and rank sums, and the largest size.
__global__ void __launch_bounds__(kThreads)
tree_reduce_kernel(const uint64_t* __restrict__ size, const int64_t* __restrict__ sum, uint64_t m,
                   unsigned long long* nontrivial, unsigned long long* nt_size,
                   unsigned long long* nt_rank, unsigned long long* max_size) {
  uint64_t c = 0, zs = 0, zr = 0, mx = 0;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; t < m;
       t += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t z = size[t];
    mx = z > mx ? z : mx;
    if (z > 1) {
      c++;
      zs += z;
      zr += static_cast<uint64_t>(sum[t]);
    }
  }
  c = wave_sum_u(c);
  zs = wave_sum_u(zs);
  zr = wave_sum_u(zr);
  mx = wave_max_u(mx);
  if (__lane_id() == 0) {
    if (c) atomicAdd(nontrivial, static_cast<unsigned long long>(c));
    if (zs) atomicAdd(nt_size, static_cast<unsigned long long>(zs));
    atomicAdd(nt_rank, static_cast<unsigned long long>(zr));
    atomicMax(max_size, static_cast<unsigned long long>(mx));
  }
}

// Per-tree output for the nontrivial trees, in ascending root order.
__global__ void __launch_bounds__(kThreads)
tree_emit_kernel(const uint64_t* __restrict__ sel, uint64_t k, const uint64_t* __restrict__ tree_root,
                 const uint64_t* __restrict__ size, const int64_t* __restrict__ sum,
                 const int64_t* __restrict__ mx, uint64_t* __restrict__ out_root,
                 uint64_t* __restrict__ out_size, int64_t* __restrict__ out_max,
                 int64_t* __restrict__ out_sum) {
  const uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= k) return;
  const uint64_t t = sel[j];
  if (out_root) out_root[j] = tree_root[t];
  if (out_size) out_size[j] = size[t];
  if (out_max) out_max[j] = mx[t];
  if (out_sum) out_sum[j] = sum[t];
}

template <class P>
uint64_t compact(Context* ctx, P pred, uint64_t n, uint64_t* out, uint64_t cap, uint64_t* pos,
                 hipStream_t st) {
  if (n == 0) return 0;
  Scratch scratch(ctx);
  const uint64_t tiles = blocks_for(n, kThreads);
  uint64_t* counts = scratch.get<uint64_t>(tiles + 1);
  uint64_t* offs = scratch.get<uint64_t>(tiles + 1);
  hipLaunchKernelGGL(tile_count_kernel<P>, dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0, st, pred,
                     n, counts);
  // one more zero entry so that offs[tiles] is the total
  KC_HIP(hipMemsetAsync(counts + tiles, 0, sizeof(uint64_t), st));
  exclusive_scan_u64(ctx, counts, offs, tiles + 1, st);
  hipLaunchKernelGGL(tile_write_kernel<P>, dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0, st, pred,
                     n, offs, out, cap, pos);
  uint64_t total = 0;
  KC_HIP(hipMemcpyAsync(&total, offs + tiles, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  KC_HIP(hipStreamSynchronize(st));
  return total;
}

}  // namespace

void pred_bitmap(Context* ctx, const uint64_t* succ, uint64_t n, uint32_t* bits, bool* dup_out,
                 hipStream_t st) {
  const uint64_t words = (n + 31) / 32;
  hipLaunchKernelGGL(fill_u32, dim3(grid_for(ctx, words)), dim3(kThreads), 0, st, bits, words, 0u);
  KC_HIP(hipMemsetAsync(ctx->d_counters, 0, sizeof(uint64_t), st));
  hipLaunchKernelGGL(pred_bits_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, succ, n, bits,
                     ctx->d_err, reinterpret_cast<unsigned long long*>(ctx->d_counters));
  uint64_t dup = 0;
  read_counters(ctx, st, 0, 1, &dup);  // raises range errors
  if (dup_out) *dup_out = dup != 0;
}

bool is_list_device(Context* ctx, const uint64_t* succ, uint64_t n, hipStream_t st) {
  if (n == 0) return true;
  Scratch scratch(ctx);
  uint32_t* bits = scratch.get<uint32_t>((n + 31) / 32);
  bool dup = false;
  pred_bitmap(ctx, succ, n, bits, &dup, st);
  return !dup;
}

uint64_t leaf_info_device(Context* ctx, const uint64_t* succ, uint64_t n, uint8_t* is_leaf,
                          hipStream_t st) {
  if (n == 0) return 0;
  Scratch scratch(ctx);
  uint32_t* bits = scratch.get<uint32_t>((n + 31) / 32);
  pred_bitmap(ctx, succ, n, bits, nullptr, st);
  KC_HIP(hipMemsetAsync(ctx->d_counters, 0, sizeof(uint64_t), st));
  hipLaunchKernelGGL(leaf_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, bits, n, is_leaf,
                     reinterpret_cast<unsigned long long*>(ctx->d_counters));
  uint64_t count = 0;
  read_counters(ctx, st, 0, 1, &count);
  return count;
}

uint64_t reverse_list_device(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n,
                             uint64_t* pred, int64_t* dpred, uint64_t* roots, uint64_t roots_cap,
                             hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(self_pred_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, pred, dpred, n);
  hipLaunchKernelGGL(reverse_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, succ, w, n, pred,
                     dpred, ctx->d_err);
  check_err(ctx, st);
  return compact(ctx, RootPred{succ}, n, roots, roots_cap, nullptr, st);
}

void compute_stats_device(Context* ctx, const uint64_t* root, const int64_t* rank, uint64_t n, int level,
                          kc_stats* out, uint64_t* tree_root, uint64_t* tree_size, int64_t* tree_max,
                          int64_t* tree_sum, uint64_t tree_cap, hipStream_t st) {
  *out = kc_stats{};
  if (n == 0 || level <= KC_STATS_NONE) return;
  Scratch scratch(ctx);
  // counters: [0] roots, [1] max rank, [2] rank sum
  uint64_t* c = scratch.get<uint64_t>(8);
  hipLaunchKernelGGL(counters_init_kernel, dim3(1), dim3(64), 0, st, c, 1);
  hipLaunchKernelGGL(basic_stats_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, root, rank, n,
                     reinterpret_cast<unsigned long long*>(c), reinterpret_cast<long long*>(c + 1),
                     reinterpret_cast<unsigned long long*>(c + 2));
  uint64_t h[8];
  KC_HIP(hipMemcpyAsync(h, c, sizeof(h), hipMemcpyDeviceToHost, st));
  KC_HIP(hipStreamSynchronize(st));
  out->num_trees = h[0];
  out->max_rank = static_cast<int64_t>(h[1]);
  out->avg_rank = static_cast<double>(static_cast<int64_t>(h[2])) / static_cast<double>(n);
  if (level == KC_STATS_BASIC) return;

  // tree index of every root (ascending root id), then per-tree accumulators
  uint64_t* tid = scratch.get<uint64_t>(n);
  const uint64_t m = compact(ctx, RootPred{root}, n, nullptr, 0, tid, st);
  uint64_t* roots_of = scratch.get<uint64_t>(std::max<uint64_t>(m, 1));
  compact(ctx, RootPred{root}, n, roots_of, m, nullptr, st);
  uint64_t* size = scratch.get<uint64_t>(std::max<uint64_t>(m, 1));
  int64_t* sum = scratch.get<int64_t>(std::max<uint64_t>(m, 1));
  int64_t* mx = scratch.get<int64_t>(std::max<uint64_t>(m, 1));
  hipLaunchKernelGGL(tree_init_kernel, dim3(grid_for(ctx, m)), dim3(kThreads), 0, st, size, sum, mx, m);
  hipLaunchKernelGGL(tree_accumulate_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, root, rank, n,
                     tid, size, sum, mx, ctx->d_err);
  check_err(ctx, st);
  hipLaunchKernelGGL(counters_init_kernel, dim3(1), dim3(64), 0, st, c, 0);
  hipLaunchKernelGGL(tree_reduce_kernel, dim3(grid_for(ctx, m)), dim3(kThreads), 0, st, size, sum, m,
                     reinterpret_cast<unsigned long long*>(c), reinterpret_cast<unsigned long long*>(c + 1),
                     reinterpret_cast<unsigned long long*>(c + 2), reinterpret_cast<unsigned long long*>(c + 3));
  KC_HIP(hipMemcpyAsync(h, c, sizeof(h), hipMemcpyDeviceToHost, st));
  KC_HIP(hipStreamSynchronize(st));
  out->num_trees = m;
  out->num_nontrivial_trees = h[0];
  out->nontrivial_size_sum = h[1];
  out->nontrivial_rank_sum = static_cast<int64_t>(h[2]);
  out->max_size = h[3];
  // safe_division (stats.cpp:18-22)
  auto div = [](double a, double b) { return b == 0 ? 0.0 : a / b; };
  const uint64_t trivial = m - out->num_nontrivial_trees;
  out->avg_size = div(static_cast<double>(out->nontrivial_size_sum + trivial), static_cast<double>(m));
  out->nontrivial_avg_size =
      div(static_cast<double>(out->nontrivial_size_sum), static_cast<double>(out->num_nontrivial_trees));
  out->nontrivial_avg_rank =
      div(static_cast<double>(out->nontrivial_rank_sum), static_cast<double>(out->nontrivial_size_sum));
  if (level != KC_STATS_EXTENSIVE || tree_cap == 0) return;
  uint64_t* sel = scratch.get<uint64_t>(std::max<uint64_t>(out->num_nontrivial_trees, 1));
  const uint64_t k = compact(ctx, NontrivialPred{size}, m, sel, out->num_nontrivial_trees, nullptr, st);
  const uint64_t kk = std::min(k, tree_cap);
  if (kk)
    hipLaunchKernelGGL(tree_emit_kernel, dim3(blocks_for(kk, kThreads)), dim3(kThreads), 0, st, sel, kk,
                       roots_of, size, sum, mx, tree_root, tree_size, tree_max, tree_sum);
  KC_HIP(hipGetLastError());
}

}  // namespace kc
