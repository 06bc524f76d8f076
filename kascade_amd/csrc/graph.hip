// Graph inputs for tree rooting: seeded random graphs, their CSR, and the
// BFS forest of bench/input/bfs_forest_generation.cpp:52-76 (the reference's
// "bfs" and "eulertour" input processing, bench/input/generation.cpp:60-96).
//
// The BFS forest: isolated vertices are their own parents; every other
// connected component becomes a BFS tree rooted at its smallest vertex id
// (the reference starts each BFS at the smallest unvisited id).  (root, depth)
// of every vertex is therefore unique - the component minimum and the BFS
// distance from it - while the parent among equally near neighbours depends on
// visiting order.  On the device: connected components by min-label hooking
// with pointer jumping, then one level-synchronous BFS from all component
// minima at once.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;
constexpr uint64_t kUnvisited = ~uint64_t{0};

unsigned grid_for(const Context* ctx, uint64_t n) {
  const uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 8;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(cap, blocks_for(n, kThreads))));
}

// Edge e of a seeded graph (the same arithmetic as orc_gen_graph).
//   GNM: both endpoints uniform in [0, n); a self loop moves its head to the
//        next vertex.
//   RMAT: one quadrant per bit of ceil(log2 n) bits with probabilities
//        (0.57, 0.19, 0.19, 0.05) (the Graph500 parameters), ids >= n folded
//        by modulo; a self loop moves its head to the next vertex.
__device__ inline void graph_edge(int kind, uint64_t n, uint64_t seed, uint64_t e, uint64_t* u,
                                           uint64_t* v) {
  uint64_t a, b;
  if (kind == KC_GRAPH_GNM) {
    a = splitmix64(seed ^ (2 * e + 1) * 0x9E3779B97F4A7C15ull) % n;
    b = splitmix64(seed ^ (2 * e + 2) * 0xD1B54A32D192ED03ull) % n;
  } else {
    uint32_t bits = 0;
    while (bits < 63 && (uint64_t{1} << bits) < n) bits++;
    a = 0;
    b = 0;
    for (uint32_t k = 0; k < bits; k++) {
      // 16-bit uniform draw per level: thresholds 0.57, 0.76, 0.95 of 65536
      const uint32_t r = static_cast<uint32_t>(splitmix64(seed ^ (e * 64 + k + 1) * 0x94D049BB133111EBull) & 0xFFFF);
      const uint32_t q = r < 37356 ? 0 : r < 49807 ? 1 : r < 62259 ? 2 : 3;
      a = (a << 1) | (q >> 1);
      b = (b << 1) | (q & 1);
    }
    a %= n;
    b %= n;
  }
  if (a == b) b = (b + 1) % n;
  *u = a;
  *v = b;
}

__global__ void __launch_bounds__(kThreads)
gen_graph_kernel(int kind, uint64_t n, uint64_t m, uint64_t seed, uint64_t* __restrict__ src,
                 uint64_t* __restrict__ dst) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * kThreads)
    graph_edge(kind, n, seed, e, &src[e], &dst[e]);
}

__global__ void __launch_bounds__(kThreads)
fill_u64_kernel(uint64_t* __restrict__ p, uint64_t n, uint64_t v) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads)
    p[i] = v;
}

__global__ void __launch_bounds__(kThreads)
iota_kernel(uint64_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads)
    p[i] = i;
}

__global__ void __launch_bounds__(kThreads)
degree_kernel(const uint64_t* __restrict__ src, const uint64_t* __restrict__ dst, uint64_t n, uint64_t m,
              unsigned long long* __restrict__ deg, uint32_t* err) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t u = src[e], v = dst[e];
    if (u >= n || v >= n) {
      raise_err(err, kErrRange);
      continue;
    }
    atomicAdd(&deg[u], 1ull);
    atomicAdd(&deg[v], 1ull);
  }
}

__global__ void __launch_bounds__(kThreads)
scatter_edges_kernel(const uint64_t* __restrict__ src, const uint64_t* __restrict__ dst, uint64_t n, uint64_t m,
                     unsigned long long* __restrict__ cursor, uint64_t* __restrict__ adj) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t u = src[e], v = dst[e];
    if (u >= n || v >= n) continue;
    adj[atomicAdd(&cursor[u], 1ull)] = v;
    adj[atomicAdd(&cursor[v], 1ull)] = u;
  }
}

// xadj[n] = 2m after the exclusive scan of the degrees
__global__ void xadj_end_kernel(uint64_t* xadj, uint64_t n, uint64_t total) {
  if (threadIdx.x == 0) xadj[n] = total;
}

// Hooking: for every edge the root of the larger label is pointed at the
// smaller label (labels only decrease; a label is always a vertex of the same
// component), so the fixed point is the component minimum.
__global__ void __launch_bounds__(kThreads)
cc_hook_kernel(const uint64_t* __restrict__ xadj, const uint64_t* __restrict__ adj, uint64_t n,
               unsigned long long* __restrict__ label, unsigned long long* changed) {
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t b = xadj[v], e = xadj[v + 1];
    for (uint64_t k = b; k < e; k++) {
      const uint64_t u = adj[k];
      const unsigned long long lv = label[v], lu = label[u];
      if (lu < lv) {
        if (atomicMin(&label[lv], lu) > lu) *changed = 1;
      } else if (lv < lu) {
        if (atomicMin(&label[lu], lv) > lv) *changed = 1;
      }
    }
  }
}

// Pointer jumping: label[v] = label[label[v]] until no label moves.
__global__ void __launch_bounds__(kThreads)
cc_jump_kernel(uint64_t n, unsigned long long* __restrict__ label, unsigned long long* changed) {
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const unsigned long long l = label[v];
    const unsigned long long ll = label[l];
    if (ll < l) {
      label[v] = ll;
      *changed = 1;
    }
  }
}

// Level 0 of the BFS: component minima (label[v] == v) are roots.
__global__ void __launch_bounds__(kThreads)
bfs_roots_kernel(const unsigned long long* __restrict__ label, uint64_t n, uint64_t* __restrict__ parent,
                 uint64_t* __restrict__ frontier, unsigned long long* count) {
  const uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const bool root = v < n && label[v] == v;
  if (v < n) parent[v] = root ? v : kUnvisited;
  const uint64_t k = wave_append(count, root);
  if (root) frontier[k] = v;
}

// One BFS level: every frontier vertex claims its unvisited neighbours
// (compare-and-swap on parent) and appends them to the next frontier.
__global__ void __launch_bounds__(kThreads)
bfs_level_kernel(const uint64_t* __restrict__ xadj, const uint64_t* __restrict__ adj,
                 const uint64_t* __restrict__ frontier, uint64_t nf, unsigned long long* __restrict__ parent,
                 uint64_t* __restrict__ next, unsigned long long* next_count) {
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kThreads; base < nf;
       base += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t f = base + threadIdx.x;
    const bool live = f < nf;
    const uint64_t u = live ? frontier[f] : 0;
    const uint64_t b = live ? xadj[u] : 0, e = live ? xadj[u + 1] : 0;
    // neighbours in rounds so that every lane of the wave appends together
    uint64_t k = b;
    for (;;) {
      const bool more = k < e;
      if (__ballot(more) == 0) break;
      bool claimed = false;
      uint64_t v = 0;
      if (more) {
        v = adj[k++];
        claimed = parent[v] == kUnvisited &&
                  atomicCAS(&parent[v], static_cast<unsigned long long>(kUnvisited),
                            static_cast<unsigned long long>(u)) == kUnvisited;
      }
      const uint64_t slot = wave_append(next_count, claimed);
      if (claimed) next[slot] = v;
    }
  }
}

// ---- random geometric graph (the role of KaGen's rgg2d) ---------------------
// Point i: 31-bit fixed-point coordinates from the seed (the same arithmetic as
// orc_gen_rgg2d).  Cells of a g x g grid are w >= r wide, so every pair within
// r lies in the same or an adjacent cell.
__device__ __forceinline__ uint32_t rgg_x(uint64_t seed, uint64_t i) {
  return static_cast<uint32_t>(splitmix64(seed ^ (2 * i + 1) * 0xA0761D6478BD642Full) >> 33);
}
__device__ __forceinline__ uint32_t rgg_y(uint64_t seed, uint64_t i) {
  return static_cast<uint32_t>(splitmix64(seed ^ (2 * i + 2) * 0xE7037ED1A0B428DBull) >> 33);
}

__global__ void __launch_bounds__(kThreads)
rgg_cells_kernel(uint64_t n, uint64_t seed, uint64_t w, uint64_t g, uint32_t* __restrict__ cell,
                 uint32_t* __restrict__ idx, unsigned long long* __restrict__ count) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t c = (rgg_y(seed, i) / w) * g + rgg_x(seed, i) / w;
    cell[i] = static_cast<uint32_t>(c);
    idx[i] = static_cast<uint32_t>(i);
    atomicAdd(&count[c], 1ull);
  }
}

// Neighbours j > i within r of point i, visited in the order of the oracle:
// neighbour cells row by row, points of a cell by ascending id (the stable
// sort by cell keeps that order).  kWrite: write the edges at out_off, else
// count them.
template <bool kWrite>
__global__ void __launch_bounds__(kThreads)
rgg_edges_kernel(uint64_t n, uint64_t seed, uint64_t w, uint64_t g, uint64_t r2,
                 const uint64_t* __restrict__ start, const uint32_t* __restrict__ sorted_idx,
                 uint64_t* __restrict__ cnt, const uint64_t* __restrict__ out_off, uint64_t* __restrict__ src,
                 uint64_t* __restrict__ dst) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const uint64_t xi = rgg_x(seed, i), yi = rgg_y(seed, i);
    const int64_t cx = static_cast<int64_t>(xi / w), cy = static_cast<int64_t>(yi / w);
    uint64_t k = kWrite ? out_off[i] : 0;
    for (int64_t dy = -1; dy <= 1; dy++) {
      const int64_t ay = cy + dy;
      if (ay < 0 || ay >= static_cast<int64_t>(g)) continue;
      for (int64_t dx = -1; dx <= 1; dx++) {
        const int64_t ax = cx + dx;
        if (ax < 0 || ax >= static_cast<int64_t>(g)) continue;
        const uint64_t c = static_cast<uint64_t>(ay) * g + static_cast<uint64_t>(ax);
        for (uint64_t p = start[c]; p < start[c + 1]; p++) {
          const uint64_t j = sorted_idx[p];
          if (j <= i) continue;
          const uint64_t xj = rgg_x(seed, j), yj = rgg_y(seed, j);
          const uint64_t ddx = xi > xj ? xi - xj : xj - xi;
          const uint64_t ddy = yi > yj ? yi - yj : yj - yi;
          if (ddx * ddx + ddy * ddy > r2) continue;
          if (kWrite) {
            src[k] = i;
            dst[k] = j;
          }
          k++;
        }
      }
    }
    if (!kWrite) cnt[i] = k;
  }
}

// The end of the last cell (after the exclusive scan of the cell counts).
__global__ void rgg_cells_end_kernel(uint64_t* __restrict__ start, uint64_t cells, uint64_t n) {
  if (threadIdx.x == 0) start[cells] = n;
}

__global__ void rgg_total_kernel(uint64_t n, const uint64_t* __restrict__ off, const uint64_t* __restrict__ cnt,
                                 uint64_t* __restrict__ total) {
  if (threadIdx.x == 0) total[0] = n ? off[n - 1] + cnt[n - 1] : 0;
}

}  // namespace

uint64_t rgg2d_radius(uint64_t n, uint64_t m) {
  if (n < 2 || m == 0) return 0;
  const double r = std::sqrt(2.0 * static_cast<double>(m) / (3.141592653589793 * static_cast<double>(n) *
                                                               static_cast<double>(n)));
  if (r >= 1.0) return uint64_t{1} << 31;
  return static_cast<uint64_t>(r * 2147483648.0);
}

uint64_t gen_rgg2d_device(Context* ctx, uint64_t n, uint64_t m, uint64_t seed, uint64_t* src, uint64_t* dst,
                          uint64_t cap, hipStream_t st) {
  if (n >= (uint64_t{1} << 32)) throw KcError{KC_EINVAL, "rgg2d supports n < 2^32"};
  const uint64_t r_fp = rgg2d_radius(n, m);
  if (r_fp == 0) return 0;
  uint64_t g = (uint64_t{1} << 31) / r_fp;
  uint64_t side = 1;
  while (side * side < n) side++;
  g = std::max<uint64_t>(1, std::min<uint64_t>({g, side, 65535}));
  const uint64_t w = ((uint64_t{1} << 31) + g - 1) / g;
  const uint64_t cells = g * g;
  Scratch scratch(ctx);
  uint32_t* cell = scratch.get<uint32_t>(n);
  uint32_t* idx = scratch.get<uint32_t>(n);
  uint64_t* count = scratch.get<uint64_t>(cells);
  uint64_t* start = scratch.get<uint64_t>(cells + 1);
  uint64_t* cnt = scratch.get<uint64_t>(n);
  uint64_t* off = scratch.get<uint64_t>(n);
  uint64_t* total = scratch.get<uint64_t>(1);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(ctx, cells)), dim3(kThreads), 0, st, count, cells, uint64_t{0});
  hipLaunchKernelGGL(rgg_cells_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, n, seed, w, g, cell, idx,
                     reinterpret_cast<unsigned long long*>(count));
  unsigned bits = 1;
  while (bits < 32 && (uint64_t{1} << bits) < cells) bits++;
  radix_sort_pairs_u32(ctx, cell, idx, n, bits, st);  // stable: ascending ids within a cell
  exclusive_scan_u64(ctx, count, start, cells, st);
  hipLaunchKernelGGL(rgg_cells_end_kernel, dim3(1), dim3(64), 0, st, start, cells, n);
  hipLaunchKernelGGL(rgg_edges_kernel<false>, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, n, seed, w, g,
                     r_fp * r_fp, start, idx, cnt, nullptr, nullptr, nullptr);
  exclusive_scan_u64(ctx, cnt, off, n, st);
  hipLaunchKernelGGL(rgg_total_kernel, dim3(1), dim3(64), 0, st, n, off, cnt, total);
  uint64_t edges = 0;
  KC_HIP(hipMemcpyAsync(&edges, total, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  KC_HIP(hipStreamSynchronize(st));
  if (!src || !dst) return edges;
  if (cap < edges) throw KcError{KC_EINVAL, "rgg2d: the edge arrays are smaller than the edge count"};
  hipLaunchKernelGGL(rgg_edges_kernel<true>, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, n, seed, w, g,
                     r_fp * r_fp, start, idx, cnt, off, src, dst);
  KC_HIP(hipGetLastError());
  return edges;
}

void gen_graph_device(Context* ctx, int kind, uint64_t n, uint64_t m, uint64_t seed, uint64_t* src, uint64_t* dst,
                      hipStream_t st) {
  if (m == 0) return;
  hipLaunchKernelGGL(gen_graph_kernel, dim3(grid_for(ctx, m)), dim3(kThreads), 0, st, kind, n, m, seed, src, dst);
  KC_HIP(hipGetLastError());
}

void build_csr_device(Context* ctx, const uint64_t* src, const uint64_t* dst, uint64_t n, uint64_t m,
                      uint64_t* xadj, uint64_t* adj, hipStream_t st) {
  Scratch scratch(ctx);
  uint64_t* deg = scratch.get<uint64_t>(n);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, deg, n, uint64_t{0});
  if (m)
    hipLaunchKernelGGL(degree_kernel, dim3(grid_for(ctx, m)), dim3(kThreads), 0, st, src, dst, n, m,
                       reinterpret_cast<unsigned long long*>(deg), ctx->d_err);
  check_err(ctx, st);
  exclusive_scan_u64(ctx, deg, xadj, n, st);
  hipLaunchKernelGGL(xadj_end_kernel, dim3(1), dim3(64), 0, st, xadj, n, 2 * m);
  if (m == 0) return;
  uint64_t* cursor = scratch.get<uint64_t>(n);
  KC_HIP(hipMemcpyAsync(cursor, xadj, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(scatter_edges_kernel, dim3(grid_for(ctx, m)), dim3(kThreads), 0, st, src, dst, n, m,
                     reinterpret_cast<unsigned long long*>(cursor), adj);
  KC_HIP(hipGetLastError());
}

void bfs_forest_device(Context* ctx, const uint64_t* xadj, const uint64_t* adj, uint64_t n, uint64_t* parent,
                       hipStream_t st) {
  if (n == 0) return;
  Phase ph(ctx, "bfs_forest", st);
  Scratch scratch(ctx);
  auto* label = reinterpret_cast<unsigned long long*>(scratch.get<uint64_t>(n));
  auto* flag = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st,
                     reinterpret_cast<uint64_t*>(label), n);
  // connected components; each round either lowers some label or ends
  for (uint64_t round = 0;; round++) {
    if (round > 4 * 64 + n) throw KcError{KC_EDEVICE, "connected components did not converge"};
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, sizeof(uint64_t), st));
    hipLaunchKernelGGL(cc_hook_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, xadj, adj, n, label, flag);
    for (;;) {
      KC_HIP(hipMemsetAsync(ctx->d_counters + 1, 0, sizeof(uint64_t), st));
      hipLaunchKernelGGL(cc_jump_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, st, n, label, flag + 1);
      uint64_t moved = 0;
      read_counters(ctx, st, 1, 1, &moved);
      if (!moved) break;
    }
    uint64_t hooked = 0;
    read_counters(ctx, st, 0, 1, &hooked);
    if (!hooked) break;
  }
  // multi-source BFS from every component minimum
  uint64_t* cur = scratch.get<uint64_t>(n);
  uint64_t* next = scratch.get<uint64_t>(n);
  KC_HIP(hipMemsetAsync(ctx->d_counters, 0, sizeof(uint64_t), st));
  hipLaunchKernelGGL(bfs_roots_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, label, n, parent, cur,
                     flag);
  uint64_t nf = 0;
  read_counters(ctx, st, 0, 1, &nf);
  uint64_t visited = nf;
  while (nf) {
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, sizeof(uint64_t), st));
    hipLaunchKernelGGL(bfs_level_kernel, dim3(grid_for(ctx, nf)), dim3(kThreads), 0, st, xadj, adj, cur, nf,
                       reinterpret_cast<unsigned long long*>(parent), next, flag);
    read_counters(ctx, st, 0, 1, &nf);
    visited += nf;
    if (visited > n) throw KcError{KC_EDEVICE, "BFS visited a vertex twice"};
    std::swap(cur, next);
  }
  if (visited != n) throw KcError{KC_EINVAL, "adjacency is not symmetric: BFS left vertices unvisited"};
}

}  // namespace kc
