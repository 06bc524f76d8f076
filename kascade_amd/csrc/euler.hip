// Euler-tour tree rooting (lib/kascade/eulertour.cpp:33-127, :235-278,
// :337-377) on one device.
//
// Element numbering of the tour (2n elements, built without a hash map; the
// two edges of a vertex share one 32-byte pair, so a leaf's down -> up step
// stays inside the cache line the down edge was read from):
//   element 2v     (v non-root): upward edge   (v -> parent v), weight +1
//   element 2v + 1 (v non-root): downward edge (parent v -> v), weight -1
//   element 2r     (r a root)  : the tour's terminal dummy (rank 0, self loop),
//                                the per-root dummy of eulertour.cpp:92-104
//   element 2r + 1 (r a root)  : unused, an isolated self loop
// Children are grouped by parent either with a stable device radix sort (the
// CSR of reverse_rooted_tree / make_graph, successor_utils.cpp:236-300) or by
// atomic exchange into per-parent lists.  With the cyclic adjacency order
// [parent, children...] of each vertex the tour links are
// (compute_euler_tour_impl, succ[(u_prev, v)] = (v, u_next)):
//   down(c)   -> down(first child of c), or up(c) if c is a leaf
//   up(c)     -> down(next sibling of c), or up(parent) after the last child;
//                for a root parent that is the root's dummy (element 2 root)
// Ranking the tour gives, for element 2v, root = 2 x the tree root (the
// dummy's element id) and rank = depth(v): exactly the (root, depth) that
// map_euler_tour_back writes; the ranking's output map keeps the even
// elements and halves the root ids.  Odd elements are never written out.
#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;

// R: the tour's record width (Rec, or the narrow Rec8 for tours below 2^30
// records).
template <class R>
__global__ void __launch_bounds__(kThreads)
tree_keys_kernel(const uint64_t* __restrict__ parent, uint64_t n, uint32_t* __restrict__ keys,
                 uint32_t* __restrict__ vals, R* __restrict__ tour, uint32_t* err,
                 unsigned long long* roots) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool is_root = false;
  if (v < n) {
    uint64_t p = parent[v];
    if (p >= n) {
      raise_err(err, kErrRange);
      p = v;
    }
    is_root = (p == v);
    keys[v] = is_root ? static_cast<uint32_t>(n) : static_cast<uint32_t>(p);
    vals[v] = static_cast<uint32_t>(v);
    if (is_root) {
      tour[2 * v] = RecOps<R>::encode(2 * v, 0, 2 * v);          // terminal dummy
      tour[2 * v + 1] = RecOps<R>::encode(2 * v + 1, 0, 2 * v + 1);  // unused slot
    } else {
      tour[2 * v + 1] = RecOps<R>::encode(2 * v, -1, 2 * v + 1);     // leaf default: down(v) -> up(v)
    }
  }
  wave_append(roots, is_root);
}

template <class R>
__global__ void __launch_bounds__(kThreads)
tour_link_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                 const uint64_t* __restrict__ parent, uint64_t n, R* __restrict__ tour,
                 uint64_t* __restrict__ heads, unsigned long long* nheads) {
  uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool head = false;
  uint32_t c = 0;
  if (j < n) {
    const uint32_t p = keys[j];
    if (p != n) {  // roots sort last
      c = vals[j];
      const bool first = (j == 0) || keys[j - 1] != p;
      const bool last = (j + 1 == n) || keys[j + 1] != p;
      // up(c): next sibling's down edge, or up(p) / p's dummy after the last child
      tour[2 * uint64_t{c}] = RecOps<R>::encode(last ? 2 * uint64_t{p} : 2 * uint64_t{vals[j + 1]} + 1, 1, 2 * uint64_t{c});
      // down(p) -> down(first child); a root p has no down edge, and its
      // first child's down edge is where its tour starts
      if (first) {
        if (parent[p] != p)
          tour[2 * uint64_t{p} + 1] = RecOps<R>::encode(2 * uint64_t{c} + 1, -1, 2 * uint64_t{p} + 1);
        else
          head = true;
      }
    }
  }
  const uint64_t k = wave_append(nheads, head);
  if (head) heads[k] = 2 * uint64_t{c} + 1;
}

// Linked build: children lists by atomic exchange on the parent's head slot
// (one random atomic per vertex, one per group of a wave's lanes that share
// a parent); sib[v] = the sibling inserted before v.  The
// child order is whatever order the exchanges happened in; any cyclic order of
// a vertex's neighbours gives a valid tour and the same (root, depth).
__global__ void __launch_bounds__(kThreads)
tour_link_atomic_kernel(const uint64_t* __restrict__ parent, uint64_t n, uint32_t* __restrict__ head,
                        uint32_t* __restrict__ sib, uint32_t* err, unsigned long long* roots) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool is_root = false;
  uint64_t p = 0;
  if (v < n) {
    p = parent[v];
    if (p >= n) raise_err(err, kErrRange);
    is_root = (p == v) || p >= n;
  }
  link_child(head, sib, p, static_cast<uint32_t>(v), v < n && !is_root);
  wave_append(roots, is_root);
}

// Writes both tour elements of v as one 32-byte pair (coalesced):
//   up(v)   -> down(sibling inserted before v), or up(p) / p's dummy
//   down(v) -> down(head child of v), or up(v) for a leaf
// and for a root the terminal dummy plus the unused slot.
template <class R>
__global__ void __launch_bounds__(kThreads)
tour_pairs_kernel(const uint64_t* __restrict__ parent, uint64_t n, const uint32_t* __restrict__ head,
                  const uint32_t* __restrict__ sib, R* __restrict__ tour,
                  uint64_t* __restrict__ heads, unsigned long long* nheads) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool tour_head = false;
  uint32_t h = 0xFFFFFFFFu;
  if (v < n) {
    const uint64_t p = parent[v];
    h = head[v];
    Rec up, down;
    if (p == v || p >= n) {
      up = Rec{2 * v, 0};
      down = Rec{2 * v + 1, 0};
      tour_head = h != 0xFFFFFFFFu;  // a root's tour starts at its head child's down edge
    } else {
      const uint32_t s = sib[v];
      up = Rec{s != 0xFFFFFFFFu ? 2 * uint64_t{s} + 1 : 2 * p, 1};
      down = Rec{h != 0xFFFFFFFFu ? 2 * uint64_t{h} + 1 : 2 * v, -1};
    }
    tour[2 * v] = RecOps<R>::encode(up.s, up.w, 2 * v);
    tour[2 * v + 1] = RecOps<R>::encode(down.s, down.w, 2 * v + 1);
  }
  const uint64_t k = wave_append(nheads, tour_head);
  if (tour_head) heads[k] = 2 * uint64_t{h} + 1;
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t value) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) p[i] = value;
}

unsigned bits_for(uint64_t x) {
  unsigned b = 1;
  while (b < 64 && (uint64_t{1} << b) <= x) b++;
  return b;
}

}  // namespace

// Builds the 2n-element tour into `tour`; returns the number of roots.
// sorted: children grouped by a stable radix sort of (parent, child) pairs, so
// the tour is the same on every run; otherwise children are linked by atomic
// exchange (faster; the tour's child order may differ between runs, the ranks
// never do).
// heads (capacity n) receives the element where each rooted tour starts (the
// down edge of the root's first child); *nheads their number.
template <class R>
static uint64_t build_tour(Context* ctx, const uint64_t* parent, uint64_t n, R* tour, bool sorted,
                           uint64_t* heads, uint64_t* nheads, hipStream_t st) {
  if (n >= (uint64_t{1} << 32) - 1) throw KcError{KC_EINVAL, "tree rooting supports n < 2^32 - 1"};
  Phase ph(ctx, "tour_build", st);
  Scratch scratch(ctx);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  uint64_t counts[2] = {0, 0};
  if (!sorted) {
    uint32_t* head = scratch.get<uint32_t>(n);
    uint32_t* sib = scratch.get<uint32_t>(n);
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, 2 * sizeof(uint64_t), st));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, head, n,
                       0xFFFFFFFFu);
    hipLaunchKernelGGL(tour_link_atomic_kernel, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st,
                       parent, n, head, sib, ctx->d_err, ctr);
    hipLaunchKernelGGL(tour_pairs_kernel<R>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, parent, n,
                       head, sib, tour, heads, ctr + 1);
    read_counters(ctx, st, 0, 2, counts);
    *nheads = counts[1];
    return counts[0];
  }
  uint32_t* keys = scratch.get<uint32_t>(n);
  uint32_t* vals = scratch.get<uint32_t>(n);
  KC_HIP(hipMemsetAsync(ctx->d_counters, 0, 2 * sizeof(uint64_t), st));
  hipLaunchKernelGGL(tree_keys_kernel<R>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, parent,
                     n, keys, vals, tour, ctx->d_err, ctr);
  radix_sort_pairs_u32(ctx, keys, vals, n, bits_for(n), st);
  hipLaunchKernelGGL(tour_link_kernel<R>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, keys,
                     vals, parent, n, tour, heads, ctr + 1);
  read_counters(ctx, st, 0, 2, counts);
  *nheads = counts[1];
  return counts[0];
}

// ---------------------------------------------------------------------------
// compute_euler_tour (eulertour.cpp:379-392) in the reference's own element
// order, so that a tour handed over as an input list is the same list.  On
// one PE the reference builds the CSR of reverse_rooted_tree / make_graph
// (successor_utils.cpp:236-300): the edges of vertex v are emitted in the
// order of the loop index that creates them (child c at index c, the parent
// at index v) and filled back to front, so v's neighbours are its children
// above v in descending order, then its parent, then its children below v in
// descending order.  Edge (v, nb_v[i]) is element off(v) + i, off = exclusive
// scan of the degrees; compute_euler_tour_impl (eulertour.cpp:76-111) links
// succ[(x, y)] = (y, the neighbour after x in y's cyclic order), except that
// the wrap to a root's first neighbour goes to the root's dummy, element
// E + (roots with children before it), E = 2(n - roots); dummies and the
// slots of childless roots are self loops (:116).  Weights (:45-53): +1 on an
// upward edge, -1 on a downward edge, 0 on the dummies.
// Here the neighbour positions follow from each vertex's child count k, the
// number L of children below it, and a child's rank m among its siblings
// (ascending ids, from a stable radix sort of (parent, child)):
//   pos of child c_m in p's list = k-1-m if m >= L, else k-m with a parent
//   slot (p not a root) or k-1-m without; the parent sits at k - L.
// ---------------------------------------------------------------------------
namespace {

__global__ void __launch_bounds__(kThreads)
ref_keys_kernel(const uint64_t* __restrict__ parent, uint64_t n, uint32_t* __restrict__ keys,
                uint32_t* __restrict__ vals, uint32_t* err, unsigned long long* roots) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool is_root = false;
  if (v < n) {
    uint64_t p = parent[v];
    if (p >= n) {
      raise_err(err, kErrRange);
      p = v;
    }
    is_root = (p == v);
    keys[v] = is_root ? static_cast<uint32_t>(n) : static_cast<uint32_t>(p);
    vals[v] = static_cast<uint32_t>(v);
  }
  wave_append(roots, is_root);
}

// Over the sorted pairs: first / one-past-last position of each parent's
// children, and one past the last child below the parent (0 = none).
__global__ void __launch_bounds__(kThreads)
ref_groups_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint64_t n,
                  uint64_t* __restrict__ start, uint64_t* __restrict__ end, uint64_t* __restrict__ lpos) {
  uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= n) return;
  const uint32_t p = keys[j];
  if (p == n) return;  // roots sort last
  const bool first = (j == 0) || keys[j - 1] != p;
  const bool last = (j + 1 == n) || keys[j + 1] != p;
  if (first) start[p] = j;
  if (last) end[p] = j + 1;
  const uint32_t c = vals[j];
  if (c < p && (last || vals[j + 1] > p)) lpos[p] = j + 1;
}

// deg(v) = children + (v not a root); droot(v) = 1 for a root with children.
__global__ void __launch_bounds__(kThreads)
ref_degree_kernel(const uint64_t* __restrict__ parent, uint64_t n, const uint64_t* __restrict__ start,
                  const uint64_t* __restrict__ end, uint64_t* __restrict__ deg, uint64_t* __restrict__ droot) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (v >= n) return;
  const uint64_t k = end[v] ? end[v] - start[v] : 0;
  const bool root = parent[v] == v || parent[v] >= n;
  deg[v] = k + (root ? 0 : 1);
  droot[v] = (root && k > 0) ? 1 : 0;
}

// Both elements of every tree edge (c, p): (c, p) = element off(c) + k_c - L_c
// and (p, c) = element off(p) + pos_p(c).
__global__ void __launch_bounds__(kThreads)
ref_link_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                const uint64_t* __restrict__ parent, uint64_t n, const uint64_t* __restrict__ start,
                const uint64_t* __restrict__ end, const uint64_t* __restrict__ lpos,
                const uint64_t* __restrict__ off, const uint64_t* __restrict__ droot, uint64_t E,
                uint64_t* __restrict__ tsucc, int64_t* __restrict__ tw) {
  uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= n) return;
  const uint64_t p = keys[j];
  if (p == n) return;
  const uint64_t c = vals[j];
  const uint64_t m = j - start[p];
  const uint64_t kp = end[p] - start[p];
  const uint64_t lp = lpos[p] ? lpos[p] - start[p] : 0;
  const bool proot = parent[p] == p;
  const uint64_t kc = end[c] ? end[c] - start[c] : 0;
  const uint64_t lc = lpos[c] ? lpos[c] - start[c] : 0;
  const uint64_t degp = kp + (proot ? 0 : 1);
  const uint64_t degc = kc + 1;
  const uint64_t pos_pc = m >= lp ? kp - 1 - m : kp - m - (proot ? 1 : 0);
  // (c, p) -> (p, the neighbour after c in p's list), or p's dummy on the wrap
  const uint64_t a = off[c] + (kc - lc);
  const uint64_t ip = (pos_pc + 1) % degp;
  tsucc[a] = (proot && ip == 0) ? E + droot[p] : off[p] + ip;
  // (p, c) -> (c, the neighbour after p in c's list)
  const uint64_t b = off[p] + pos_pc;
  tsucc[b] = off[c] + (kc - lc + 1) % degc;
  if (tw) {
    tw[a] = 1;
    tw[b] = -1;
  }
}

__global__ void __launch_bounds__(kThreads)
ref_dummies_kernel(uint64_t E, uint64_t roots, uint64_t* __restrict__ tsucc, int64_t* __restrict__ tw) {
  uint64_t t = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (t >= roots) return;
  tsucc[E + t] = E + t;
  if (tw) tw[E + t] = 0;
}

}  // namespace

uint64_t compute_euler_tour_device(Context* ctx, const uint64_t* parent, uint64_t n, uint64_t* tour_succ,
                                   int64_t* tour_w, hipStream_t st) {
  if (n == 0) return 0;
  check_records(2 * n);
  Scratch scratch(ctx);
  uint32_t* keys = scratch.get<uint32_t>(n);
  uint32_t* vals = scratch.get<uint32_t>(n);
  uint64_t* start = scratch.get<uint64_t>(n);
  uint64_t* end = scratch.get<uint64_t>(n);
  uint64_t* lpos = scratch.get<uint64_t>(n);
  uint64_t* deg = scratch.get<uint64_t>(n);
  uint64_t* droot = scratch.get<uint64_t>(n);
  uint64_t* off = scratch.get<uint64_t>(n);
  uint64_t* droot_off = scratch.get<uint64_t>(n);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  const unsigned blocks = blocks_for(n, kThreads);
  KC_HIP(hipMemsetAsync(ctx->d_counters, 0, sizeof(uint64_t), st));
  KC_HIP(hipMemsetAsync(end, 0, n * sizeof(uint64_t), st));
  KC_HIP(hipMemsetAsync(lpos, 0, n * sizeof(uint64_t), st));
  hipLaunchKernelGGL(ref_keys_kernel, dim3(blocks), dim3(kThreads), 0, st, parent, n, keys, vals, ctx->d_err, ctr);
  uint64_t roots = 0;
  read_counters(ctx, st, 0, 1, &roots);  // also raises an out-of-range parent
  radix_sort_pairs_u32(ctx, keys, vals, n, bits_for(n), st);
  hipLaunchKernelGGL(ref_groups_kernel, dim3(blocks), dim3(kThreads), 0, st, keys, vals, n, start, end, lpos);
  hipLaunchKernelGGL(ref_degree_kernel, dim3(blocks), dim3(kThreads), 0, st, parent, n, start, end, deg, droot);
  exclusive_scan_u64(ctx, deg, off, n, st);
  exclusive_scan_u64(ctx, droot, droot_off, n, st);
  const uint64_t E = 2 * (n - roots);
  hipLaunchKernelGGL(ref_link_kernel, dim3(blocks), dim3(kThreads), 0, st, keys, vals, parent, n, start, end, lpos,
                     off, droot_off, E, tour_succ, tour_w);
  hipLaunchKernelGGL(ref_dummies_kernel, dim3(blocks_for(roots, kThreads)), dim3(kThreads), 0, st, E, roots,
                     tour_succ, tour_w);
  check_err(ctx, st);
  return E + roots;
}

void root_tree_device(Context* ctx, const uint64_t* parent, uint64_t n, uint64_t* root,
                      int64_t* depth, const kc_tree_cfg* cfg, hipStream_t st) {
  if (n == 0) return;
  check_records(2 * n);
  Scratch scratch(ctx);
  uint64_t* heads = scratch.get<uint64_t>(n);
  const bool sorted = cfg && cfg->tour_build == KC_TOUR_SORTED;
  const kc_rank_cfg* rc = cfg ? &cfg->rank : nullptr;
  RankTuning tune = tuning_for(ctx, n, rc, KC_TUNING_TREE);
  const int algo = resolve_algorithm(rc, n, true);
  if (narrow_wanted(ctx, 2 * n, algo, tune)) {
    // an 8-byte tour (weights +-1 and distances below 2n always fit)
    Scratch narrow(ctx);
    Rec8* tour8 = narrow.get<Rec8>(2 * n);
    uint64_t nheads = 0;
    const uint64_t roots = build_tour(ctx, parent, n, tour8, sorted, heads, &nheads, st);
    if (narrow_ids_fit(2 * n, 2 * roots)) {
      const TourRegion tr{0, heads, nheads};
      const uint32_t redo = rank_records_narrow(ctx, tour8, 2 * n, n, root, depth, 2 * roots, tune, st, 1, &tr);
      if (redo == 0) {
        check_err(ctx, st);
        ctx->narrow_calls++;
        return;
      }
      ctx->narrow_redos++;
      if (redo & kErrOverflow) tune.launch = KC_LAUNCH_HOST;
    }
  }
  Rec* tour = scratch.get<Rec>(2 * n);
  for (;;) {
    uint64_t nheads = 0;
    const uint64_t roots = build_tour(ctx, parent, n, tour, sorted, heads, &nheads, st);
    // rank_via_euler_tour_select_algorithm (eulertour.cpp:282-321)
    const TourRegion tr{0, heads, nheads};
    if (rank_records(ctx, tour, 2 * n, n, root, depth, 2 * roots, tune, algo, st, /*out_shift=*/1, &tr)) break;
    tune.launch = KC_LAUNCH_HOST;  // a capacity overflowed: redo host-sized
  }
  check_err(ctx, st);
}

}  // namespace kc
