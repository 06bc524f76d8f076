// List ranking on one MI355X: sync pointer doubling and the sparse ruling set.
//
// Reference algorithms (niklas-uhl/kascade):
//   pointer doubling     lib/kascade/pointer_doubling_generic.hpp:608-700
//   sparse ruling set    lib/kascade/sparse_ruling_set.cpp:45-393
//
// Layout: the level-0 pass packs succ[] / weight[] into 16-byte records so that
// every random access of the chase is one 16-byte load.  A claimed record is
// overwritten in place with {ruler id | root flag, distance from the ruler}; an
// unclaimed record keeps {successor, weight}.
//
// Two drivers run the same kernels.  Host-sized levels (srs_level,
// rank_records_impl) read each level's counts on the host to size the next;
// device-sized levels (ds_srs_level, ds_rank) fix capacities up front and keep
// the counts in device memory, so a whole ranking is one launch sequence that
// rank_records can capture and replay as a HIP graph.
#include <algorithm>
#include <cstdio>

#include "common.hpp"
#include "device.hpp"

namespace kc {

namespace {

constexpr unsigned kThreads = 256;

// Counter slots in Context::d_counters.
enum Slot : int {
  kSlotTails = 0,
  kSlotWork = 1,
  kSlotExtra = 2,
  kSlotUnclaimed = 3,
  kSlotActive = 4,
  // chase chain lengths: max, max of the complement (= ~min), sum
  kSlotChainMax = 5,
  kSlotChainMinC = 6,
  kSlotChainSum = 7,
  // ruler spawn: resumed chases handed out, idle lanes, spawns, continuations, stop flag
  kSlotSpawnWork = 8,
  kSlotSpawnIdle = 9,
  kSlotSpawned = 10,
  kSlotSpawnCont = 11,
  kSlotSpawnStop = 12,
  kSlotCount = 13,
};

// Writes element e's final (root, rank): as a segment-local u32 root (root_id
// is then already local), through the job table of a mixed batch, or at the
// OutMap position (non-temporal: written once, never read).
__device__ __forceinline__ void emit(const JobOut& jo, OutMap om, uint64_t n_out, uint64_t* __restrict__ root,
                                     int64_t* __restrict__ rank, uint64_t e, uint64_t root_id, int64_t r) {
  if (jo.seg_root) {
    store_stream(&jo.seg_root[e], static_cast<uint32_t>(root_id));
    store_stream(&rank[e], r);
    return;
  }
  if (jo.lists == nullptr && jo.trees == nullptr) {
    if (!out_selected(om, e, n_out)) return;
    store_stream(&root[out_pos(om, e)], root_id >> om.shift);
    store_stream(&rank[out_pos(om, e)], r);
    return;
  }
  if (e < jo.nl) {
    uint32_t j = jo.list_block[e >> 8];
    while (j + 1 < jo.nlists && jo.lists[j + 1].base <= e) j++;
    const uint64_t base = jo.lists[j].base;
    store_stream(&jo.lists[j].root_out[e - base], root_id - base);
    store_stream(&jo.lists[j].rank_out[e - base], r);
    return;
  }
  const uint64_t x = e - jo.nl;
  if (x & 1) return;  // downward tour edge
  const uint64_t t = x >> 1;
  uint32_t j = jo.tree_block[t >> 8];
  while (j + 1 < jo.ntrees && jo.trees[j + 1].base <= t) j++;
  const uint64_t base = jo.trees[j].base;
  store_stream(&jo.trees[j].root_out[t - base], ((root_id - jo.nl) >> 1) - base);
  store_stream(&jo.trees[j].rank_out[t - base], r);
}

__global__ void zero_u64_kernel(unsigned long long* p, int count) {
  int i = threadIdx.x;
  if (i < count) p[i] = 0;
}

// set_initial_ranking_state semantics for NULL weights (list_ranking.cpp:60-80);
// range and tail-weight validation; tail count.  Narrow records: a weight
// that does not fit sets *wide (the caller then packs wide records).
template <class R>
__global__ void __launch_bounds__(kThreads)
pack_kernel(const uint64_t* __restrict__ succ, const int64_t* __restrict__ w, uint64_t n,
            R* __restrict__ rec, uint32_t* err, unsigned long long* tails, unsigned long long* wide) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool tail = false;
  if (v < n) {
    uint64_t s = succ[v];
    int64_t wt;
    if (s >= n) {
      raise_err(err, kErrRange);
      s = v;
    }
    tail = (s == v);
    if (w) {
      wt = w[v];
      if (tail && wt != 0) raise_err(err, kErrRootWeight);
      if (!RecOps<R>::fits(wt)) *wide = 1;
    } else {
      wt = tail ? 0 : 1;
    }
    rec[v] = RecOps<R>::encode(s, wt, v);
  }
  wave_append(tails, tail);
}

// ---------------------------------------------------------------------------
// Ruler chase (sparse_ruling_set.cpp:157-265, sync engine
// ruler_chasing_engine.hpp:57-101).  Persistent lanes: each lane chases one
// ruler's segment and, when done, takes the next ruler from a global counter
// (wave-aggregated), so lanes stay busy however uneven the ruler gaps are.
// A chase claims every element it passes; it stops at the next ruler (a
// computed predicate) or at a tail, which gets a fresh sublist id.  The
// sublist link (ruler -> next ruler, gap) is the reversed-orientation message
// of work_on_item :242-251, written directly in sublist ids.
// ---------------------------------------------------------------------------
// Ruler spawn (sparse_ruling_set.cpp:186-231, :250-256) in three launches.
// The reference starts a new ruler at an unreached element whenever a chase
// ends, so its last long gaps are split while chases still run.  Here:
//  1. the chase (kChaseSuspend) runs as usual until fewer than a set number
//     of lanes still chase (the lanes report in quanta once the ruler pool is
//     empty); then every wave writes the state of its unfinished chases (and
//     rulers it took but did not start) into a continuation list and exits;
//  2. spawn_mark_kernel probes candidates, one per block of 2^logC ids at a
//     hashed offset (computed, like the rulers): an unreached candidate c
//     becomes a new ruler j (a fresh sublist id, as a reached tail gets),
//     its record {j | root flag | aux flag, 0}, and its chase is appended to
//     the continuation list;
//  3. the chase (kChaseResume) runs the continuation list; a chase that
//     reaches a spawned ruler links to it, as to a computed one.
// The launch boundaries order every claim before the next pass, so no lane
// ever waits on another and no access needs to be atomic.
struct Cont {
  uint64_t ruler;  // sublist id of the chase's ruler (bit 63: its first step is still to do)
  uint64_t cur;    // next element to look at
  int64_t dist;    // distance from the ruler to cur
  uint64_t pad;
};
constexpr uint64_t kContFirst = uint64_t{1} << 63;
constexpr uint64_t kSpawnBits = kRootFlag | kAuxFlag;

struct SpawnArgs {
  unsigned long long* idle_ctr = nullptr;  // lanes idle after the ruler pool (in quanta)
  uint64_t idle_goal = 0;                  // suspend once this many are idle
  unsigned long long* stop = nullptr;      // set when the goal is reached
  Cont* cont = nullptr;                    // continuation list
  unsigned long long* cont_ctr = nullptr;  // its length
  uint64_t cont_cap = 0;
};

enum ChaseMode : int { kChasePlain = 0, kChaseSuspend = 1, kChaseResume = 2 };

// The last launch of a device-sized level's chase writes the size of the
// ruler sublist (rulers + tour heads + tails reached and spawned rulers) from
// its last workgroup to finish, so no separate launch has to.
struct LevelEnd {
  uint64_t* m_out = nullptr;                 // where the sublist size goes (nullptr: not this launch)
  unsigned long long* blocks_done = nullptr;  // workgroups finished (zeroed with the level's counters)
};
constexpr int kIdleQuantum = 16;  // lanes per idle report
// Iterations between two reads of the stop flag.  Each read is an agent-scope
// load that goes past L2; in the tail, where few waves remain, reads every 32
// steps cost 1 % of a 2^27 ranking and every 8 steps 3 % (128: the best of
// 32 / 128 / 256 / 512 measured).
constexpr uint32_t kStopPollMask = 127;

// kTrace: also count chain lengths (the per-level trace); off in plain calls.
// R: the level's record width (Rec, or Rec8 at a narrow top level).
template <bool kTrace, int kMode, class R>
__global__ void __launch_bounds__(kThreads)
chase_kernel(R* __restrict__ rec, uint64_t n, uint32_t logK, uint32_t seed, uint64_t nR,
             const uint64_t* __restrict__ forced, uint64_t nF, uint64_t tour_lo,
             uint64_t cap, Rec* __restrict__ sub, uint64_t* __restrict__ sub_orig,
             unsigned long long* work_ctr, unsigned long long* extra_ctr, uint32_t* err,
             uint32_t chunk, unsigned long long* chain, const uint64_t* __restrict__ n_dev, SpawnArgs sp,
             LevelEnd le) {
  // an earlier kernel found invalid input: stop, as the host-sized path does
  // (decided once per workgroup, so every wave reaches the barrier at the end)
  if (__syncthreads_or(threadIdx.x == 0 && *err != 0)) return;
  using Ops = RecOps<R>;
  bool narrow_over = false;  // a claim distance did not fit the narrow record
  const uint64_t lt = lanemask_lt();
  if (n_dev) {  // device-sized level: n is a capacity, the size is in memory
    n = *n_dev < n ? *n_dev : n;
    nR = (n + (uint64_t{1} << logK) - 1) >> logK;
  }
  // rulers [0, nR) sit at computed places; [nR, nRF) start at the tour heads
  const uint64_t nRF = nR + nF;
  // work items: rulers, or (resume) continuations
  uint64_t nwork = nRF;
  if constexpr (kMode == kChaseResume) {
    nwork = *sp.cont_ctr;
    if (nwork > sp.cont_cap) nwork = sp.cont_cap;
  }
  // downward tour edges carry no claim (their results are never read)
  auto claims = [&](uint64_t e) { return e < tour_lo || ((e - tour_lo) & 1) == 0; };
  // chain lengths of this lane (trace statistics, one reduction per wave at exit)
  uint64_t steps = 0, len_max = 0, len_minc = 0, len_sum = 0;
  auto chain_end = [&]() {
    if constexpr (kTrace) {
      len_max = steps > len_max ? steps : len_max;
      len_minc = ~steps > len_minc ? ~steps : len_minc;
      len_sum += steps;
    }
  };
  // Wave-uniform chunk of work ids [cbase, cend): one atomic per `chunk`
  // items, so idle lanes refill every step without touching the counter.
  uint64_t cbase = 0, cend = 0;
  bool wave_done = false;
  bool active = false;
  bool first = false;
  uint64_t ruler = 0, cur = 0;
  int64_t dist = 0;
  uint32_t idle_reported = 0;  // (suspend) lanes of this wave reported idle
  uint32_t iter = 0;
  auto flush_trace = [&]() {
    if constexpr (kTrace) {
      for (int o = 32; o > 0; o >>= 1) {
        const uint64_t a = __shfl_xor(len_max, o), b = __shfl_xor(len_minc, o);
        len_max = a > len_max ? a : len_max;
        len_minc = b > len_minc ? b : len_minc;
        len_sum += __shfl_xor(len_sum, o);
      }
      if (__lane_id() == 0 && len_sum) {
        atomicMax(&chain[0], static_cast<unsigned long long>(len_max));
        atomicMax(&chain[1], static_cast<unsigned long long>(len_minc));
        atomicAdd(&chain[2], static_cast<unsigned long long>(len_sum));
      }
    }
  };
  for (;;) {
    uint64_t need = __ballot(!active);
    if (need != 0 && !wave_done) {
      if (cbase == cend) {
        unsigned long long b = 0;
        if (__lane_id() == 0) b = atomicAdd(work_ctr, static_cast<unsigned long long>(chunk));
        b = __shfl(b, 0);
        if (b >= nwork) {
          wave_done = true;
        } else {
          cbase = b;
          cend = b + chunk < nwork ? b + chunk : nwork;
        }
      }
      if (!wave_done) {
        const uint64_t avail = cend - cbase;
        const uint64_t r = __popcll(need & lt);
        if (!active && r < avail) {
          if constexpr (kMode == kChaseResume) {
            const Cont c = sp.cont[cbase + r];
            ruler = c.ruler & ~kContFirst;
            first = (c.ruler & kContFirst) != 0;
            cur = c.cur;
            dist = c.dist;
            if constexpr (kTrace) steps = first ? 0 : 1;
          } else {
            ruler = cbase + r;
            cur = ruler < nR ? ruler_of_block(ruler, n, logK, seed) : forced[ruler - nR];
            first = true;
          }
          active = true;
        }
        const uint64_t cnt = __popcll(need);
        cbase += cnt < avail ? cnt : avail;
      }
    }
    if constexpr (kMode == kChaseSuspend) {
      if (wave_done) {
        // report idle lanes in quanta (a handful of atomics per wave); the
        // report that reaches the goal raises the stop flag
        const uint32_t idle = static_cast<uint32_t>(__popcll(__ballot(!active)));
        const uint32_t q = idle / kIdleQuantum * kIdleQuantum;
        if (q > idle_reported) {
          if (__lane_id() == 0) {
            const unsigned long long add = q - idle_reported;
            const unsigned long long before = atomicAdd(sp.idle_ctr, add);
            if (before < sp.idle_goal && before + add >= sp.idle_goal)
              __hip_atomic_store(sp.stop, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          idle_reported = q;
        }
      }
      if ((iter & kStopPollMask) == 0) {
        unsigned long long stop = 0;
        if (__lane_id() == 0) stop = __hip_atomic_load(sp.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__shfl(stop, 0) != 0) {
          // hand over the unfinished chases and the rulers taken but not started
          const uint64_t k = wave_append(sp.cont_ctr, active);
          if (active && k < sp.cont_cap)
            sp.cont[k] = Cont{ruler | (first ? kContFirst : 0), cur, dist, 0};
          for (uint64_t b = cbase; b < cend; b += 64) {
            const uint64_t id = b + __lane_id();
            const bool mine = id < cend;
            const uint64_t k2 = wave_append(sp.cont_ctr, mine);
            if (mine && k2 < sp.cont_cap) {
              const uint64_t at = id < nR ? ruler_of_block(id, n, logK, seed) : forced[id - nR];
              sp.cont[k2] = Cont{id | kContFirst, at, 0, 0};
            }
          }
          flush_trace();
          break;
        }
      }
      iter++;
    }
    if (__ballot(active) == 0) {
      if (wave_done) {
        flush_trace();
        break;
      }
      continue;
    }
    if (!active) continue;
    if (first) {
      first = false;
      sub_orig[ruler] = cur;
      if (ruler >= nR && is_ruler(cur, n, logK, seed)) {
        // a tour head that is also a computed ruler is chased as that one;
        // this id precedes it in the sublist at distance 0
        sub[ruler] = Rec{cur >> logK, 0};
        active = false;
        continue;
      }
      // the ruler's own record: claim it with distance 0
      const Rec rr = Ops::decode(rec[cur], cur);
      if (claims(cur)) rec[cur] = Ops::encode(ruler | kRootFlag, 0, cur);
      if constexpr (kTrace) steps = 1;
      if (rr.s == cur) {
        sub[ruler] = Rec{ruler, 0};  // the ruler is a tail
        active = false;
        chain_end();
      } else {
        dist = rr.w;
        cur = rr.s;
      }
      continue;
    }
    if (is_ruler(cur, n, logK, seed)) {
      const uint64_t id = cur >> logK;
      if (id == ruler) raise_err(err, kErrCycle);  // came back to its own ruler
      sub[ruler] = Rec{id, dist};
      active = false;
      chain_end();
      continue;
    }
    const Rec r = Ops::decode(rec[cur], cur);
    if (r.s & kSpawnBits) {
      if (kMode == kChaseResume && (r.s & kSpawnBits) == kSpawnBits) {
        const uint64_t id = r.s & kIdMask;
        if (id == ruler) raise_err(err, kErrCycle);  // came back to its own spawned ruler
        sub[ruler] = Rec{id, dist};  // reached a spawned ruler
        active = false;
        chain_end();
        continue;
      }
      raise_err(err, kErrNotList);  // reached twice: in-degree > 1
      active = false;
    } else if (r.s == cur) {
      const uint64_t j = nRF + atomicAdd(extra_ctr, 1ull);
      if (j < cap) {
        sub[j] = Rec{j, 0};
        sub_orig[j] = cur;
        rec[cur] = Ops::encode(j | kRootFlag, 0, cur);
        sub[ruler] = Rec{j, dist};
      } else {
        raise_err(err, kErrNotList);
      }
      active = false;
      chain_end();
    } else {
      if (claims(cur)) {
        if constexpr (Ops::kNarrow) narrow_over |= !Ops::fits(dist);
        rec[cur] = Ops::encode(ruler | kRootFlag, dist, cur);
      }
      dist += r.w;
      cur = r.s;
      if constexpr (kTrace) steps++;
    }
  }
  if constexpr (Ops::kNarrow)
    if (narrow_over) raise_err(err, kErrNarrow);
  if (le.m_out) {
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();  // this workgroup's sublist links and extra ids before its report
      const unsigned long long done = atomicAdd(le.blocks_done, 1ull);
      if (done + 1 == gridDim.x) {
        __threadfence();
        const unsigned long long extra = __hip_atomic_load(extra_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        le.m_out[0] = nR + nF + extra;
      }
    }
  }
}

// Step 2 of the spawn: one thread per candidate block.  An unreached
// candidate (not a computed ruler, record unclaimed) becomes spawned ruler j
// while spawn slots last; its chase starts at its successor.  In a tour
// region only upward edges (even offsets from tour_lo) carry claims, so a
// candidate there is moved to the upward edge of its vertex: a downward
// edge's record stays unclaimed after a chase has passed it.
template <class R>
__global__ void __launch_bounds__(kThreads)
spawn_mark_kernel(R* __restrict__ rec, uint64_t n, uint32_t logK, uint32_t seed, uint64_t nR, uint64_t nF,
                  uint32_t logC, uint32_t cseed, uint64_t cap, uint64_t budget, Rec* __restrict__ sub,
                  uint64_t* __restrict__ sub_orig, unsigned long long* extra_ctr, unsigned long long* spawn_ctr,
                  SpawnArgs sp, const uint64_t* __restrict__ n_dev, const uint32_t* __restrict__ err,
                  uint64_t tour_lo) {
  if (*err) return;
  if (n_dev) {
    n = *n_dev < n ? *n_dev : n;
    nR = (n + (uint64_t{1} << logK) - 1) >> logK;
  }
  const uint64_t nC = (n + (uint64_t{1} << logC) - 1) >> logC;
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool want = false;
  uint64_t c = 0;
  Rec r{0, 0};
  if (b < nC) {
    c = ruler_of_block(b, n, logC, cseed);
    bool ok = true;
    if (c >= tour_lo && ((c - tour_lo) & 1)) {
      // a downward edge: take the upward edge of the same vertex (c - 1), or
      // the next one when c - 1 lies in the previous candidate block
      if (c > (b << logC)) {
        c--;
      } else {
        c++;
        ok = c < n && c < ((b + 1) << logC);
      }
    }
    if (ok && !is_ruler(c, n, logK, seed)) {
      r = RecOps<R>::decode(rec[c], c);
      want = !(r.s & kSpawnBits);
    }
  }
  const uint64_t slot = wave_append(spawn_ctr, want);
  const bool go = want && slot < budget;
  const uint64_t j = nR + nF + wave_append(extra_ctr, go);
  if (!go) return;
  if (j >= cap) return;  // cannot happen: cap holds every tail plus the budget
  rec[c] = RecOps<R>::encode(j | kSpawnBits, 0, c);
  sub_orig[j] = c;
  if (r.s == c) {
    sub[j] = Rec{j, 0};  // a tail: the spawned ruler ends its list
    return;
  }
  const uint64_t k = wave_append(sp.cont_ctr, true);
  if (k < sp.cont_cap) sp.cont[k] = Cont{j, r.s, r.w, 0};
}

// The rulers' final (root, rank), as propagation gathers them.  root and rank
// hold every value; when the values fit 32 bits, rank4 (the rank as int32)
// and packed (the root id in the low, the rank in the high 32 bits) let
// propagation gather 4 bytes per element (one root) or 8 (several) instead
// of 8 or 16, from a table half the size.  flags (kTableFlags words, zeroed
// before map_roots_kernel): [0] two rulers have different roots, [1] a rank
// does not fit int32, [2] a root id does not fit 32 bits.
constexpr int kTableFlags = 3;
struct RulerTable {
  uint64_t* root;
  int64_t* rank;
  int32_t* rank4;
  uint64_t* packed;
  uint32_t* flags;
};

// rank4 and packed take 12 bytes per ruler; they live in the memory of the
// ruler sublist's records (16 bytes per ruler), which are dead once the
// sublist has been ranked.
template <class R>
RulerTable ruler_table(uint64_t* root, int64_t* rank, R* dead_sub, uint64_t cap, uint32_t* flags) {
  static_assert(sizeof(R) >= 12, "rank4 and packed need 12 bytes per ruler");
  uint64_t* packed = reinterpret_cast<uint64_t*>(dead_sub);
  return RulerTable{root, rank, reinterpret_cast<int32_t*>(packed + cap), packed, flags};
}

// The ranked sublist T (root as a sublist id, rank) -> the ruler table.  T
// itself is only read, so every thread compares against the same T[0].s;
// the flags are set with plain stores of 1 (every writer writes the same).
__global__ void __launch_bounds__(kThreads)
map_roots_kernel(const Rec* __restrict__ T, const uint64_t* __restrict__ sub_orig, uint64_t m, RulerTable tab,
                 const uint64_t* __restrict__ m_dev, const uint32_t* __restrict__ err) {
  if (*err) return;  // the sublist ranking failed: T is not valid
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (m_dev && *m_dev < m) m = *m_dev;
  if (i >= m) return;
  const Rec t = T[i];
  const uint64_t root_id = sub_orig[t.s];
  const int64_t r = t.w;
  if (t.s != T[0].s && tab.flags[0] == 0) tab.flags[0] = 1;
  if (r != static_cast<int32_t>(r) && tab.flags[1] == 0) tab.flags[1] = 1;
  if ((root_id >> 32) && tab.flags[2] == 0) tab.flags[2] = 1;
  tab.root[i] = root_id;
  tab.rank[i] = r;
  tab.rank4[i] = static_cast<int32_t>(r);
  tab.packed[i] = (root_id & 0xFFFFFFFFull) | (static_cast<uint64_t>(static_cast<uint32_t>(r)) << 32);
}

// Segmented batches: the tails of the ruler sublist (self loops; the roots of
// every ruler) get their segment-local id in place of the global one, so the
// roots that map_roots_kernel hands to propagation are already local.  Only
// root ids of the sublist are ever looked up in sub_orig.
__global__ void __launch_bounds__(kThreads)
seg_local_tails_kernel(const Rec* __restrict__ sub, uint64_t* __restrict__ sub_orig, uint64_t m,
                       const uint64_t* __restrict__ m_dev, const uint64_t* __restrict__ seg_off, uint64_t nseg,
                       const uint32_t* __restrict__ err) {
  if (*err) return;
  uint64_t j = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (m_dev && *m_dev < m) m = *m_dev;
  if (j >= m || sub[j].s != j) return;
  const uint64_t g = sub_orig[j];
  sub_orig[j] = g - seg_off[segment_of(seg_off, nseg, g)];
}

// Rank propagation (ruler_propagation.hpp:148-257, pull mode): every claimed
// element reads its ruler's (root, rank) and subtracts its distance from the
// ruler.  Unclaimed elements (list prefixes no ruler reaches, the analogue of
// fixup_unreached, post_invert.hpp:195-264) are appended to a compact list and
// their record is re-marked with that compact id.
//
// Two launch shapes.  One element per thread, one counter atomic per wave
// that has an unclaimed element: the fastest when few waves have one.
// Grouped (kGrouped): a grid of resident size strides over the level, and
// each wave appends the unclaimed elements of kPropGroup strides with one
// atomic.  A forest of many short lists has an unreached prefix in most
// waves, and their atomics on the single counter then dominate (MI355X,
// 200,000 lists of 2^6..2^10 elements, 6.9e7 in all: 10.5 ms one per
// thread, 0.79 ms grouped); where they are rare every looped shape measured
// slower (20,000 lists of 2^8..2^20 with spawn, 2.5e9 elements: 11.3 ms one
// per thread, 18.9 grouped, 13.4-15.1 with 4-16 elements per thread), most
// likely because each load's wait also waits for the previous element's
// streaming stores (one wait counter covers both on this GPU).
constexpr unsigned kPropGroup = 32;

template <class R>
__device__ __forceinline__ bool propagate_element(R* __restrict__ rec, uint64_t v, const RulerTable& tab, int table,
                                                  uint64_t root0, uint64_t n_out, uint64_t* __restrict__ root,
                                                  int64_t* __restrict__ rank, OutMap om, uint64_t tour_lo,
                                                  uint32_t* err, const JobOut& jo) {
  const Rec r = RecOps<R>::decode(load_stream(&rec[v]), v);
  if (r.s & kRootFlag) {
    const uint64_t id = r.s & kIdMask;
    if (table == 0) {
      emit(jo, om, n_out, root, rank, v, root0, tab.rank4[id] - r.w);
    } else if (table == 1) {
      emit(jo, om, n_out, root, rank, v, root0, tab.rank[id] - r.w);
    } else if (table == 2) {
      const uint64_t e = tab.packed[id];
      emit(jo, om, n_out, root, rank, v, e & 0xFFFFFFFFull, static_cast<int32_t>(e >> 32) - r.w);
    } else {
      emit(jo, om, n_out, root, rank, v, tab.root[id], tab.rank[id] - r.w);
    }
    return false;
  }
  if (v < tour_lo) return true;
  if (((v - tour_lo) & 1) == 0) {
    // an unreached upward edge of a tour: an isolated root, else a cycle
    if (r.s == v)
      emit(jo, om, n_out, root, rank, v, v, 0);
    else
      raise_err(err, kErrCycle);
  }
  return false;
}

// Appends element v (unclaimed) at compact slot k.
template <class R>
__device__ __forceinline__ void collect_element(R* __restrict__ rec, uint64_t v, uint64_t k,
                                                uint64_t* __restrict__ u_ids, Rec* __restrict__ u_orig,
                                                uint64_t ucap, uint32_t* err, bool device_sized) {
  if (k < ucap) {
    u_ids[k] = v;
    u_orig[k] = RecOps<R>::decode(rec[v], v);  // unclaimed: unchanged since read
    RecOps<R>::set_s(&rec[v], k | kAuxFlag);
  } else if (device_sized) {
    // beyond the capacity: host-sized levels collect the rest in a second
    // pass; device-sized ones flag the call to be redone host-sized
    raise_err(err, kErrOverflow);
  }
}

template <class R, bool kGrouped>
__global__ void __launch_bounds__(kThreads)
propagate_kernel(R* __restrict__ rec, uint64_t n, uint64_t n_out, RulerTable tab,
                 uint64_t* __restrict__ root, int64_t* __restrict__ rank, OutMap om,
                 uint64_t* __restrict__ u_ids, Rec* __restrict__ u_orig, uint64_t ucap,
                 unsigned long long* u_ctr, uint64_t tour_lo, uint32_t* err,
                 const uint64_t* __restrict__ n_dev, bool device_sized, JobOut jo) {
  if (*err) return;
  if (n_dev && *n_dev < n) n = *n_dev;
  // All rulers in one list (one root): gather only the rank, 4 bytes when
  // every rank fits (table 0; else 1).  Several roots: one 8-byte (root,
  // rank) entry when both fit (2), else the two 8-byte tables (3).  Smaller
  // tables keep more in L2.
  const bool rank32 = tab.flags[1] == 0;
  const int table = tab.flags[0] == 0 ? (rank32 ? 0 : 1) : (rank32 && tab.flags[2] == 0 ? 2 : 3);
  const uint64_t root0 = tab.root[0];
  if constexpr (!kGrouped) {
    const uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const bool unclaimed =
        v < n && propagate_element(rec, v, tab, table, root0, n_out, root, rank, om, tour_lo, err, jo);
    const uint64_t k = wave_append(u_ctr, unclaimed);
    if (unclaimed) collect_element(rec, v, k, u_ids, u_orig, ucap, err, device_sized);
  } else {
    const unsigned lane = __lane_id();
    const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x) / 64;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads;
    for (uint64_t g = wave * 64; g < n; g += kPropGroup * stride) {
      uint32_t mask = 0;  // bit q: element g + q * stride + lane is unclaimed
      for (unsigned q = 0; q < kPropGroup; q++) {
        const uint64_t wb = g + q * stride;
        if (wb >= n) break;  // uniform over the wave
        if (wb + lane < n &&
            propagate_element(rec, wb + lane, tab, table, root0, n_out, root, rank, om, tour_lo, err, jo))
          mask |= 1u << q;
      }
      const uint32_t cnt = static_cast<uint32_t>(__popc(mask));
      uint32_t inc = cnt;  // inclusive prefix sum over the wave
      for (unsigned o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
      }
      const uint32_t total = __shfl(inc, 63);
      if (total == 0) continue;
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(u_ctr, static_cast<unsigned long long>(total));
      base = __shfl(base, 0);
      uint64_t k = base + inc - cnt;
      for (; mask != 0; mask &= mask - 1, k++)
        collect_element(rec, g + uint64_t(__ffs(mask) - 1) * stride + lane, k, u_ids, u_orig, ucap, err,
                        device_sized);
    }
  }
}

// Second collection pass after the unclaimed list outgrew its first capacity:
// appends the still unmarked unclaimed elements after slot `first`.
__global__ void __launch_bounds__(kThreads)
collect_unclaimed_kernel(Rec* __restrict__ rec, uint64_t n, uint64_t* __restrict__ u_ids,
                         Rec* __restrict__ u_orig, uint64_t first, uint64_t cap,
                         unsigned long long* ctr) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  Rec r{0, 0};
  bool mine = false;
  if (v < n) {  // n: the end of the list region (tours are never collected)
    r = rec[v];
    mine = !(r.s & (kRootFlag | kAuxFlag));
  }
  const uint64_t k = first + wave_append(ctr, mine);
  if (mine && k < cap) {
    u_ids[k] = v;
    u_orig[k] = r;
    rec[v].s = k | kAuxFlag;
  }
}

// Unclaimed sub-problem: an unclaimed element's successor is unclaimed (link
// inside the sub-problem) or already final (it becomes a local tail whose
// offset/root come from that final element), or it is a true tail.
template <class R>
__global__ void __launch_bounds__(kThreads)
unclaimed_build_kernel(const R* __restrict__ rec, const uint64_t* __restrict__ u_ids,
                       const Rec* __restrict__ u_orig, uint64_t nu, const uint64_t* __restrict__ Troot,
                       const int64_t* __restrict__ T8, Rec* __restrict__ urec, Rec* __restrict__ ext,
                       const unsigned long long* __restrict__ nu_dev, const uint32_t* __restrict__ err,
                       const uint64_t* __restrict__ seg_off, uint64_t nseg) {
  if (*err) return;
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (nu_dev && *nu_dev < nu) nu = *nu_dev;
  if (k >= nu) return;
  const uint64_t v = u_ids[k];
  const Rec r0 = u_orig[k];
  const uint64_t s = r0.s;
  if (s == v) {
    // a tail no ruler reached: its own root (segment-local in a segmented batch)
    urec[k] = Rec{k, 0};
    ext[k] = Rec{seg_off ? v - seg_off[segment_of(seg_off, nseg, v)] : v, 0};
    return;
  }
  const Rec ms = RecOps<R>::decode(rec[s], s);
  if ((ms.s & kAuxFlag) && !(ms.s & kRootFlag)) {
    urec[k] = Rec{ms.s & kIdMask, r0.w};
    ext[k] = Rec{0, 0};
  } else if (ms.s & kRootFlag) {
    // claimed: its final root / rank follow from its ruler's entry
    const uint64_t id = ms.s & kIdMask;
    urec[k] = Rec{k, 0};
    ext[k] = Rec{Troot[id], r0.w + T8[id] - ms.w};
  } else {
    // not collected: a device-sized level past its capacity, which is redone
    // host-sized; any in-range values do
    urec[k] = Rec{k, 0};
    ext[k] = Rec{0, 0};
  }
}

__global__ void __launch_bounds__(kThreads)
unclaimed_fix_kernel(const uint64_t* __restrict__ u_ids, uint64_t nu, uint64_t n_out,
                     const Rec* __restrict__ ures, const Rec* __restrict__ ext,
                     uint64_t* __restrict__ root, int64_t* __restrict__ rank, OutMap om,
                     const unsigned long long* __restrict__ nu_dev, const uint32_t* __restrict__ err,
                     JobOut jo) {
  if (*err) return;  // the prefix ranking failed: ures is not valid
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (nu_dev && *nu_dev < nu) nu = *nu_dev;
  if (k >= nu) return;
  const uint64_t v = u_ids[k];
  const Rec res = ures[k];
  const Rec e = ext[res.s];
  emit(jo, om, n_out, root, rank, v, e.s, res.w + e.w);
}

// ---------------------------------------------------------------------------
// One-workgroup solver for small levels: Jacobi pointer doubling in LDS.
// ---------------------------------------------------------------------------
constexpr uint32_t kSmallMax = 8192;
constexpr unsigned kSmallThreads = 1024;
constexpr uint32_t kSmallFlag = 0x80000000u;

// Jacobi doubling over s_succ / s_rank (n <= kSmallMax entries; a tail carries
// kSmallFlag); every thread of the workgroup calls it.  Afterwards s_succ[i]
// is i's tail (flagged) and s_rank[i] its distance to it.  More than 16
// rounds mean a cycle without a tail.
__device__ void lds_doubling(uint32_t* s_succ, int64_t* s_rank, uint32_t n, uint32_t* err) {
  constexpr int kPer = kSmallMax / kSmallThreads;
  uint32_t ns[kPer];
  int64_t nr[kPer];
  int rounds = 0;
  for (;;) {
    int any = 0;
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const uint32_t i = threadIdx.x + q * kSmallThreads;
      if (i < n) {
        const uint32_t s = s_succ[i];
        ns[q] = s;
        nr[q] = s_rank[i];
        if (!(s & kSmallFlag)) {
          ns[q] = s_succ[s];
          nr[q] += s_rank[s];
          any |= !(ns[q] & kSmallFlag);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const uint32_t i = threadIdx.x + q * kSmallThreads;
      if (i < n) {
        s_succ[i] = ns[q];
        s_rank[i] = nr[q];
      }
    }
    rounds++;
    const int more = __syncthreads_or(any);
    if (!more) break;
    if (rounds > 15) {
      if (threadIdx.x == 0) raise_err(err, kErrCycle);
      break;
    }
  }
}

__global__ void __launch_bounds__(kSmallThreads)
small_rank_kernel(const Rec* __restrict__ rec, uint32_t n, uint32_t n_out, uint64_t* __restrict__ root,
                  int64_t* __restrict__ rank, OutMap om, uint32_t* err, const uint64_t* __restrict__ n_dev) {
  if (*err) return;
  if (n_dev && *n_dev < n) {  // device-sized: every element is an output
    n = static_cast<uint32_t>(*n_dev);
    n_out = n;
  }
  __shared__ uint32_t s_succ[kSmallMax];
  __shared__ int64_t s_rank[kSmallMax];
  for (uint32_t i = threadIdx.x; i < n; i += kSmallThreads) {
    const Rec r = rec[i];
    uint32_t s = static_cast<uint32_t>(r.s & kIdMask);
    if ((r.s & (kRootFlag | kAuxFlag)) || r.s >= n) {
      raise_err(err, kErrRange);
      s = i;
    }
    if (s == i) {
      if (r.w != 0) raise_err(err, kErrCycle);  // a self-loop carrying weight closes a cycle
      s |= kSmallFlag;
    }
    s_succ[i] = s;
    s_rank[i] = r.w;
  }
  __syncthreads();
  lds_doubling(s_succ, s_rank, n, err);
  for (uint32_t i = threadIdx.x; i < n; i += kSmallThreads) {
    if (!out_selected(om, i, n_out)) continue;
    root[out_pos(om, i)] = (s_succ[i] & ~kSmallFlag) >> om.shift;
    rank[out_pos(om, i)] = s_rank[i];
  }
}

// The unclaimed sub-problem of a level in one workgroup, when it has at most
// kSmallMax elements: unclaimed_build_kernel's links, the LDS solver and
// unclaimed_fix_kernel's outputs in one launch instead of three.  ext (nu
// records of global scratch) is written and read back by this workgroup only.
template <class R>
__global__ void __launch_bounds__(kSmallThreads)
unclaimed_small_kernel(const R* __restrict__ rec, const uint64_t* __restrict__ u_ids,
                       const Rec* __restrict__ u_orig, uint64_t nu, const uint64_t* __restrict__ Troot,
                       const int64_t* __restrict__ T8, Rec* ext,
                       uint64_t n_out, uint64_t* __restrict__ root, int64_t* __restrict__ rank, OutMap om,
                       const unsigned long long* __restrict__ nu_dev, uint32_t* err, JobOut jo,
                       const uint64_t* __restrict__ seg_off, uint64_t nseg) {
  if (*err) return;  // the level failed (or overflowed): its records are not valid
  if (nu_dev && *nu_dev < nu) nu = *nu_dev;
  __shared__ uint32_t s_succ[kSmallMax];
  __shared__ int64_t s_rank[kSmallMax];
  const uint32_t m = static_cast<uint32_t>(nu);
  for (uint32_t k = threadIdx.x; k < m; k += kSmallThreads) {
    const uint64_t v = u_ids[k];
    const Rec r0 = u_orig[k];
    uint32_t link = k;
    int64_t w = 0;
    Rec e{0, 0};
    if (r0.s == v) {
      e = Rec{seg_off ? v - seg_off[segment_of(seg_off, nseg, v)] : v, 0};
    } else {
      const Rec ms = RecOps<R>::decode(rec[r0.s], r0.s);
      if ((ms.s & kAuxFlag) && !(ms.s & kRootFlag)) {
        link = static_cast<uint32_t>(ms.s & kIdMask);
        w = r0.w;
      } else if (ms.s & kRootFlag) {
        const uint64_t id = ms.s & kIdMask;
        e = Rec{Troot[id], r0.w + T8[id] - ms.w};
      }
    }
    ext[k] = e;
    s_succ[k] = link == k ? (k | kSmallFlag) : link;
    s_rank[k] = w;
  }
  __syncthreads();
  lds_doubling(s_succ, s_rank, m, err);
  for (uint32_t k = threadIdx.x; k < m; k += kSmallThreads) {
    const Rec e = ext[s_succ[k] & ~kSmallFlag];
    emit(jo, om, n_out, root, rank, u_ids[k], e.s, s_rank[k] + e.w);
  }
}

// ---------------------------------------------------------------------------
// Sync pointer doubling (Wyllie), Jacobi over double-buffered records: every
// reply is taken from the state before the round, as the reference builds all
// replies before any update (request_aggregation_scheme.hpp:156-179 then
// update_via_read_offsets, pointer_doubling_generic.hpp:62-81).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
pd_flag_roots_kernel(Rec* __restrict__ a, uint64_t n, uint32_t* err, const uint64_t* __restrict__ n_dev) {
  if (*err) return;
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (n_dev && *n_dev < n) n = *n_dev;
  if (v >= n) return;
  Rec r = a[v];
  if ((r.s & (kRootFlag | kAuxFlag)) || r.s >= n) {
    raise_err(err, kErrRange);
    r.s = v;
  }
  if (r.s == v) {  // initialize_active_vertices :418-438
    if (r.w != 0) raise_err(err, kErrCycle);
    r.s = v | kRootFlag;
    a[v] = r;
  }
}

// `active` is a flag (any element still unfinished), written by one lane per
// wave that has one: a plain store of 1 into the XCD's L2, read only by later
// launches (kernel boundaries write it back).  Agent-scope atomics here (a
// load, then a store, per wave) went past L2 to one address and cost about
// 85 us per round at 2^20.
__global__ void __launch_bounds__(kThreads)
pd_round_kernel(const Rec* __restrict__ a, Rec* __restrict__ b, uint64_t n,
                unsigned long long* active) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool still = false;
  if (v < n) {
    const Rec r = a[v];
    if (r.s & kRootFlag) {
      b[v] = r;
    } else {
      const Rec t = a[r.s];
      b[v] = Rec{t.s, r.w + t.w};
      still = !(t.s & kRootFlag);
    }
  }
  const uint64_t m = __ballot(still);
  if (m != 0 && __lane_id() == static_cast<unsigned>(__ffsll(static_cast<long long>(m)) - 1)) *active = 1ull;
}

__global__ void __launch_bounds__(kThreads)
pd_write_kernel(const Rec* __restrict__ a, uint64_t n, uint64_t n_out, uint64_t* __restrict__ root,
                int64_t* __restrict__ rank, OutMap om) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (v >= n || !out_selected(om, v, n_out)) return;
  const Rec r = a[v];
  root[out_pos(om, v)] = (r.s & ~kRootFlag) >> om.shift;
  rank[out_pos(om, v)] = r.w;
}

// Device-sized doubling: a fixed number of rounds, buf[r & 1] -> buf[(r + 1) & 1].
// done[r] = 1 when round r left no element unfinished; later rounds return
// at once, and the result is the output of the first such round.
__global__ void __launch_bounds__(kThreads)
pd_round_ds_kernel(const Rec* __restrict__ a, Rec* __restrict__ b, uint64_t n,
                   const uint64_t* __restrict__ n_dev, uint32_t* __restrict__ active, uint32_t r,
                   const uint32_t* __restrict__ err) {
  if (*err) return;
  if (r > 0 && active[r - 1] == 0) return;  // finished earlier
  if (n_dev && *n_dev < n) n = *n_dev;
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool still = false;
  if (v < n) {
    const Rec x = a[v];
    if (x.s & kRootFlag) {
      b[v] = x;
    } else {
      const Rec t = a[x.s];
      b[v] = Rec{t.s, x.w + t.w};
      still = !(t.s & kRootFlag);
    }
  }
  const uint64_t m = __ballot(still);
  if (m != 0 && __lane_id() == static_cast<unsigned>(__ffsll(static_cast<long long>(m)) - 1)) active[r] = 1u;
}

__global__ void __launch_bounds__(kThreads)
pd_write_ds_kernel(const Rec* __restrict__ buf0, const Rec* __restrict__ buf1, uint64_t n,
                   const uint64_t* __restrict__ n_dev, uint64_t n_out, uint64_t* __restrict__ root,
                   int64_t* __restrict__ rank, OutMap om, const uint32_t* __restrict__ active,
                   uint32_t rounds, uint32_t* err) {
  if (*err) return;
  uint32_t last = rounds;
  for (uint32_t r = 0; r < rounds; r++)
    if (active[r] == 0) {
      last = r;
      break;
    }
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (last == rounds) {  // still unfinished after every round: a cycle
    if (v == 0) raise_err(err, kErrCycle);
    return;
  }
  if (n_dev) {
    if (*n_dev < n) n = *n_dev;
    n_out = n;
  }
  if (v >= n || !out_selected(om, v, n_out)) return;
  const Rec x = ((last + 1) & 1) ? buf1[v] : buf0[v];
  root[out_pos(om, v)] = (x.s & ~kRootFlag) >> om.shift;
  rank[out_pos(om, v)] = x.w;
}

// The doubling state after some rounds, root flags kept (bits.hpp:9-21).
__global__ void __launch_bounds__(kThreads)
pd_state_kernel(const Rec* __restrict__ a, uint64_t n, uint64_t* __restrict__ succ,
                int64_t* __restrict__ rank) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (v >= n) return;
  const Rec r = a[v];
  succ[v] = r.s;
  rank[v] = r.w;
}

void zero_counters(Context* ctx, hipStream_t st) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st,
                     reinterpret_cast<unsigned long long*>(ctx->d_counters), kSlotCount);
}

const char* const kChaseNames[] = {"chase_ruler_L0", "chase_ruler_L1", "chase_ruler_L2",
                                   "chase_ruler_L3", "chase_ruler_L4", "chase_ruler_L5",
                                   "chase_ruler_L6", "chase_ruler_L7", "chase_ruler_L8"};
const char* const kPropNames[] = {"ruler_propagation_L0", "ruler_propagation_L1",
                                  "ruler_propagation_L2", "ruler_propagation_L3",
                                  "ruler_propagation_L4", "ruler_propagation_L5",
                                  "ruler_propagation_L6", "ruler_propagation_L7",
                                  "ruler_propagation_L8"};

unsigned log2_floor(uint64_t x) {
  unsigned r = 0;
  while (x > 1) {
    x >>= 1;
    r++;
  }
  return r;
}

void rank_records_impl(Context* ctx, Rec* rec, uint64_t n, uint64_t n_out, uint64_t* root,
                       int64_t* rank, OutMap om, uint64_t max_tails, const RankTuning& tune,
                       int algorithm, int level, hipStream_t st, const TourRegion* tours = nullptr,
                       const JobOut& jo = JobOut{});

uint32_t pd_records_impl(Context* ctx, Rec* rec, uint64_t n, uint64_t n_out, uint64_t* root,
                         int64_t* rank, OutMap om, hipStream_t st) {
  Phase ph(ctx, "pointer_doubling", st);
  Scratch scratch(ctx);
  Rec* other = scratch.get<Rec>(n);
  const unsigned blocks = blocks_for(n, kThreads);
  hipLaunchKernelGGL(pd_flag_roots_kernel, dim3(blocks), dim3(kThreads), 0, st, rec, n, ctx->d_err, nullptr);
  Rec* a = rec;
  Rec* b = other;
  unsigned long long* active = reinterpret_cast<unsigned long long*>(ctx->d_counters) + kSlotActive;
  const unsigned max_rounds = log2_floor(n) + 2;
  unsigned rounds = 0;
  for (;;) {
    zero_counters(ctx, st);
    hipLaunchKernelGGL(pd_round_kernel, dim3(blocks), dim3(kThreads), 0, st, a, b, n, active);
    std::swap(a, b);
    rounds++;
    uint64_t left = 0;
    read_counters(ctx, st, kSlotActive, 1, &left);
    if (left == 0) break;
    if (rounds > max_rounds) {
      check_err(ctx, st);
      throw KcError{KC_ECYCLE, "pointer doubling did not terminate: cycle without a root"};
    }
  }
  check_err(ctx, st);
  const uint64_t n_sel = std::min<uint64_t>(n, n_out << om.shift);
  hipLaunchKernelGGL(pd_write_kernel, dim3(blocks_for(n_sel, kThreads)), dim3(kThreads), 0, st, a,
                     n_sel, n_out, root, rank, om);
  return rounds;
}

// The MI355X's Infinity Cache (MALL), shared by its eight XCDs.
constexpr uint64_t kMallBytes = uint64_t{256} << 20;

// Ruler spacing of a level: the tuned K, unless the caller fixed it, capped
// for a forest of max_tails lists at the first power of two above
// (mean list length) / 16, so the unreached head prefixes (about K per list)
// stay within about an eighth of the level.  The table's spacings above 128
// (sizes beyond 2^27) were tuned on single lists.  A top level of more than
// 64 lists takes its rulers in id order, so the chases in flight (`lanes` of
// them, K ids apart) work in a window of about lanes x K records of
// record_bytes each, and the lists inside that window are all they touch:
// the spacing is capped so the window fits the MALL (20,000 lists, 2.5e9
// elements, 16-byte records, 3 chase workgroups per CU: 116 ms at K = 128
// (a 402 MB window), 106 ms at 64 (201 MB), 124 ms at 32; with 2 workgroups
// per CU 106.6 ms at 128 (268 MB)).
uint32_t level_spacing(const RankTuning& tune, int level, uint64_t n, uint64_t max_tails, uint64_t lanes,
                       uint64_t record_bytes) {
  uint32_t spacing = level == 0 ? tune.spacing_top : tune.spacing_deep;
  if (!(level == 0 ? tune.spacing_top_fixed : tune.spacing_deep_fixed)) {
    const uint64_t mean = n / std::max<uint64_t>(1, max_tails);
    uint32_t cap = 2;
    while (cap < spacing && uint64_t{cap} * 16 <= mean) cap *= 2;
    spacing = std::min(spacing, cap);
    if (level == 0 && max_tails > 64 && tune.window_cap) {
      uint32_t window = 128;
      while (window > 2 && lanes * window * record_bytes > kMallBytes) window /= 2;
      spacing = std::min(spacing, window);
    }
  }
  return spacing;
}

// Ruler spawn for a level: candidates one per block of K / 8 ids (at least
// 2), the chase suspended once fewer than lanes / 16 lanes still chase, at
// most one spawn per two rulers (`budget`: the spawned rulers' share of the
// sublist capacity).  Automatic for list levels (no tours) of at least 2^20
// elements with K >= 256: the tail it cuts grows with K, and below that its
// two extra launches cost more than they save (2^27 elements, 6 runs each:
// K = 128 chase 7.15 ms without spawn, 7.37-7.50 with it; K = 256 7.96
// without, 7.49-7.54 with it, whose 4 MB rank table then makes propagation
// 0.45 ms shorter than at K = 128).  KC_SPAWN_ON forces it on any level
// whose K exceeds 8, tour levels included (candidates on upward edges); AUTO
// keeps it off for tours, whose chase has no long tail at the tree rows'
// K of 128 and below (2^26-vertex tree at K = 128: level-0 chase 4.14 ms
// with or without spawn, the whole call 8.58 against 8.38 ms; at K = 256
// spawn helps, 4.50 against 4.97 ms, but K = 128 without it stays faster).
constexpr uint32_t kSpawnShift = 3;    // candidates per ruler: 2^kSpawnShift
constexpr uint64_t kSpawnActive = 16;  // suspend when fewer than lanes / this still chase
struct SpawnPlan {
  bool on = false;
  uint32_t logC = 0;
  uint64_t budget = 0;
  uint64_t idle_goal = 0;
  uint64_t cont_cap = 0;
};
SpawnPlan spawn_plan(const RankTuning& tune, uint64_t n, uint32_t logK, bool has_tours, unsigned chase_blocks) {
  SpawnPlan p;
  if (tune.spawn == KC_SPAWN_OFF || logK <= kSpawnShift) return p;
  if (tune.spawn != KC_SPAWN_ON && (has_tours || n < (uint64_t{1} << 20) || logK < 8)) return p;
  p.on = true;
  p.logC = logK - kSpawnShift;
  p.budget = std::max<uint64_t>(((n + (uint64_t{1} << logK) - 1) >> logK) / 2, 64);
  const uint64_t lanes = uint64_t{chase_blocks} * kThreads;
  p.idle_goal = lanes - lanes / kSpawnActive;
  // every lane's chase, the rulers a wave took but did not start, the spawns
  p.cont_cap = lanes + uint64_t{chase_blocks} * (kThreads / 64) * tune.chase_chunk + p.budget;
  return p;
}

// Launches propagation in the shape that suits the expected number of
// unreached elements: about `lists` prefixes of K elements each, or K / 8
// when spawn starts rulers inside them.  Grouped when they would put an
// unclaimed element into a quarter of the waves or more.
template <class R>
void launch_propagate(Context* ctx, hipStream_t st, uint64_t n, uint64_t lists, uint32_t spacing, bool spawn,
                      R* rec, uint64_t n_out, const RulerTable& tab, uint64_t* root, int64_t* rank, OutMap om,
                      uint64_t* u_ids, Rec* u_orig, uint64_t ucap, unsigned long long* u_ctr, uint64_t tour_lo,
                      const uint64_t* n_dev, bool device_sized, const JobOut& jo) {
  const uint64_t list_end = std::min(n, tour_lo);
  const uint64_t expected = std::min<uint64_t>(list_end, lists * (spawn ? spacing >> kSpawnShift : spacing));
  if (expected * 256 >= n && n > 0) {
    const unsigned grid = static_cast<unsigned>(
        std::min<uint64_t>(blocks_for(n, kThreads), static_cast<uint64_t>(ctx->num_cus) * 8));
    hipLaunchKernelGGL((propagate_kernel<R, true>), dim3(grid), dim3(kThreads), 0, st, rec, n, n_out, tab, root,
                       rank, om, u_ids, u_orig, ucap, u_ctr, tour_lo, ctx->d_err, n_dev, device_sized, jo);
  } else {
    hipLaunchKernelGGL((propagate_kernel<R, false>), dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, rec, n,
                       n_out, tab, root, rank, om, u_ids, u_orig, ucap, u_ctr, tour_lo, ctx->d_err, n_dev,
                       device_sized, jo);
  }
}


// Counter slots of one level's chase.
struct ChaseSlots {
  unsigned long long* work;
  unsigned long long* extra;
  unsigned long long* chain;
  unsigned long long* resume_work;
  unsigned long long* idle;
  unsigned long long* spawned;
  unsigned long long* cont_ctr;
  unsigned long long* stop;
};

// The level's chase: one launch, or with spawn the suspend, mark and resume
// launches.  The caller zeroed the counters.
template <class R>
void launch_chase(Context* ctx, Scratch& scratch, const SpawnPlan& sp, bool trace, unsigned chase_blocks, R* rec,
                  uint64_t n, uint32_t logK, uint32_t seed, uint64_t nR, const uint64_t* forced, uint64_t nF,
                  uint64_t tour_lo, uint64_t cap, Rec* sub, uint64_t* sub_orig, const ChaseSlots& cs,
                  uint32_t chunk, const uint64_t* n_dev, int level, hipStream_t st, LevelEnd le = LevelEnd{}) {
  Phase ph(ctx, kChaseNames[std::min(level, 8)], st);
  SpawnArgs sa;
  if (!sp.on) {
    auto kern = trace ? chase_kernel<true, kChasePlain, R> : chase_kernel<false, kChasePlain, R>;
    hipLaunchKernelGGL(kern, dim3(chase_blocks), dim3(kThreads), 0, st, rec, n, logK, seed, nR, forced, nF,
                       tour_lo, cap, sub, sub_orig, cs.work, cs.extra, ctx->d_err, chunk, cs.chain, n_dev, sa, le);
    return;
  }
  sa.idle_ctr = cs.idle;
  sa.idle_goal = sp.idle_goal;
  sa.stop = cs.stop;
  sa.cont = scratch.get<Cont>(sp.cont_cap);
  sa.cont_ctr = cs.cont_ctr;
  sa.cont_cap = sp.cont_cap;
  auto k1 = trace ? chase_kernel<true, kChaseSuspend, R> : chase_kernel<false, kChaseSuspend, R>;
  hipLaunchKernelGGL(k1, dim3(chase_blocks), dim3(kThreads), 0, st, rec, n, logK, seed, nR, forced, nF, tour_lo,
                     cap, sub, sub_orig, cs.work, cs.extra, ctx->d_err, chunk, cs.chain, n_dev, sa, LevelEnd{});
  const uint64_t nC = (n + (uint64_t{1} << sp.logC) - 1) >> sp.logC;
  hipLaunchKernelGGL(spawn_mark_kernel<R>, dim3(blocks_for(nC, kThreads)), dim3(kThreads), 0, st, rec, n, logK, seed,
                     nR, nF, sp.logC, seed ^ 0x5bd1e995u, cap, sp.budget, sub, sub_orig, cs.extra, cs.spawned, sa,
                     n_dev, ctx->d_err, tour_lo);
  auto k3 = trace ? chase_kernel<true, kChaseResume, R> : chase_kernel<false, kChaseResume, R>;
  hipLaunchKernelGGL(k3, dim3(chase_blocks), dim3(kThreads), 0, st, rec, n, logK, seed, nR, forced, nF, tour_lo,
                     cap, sub, sub_orig, cs.resume_work, cs.extra, ctx->d_err, chunk, cs.chain, n_dev, sa, le);
}

void srs_level(Context* ctx, Rec* rec, uint64_t n, uint64_t n_out, uint64_t* root,
               int64_t* rank, OutMap om, uint64_t max_tails, const RankTuning& tune,
               int level, hipStream_t st, const TourRegion* tours, const JobOut& jo) {
  const uint64_t tour_lo = tours ? tours->lo : ~uint64_t{0};
  const uint64_t nF = tours ? tours->nheads : 0;
  const uint64_t list_end = std::min(n, tour_lo);
  Scratch scratch(ctx);
  const uint64_t max_blocks = static_cast<uint64_t>(ctx->num_cus) * tune.chase_blocks_per_cu;
  const uint32_t spacing = level_spacing(tune, level, n, max_tails, max_blocks * kThreads, sizeof(Rec));
  const uint32_t logK = log2_floor(spacing);
  const uint64_t nR = (n + (uint64_t{1} << logK) - 1) >> logK;
  const uint64_t lanes_wanted = nR + nF;
  const unsigned chase_blocks = static_cast<unsigned>(
      std::max<uint64_t>(1, std::min<uint64_t>(max_blocks, (lanes_wanted + kThreads - 1) / kThreads)));
  const SpawnPlan sp = spawn_plan(tune, n, logK, tours != nullptr, chase_blocks);
  const uint64_t cap = nR + nF + std::min<uint64_t>(max_tails, n) + sp.budget;
  Rec* sub = scratch.get<Rec>(cap);
  Rec* T = scratch.get<Rec>(cap);
  uint64_t* sub_orig = scratch.get<uint64_t>(cap);
  const uint32_t seed = static_cast<uint32_t>(splitmix64(0xC0FFEEull + level * 0x1234567ull + n));

  zero_counters(ctx, st);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  const ChaseSlots cs{ctr + kSlotWork,      ctr + kSlotExtra,   ctr + kSlotChainMax, ctr + kSlotSpawnWork,
                      ctr + kSlotSpawnIdle, ctr + kSlotSpawned, ctr + kSlotSpawnCont, ctr + kSlotSpawnStop};
  launch_chase(ctx, scratch, sp, ctx->trace_on, chase_blocks, rec, n, logK, seed, nR, tours ? tours->heads : nullptr,
               nF, tour_lo, cap, sub, sub_orig, cs, tune.chase_chunk, nullptr, level, st);
  uint64_t extra = 0;
  read_counters(ctx, st, kSlotExtra, 1, &extra);  // also raises device errors
  const uint64_t spawned = std::min<uint64_t>(ctx->h_pinned[kSlotSpawned], sp.budget);  // copied by read_counters
  const uint64_t m = nR + nF + extra;
  if (jo.seg_off)
    hipLaunchKernelGGL(seg_local_tails_kernel, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, st, sub, sub_orig,
                       m, nullptr, jo.seg_off, jo.nseg, ctx->d_err);
  size_t trace_at = 0;
  if (ctx->trace_on) {
    // the counter block was copied to the pinned mirror by read_counters
    kc_trace_level t{};
    t.level = level;
    t.kind = KC_TRACE_SRS;
    t.subproblem = ctx->trace_subproblem;
    t.n = n;
    t.spacing = spacing;
    t.rulers = nR + nF;
    t.tail_rulers = extra - spawned;
    t.spawned = spawned;
    t.chain_length_max = ctx->h_pinned[kSlotChainMax];
    t.chain_length_min = nR + nF ? ~ctx->h_pinned[kSlotChainMinC] : 0;
    t.chain_length_avg = nR + nF + spawned
                             ? static_cast<double>(ctx->h_pinned[kSlotChainSum]) / static_cast<double>(nR + nF + spawned)
                             : 0;
    trace_at = ctx->trace.size();
    ctx->trace.push_back(t);
  }

  // rank the ruler sublist (recursion, sparse_ruling_set.cpp:300-309)
  const int role = ctx->trace_subproblem;  // the sublist recursion inherits it
  rank_records_impl(ctx, sub, m, m, &T[0].s, &T[0].w, OutMap{2, 0}, std::min(max_tails, m), tune, m > n / 2 ? KC_ALGO_POINTER_DOUBLING : KC_ALGO_SPARSE_RULING_SET,
                    level + 1, st);
  int64_t* T8 = scratch.get<int64_t>(m);
  uint64_t* Troot = scratch.get<uint64_t>(m);
  uint32_t* flags = scratch.get<uint32_t>(kTableFlags);
  KC_HIP(hipMemsetAsync(flags, 0, kTableFlags * sizeof(uint32_t), st));
  const RulerTable tab = ruler_table(Troot, T8, sub, m, flags);
  hipLaunchKernelGGL(map_roots_kernel, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, st, T,
                     sub_orig, m, tab, nullptr, ctx->d_err);

  // propagation + collection of unclaimed elements
  zero_counters(ctx, st);
  // unclaimed elements are list prefixes: few unless the lists are short
  const uint64_t ucap = std::min<uint64_t>(list_end, list_end / 8 + 65536);
  uint64_t* u_ids = scratch.get<uint64_t>(ucap);
  Rec* u_orig = scratch.get<Rec>(ucap);
  {
    Phase ph(ctx, kPropNames[std::min(level, 8)], st);
    launch_propagate(ctx, st, n, std::min(max_tails, n), spacing, sp.on, rec, n_out, tab, root, rank, om, u_ids,
                     u_orig, ucap, ctr + kSlotUnclaimed, tour_lo, nullptr, false, jo);
  }
  uint64_t nu = 0;
  read_counters(ctx, st, kSlotUnclaimed, 1, &nu);
  if (ctx->trace_on) ctx->trace[trace_at].unreached = nu;
  if (nu == 0) return;
  if (nu > ucap) {
    uint64_t* ids2 = scratch.get<uint64_t>(nu);
    Rec* orig2 = scratch.get<Rec>(nu);
    KC_HIP(hipMemcpyAsync(ids2, u_ids, ucap * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
    KC_HIP(hipMemcpyAsync(orig2, u_orig, ucap * sizeof(Rec), hipMemcpyDeviceToDevice, st));
    zero_counters(ctx, st);
    hipLaunchKernelGGL(collect_unclaimed_kernel, dim3(blocks_for(list_end, kThreads)), dim3(kThreads), 0, st,
                       rec, list_end, ids2, orig2, ucap, nu, ctr + kSlotUnclaimed);
    u_ids = ids2;
    u_orig = orig2;
  }
  Phase ph(ctx, "unclaimed", st);
  Rec* ext = scratch.get<Rec>(nu);
  if (nu <= kSmallMax && nu <= tune.small_n && !ctx->trace_on) {
    // small enough for one workgroup (the per-level trace keeps the
    // sub-problem as its own ranking)
    hipLaunchKernelGGL(unclaimed_small_kernel<Rec>, dim3(1), dim3(kSmallThreads), 0, st, rec, u_ids, u_orig, nu,
                       Troot, T8, ext, n_out, root, rank, om, nullptr, ctx->d_err, jo, jo.seg_off, jo.nseg);
    return;
  }
  Rec* urec = scratch.get<Rec>(nu);
  Rec* ures = scratch.get<Rec>(nu);
  hipLaunchKernelGGL(unclaimed_build_kernel<Rec>, dim3(blocks_for(nu, kThreads)), dim3(kThreads), 0, st,
                     rec, u_ids, u_orig, nu, Troot, T8, urec, ext, nullptr, ctx->d_err, jo.seg_off, jo.nseg);
  // every unclaimed prefix belongs to a different list of this level
  ctx->trace_subproblem = 1;
  rank_records_impl(ctx, urec, nu, nu, &ures[0].s, &ures[0].w, OutMap{2, 0}, std::min(nu, max_tails), tune,
                    nu > n / 2 ? KC_ALGO_POINTER_DOUBLING : KC_ALGO_SPARSE_RULING_SET, level + 1,
                    st);
  ctx->trace_subproblem = role;
  hipLaunchKernelGGL(unclaimed_fix_kernel, dim3(blocks_for(nu, kThreads)), dim3(kThreads), 0, st,
                     u_ids, nu, n_out, ures, ext, root, rank, om, nullptr, ctx->d_err, jo);
}

void rank_records_impl(Context* ctx, Rec* rec, uint64_t n, uint64_t n_out, uint64_t* root,
                       int64_t* rank, OutMap om, uint64_t max_tails, const RankTuning& tune,
                       int algorithm, int level, hipStream_t st, const TourRegion* tours, const JobOut& jo) {
  if (n == 0) return;
  if (n <= tune.small_n && n <= kSmallMax) {
    if (ctx->trace_on) {
      kc_trace_level t{};
      t.level = level;
      t.kind = KC_TRACE_SMALL;
      t.subproblem = ctx->trace_subproblem;
      t.n = n;
      ctx->trace.push_back(t);
    }
    Phase ph(ctx, "small_rank", st);
    hipLaunchKernelGGL(small_rank_kernel, dim3(1), dim3(kSmallThreads), 0, st, rec,
                       static_cast<uint32_t>(n), static_cast<uint32_t>(std::min<uint64_t>(n_out, n)), root, rank, om,
                       ctx->d_err, nullptr);
    return;
  }
  if (algorithm == KC_ALGO_POINTER_DOUBLING || level >= static_cast<int>(tune.max_levels)) {
    const uint32_t rounds = pd_records_impl(ctx, rec, n, n_out, root, rank, om, st);
    if (ctx->trace_on) {
      kc_trace_level t{};
      t.level = level;
      t.kind = KC_TRACE_PD;
      t.subproblem = ctx->trace_subproblem;
      t.n = n;
      t.rounds = rounds;
      ctx->trace.push_back(t);
    }
    return;
  }
  srs_level(ctx, rec, n, n_out, root, rank, om, max_tails, tune, level, st,
            tours && tours->lo < n ? tours : nullptr, jo);
}

// ---------------------------------------------------------------------------
// Device-sized levels.  The host fixes every level's capacity from (n, number
// of lists) alone; the sizes themselves stay in device memory (n_dev: the
// level's element count, at most n), so the whole ranking is one launch
// sequence with no host round trip, which a HIP graph can capture.  Each
// level zeroes its own counters with a kernel (no memset nodes).
// ---------------------------------------------------------------------------
enum DsSlot : int {
  kDsWork = 0, kDsExtra = 1, kDsUnclaimed = 2, kDsM = 3, kDsChain = 4,  // chain: 3 slots
  kDsSpawnWork = 7, kDsSpawnIdle = 8, kDsSpawned = 9, kDsSpawnCont = 10, kDsSpawnStop = 11,
  kDsBlocksDone = 12, kDsTableFlags = 13, kDsSlots = 15  // table flags: 2 slots (kTableFlags words)
};

template <class R>
void ds_rank(Context* ctx, R* rec, uint64_t n, const uint64_t* n_dev, uint64_t n_out, uint64_t* root,
             int64_t* rank, OutMap om, uint64_t max_tails, const RankTuning& tune, int algorithm, int level,
             hipStream_t st, const TourRegion* tours, const JobOut& jo = JobOut{});

void ds_pd(Context* ctx, Rec* rec, uint64_t n, const uint64_t* n_dev, uint64_t n_out, uint64_t* root,
           int64_t* rank, OutMap om, hipStream_t st) {
  Phase ph(ctx, "pointer_doubling", st);
  Scratch scratch(ctx);
  Rec* other = scratch.get<Rec>(n);
  const uint32_t rounds = log2_floor(n) + 2;
  // one flag per round, zeroed as (rounds + 1) / 2 u64 words
  uint32_t* active = scratch.get<uint32_t>(rounds + 1);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned long long*>(active),
                     static_cast<int>((rounds + 1) / 2));
  const unsigned blocks = blocks_for(n, kThreads);
  hipLaunchKernelGGL(pd_flag_roots_kernel, dim3(blocks), dim3(kThreads), 0, st, rec, n, ctx->d_err, n_dev);
  Rec* buf[2] = {rec, other};
  for (uint32_t r = 0; r < rounds; r++)
    hipLaunchKernelGGL(pd_round_ds_kernel, dim3(blocks), dim3(kThreads), 0, st, buf[r & 1], buf[(r + 1) & 1], n,
                       n_dev, active, r, ctx->d_err);
  hipLaunchKernelGGL(pd_write_ds_kernel, dim3(blocks_for(std::min<uint64_t>(n, n_out << om.shift), kThreads)),
                     dim3(kThreads), 0, st, rec, other, n, n_dev, n_out, root, rank, om, active, rounds,
                     ctx->d_err);
}

// A device-sized ruling-set level's shape, fixed on the host from its
// capacity n, the number of lists and the tuning: the ruler spacing, the
// chase grid, the spawn plan, the sublist capacity (every id a claim can
// carry is below cap) and the unreached-prefix capacity (every collected id
// is below ucap).
struct DsLevel {
  uint32_t spacing, logK;
  uint64_t nR, lists;
  unsigned chase_blocks;
  SpawnPlan sp;
  uint64_t cap, ucap;
};

DsLevel ds_level(const Context* ctx, uint64_t n, uint64_t max_tails, const RankTuning& tune, int level,
                 const TourRegion* tours, size_t record_bytes) {
  DsLevel L{};
  const uint64_t tour_lo = tours ? tours->lo : ~uint64_t{0};
  const uint64_t nF = tours ? tours->nheads : 0;
  const uint64_t list_end = std::min(n, tour_lo);
  const uint64_t max_blocks = static_cast<uint64_t>(ctx->num_cus) * tune.chase_blocks_per_cu;
  L.spacing = level_spacing(tune, level, n, max_tails, max_blocks * kThreads, record_bytes);
  L.logK = log2_floor(L.spacing);
  L.nR = (n + (uint64_t{1} << L.logK) - 1) >> L.logK;
  L.lists = std::min<uint64_t>(max_tails, n);
  L.chase_blocks = static_cast<unsigned>(
      std::max<uint64_t>(1, std::min<uint64_t>(max_blocks, (L.nR + nF + kThreads - 1) / kThreads)));
  L.sp = spawn_plan(tune, n, L.logK, tours != nullptr, L.chase_blocks);
  L.cap = L.nR + nF + L.lists + L.sp.budget;
  // Unreached head prefixes: about K elements per list; the capacity leaves
  // a wide margin (a prefix longer than 24 K has probability e^-24 on a
  // random list) and an overflow is redone host-sized.
  L.ucap = std::min<uint64_t>(list_end, 2 * uint64_t{L.spacing} * L.lists + 24 * uint64_t{L.spacing} + 4096);
  return L;
}

template <class R>
void ds_srs_level(Context* ctx, R* rec, uint64_t n, const uint64_t* n_dev, uint64_t n_out, uint64_t* root,
                  int64_t* rank, OutMap om, uint64_t max_tails, const RankTuning& tune, int level,
                  hipStream_t st, const TourRegion* tours, const JobOut& jo) {
  const uint64_t tour_lo = tours ? tours->lo : ~uint64_t{0};
  const uint64_t nF = tours ? tours->nheads : 0;
  Scratch scratch(ctx);
  const DsLevel L = ds_level(ctx, n, max_tails, tune, level, tours, sizeof(R));
  const uint32_t logK = L.logK;
  const uint64_t nR = L.nR;
  const uint64_t lists = L.lists;
  const unsigned chase_blocks = L.chase_blocks;
  const SpawnPlan& sp = L.sp;
  const uint64_t cap = L.cap;
  Rec* sub = scratch.get<Rec>(cap);
  Rec* T = scratch.get<Rec>(cap);
  uint64_t* sub_orig = scratch.get<uint64_t>(cap);
  int64_t* T8 = scratch.get<int64_t>(cap);
  uint64_t* Troot = scratch.get<uint64_t>(cap);
  unsigned long long* ctr = scratch.get<unsigned long long>(kDsSlots);
  const RulerTable tab = ruler_table(Troot, T8, sub, cap, reinterpret_cast<uint32_t*>(ctr + kDsTableFlags));
  uint64_t* m_dev = reinterpret_cast<uint64_t*>(ctr + kDsM);
  const uint32_t seed = static_cast<uint32_t>(splitmix64(0xC0FFEEull + level * 0x1234567ull + n));
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, ctr, kDsSlots);
  const ChaseSlots cs{ctr + kDsWork,      ctr + kDsExtra,   ctr + kDsChain,     ctr + kDsSpawnWork,
                      ctr + kDsSpawnIdle, ctr + kDsSpawned, ctr + kDsSpawnCont, ctr + kDsSpawnStop};
  // the chase's last workgroup writes the sublist's size to m_dev
  launch_chase(ctx, scratch, sp, false, chase_blocks, rec, n, logK, seed, nR, tours ? tours->heads : nullptr, nF,
               tour_lo, cap, sub, sub_orig, cs, tune.chase_chunk, n_dev, level, st,
               LevelEnd{m_dev, ctr + kDsBlocksDone});
  if (jo.seg_off)
    hipLaunchKernelGGL(seg_local_tails_kernel, dim3(blocks_for(cap, kThreads)), dim3(kThreads), 0, st, sub, sub_orig,
                       cap, m_dev, jo.seg_off, jo.nseg, ctx->d_err);
  // rank the ruler sublist (recursion, sparse_ruling_set.cpp:300-309)
  ds_rank(ctx, sub, cap, m_dev, cap, &T[0].s, &T[0].w, OutMap{2, 0}, lists, tune,
          cap > n / 2 ? KC_ALGO_POINTER_DOUBLING : KC_ALGO_SPARSE_RULING_SET, level + 1, st, nullptr);
  hipLaunchKernelGGL(map_roots_kernel, dim3(blocks_for(cap, kThreads)), dim3(kThreads), 0, st, T, sub_orig, cap,
                     tab, m_dev, ctx->d_err);

  const uint64_t ucap = L.ucap;
  uint64_t* u_ids = scratch.get<uint64_t>(std::max<uint64_t>(ucap, 1));
  Rec* u_orig = scratch.get<Rec>(std::max<uint64_t>(ucap, 1));
  {
    Phase ph(ctx, kPropNames[std::min(level, 8)], st);
    launch_propagate(ctx, st, n, lists, L.spacing, sp.on, rec, n_out, tab, root, rank, om, u_ids, u_orig, ucap,
                     ctr + kDsUnclaimed, tour_lo, n_dev, true, jo);
  }
  if (ucap == 0) return;  // tours only: nothing is left unreached
  Phase ph(ctx, "unclaimed", st);
  Rec* ext = scratch.get<Rec>(ucap);
  const unsigned long long* nu_dev = ctr + kDsUnclaimed;
  if (ucap <= kSmallMax && ucap <= tune.small_n) {
    hipLaunchKernelGGL(unclaimed_small_kernel<R>, dim3(1), dim3(kSmallThreads), 0, st, rec, u_ids, u_orig, ucap,
                       Troot, T8, ext, n_out, root, rank, om, nu_dev, ctx->d_err, jo, jo.seg_off, jo.nseg);
    return;
  }
  Rec* urec = scratch.get<Rec>(ucap);
  Rec* ures = scratch.get<Rec>(ucap);
  hipLaunchKernelGGL(unclaimed_build_kernel<R>, dim3(blocks_for(ucap, kThreads)), dim3(kThreads), 0, st, rec, u_ids,
                     u_orig, ucap, Troot, T8, urec, ext, nu_dev, ctx->d_err, jo.seg_off, jo.nseg);
  ds_rank(ctx, urec, ucap, reinterpret_cast<const uint64_t*>(nu_dev), ucap, &ures[0].s, &ures[0].w, OutMap{2, 0},
          std::min(ucap, max_tails), tune, ucap > n / 2 ? KC_ALGO_POINTER_DOUBLING : KC_ALGO_SPARSE_RULING_SET,
          level + 1, st, nullptr);
  hipLaunchKernelGGL(unclaimed_fix_kernel, dim3(blocks_for(ucap, kThreads)), dim3(kThreads), 0, st, u_ids, ucap,
                     n_out, ures, ext, root, rank, om, nu_dev, ctx->d_err, jo);
}

template <class R>
void ds_rank(Context* ctx, R* rec, uint64_t n, const uint64_t* n_dev, uint64_t n_out, uint64_t* root,
             int64_t* rank, OutMap om, uint64_t max_tails, const RankTuning& tune, int algorithm, int level,
             hipStream_t st, const TourRegion* tours, const JobOut& jo) {
  if (n == 0) return;
  if constexpr (RecOps<R>::kNarrow) {
    // a narrow top level is always a ruling-set level (rank_records_narrow)
    (void)algorithm;
    ds_srs_level(ctx, rec, n, n_dev, n_out, root, rank, om, max_tails, tune, level, st,
                 tours && tours->lo < n ? tours : nullptr, jo);
  } else {
    if (n <= tune.small_n && n <= kSmallMax) {
      Phase ph(ctx, "small_rank", st);
      hipLaunchKernelGGL(small_rank_kernel, dim3(1), dim3(kSmallThreads), 0, st, rec, static_cast<uint32_t>(n),
                         static_cast<uint32_t>(std::min<uint64_t>(n_out, n)), root, rank, om, ctx->d_err, n_dev);
      return;
    }
    if (algorithm == KC_ALGO_POINTER_DOUBLING || level >= static_cast<int>(tune.max_levels)) {
      ds_pd(ctx, rec, n, n_dev, n_out, root, rank, om, st);
      return;
    }
    ds_srs_level(ctx, rec, n, n_dev, n_out, root, rank, om, max_tails, tune, level, st,
                 tours && tours->lo < n ? tours : nullptr, jo);
  }
}

}  // namespace

// Built-in table (the autotuner's output for MI355X at the time of the last
// sweep; a tuning_table.txt next to the library overrides it).
const TuneRow kBuiltinTuning[] = {
    {10, 4, 8, 4, 64}, {12, 4, 2, 8, 64}, {14, 4, 32, 2, 64},
    {16, 16, 16, 8, 64}, {18, 4, 4, 4, 64}, {20, 16, 4, 1, 64},
    {22, 16, 8, 1, 64}, {24, 16, 16, 2, 64}, {25, 32, 8, 2, 64}, {26, 64, 8, 2, 64},
    {27, 256, 8, 2, 64}, {28, 512, 8, 3, 64}, {30, 512, 8, 3, 64},
};

// Tree rooting, by vertices (the tour has twice as many records).
const TuneRow kBuiltinTreeTuning[] = {
    {10, 4, 16, 8, 64}, {12, 4, 32, 4, 64}, {14, 8, 32, 4, 64},
    {16, 4, 8, 1, 64}, {18, 8, 4, 1, 64}, {20, 16, 4, 1, 64},
    {22, 32, 8, 2, 64}, {24, 32, 16, 3, 64}, {26, 64, 8, 2, 64},
    {28, 256, 16, 2, 64},
};

// Segmented batches, by total elements.  Sizes below the first class use the
// list table, whose level-0 spacing a batch of many lists caps by a cache
// window (level_spacing); a row of this table is used as measured.  MI355X,
// batches of lists log-uniform in [2^8, 2^20] with 8-byte records: 2000
// lists (2.5e8 elements) 11.11 ms against 12.55 with the list table, 8000
// (1.0e9) 39.9 against 43.8, 20,000 (2.5e9) 94.6 against 105.2.
const TuneRow kBuiltinSegTuning[] = {
    {27, 256, 16, 2, 64},
};

RankTuning tuning_for(const Context* ctx, uint64_t n, const kc_rank_cfg* cfg, int kind) {
  RankTuning t{};
  t.max_levels = 8;
  t.small_n = kSmallMax;
  t.window_cap = true;
  uint32_t lg = 0;
  while (lg < 63 && (uint64_t{2} << lg) <= n) lg++;
  const TuneRow* rows = kBuiltinTuning;
  size_t nrows = sizeof(kBuiltinTuning) / sizeof(kBuiltinTuning[0]);
  const std::vector<TuneRow>* loaded = ctx ? &ctx->tuning : nullptr;
  if (kind == KC_TUNING_TREE) {
    rows = kBuiltinTreeTuning;
    nrows = sizeof(kBuiltinTreeTuning) / sizeof(kBuiltinTreeTuning[0]);
    loaded = ctx ? &ctx->tuning_tree : nullptr;
  } else if (kind == KC_TUNING_SEGMENTED) {
    rows = kBuiltinSegTuning;
    nrows = sizeof(kBuiltinSegTuning) / sizeof(kBuiltinSegTuning[0]);
    loaded = ctx ? &ctx->tuning_seg : nullptr;
  }
  if (loaded && !loaded->empty()) {
    rows = loaded->data();
    nrows = loaded->size();
  }
  if (kind == KC_TUNING_SEGMENTED) {
    if (rows[0].log2n > lg) return tuning_for(ctx, n, cfg, KC_TUNING_LIST);
    t.window_cap = false;
  }
  // the row of the largest size class not above log2(n), else the smallest
  const TuneRow* pick = &rows[0];
  for (size_t i = 0; i < nrows; i++)
    if (rows[i].log2n <= lg) pick = &rows[i];
  t.spacing_top = pick->spacing_top;
  t.spacing_deep = pick->spacing_deep;
  t.chase_blocks_per_cu = pick->chase_blocks_per_cu;
  t.chase_chunk = pick->chase_chunk;
  if (cfg) {
    if (cfg->chase_blocks_per_cu) t.chase_blocks_per_cu = std::min(cfg->chase_blocks_per_cu, 64u);
    if (cfg->chase_chunk) t.chase_chunk = std::min(cfg->chase_chunk, 1u << 16);
    auto pow2 = [](uint32_t x) { return x >= 2 && (x & (x - 1)) == 0; };
    if (cfg->ruler_spacing) {
      if (!pow2(cfg->ruler_spacing) || cfg->ruler_spacing > (1u << 20))
        throw KcError{KC_EINVAL, "ruler_spacing must be a power of two in [2, 2^20]"};
      t.spacing_top = cfg->ruler_spacing;
      t.spacing_top_fixed = true;
    }
    if (cfg->ruler_spacing_deep) {
      if (!pow2(cfg->ruler_spacing_deep) || cfg->ruler_spacing_deep > (1u << 20))
        throw KcError{KC_EINVAL, "ruler_spacing_deep must be a power of two in [2, 2^20]"};
      t.spacing_deep = cfg->ruler_spacing_deep;
      t.spacing_deep_fixed = true;
    }
    if (cfg->max_levels) t.max_levels = cfg->max_levels;
    if (cfg->launch > KC_LAUNCH_HOST) throw KcError{KC_EINVAL, "invalid launch mode"};
    t.launch = cfg->launch;
    if (cfg->spawn > KC_SPAWN_ON) throw KcError{KC_EINVAL, "invalid spawn mode"};
    t.spawn = cfg->spawn;
    if (cfg->records > KC_RECORDS_WIDE) throw KcError{KC_EINVAL, "invalid record width"};
    t.records = cfg->records;
  }
  return t;
}

void pack_records(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n, Rec* rec,
                  hipStream_t st) {
  Phase ph(ctx, "pack", st);
  zero_counters(ctx, st);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  hipLaunchKernelGGL(pack_kernel<Rec>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, succ, w, n, rec,
                     ctx->d_err, ctr + kSlotTails, ctr + kSlotExtra);
}

namespace {

// Packs narrow records; counter kSlotExtra becomes nonzero when a weight does
// not fit int32 (the records are then not usable).
void pack_records_narrow(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n, Rec8* rec,
                         hipStream_t st) {
  Phase ph(ctx, "pack", st);
  zero_counters(ctx, st);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  hipLaunchKernelGGL(pack_kernel<Rec8>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, succ, w, n, rec,
                     ctx->d_err, ctr + kSlotTails, ctr + kSlotExtra);
}

}  // namespace

bool narrow_wanted(const Context* ctx, uint64_t records, int algo, const RankTuning& tune, uint64_t max_records) {
  return tune.records != KC_RECORDS_WIDE && records >= kNarrowMin && records < max_records &&
         algo == KC_ALGO_SPARSE_RULING_SET &&
         tune.launch != KC_LAUNCH_HOST && !ctx->trace_on && tune.max_levels > 0;
}

bool narrow_ids_fit(uint64_t records, uint64_t tails) {
  return records / 2 + records / 4 + tails < kNarrowMaxIds;
}

namespace {

// Everything the device-sized launch sequence depends on: equal keys give the
// same kernels with the same arguments.
std::vector<uint64_t> launch_key(Context* ctx, const void* rec, uint64_t n, uint64_t n_out, uint64_t* root,
                                 int64_t* rank, uint64_t max_tails, const RankTuning& t, int algorithm,
                                 uint32_t out_shift, const TourRegion* tours, const JobOut& jo) {
  std::vector<uint64_t> k = {reinterpret_cast<uint64_t>(rec), n, n_out, reinterpret_cast<uint64_t>(root),
                             reinterpret_cast<uint64_t>(rank), max_tails, static_cast<uint64_t>(algorithm),
                             out_shift, t.spacing_top, t.spacing_deep, t.max_levels, t.small_n,
                             t.chase_blocks_per_cu, t.chase_chunk, t.spacing_top_fixed, t.spacing_deep_fixed, t.spawn,
                             reinterpret_cast<uint64_t>(ctx->d_err), static_cast<uint64_t>(ctx->num_cus)};
  if (tours) {
    k.push_back(tours->lo);
    k.push_back(reinterpret_cast<uint64_t>(tours->heads));
    k.push_back(tours->nheads);
  } else {
    k.push_back(~uint64_t{0});
  }
  for (const void* p : {static_cast<const void*>(jo.lists), static_cast<const void*>(jo.trees),
                        static_cast<const void*>(jo.list_block), static_cast<const void*>(jo.tree_block)})
    k.push_back(reinterpret_cast<uint64_t>(p));
  k.push_back(jo.nlists);
  k.push_back(jo.ntrees);
  k.push_back(jo.nl);
  k.push_back(reinterpret_cast<uint64_t>(jo.seg_root));
  k.push_back(reinterpret_cast<uint64_t>(jo.seg_off));
  k.push_back(jo.nseg);
  ctx->arena.signature(k);
  return k;
}

}  // namespace

int resolve_algorithm(const kc_rank_cfg* cfg, uint64_t n, bool tree) {
  const int algo = cfg ? cfg->algorithm : KC_ALGO_AUTO;
  if (algo != KC_ALGO_AUTO && algo != KC_ALGO_POINTER_DOUBLING && algo != KC_ALGO_SPARSE_RULING_SET)
    throw KcError{KC_EINVAL, "invalid algorithm"};
  if (algo != KC_ALGO_AUTO) return algo;
  // a configuration that sets ruling-set knobs asks for the ruling set
  if (cfg && (cfg->ruler_spacing || cfg->ruler_spacing_deep || cfg->max_levels || cfg->spawn == KC_SPAWN_ON))
    return KC_ALGO_SPARSE_RULING_SET;
  return n <= (tree ? kAutoPdMaxTree : kAutoPdMaxList) ? KC_ALGO_POINTER_DOUBLING : KC_ALGO_SPARSE_RULING_SET;
}

bool top_level_is_ruling_set(uint64_t n, const RankTuning& tune, int algorithm) {
  return !(n <= tune.small_n && n <= kSmallMax) && algorithm == KC_ALGO_SPARSE_RULING_SET &&
         tune.max_levels > 0;
}

namespace {

// rank_records for either record width; returns the redo bits (0: done).
template <class R>
uint32_t rank_records_t(Context* ctx, R* rec, uint64_t n, uint64_t n_out, uint64_t* root,
                        int64_t* rank, uint64_t max_tails, const RankTuning& tune, int algorithm,
                        hipStream_t st, uint32_t out_shift, const TourRegion* tours, const JobOut* jobs_out) {
  const OutMap om{1, out_shift};
  const JobOut jo = jobs_out ? *jobs_out : JobOut{};
  if (jobs_out && !top_level_is_ruling_set(n, tune, algorithm))
    throw KcError{KC_EINVAL, "per-job outputs need the sparse ruling set at the top level"};
  // the per-level trace reads every level's counts on the host
  if (tune.launch == KC_LAUNCH_HOST || ctx->trace_on) {
    if constexpr (RecOps<R>::kNarrow) {
      throw KcError{KC_EINVAL, "narrow records need device-sized levels"};
    } else {
      rank_records_impl(ctx, rec, n, n_out, root, rank, om, max_tails, tune, algorithm, 0, st, tours, jo);
      return 0;
    }
  }
  auto launch_all = [&] {
    ds_rank(ctx, rec, n, nullptr, n_out, root, rank, om, max_tails, tune, algorithm, 0, st, tours, jo);
  };
  if (tune.launch == KC_LAUNCH_AUTO && !ctx->prof.on) {
    // A call repeated with the same arguments (and scratch layout) is
    // captured once as a HIP graph and replayed: one launch instead of one
    // per kernel.  The first occurrence runs directly (its scratch is then
    // allocated, so the capture allocates nothing).
    std::vector<uint64_t> key = launch_key(ctx, rec, n, n_out, root, rank, max_tails, tune, algorithm, out_shift,
                                           tours, jo);
    key.push_back(RecOps<R>::kTag);  // Rec8 and Rec8D have the same size
    if (ctx->graph_exec && key == ctx->graph_key) {
      KC_HIP(hipGraphLaunch(ctx->graph_exec, st));
      ctx->graph_replays++;
    } else if (key == ctx->last_key) {
      // Capture, instantiate, launch.  A capture or instantiation the runtime
      // refuses leaves nothing executed: the sequence then runs directly.
      hipGraph_t g = nullptr;
      hipGraphExec_t exec = nullptr;
      hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        try {
          launch_all();
        } catch (...) {
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        e = hipStreamEndCapture(st, &g);
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
      }
      if (e != hipSuccess) {
        (void)hipGetLastError();
        if (exec) (void)hipGraphExecDestroy(exec);
        launch_all();
      } else {
        if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = exec;
        ctx->graph_key = key;
        ctx->graph_captures++;
        KC_HIP(hipGraphLaunch(exec, st));
      }
    } else {
      launch_all();
    }
    ctx->last_key = std::move(key);
  } else {
    launch_all();
  }
  const uint32_t redo = check_err_overflow(ctx, st);
  if (redo & kErrOverflow) ctx->host_redos++;
  return redo;
}

}  // namespace

bool rank_records(Context* ctx, Rec* rec, uint64_t n, uint64_t n_out, uint64_t* root,
                  int64_t* rank, uint64_t max_tails, const RankTuning& tune, int algorithm,
                  hipStream_t st, uint32_t out_shift, const TourRegion* tours, const JobOut* jobs_out) {
  return rank_records_t(ctx, rec, n, n_out, root, rank, max_tails, tune, algorithm, st, out_shift, tours,
                        jobs_out) == 0;
}

uint32_t rank_records_narrow(Context* ctx, Rec8* rec, uint64_t n, uint64_t n_out, uint64_t* root, int64_t* rank,
                             uint64_t max_tails, const RankTuning& tune, hipStream_t st, uint32_t out_shift,
                             const TourRegion* tours, const JobOut* jobs_out) {
  return rank_records_t(ctx, rec, n, n_out, root, rank, max_tails, tune, KC_ALGO_SPARSE_RULING_SET, st, out_shift,
                        tours, jobs_out);
}

uint32_t rank_records_narrow(Context* ctx, Rec8D* rec, uint64_t n, uint64_t n_out, uint64_t* root, int64_t* rank,
                             uint64_t max_tails, const RankTuning& tune, hipStream_t st, const JobOut* jobs_out) {
  return rank_records_t(ctx, rec, n, n_out, root, rank, max_tails, tune, KC_ALGO_SPARSE_RULING_SET, st, 0, nullptr,
                        jobs_out);
}

bool narrow_level0_fits(const Context* ctx, uint64_t n, uint64_t tails, const RankTuning& tune) {
  const DsLevel L = ds_level(ctx, n, tails, tune, 0, nullptr, sizeof(Rec8D));
  return L.cap < kNarrowMaxIds && L.ucap < kNarrowMaxIds;
}

void rank_list_device(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n,
                      uint64_t* root, int64_t* rank, const kc_rank_cfg* cfg, hipStream_t st) {
  if (n == 0) return;
  check_records(n);
  RankTuning tune = tuning_for(ctx, n, cfg);
  const int algo = resolve_algorithm(cfg, n, false);
  if (cfg && cfg->validate) validate_list(ctx, succ, n, st);
  uint32_t contract = contraction_mode(cfg, n);
  ContractOut out;
  out.root = root;
  out.rank = rank;
  if (contract == KC_CONTRACT_AUTO && n >= kContractEarlySample)
    contract = contraction_sample_list(ctx, succ, n, st) ? KC_CONTRACT_ON : KC_CONTRACT_OFF;
  if (contract == KC_CONTRACT_ON) {
    rank_contracted_list(ctx, succ, w, n, cfg, algo, out, st);
    check_err(ctx, st);
    return;
  }
  Scratch scratch(ctx);
  if (contract != KC_CONTRACT_ON && narrow_wanted(ctx, n, algo, tune)) {
    Scratch narrow(ctx);  // released before the wide records are taken
    Rec8* rec8 = narrow.get<Rec8>(n);
    pack_records_narrow(ctx, succ, w, n, rec8, st);
    if (contract == KC_CONTRACT_AUTO)
      contraction_sample(ctx, rec8, n, reinterpret_cast<unsigned long long*>(ctx->d_counters) + kSlotWork, st);
    uint64_t c[3] = {0, 0, 0};
    read_counters(ctx, st, kSlotTails, 3, c);  // tails, sample, "a weight does not fit"
    if (contract == KC_CONTRACT_AUTO && contraction_sample_wins(c[1])) {
      rank_contracted_list(ctx, succ, w, n, cfg, algo, out, st);
      check_err(ctx, st);
      return;
    }
    if (c[2] == 0 && narrow_ids_fit(n, c[0])) {
      const uint32_t redo = rank_records_narrow(ctx, rec8, n, n, root, rank, c[0], tune, st);
      if (redo == 0) {
        check_err(ctx, st);
        ctx->narrow_calls++;
        return;
      }
      ctx->narrow_redos++;
      if (redo & kErrOverflow) tune.launch = KC_LAUNCH_HOST;  // and wide
    }
  }
  Rec* rec = scratch.get<Rec>(n);
  pack_records(ctx, succ, w, n, rec, st);
  if (contract == KC_CONTRACT_AUTO)
    contraction_sample(ctx, rec, n, reinterpret_cast<unsigned long long*>(ctx->d_counters) + kSlotWork, st);
  uint64_t c[2] = {0, 0};
  read_counters(ctx, st, kSlotTails, 2, c);
  const uint64_t tails = c[0];
  if (contract == KC_CONTRACT_AUTO && contraction_sample_wins(c[1])) {
    rank_contracted(ctx, rec, n, w == nullptr, cfg, algo, out, st);
    check_err(ctx, st);
    return;
  }
  while (!rank_records(ctx, rec, n, n, root, rank, tails, tune, algo, st)) {
    tune.launch = KC_LAUNCH_HOST;  // a capacity overflowed: redo host-sized
    pack_records(ctx, succ, w, n, rec, st);
  }
  check_err(ctx, st);
}

void pd_rounds_device(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n,
                      uint32_t rounds, uint64_t* succ_out, int64_t* rank_out, hipStream_t st) {
  if (n == 0) return;
  Scratch scratch(ctx);
  Rec* a = scratch.get<Rec>(n);
  Rec* b = scratch.get<Rec>(n);
  pack_records(ctx, succ, w, n, a, st);
  check_err(ctx, st);
  const unsigned blocks = blocks_for(n, kThreads);
  hipLaunchKernelGGL(pd_flag_roots_kernel, dim3(blocks), dim3(kThreads), 0, st, a, n, ctx->d_err, nullptr);
  unsigned long long* active = reinterpret_cast<unsigned long long*>(ctx->d_counters) + kSlotActive;
  for (uint32_t r = 0; r < rounds; r++) {
    hipLaunchKernelGGL(pd_round_kernel, dim3(blocks), dim3(kThreads), 0, st, a, b, n, active);
    std::swap(a, b);
  }
  hipLaunchKernelGGL(pd_state_kernel, dim3(blocks), dim3(kThreads), 0, st, a, n, succ_out, rank_out);
  check_err(ctx, st);
}

}  // namespace kc
