// Mixed batches: list-ranking and tree-rooting jobs coalesced into one
// ranking pass (the serving path).
//
// Record layout of a batch with N_l list elements and N_t tree vertices:
//   [0, N_l)                 list elements, ids = list base + local id
//   N_l + 2t                 upward tour edge / per-root dummy of tree vertex t
//   N_l + 2t + 1             downward tour edge into tree vertex t
// (the same interleaved tour numbering as the single-tree path, shifted by
// N_l; t = tree base + local id).  One ranking of all N_l + 2N_t records
// serves every job; outputs are copied back to each job's arrays with local
// root ids.
#include <algorithm>

#include "batch.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;

// job containing element e: largest j with jobs[j].base <= e (bases ascending)
__device__ __forceinline__ uint32_t job_of(const BatchJob* __restrict__ jobs, uint32_t lo, uint32_t hi,
                                           uint64_t e) {
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) / 2;
    if (jobs[mid].base <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void block_job_range(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                uint64_t first, uint64_t last, uint32_t* s_range) {
  if (threadIdx.x == 0) s_range[0] = job_of(jobs, 0, njobs - 1, first);
  if (threadIdx.x == 1) s_range[1] = job_of(jobs, 0, njobs - 1, last);
  __syncthreads();
}

// A block covers kPer x kThreads consecutive items, so the two searches for
// its job range (dependent loads over the whole job table) are paid once per
// 4096 items; each thread then moves forward from its previous item's job.
constexpr unsigned kPer = 16;
__device__ __forceinline__ uint64_t chunk_first() {
  return static_cast<uint64_t>(blockIdx.x) * kPer * kThreads;
}
__device__ __forceinline__ uint64_t chunk_last(uint64_t count) {
  return std::min<uint64_t>(chunk_first() + uint64_t{kPer} * kThreads, count) - 1;
}

// The job of a thread's current item, its fields kept in registers: a
// thread's items ascend, so it only moves forward, reloading the job when an
// item passes the next job's base (jobs are at least one item long).
struct JobCursor {
  const BatchJob* jobs;
  uint32_t njobs;
  uint32_t j;
  uint64_t next_base;
  BatchJob job;
  __device__ void start(const BatchJob* js, uint32_t nj, uint32_t j0) {
    jobs = js;
    njobs = nj;
    j = j0;
    job = jobs[j];
    next_base = j + 1 < njobs ? jobs[j + 1].base : ~uint64_t{0};
  }
  __device__ void seek(uint64_t e) {
    while (e >= next_base) {
      j++;
      job = jobs[j];
      next_base = j + 1 < njobs ? jobs[j + 1].base : ~uint64_t{0};
    }
  }
};

// R: Rec, or Rec8 for a narrow batch (a weight outside int32 then sets *wide
// and the batch is packed again with wide records).
template <class R>
__global__ void __launch_bounds__(kThreads)
batch_list_pack_kernel(const BatchJob* __restrict__ jobs, uint32_t njobs, uint64_t nl,
                       R* __restrict__ rec, uint32_t* err, unsigned long long* tails,
                       unsigned long long* wide) {
  __shared__ uint32_t s_range[2];
  block_job_range(jobs, njobs, chunk_first(), chunk_last(nl), s_range);
  JobCursor jc;
  jc.start(jobs, njobs, s_range[0]);
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t e = chunk_first() + q * kThreads + threadIdx.x;
    bool tail = false;
    if (e < nl) {
      jc.seek(e);
      const BatchJob& j = jc.job;
      const uint64_t i = e - j.base;
      uint64_t s = j.succ[i];
      if (s >= j.n) {
        raise_err(err, kErrRange);
        s = i;
      }
      tail = (s == i);
      int64_t w;
      if (j.w) {
        w = j.w[i];
        if (tail && w != 0) raise_err(err, kErrRootWeight);
        if (!RecOps<R>::fits(w)) *wide = 1;
      } else {
        w = tail ? 0 : 1;
      }
      rec[e] = RecOps<R>::encode(j.base + s, w, e);
    }
    wave_append(tails, tail);
  }
}

__global__ void __launch_bounds__(kThreads)
batch_tree_keys_kernel(const BatchJob* __restrict__ jobs, uint32_t njobs, uint64_t nt, uint64_t nl,
                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                       uint8_t* __restrict__ is_root, Rec* __restrict__ rec, uint32_t* err,
                       unsigned long long* roots) {
  __shared__ uint32_t s_range[2];
  block_job_range(jobs, njobs, chunk_first(), chunk_last(nt), s_range);
  JobCursor jc;
  jc.start(jobs, njobs, s_range[0]);
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t t = chunk_first() + q * kThreads + threadIdx.x;
    bool root = false;
    if (t < nt) {
      jc.seek(t);
      const BatchJob& j = jc.job;
      const uint64_t i = t - j.base;
      uint64_t p = j.succ[i];
      if (p >= j.n) {
        raise_err(err, kErrRange);
        p = i;
      }
      root = (p == i);
      keys[t] = root ? static_cast<uint32_t>(nt) : static_cast<uint32_t>(j.base + p);
      vals[t] = static_cast<uint32_t>(t);
      is_root[t] = root;
      if (root) {
        rec[nl + 2 * t] = Rec{nl + 2 * t, 0};
        rec[nl + 2 * t + 1] = Rec{nl + 2 * t + 1, 0};
      } else {
        rec[nl + 2 * t + 1] = Rec{nl + 2 * t, -1};
      }
    }
    wave_append(roots, root);
  }
}

__global__ void __launch_bounds__(kThreads)
batch_tour_link_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                       const uint8_t* __restrict__ is_root, uint64_t nt, uint64_t nl,
                       Rec* __restrict__ rec, uint64_t* __restrict__ heads, unsigned long long* nheads) {
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  bool head = false;
  uint32_t c = 0;
  if (k < nt) {
    const uint32_t p = keys[k];
    if (p != nt) {
      c = vals[k];
      const bool first = (k == 0) || keys[k - 1] != p;
      const bool last = (k + 1 == nt) || keys[k + 1] != p;
      rec[nl + 2 * uint64_t{c}] = Rec{last ? nl + 2 * uint64_t{p} : nl + 2 * uint64_t{vals[k + 1]} + 1, 1};
      if (first) {
        if (!is_root[p])
          rec[nl + 2 * uint64_t{p} + 1] = Rec{nl + 2 * uint64_t{c} + 1, -1};
        else
          head = true;  // the root's tour starts at this down edge
      }
    }
  }
  const uint64_t h = wave_append(nheads, head);
  if (head) heads[h] = nl + 2 * uint64_t{c} + 1;
}

// Linked tour build for a batch of trees (child lists by atomic exchange;
// sib[t] = the sibling inserted before t).
__global__ void __launch_bounds__(kThreads)
batch_tree_link_kernel(const BatchJob* __restrict__ jobs, uint32_t njobs, uint64_t nt,
                       uint32_t* __restrict__ head, uint32_t* __restrict__ sib, uint32_t* err,
                       unsigned long long* roots) {
  __shared__ uint32_t s_range[2];
  block_job_range(jobs, njobs, chunk_first(), chunk_last(nt), s_range);
  JobCursor jc;
  jc.start(jobs, njobs, s_range[0]);
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t t = chunk_first() + q * kThreads + threadIdx.x;
    bool root = false;
    uint64_t slot = 0;
    if (t < nt) {
      jc.seek(t);
      const BatchJob& j = jc.job;
      const uint64_t i = t - j.base;
      const uint64_t p = j.succ[i];
      if (p >= j.n) raise_err(err, kErrRange);
      root = (p == i) || p >= j.n;
      slot = j.base + p;
    }
    link_child(head, sib, slot, static_cast<uint32_t>(t), t < nt && !root);
    wave_append(roots, root);
  }
}

// Both tour elements of tree vertex t as one 32-byte pair (see tour_pairs_kernel).
template <class R>
__global__ void __launch_bounds__(kThreads)
batch_tree_pairs_kernel(const BatchJob* __restrict__ jobs, uint32_t njobs, uint64_t nt, uint64_t nl,
                        const uint32_t* __restrict__ head, const uint32_t* __restrict__ sib,
                        R* __restrict__ rec, uint64_t* __restrict__ heads, unsigned long long* nheads) {
  __shared__ uint32_t s_range[2];
  block_job_range(jobs, njobs, chunk_first(), chunk_last(nt), s_range);
  JobCursor jc;
  jc.start(jobs, njobs, s_range[0]);
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t t = chunk_first() + q * kThreads + threadIdx.x;
    bool tour_head = false;
    uint32_t h = 0xFFFFFFFFu;
    if (t < nt) {
      jc.seek(t);
      const BatchJob& j = jc.job;
      const uint64_t i = t - j.base;
      const uint64_t p = j.succ[i];
      h = head[t];
      Rec up, down;
      if (p == i || p >= j.n) {
        up = Rec{nl + 2 * t, 0};
        down = Rec{nl + 2 * t + 1, 0};
        tour_head = h != 0xFFFFFFFFu;  // a root's tour starts at its head child's down edge
      } else {
        const uint32_t sb = sib[t];
        up = Rec{sb != 0xFFFFFFFFu ? nl + 2 * uint64_t{sb} + 1 : nl + 2 * (j.base + p), 1};
        down = Rec{h != 0xFFFFFFFFu ? nl + 2 * uint64_t{h} + 1 : nl + 2 * t, -1};
      }
      rec[nl + 2 * t] = RecOps<R>::encode(up.s, up.w, nl + 2 * t);
      rec[nl + 2 * t + 1] = RecOps<R>::encode(down.s, down.w, nl + 2 * t + 1);
    }
    const uint64_t k = wave_append(nheads, tour_head);
    if (tour_head) heads[k] = nl + 2 * uint64_t{h} + 1;
  }
}

// block_job[b] = the job of item 256 b (the JobOut lookup table).
__global__ void __launch_bounds__(kThreads)
job_block_kernel(const BatchJob* __restrict__ jobs, uint32_t njobs, uint64_t count,
                 uint32_t* __restrict__ block_job) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (b >= (count + 255) / 256) return;
  block_job[b] = job_of(jobs, 0, njobs - 1, b * 256);
}

// Copies outputs back: list element e -> its job (shift 0, step 1); tree
// vertex t, whose upward edge is element nl + 2t -> its job (shift nl, step 2).
__global__ void __launch_bounds__(kThreads)
batch_scatter_kernel(const BatchJob* __restrict__ jobs, uint32_t njobs, uint64_t count,
                     uint64_t shift, uint32_t step, const uint64_t* __restrict__ root,
                     const int64_t* __restrict__ rank) {
  __shared__ uint32_t s_range[2];
  block_job_range(jobs, njobs, chunk_first(), chunk_last(count), s_range);
  uint32_t jl = s_range[0];
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t e = chunk_first() + q * kThreads + threadIdx.x;
    if (e >= count) return;
    jl = job_of(jobs, jl, s_range[1], e);
    const BatchJob j = jobs[jl];
    const uint64_t i = e - j.base;
    const uint64_t x = shift + step * e;
    j.root_out[i] = (root[x] - shift) / step - j.base;
    j.rank_out[i] = rank[x];
  }
}

unsigned bits_for(uint64_t x) {
  unsigned b = 1;
  while (b < 64 && (uint64_t{1} << b) <= x) b++;
  return b;
}

}  // namespace

void rank_batch_device(Context* ctx, const std::vector<BatchJob>& jobs_in, const kc_rank_cfg* cfg,
                       int tour_build, hipStream_t st) {
  std::vector<BatchJob> lists, trees;
  uint64_t nl = 0, nt = 0;
  for (const BatchJob& j : jobs_in) {
    if (j.n == 0) continue;
    BatchJob c = j;
    if (j.kind == KC_JOB_LIST) {
      c.base = nl;
      nl += j.n;
      lists.push_back(c);
    } else {
      c.base = nt;
      nt += j.n;
      trees.push_back(c);
    }
  }
  if (nl + nt == 0) return;
  if (nt >= (uint64_t{1} << 32) - 1) throw KcError{KC_EINVAL, "tree vertices per batch must be < 2^32 - 1"};
  Scratch scratch(ctx);
  const uint64_t total = nl + 2 * nt;
  // size class of the whole batch; the ruler spacing per level is capped by
  // the mean list length inside the ranking
  RankTuning tune = tuning_for(ctx, total, cfg);
  int algo = cfg ? cfg->algorithm : KC_ALGO_AUTO;
  if (algo == KC_ALGO_AUTO) algo = KC_ALGO_SPARSE_RULING_SET;
  // With the ruling set at the top level the ranking writes every job's
  // results in place; otherwise into batch-wide arrays that are scattered.
  const bool in_place = top_level_is_ruling_set(total, tune, algo);
  uint64_t* root = in_place ? nullptr : scratch.get<uint64_t>(total);
  int64_t* rank = in_place ? nullptr : scratch.get<int64_t>(total);
  BatchJob* d_lists = lists.empty() ? nullptr : scratch.get<BatchJob>(lists.size());
  BatchJob* d_trees = trees.empty() ? nullptr : scratch.get<BatchJob>(trees.size());
  if (d_lists)
    KC_HIP(hipMemcpyAsync(d_lists, lists.data(), lists.size() * sizeof(BatchJob), hipMemcpyHostToDevice, st));
  if (d_trees)
    KC_HIP(hipMemcpyAsync(d_trees, trees.data(), trees.size() * sizeof(BatchJob), hipMemcpyHostToDevice, st));
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  uint64_t* heads = nt ? scratch.get<uint64_t>(nt) : nullptr;
  const bool linked = tour_build != KC_TOUR_SORTED;
  uint32_t *head = nullptr, *sib = nullptr, *keys = nullptr, *vals = nullptr;
  uint8_t* is_root = nullptr;
  if (nt && linked) {
    head = scratch.get<uint32_t>(nt);
    sib = scratch.get<uint32_t>(nt);
  } else if (nt) {
    keys = scratch.get<uint32_t>(nt);
    vals = scratch.get<uint32_t>(nt);
    is_root = scratch.get<uint8_t>(nt);
  }
  JobOut jo;
  if (in_place) {
    jo.lists = d_lists;
    jo.trees = d_trees;
    jo.nlists = static_cast<uint32_t>(lists.size());
    jo.ntrees = static_cast<uint32_t>(trees.size());
    jo.nl = nl;
    if (nl) {
      uint32_t* lb = scratch.get<uint32_t>((nl + 255) / 256);
      hipLaunchKernelGGL(job_block_kernel, dim3(blocks_for((nl + 255) / 256, kThreads)), dim3(kThreads), 0, st,
                         d_lists, jo.nlists, nl, lb);
      jo.list_block = lb;
    }
    if (nt) {
      uint32_t* tb = scratch.get<uint32_t>((nt + 255) / 256);
      hipLaunchKernelGGL(job_block_kernel, dim3(blocks_for((nt + 255) / 256, kThreads)), dim3(kThreads), 0, st,
                         d_trees, jo.ntrees, nt, tb);
      jo.tree_block = tb;
    }
  }
  // Narrow (8-byte) records first, as for single lists: every id of a batch
  // fits 30 bits; a list weight outside int32, or a claim distance that
  // outgrows it, sends the batch to the wide records below.
  if (in_place && linked && narrow_wanted(ctx, total, algo, tune)) {
    Scratch narrow(ctx);  // released before the wide records are taken
    Rec8* rec8 = narrow.get<Rec8>(total);
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, 4 * sizeof(uint64_t), st));
    if (nl) {
      Phase ph(ctx, "batch_pack", st);
      hipLaunchKernelGGL(batch_list_pack_kernel<Rec8>, dim3(blocks_for(nl, kPer * kThreads)), dim3(kThreads), 0, st,
                         d_lists, static_cast<uint32_t>(lists.size()), nl, rec8, ctx->d_err, ctr, ctr + 3);
    }
    if (nt) {
      Phase ph(ctx, "batch_tour", st);
      KC_HIP(hipMemsetAsync(head, 0xFF, nt * sizeof(uint32_t), st));  // no child yet
      hipLaunchKernelGGL(batch_tree_link_kernel, dim3(blocks_for(nt, kPer * kThreads)), dim3(kThreads), 0, st,
                         d_trees, static_cast<uint32_t>(trees.size()), nt, head, sib, ctx->d_err, ctr + 1);
      hipLaunchKernelGGL(batch_tree_pairs_kernel<Rec8>, dim3(blocks_for(nt, kPer * kThreads)), dim3(kThreads), 0,
                         st, d_trees, static_cast<uint32_t>(trees.size()), nt, nl, head, sib, rec8, heads, ctr + 2);
    }
    uint64_t counts[4] = {0, 0, 0, 0};
    read_counters(ctx, st, 0, 4, counts);  // tails, roots, tour heads, a weight outside int32
    const uint64_t tails = counts[0] + 2 * counts[1];
    if (counts[3] == 0 && narrow_ids_fit(total, tails)) {
      const TourRegion tr{nl, heads, counts[2]};
      const uint32_t redo = rank_records_narrow(ctx, rec8, total, total, nullptr, nullptr, tails, tune, st, 0,
                                                nt ? &tr : nullptr, &jo);
      if (redo == 0) {
        check_err(ctx, st);
        ctx->narrow_calls++;
        return;
      }
      ctx->narrow_redos++;
      if (redo & kErrOverflow) tune.launch = KC_LAUNCH_HOST;  // and wide
    }
  }
  Rec* rec = scratch.get<Rec>(total);
  for (;;) {
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, 4 * sizeof(uint64_t), st));
    if (nl) {
      Phase ph(ctx, "batch_pack", st);
      hipLaunchKernelGGL(batch_list_pack_kernel<Rec>, dim3(blocks_for(nl, kPer * kThreads)), dim3(kThreads), 0, st,
                         d_lists, static_cast<uint32_t>(lists.size()), nl, rec, ctx->d_err, ctr, ctr + 3);
    }
    if (nt && linked) {
      Phase ph(ctx, "batch_tour", st);
      KC_HIP(hipMemsetAsync(head, 0xFF, nt * sizeof(uint32_t), st));  // no child yet
      hipLaunchKernelGGL(batch_tree_link_kernel, dim3(blocks_for(nt, kPer * kThreads)), dim3(kThreads), 0, st,
                         d_trees, static_cast<uint32_t>(trees.size()), nt, head, sib, ctx->d_err, ctr + 1);
      hipLaunchKernelGGL(batch_tree_pairs_kernel<Rec>, dim3(blocks_for(nt, kPer * kThreads)), dim3(kThreads), 0,
                         st, d_trees, static_cast<uint32_t>(trees.size()), nt, nl, head, sib, rec, heads, ctr + 2);
    } else if (nt) {
      Phase ph(ctx, "batch_tour", st);
      hipLaunchKernelGGL(batch_tree_keys_kernel, dim3(blocks_for(nt, kPer * kThreads)), dim3(kThreads), 0, st,
                         d_trees, static_cast<uint32_t>(trees.size()), nt, nl, keys, vals, is_root, rec,
                         ctx->d_err, ctr + 1);
      radix_sort_pairs_u32(ctx, keys, vals, nt, bits_for(nt), st);
      hipLaunchKernelGGL(batch_tour_link_kernel, dim3(blocks_for(nt, kThreads)), dim3(kThreads), 0, st, keys,
                         vals, is_root, nt, nl, rec, heads, ctr + 2);
    }
    uint64_t counts[3] = {0, 0, 0};
    read_counters(ctx, st, 0, 3, counts);
    const TourRegion tr{nl, heads, counts[2]};
    if (rank_records(ctx, rec, total, total, root, rank, counts[0] + 2 * counts[1], tune, algo, st, 0,
                     nt ? &tr : nullptr, in_place ? &jo : nullptr))
      break;
    tune.launch = KC_LAUNCH_HOST;  // a capacity overflowed: redo host-sized
  }
  check_err(ctx, st);
  if (in_place) return;
  Phase ph(ctx, "batch_scatter", st);
  if (nl)
    hipLaunchKernelGGL(batch_scatter_kernel, dim3(blocks_for(nl, kPer * kThreads)), dim3(kThreads), 0, st, d_lists,
                       static_cast<uint32_t>(lists.size()), nl, uint64_t{0}, 1u, root, rank);
  if (nt)
    hipLaunchKernelGGL(batch_scatter_kernel, dim3(blocks_for(nt, kPer * kThreads)), dim3(kThreads), 0, st, d_trees,
                       static_cast<uint32_t>(trees.size()), nt, nl, 2u, root, rank);
  KC_HIP(hipGetLastError());
}

}  // namespace kc
