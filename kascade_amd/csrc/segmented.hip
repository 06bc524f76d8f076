// Batched segmented ranking: nseg independent lists in one ranking pass.
//
// Segment s owns elements [seg_off[s], seg_off[s+1]) and stores successors as
// local u32 ids.  The batch is ranked as one forest over global ids
// seg_off[s] + local (the single-device analogue of the reference's block
// Distribution, distribution.hpp:16-113, with the segment offsets as the
// block offsets): one pack pass, one persistent chase launch over every
// segment, one propagation pass.  Roots are written back as local ids.
#include <algorithm>

#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;

// Segment of element i, searching only the segments that overlap this block.
__device__ __forceinline__ uint64_t find_segment(const uint64_t* __restrict__ seg_off,
                                                 uint64_t lo, uint64_t hi, uint64_t i) {
  // invariant: seg_off[lo] <= i < seg_off[hi + 1]
  while (lo < hi) {
    uint64_t mid = (lo + hi + 1) / 2;
    if (seg_off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Each block covers kPer x kThreads consecutive elements: the two searches
// for the block's segment range (dependent loads over all nseg offsets) are
// paid once per 4096 elements, and each thread then moves forward from the
// segment of its previous element.
constexpr unsigned kPer = 16;

__device__ __forceinline__ void chunk_range(const uint64_t* __restrict__ seg_off, uint64_t nseg,
                                            uint64_t total, uint64_t first, uint64_t* s_range) {
  const uint64_t lastv = std::min<uint64_t>(first + uint64_t{kPer} * kThreads, total) - 1;
  if (threadIdx.x == 0) s_range[0] = segment_of(seg_off, nseg, first);
  if (threadIdx.x == 1) s_range[1] = segment_of(seg_off, nseg, lastv);
  __syncthreads();
}

// R: Rec, or Rec8D for a narrow batch (a weight outside int32, or a
// successor 2^29 or more away from its element, then sets *wide and the
// batch is packed again with wide records).
// The block's segment offsets are staged in LDS (when at most kSegCache + 1
// of them) and the successors and weights are loaded before the first
// store, so no load waits for a store (one wait counter covers both).  With
// the offset searches in global memory, between stores, the pack took
// 6.7 ms for 2.5e9 elements (now 6.45) and 200,000 short lists 10.8 ms in
// all (now 9.2).
constexpr uint64_t kSegCache = 255;

template <class R>
__global__ void __launch_bounds__(kThreads)
seg_pack_kernel(const uint32_t* __restrict__ succ, const uint64_t* __restrict__ seg_off,
                uint64_t nseg, uint64_t total, const int64_t* __restrict__ w,
                R* __restrict__ rec, uint32_t* err, unsigned long long* tails, unsigned long long* wide) {
  __shared__ uint64_t s_range[2];
  __shared__ uint64_t s_off[kSegCache + 1];
  const uint64_t block_first = static_cast<uint64_t>(blockIdx.x) * kPer * kThreads;
  chunk_range(seg_off, nseg, total, block_first, s_range);
  const uint64_t seg_lo = s_range[0], seg_hi = s_range[1];
  const uint64_t cnt = seg_hi - seg_lo + 2;  // offsets seg_lo .. seg_hi + 1
  const bool cached = cnt <= kSegCache + 1;
  if (cached)
    for (uint64_t t = threadIdx.x; t < cnt; t += kThreads) s_off[t] = seg_off[seg_lo + t];
  const uint64_t first = block_first + threadIdx.x;
  uint32_t loc[kPer];
#pragma unroll
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t i = first + q * kThreads;
    loc[q] = i < total ? succ[i] : 0;
  }
  int64_t wt[kPer];
  if (w) {
#pragma unroll
    for (unsigned q = 0; q < kPer; q++) {
      const uint64_t i = first + q * kThreads;
      wt[q] = i < total ? w[i] : 0;
    }
  }
  __syncthreads();
  uint64_t lo = 0;  // segment index relative to seg_lo (cached) or absolute
  if (!cached) lo = seg_lo;
  uint32_t ntails = 0;
#pragma unroll
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t i = first + q * kThreads;
    if (i >= total) break;
    uint64_t base, size;
    if (cached) {
      uint64_t a = lo, b = cnt - 2;  // invariant: s_off[a] <= i < s_off[b + 1]
      while (a < b) {
        const uint64_t mid = (a + b + 1) / 2;
        if (s_off[mid] <= i) a = mid; else b = mid - 1;
      }
      lo = a;
      base = s_off[lo];
      size = s_off[lo + 1] - base;
    } else {
      lo = find_segment(seg_off, lo, seg_hi, i);
      base = seg_off[lo];
      size = seg_off[lo + 1] - base;
    }
    uint64_t l = loc[q];
    if (l >= size) {
      raise_err(err, kErrRange);
      l = i - base;
    }
    const uint64_t g = base + l;
    const bool tail = g == i;
    ntails += tail;
    int64_t x;
    if (w) {
      x = wt[q];
      if (tail && x != 0) raise_err(err, kErrRootWeight);
      if (!RecOps<R>::fits(x)) *wide = 1;
    } else {
      x = tail ? 0 : 1;
    }
    if (!RecOps<R>::fits_succ(g, i)) *wide = 1;
    rec[i] = RecOps<R>::encode(g, x, i);
  }
  // one atomic per wave for its tails
  for (int o = 32; o > 0; o >>= 1) ntails += __shfl_xor(ntails, o);
  if (__lane_id() == 0 && ntails) atomicAdd(tails, static_cast<unsigned long long>(ntails));
}

__global__ void __launch_bounds__(kThreads)
seg_root_local_kernel(const uint64_t* __restrict__ root_global, const uint64_t* __restrict__ seg_off,
                      uint64_t nseg, uint64_t total, uint32_t* __restrict__ root_local) {
  __shared__ uint64_t s_range[2];
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * kPer * kThreads;
  chunk_range(seg_off, nseg, total, first, s_range);
  uint64_t lo = s_range[0];
  const uint64_t hi = s_range[1];
  for (unsigned q = 0; q < kPer; q++) {
    const uint64_t i = first + q * kThreads + threadIdx.x;
    if (i >= total) return;
    lo = find_segment(seg_off, lo, hi, i);
    root_local[i] = static_cast<uint32_t>(root_global[i] - seg_off[lo]);
  }
}

__global__ void seg_check_kernel(const uint64_t* __restrict__ seg_off, uint64_t nseg, uint32_t* err) {
  uint64_t s = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (s >= nseg) return;
  const uint64_t a = seg_off[s], b = seg_off[s + 1];
  if (b < a || b - a > 0xFFFFFFFFull) raise_err(err, kErrRange);
}


// The whole batch as one forest with 16-byte records (ids of any size).
void rank_segments_wide(Context* ctx, const uint32_t* succ, const uint64_t* seg_off, uint64_t nseg, uint64_t total,
                        const int64_t* w, uint32_t* root, int64_t* rank, const kc_rank_cfg* cfg, RankTuning tune,
                        int algo, uint32_t contract, hipStream_t st) {
  // With the ruling set at the top level the ranking writes segment-local
  // roots directly; otherwise global roots go to a scratch array and a last
  // pass converts them.
  const bool local_roots = top_level_is_ruling_set(total, tune, algo);
  Scratch scratch(ctx);
  Rec* rec = scratch.get<Rec>(total);
  uint64_t* root_global = local_roots ? nullptr : scratch.get<uint64_t>(total);
  JobOut jo;
  jo.seg_root = root;
  jo.seg_off = seg_off;
  jo.nseg = nseg;
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  auto pack = [&] {
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, 3 * sizeof(uint64_t), st));
    hipLaunchKernelGGL(seg_pack_kernel<Rec>, dim3(blocks_for(total, kPer * kThreads)), dim3(kThreads), 0, st, succ,
                       seg_off, nseg, total, w, rec, ctx->d_err, ctr, ctr + 2);
  };
  pack();
  if (contract == KC_CONTRACT_AUTO) contraction_sample(ctx, rec, total, ctr + 1, st);
  uint64_t c[2] = {0, 0};
  read_counters(ctx, st, 0, 2, c);
  const uint64_t tails = c[0];
  if (contract == KC_CONTRACT_ON || (contract == KC_CONTRACT_AUTO && contraction_sample_wins(c[1]))) {
    ContractOut out;
    out.rank = rank;
    out.seg_root = root;
    out.seg_off = seg_off;
    out.nseg = nseg;
    rank_contracted(ctx, rec, total, w == nullptr, cfg, algo, out, st);
    check_err(ctx, st);
    return;
  }
  while (!rank_records(ctx, rec, total, total, root_global, rank, tails, tune, algo, st, 0, nullptr,
                       local_roots ? &jo : nullptr)) {
    tune.launch = KC_LAUNCH_HOST;  // a capacity overflowed: redo host-sized
    pack();
  }
  check_err(ctx, st);
  if (local_roots) return;
  hipLaunchKernelGGL(seg_root_local_kernel, dim3(blocks_for(total, kPer * kThreads)), dim3(kThreads), 0, st,
                     root_global, seg_off, nseg, total, root);
}

// 8-byte records whose successors are stored relative to their own
// position (Rec8D), so the batch's total size does not limit them: every
// segment below 2^29 elements, weights and claim distances inside int32, and
// the top level's sublist and prefix ids below 2^30; otherwise, or when a
// claim distance outgrows int32, the wide path.  Measured on MI355X: 1000
// lists (1.26e8 elements) 6.99 against 7.19 ms wide, 200 lists (2.5e7) 1.80
// against 1.97 ms (absolute 8-byte ids).  Ranking a batch beyond 2^30
// elements as groups of whole segments with absolute 8-byte ids (each group
// below 2^29 elements) was slower for 20,000 lists of 2.5e9 elements (115
// against 107 ms: the groups' head-prefix sub-problems cost 12 against
// 3.4 ms); relative ids keep such a batch one forest.
void rank_segments(Context* ctx, const uint32_t* succ, const uint64_t* seg_off, uint64_t nseg, uint64_t total,
                        const int64_t* w, uint32_t* root, int64_t* rank, const kc_rank_cfg* cfg, int algo,
                        hipStream_t st) {
  RankTuning tune = tuning_for(ctx, total, cfg, KC_TUNING_SEGMENTED);
  const uint32_t contract = contraction_mode(cfg, total);
  if (contract != KC_CONTRACT_ON && narrow_wanted(ctx, total, algo, tune, ~uint64_t{0})) {
    Scratch narrow(ctx);  // released before any wide records are taken
    Rec8D* rec8 = narrow.get<Rec8D>(total);
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, 3 * sizeof(uint64_t), st));
    hipLaunchKernelGGL(seg_pack_kernel<Rec8D>, dim3(blocks_for(total, kPer * kThreads)), dim3(kThreads), 0, st, succ,
                       seg_off, nseg, total, w, rec8, ctx->d_err, ctr, ctr + 2);
    if (contract == KC_CONTRACT_AUTO) contraction_sample(ctx, rec8, total, ctr + 1, st);
    uint64_t c[3] = {0, 0, 0};
    read_counters(ctx, st, 0, 3, c);  // tails, contraction sample, a weight or successor that does not fit
    const bool contracts = contract == KC_CONTRACT_AUTO && contraction_sample_wins(c[1]);
    if (!contracts && c[2] == 0 && narrow_level0_fits(ctx, total, c[0], tune)) {
      JobOut jo;
      jo.seg_root = root;
      jo.seg_off = seg_off;
      jo.nseg = nseg;
      const uint32_t redo = rank_records_narrow(ctx, rec8, total, total, nullptr, rank, c[0], tune, st, &jo);
      if (redo == 0) {
        check_err(ctx, st);
        ctx->narrow_calls++;
        return;
      }
      ctx->narrow_redos++;
      if (redo & kErrOverflow) tune.launch = KC_LAUNCH_HOST;  // and wide
    }
  }
  rank_segments_wide(ctx, succ, seg_off, nseg, total, w, root, rank, cfg, tune, algo, contract, st);
}

}  // namespace

void rank_segmented_device(Context* ctx, const uint32_t* succ, const uint64_t* seg_off,
                           uint64_t nseg, uint64_t total, const int64_t* w, uint32_t* root,
                           int64_t* rank, const kc_rank_cfg* cfg, hipStream_t st) {
  if (nseg == 0 || total == 0) return;
  check_records(total);
  hipLaunchKernelGGL(seg_check_kernel, dim3(blocks_for(nseg, kThreads)), dim3(kThreads), 0, st,
                     seg_off, nseg, ctx->d_err);
  check_err(ctx, st);
  // size class of the whole forest; the ruler spacing per level is capped by
  // the mean list length inside the ranking
  rank_segments(ctx, succ, seg_off, nseg, total, w, root, rank, cfg, resolve_algorithm(cfg, total, false), st);
}

}  // namespace kc
