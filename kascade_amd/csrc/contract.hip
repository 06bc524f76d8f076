// Local contraction of a list ranking's level 0 (the reference's
// local_contraction / local_uncontraction: pointer_doubling_generic.hpp:471-543
// for pointer doubling, sparse_ruling_set_detail/local_contraction.hpp:28-104
// for the ruling set).  The reference collapses chains whose successors stay
// on one PE, so that they cost no messages; on one GPU the expensive thing is
// a dependent step, and the chains worth collapsing are runs of consecutive
// ids (succ[v] = v + 1), which lie in consecutive memory.
//
// A run is a maximal sequence v, v + 1, ..., e with succ[u] = u + 1 inside
// it; a tail is a run of its own.  Element v is a run head unless v > 0,
// succ[v - 1] = v and v is not a tail.  The contracted list has one element
// per run, numbered in id order: run R with last element e links to the run
// of succ[e] (a head in a list: its only predecessor is e, and
// e + 1 != succ[e]) with weight W_R = w(head) + ... + w(e).  With P the
// exclusive prefix sum of the weights over all ids (wrapping; differences
// are exact), W_R = P(e) + w(e) - P(head), and ranking the contracted list
// gives rank(head of R), so inside the run
// rank(v) = rank(head) + P(head) - P(v); the root is the head (a tail) of
// the contracted root's run.  For unit weights P(v) = v serves: inside a run
// every element before v weighs 1 (only a run's last element can be a tail).
//
// A wave handles chunks of 64 consecutive ids, one id per lane: run heads
// are a ballot, contracted ids a popcount over the chunk's head mask plus the
// chunk's scanned head count, P a wave scan plus the chunk's scanned weight.
// Two coalesced passes over the input (count, build), one gather per run
// (link: its exit's contracted id from the exit chunk's head mask) and an
// expansion that reads one contracted record per element (and the weights,
// when given) replace the chase's dependent step per element.
#include <algorithm>

#include "common.hpp"
#include "device.hpp"

namespace kc {
namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kWaves = kThreads / 64;
constexpr uint32_t kSample = 65536;               // elements the AUTO decision reads
// AUTO contracts nothing smaller: below 2^22 elements the passes cost about
// what they save even on fully local lists (measured on MI355X)
constexpr uint64_t kAutoMin = uint64_t{1} << 22;
static_assert(kAutoMin >= kSample, "the sample needs n >= kSample");

// The caller's arrays (NULL weights: 1, and 0 on tails; list_ranking.cpp:60-80).
struct ListIn {
  const uint64_t* succ;
  const int64_t* w;
  __device__ __forceinline__ uint64_t s(uint64_t v) const { return __builtin_nontemporal_load(succ + v); }
  __device__ __forceinline__ uint64_t s_neighbour(uint64_t v) const { return succ[v]; }
  __device__ __forceinline__ int64_t wt(uint64_t v) const { return __builtin_nontemporal_load(w + v); }
};

// Narrow packed records (successor ids below 2^30, no flags yet): for the
// AUTO sample only.
struct Rec8In {
  const Rec8* rec;
  __device__ __forceinline__ uint64_t s_neighbour(uint64_t v) const { return rec[v].s; }
};

// Narrow records with position-relative successors (segmented batches).
struct Rec8DIn {
  const Rec8D* rec;
  __device__ __forceinline__ uint64_t s_neighbour(uint64_t v) const { return RecOps<Rec8D>::decode(rec[v], v).s; }
};

// Packed records (pack_records, the segmented pack).
struct RecIn {
  const Rec* rec;
  __device__ __forceinline__ uint64_t s(uint64_t v) const { return rec[v].s; }
  __device__ __forceinline__ uint64_t s_neighbour(uint64_t v) const { return rec[v].s; }
  __device__ __forceinline__ int64_t wt(uint64_t v) const { return rec[v].w; }
};

// Per chunk of 64 ids: bit l of mask marks id 64 * chunk + l as a run head;
// base is the contracted id of the chunk's first head (one 16-byte gather
// resolves an exit).
struct ChunkHeads {
  uint64_t mask;
  uint64_t base;
};

__device__ __forceinline__ uint64_t lanemask_le() { return ~0ull >> (63 - __lane_id()); }

__device__ __forceinline__ uint64_t wave_inclusive_sum(uint64_t x) {
  const unsigned lane = __lane_id();
  for (unsigned o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

// Each wave takes kPer consecutive chunks and issues all their loads before
// it works through them (the chunks are independent: every cross-chunk value
// comes from a scan), so each lane keeps kPer loads in flight.
constexpr unsigned kPer = 16;

__device__ __forceinline__ uint64_t wave_id() {
  return static_cast<uint64_t>(blockIdx.x) * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
}
__device__ __forceinline__ uint64_t wave_count() { return static_cast<uint64_t>(gridDim.x) * kWaves; }

template <class In>
__device__ __forceinline__ void load_succ(const In& in, uint64_t n, uint64_t c0, uint64_t (&s)[kPer]) {
#pragma unroll
  for (unsigned j = 0; j < kPer; j++) {
    const uint64_t v = (c0 + j) * 64 + __lane_id();
    s[j] = v < n ? in.s(v) : 0;
  }
}
template <class In>
__device__ __forceinline__ void load_weights(const In& in, uint64_t n, uint64_t c0, uint64_t (&x)[kPer]) {
#pragma unroll
  for (unsigned j = 0; j < kPer; j++) {
    const uint64_t v = (c0 + j) * 64 + __lane_id();
    x[j] = v < n ? static_cast<uint64_t>(in.wt(v)) : 0;
  }
}

// Pass 1: per chunk the head mask, head count and weight sum (weighted
// inputs); tails; and for the caller's arrays the checks pack_records makes.
template <class In>
__global__ void __launch_bounds__(kThreads)
lc_count_kernel(In in, uint64_t n, uint64_t chunks, bool unit, bool check, ChunkHeads* __restrict__ heads,
                uint64_t* __restrict__ cnt, uint64_t* __restrict__ wsum, uint32_t* err,
                unsigned long long* tails_ctr) {
  const unsigned lane = __lane_id();
  for (uint64_t c0 = wave_id() * kPer; c0 < chunks; c0 += wave_count() * kPer) {
    uint64_t s[kPer], x[kPer];
    load_succ(in, n, c0, s);
    if (!unit) load_weights(in, n, c0, x);
    uint64_t prev = c0 > 0 ? in.s_neighbour(c0 * 64 - 1) : 0;  // successor of the element before
    uint64_t tails = 0;
#pragma unroll
    for (unsigned j = 0; j < kPer; j++) {
      const uint64_t ch = c0 + j;
      if (ch >= chunks) break;
      const uint64_t v = ch * 64 + lane;
      const bool valid = v < n;
      uint64_t ps = __shfl_up(s[j], 1);
      if (lane == 0) ps = prev;
      prev = __shfl(s[j], 63);
      const bool tail = valid && s[j] == v;
      const bool head = valid && (v == 0 || ps != v || tail);
      const uint64_t mask = __ballot(head);
      tails += __popcll(__ballot(tail));
      if (check && valid) {
        if (s[j] >= n) raise_err(err, kErrRange);
        if (!unit && tail && x[j] != 0) raise_err(err, kErrRootWeight);
      }
      uint64_t sum = 0;
      if (!unit) sum = __shfl(wave_inclusive_sum(x[j]), 63);
      if (lane == 0) {
        heads[ch].mask = mask;
        cnt[ch] = __popcll(mask);
        if (!unit) wsum[ch] = sum;
      }
    }
    if (lane == 0 && tails) atomicAdd(tails_ctr, static_cast<unsigned long long>(tails));
  }
}

__global__ void lc_total_kernel(const uint64_t* __restrict__ base, const uint64_t* __restrict__ cnt,
                                uint64_t chunks, unsigned long long* m_out) {
  if (threadIdx.x == 0) m_out[0] = base[chunks - 1] + cnt[chunks - 1];
}

// Pass 2: the contracted records ({exit element, P through the run's last
// element}, or for a tail the resolved self loop marked with kAuxFlag), the
// head element and (weighted) P at the head of every run, and each chunk's
// first contracted id.
template <class In>
__global__ void __launch_bounds__(kThreads)
lc_build_kernel(In in, uint64_t n, uint64_t chunks, bool unit, ChunkHeads* __restrict__ heads,
                const uint64_t* __restrict__ base_in, const uint64_t* __restrict__ chunk_p,
                Rec* __restrict__ rec_c, uint64_t* __restrict__ head_elem, uint64_t* __restrict__ head_p) {
  const unsigned lane = __lane_id();
  for (uint64_t c0 = wave_id() * kPer; c0 < chunks; c0 += wave_count() * kPer) {
    uint64_t s[kPer], x[kPer];
    load_succ(in, n, c0, s);
    if (!unit) load_weights(in, n, c0, x);
    // the next chunk's first head bit, for the chunk's last lane
    uint64_t next_mask = c0 < chunks ? heads[c0].mask : 0;
#pragma unroll
    for (unsigned j = 0; j < kPer; j++) {
      const uint64_t ch = c0 + j;
      if (ch >= chunks) break;
      const uint64_t v = ch * 64 + lane;
      const bool valid = v < n;
      const uint64_t mask = next_mask;
      next_mask = ch + 1 < chunks ? heads[ch + 1].mask : 1;
      const uint64_t base = base_in[ch];
      const bool next_head = lane < 63 ? (mask >> (lane + 1) & 1) : (next_mask & 1);
      // v ends its run when v + 1 starts one or is beyond the last element
      const bool run_end = valid && (v + 1 >= n || next_head);
      uint64_t p, w;
      if (unit) {
        p = v;
        w = 1;
      } else {
        w = x[j];
        p = chunk_p[ch] + wave_inclusive_sum(w) - w;
      }
      const uint64_t cid = base + __popcll(mask & lanemask_le()) - 1;
      if (valid && (mask >> lane & 1)) {
        head_elem[cid] = v;
        if (!unit) head_p[cid] = p;
      }
      if (lane == 0) heads[ch].base = base;
      if (run_end) rec_c[cid] = s[j] == v ? Rec{cid | kAuxFlag, 0} : Rec{s[j], static_cast<int64_t>(p + w)};
    }
  }
}

// Pass 3 (on the m runs): the exit element becomes the exit's contracted id
// and P through the last element the run's weight.  In a list the exit is a
// run head other than the run's own; otherwise the input has an element with
// two predecessors or a cycle, which is reported (the record becomes a self
// loop, so the ranking still ends).  Tails record their root value: the
// element id, or its id local to its segment (seg_off).
__global__ void __launch_bounds__(kThreads)
lc_link_kernel(Rec* __restrict__ rec_c, uint64_t m, bool unit, const ChunkHeads* __restrict__ heads,
               const uint64_t* __restrict__ head_elem, const uint64_t* __restrict__ head_p,
               uint64_t* __restrict__ root_val, const uint64_t* __restrict__ seg_off, uint64_t nseg, uint32_t* err) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (c >= m) return;
  const Rec r = rec_c[c];
  if (r.s & kAuxFlag) {  // a tail: already its own contracted id
    const uint64_t e = head_elem[c];
    root_val[c] = seg_off ? e - seg_off[segment_of(seg_off, nseg, e)] : e;
    rec_c[c] = Rec{c, 0};
    return;
  }
  const ChunkHeads ch = heads[r.s >> 6];
  const unsigned b = static_cast<unsigned>(r.s & 63);
  const bool exit_is_head = ch.mask >> b & 1;
  const uint64_t h = ch.base + __popcll(ch.mask & ((uint64_t{1} << b) - 1));
  if (!exit_is_head || h == c) {
    raise_err(err, exit_is_head ? kErrCycle : kErrNotList);
    rec_c[c] = Rec{c, 0};
    return;
  }
  const uint64_t p_head = unit ? head_elem[c] : head_p[c];
  rec_c[c] = Rec{h, static_cast<int64_t>(static_cast<uint64_t>(r.w) - p_head)};
}

// Pass 4 (on the m runs, after the ranking): {root value, rank + P at the
// head} per run, into the consumed contracted records.
__global__ void __launch_bounds__(kThreads)
lc_result_kernel(Rec* __restrict__ out_c, uint64_t m, bool unit, const uint64_t* __restrict__ root_c,
                 const int64_t* __restrict__ rank_c, const uint64_t* __restrict__ root_val,
                 const uint64_t* __restrict__ head_elem, const uint64_t* __restrict__ head_p,
                 const uint32_t* __restrict__ err) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (c >= m || *err) return;
  const uint64_t p_head = unit ? head_elem[c] : head_p[c];
  out_c[c] = Rec{root_val[root_c[c]], static_cast<int64_t>(static_cast<uint64_t>(rank_c[c]) + p_head)};
}

// Pass 5: every element's result from its run's: rank(v) = (rank + P at the
// head) - P(v).  Roots go out as u64 element ids or u32 segment-local ids.
template <class In>
__global__ void __launch_bounds__(kThreads)
lc_expand_kernel(In in, uint64_t n, uint64_t chunks, bool unit, const ChunkHeads* __restrict__ heads,
                 const uint64_t* __restrict__ chunk_p, const Rec* __restrict__ out_c, uint64_t* __restrict__ root,
                 uint32_t* __restrict__ seg_root, int64_t* __restrict__ rank, const uint32_t* __restrict__ err) {
  if (*err) return;
  const unsigned lane = __lane_id();
  for (uint64_t c0 = wave_id() * kPer; c0 < chunks; c0 += wave_count() * kPer) {
    uint64_t x[kPer];
    if (!unit) load_weights(in, n, c0, x);
    Rec o[kPer];
#pragma unroll
    for (unsigned j = 0; j < kPer; j++) {
      const uint64_t ch = c0 + j;
      const uint64_t v = ch * 64 + lane;
      o[j] = Rec{0, 0};
      if (v < n) {
        const ChunkHeads hd = heads[ch];
        o[j] = out_c[hd.base + __popcll(hd.mask & lanemask_le()) - 1];
      }
    }
#pragma unroll
    for (unsigned j = 0; j < kPer; j++) {
      const uint64_t ch = c0 + j;
      if (ch >= chunks) break;
      const uint64_t v = ch * 64 + lane;
      uint64_t p = v;
      if (!unit) p = chunk_p[ch] + wave_inclusive_sum(x[j]) - x[j];
      if (v < n) {
        if (seg_root)
          store_stream(seg_root + v, static_cast<uint32_t>(o[j].s));
        else
          store_stream(root + v, o[j].s);
        store_stream(rank + v, static_cast<int64_t>(static_cast<uint64_t>(o[j].w) - p));
      }
    }
  }
}

// The AUTO decision: one element at a pseudo-random offset in each of
// kSample equal strides, counted when inside a run (not a head); one thread
// per sample, the count added to *out (zeroed before).  Needs n >= kSample.
template <class In>
__global__ void __launch_bounds__(kThreads)
lc_sample_kernel(In in, uint64_t n, unsigned long long* out) {
  __shared__ unsigned s_cnt[kWaves];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const uint64_t v = static_cast<uint64_t>(i) * n / kSample + splitmix64(i) % (n / kSample);
  unsigned c = v > 0 && in.s_neighbour(v - 1) == v && in.s_neighbour(v) != v;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if (__lane_id() == 0) s_cnt[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (unsigned k = 0; k < kWaves; k++) t += s_cnt[k];
    if (t) atomicAdd(out, static_cast<unsigned long long>(t));
  }
}

template <class In>
void contract_and_rank(Context* ctx, const In& in, uint64_t n, bool unit, bool check, const kc_rank_cfg* cfg,
                       int algorithm, const ContractOut& out, hipStream_t st) {
  Scratch scratch(ctx);
  const uint64_t chunks = (n + 63) / 64;
  // streaming passes: kPer chunks per wave, up to 2^20 blocks (2^32 ids),
  // a grid-stride loop beyond
  const unsigned blocks =
      static_cast<unsigned>(std::min<uint64_t>(blocks_for(chunks, kWaves * kPer), uint64_t{1} << 20));
  ChunkHeads* heads = scratch.get<ChunkHeads>(chunks);
  uint64_t* cnt = scratch.get<uint64_t>(chunks);
  uint64_t* base = scratch.get<uint64_t>(chunks);
  uint64_t* wsum = unit ? nullptr : scratch.get<uint64_t>(chunks);
  uint64_t* chunk_p = unit ? nullptr : scratch.get<uint64_t>(chunks);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  {
    Phase ph(ctx, "local_contraction", st);
    KC_HIP(hipMemsetAsync(ctx->d_counters, 0, 2 * sizeof(uint64_t), st));
    hipLaunchKernelGGL(lc_count_kernel<In>, dim3(blocks), dim3(kThreads), 0, st, in, n, chunks, unit, check, heads,
                       cnt, wsum, ctx->d_err, ctr);
    exclusive_scan_u64(ctx, cnt, base, chunks, st);
    if (!unit) exclusive_scan_u64(ctx, wsum, chunk_p, chunks, st);
    hipLaunchKernelGGL(lc_total_kernel, dim3(1), dim3(64), 0, st, base, cnt, chunks, ctr + 1);
  }
  uint64_t c[2] = {0, 0};
  read_counters(ctx, st, 0, 2, c);  // raises the checks' errors
  const uint64_t tails = c[0], m = c[1];
  Rec* rec_c = scratch.get<Rec>(m);
  uint64_t* head_elem = scratch.get<uint64_t>(m);
  uint64_t* head_p = unit ? nullptr : scratch.get<uint64_t>(m);
  uint64_t* root_val = scratch.get<uint64_t>(m);
  uint64_t* root_c = scratch.get<uint64_t>(m);
  int64_t* rank_c = scratch.get<int64_t>(m);
  RankTuning tune = tuning_for(ctx, m, cfg);
  for (;;) {
    {
      Phase ph(ctx, "local_contraction", st);
      hipLaunchKernelGGL(lc_build_kernel<In>, dim3(blocks), dim3(kThreads), 0, st, in, n, chunks, unit, heads, base,
                         chunk_p, rec_c, head_elem, head_p);
      hipLaunchKernelGGL(lc_link_kernel, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, st, rec_c, m, unit,
                         heads, head_elem, head_p, root_val, out.seg_root ? out.seg_off : nullptr, out.nseg,
                         ctx->d_err);
    }
    if (rank_records(ctx, rec_c, m, m, root_c, rank_c, tails, tune, algorithm, st)) break;
    tune.launch = KC_LAUNCH_HOST;  // a capacity overflowed: redo host-sized
  }
  Phase ph(ctx, "local_uncontraction", st);
  hipLaunchKernelGGL(lc_result_kernel, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, st, rec_c, m, unit, root_c,
                     rank_c, root_val, head_elem, head_p, ctx->d_err);
  hipLaunchKernelGGL(lc_expand_kernel<In>, dim3(blocks), dim3(kThreads), 0, st, in, n, chunks, unit, heads, chunk_p,
                     rec_c, out.root, out.seg_root, out.rank, ctx->d_err);
}

}  // namespace

// worth it when at least half of the elements are inside runs
bool contraction_sample_wins(uint64_t sampled_interior) { return 2 * sampled_interior >= kSample; }

uint32_t contraction_mode(const kc_rank_cfg* cfg, uint64_t n) {
  const uint32_t mode = cfg ? cfg->local_contraction : KC_CONTRACT_AUTO;
  if (mode > KC_CONTRACT_ON) throw KcError{KC_EINVAL, "invalid local_contraction mode"};
  return n < kAutoMin && mode == KC_CONTRACT_AUTO ? KC_CONTRACT_OFF : mode;
}

void contraction_sample(Context* ctx, const Rec* rec, uint64_t n, unsigned long long* out, hipStream_t st) {
  (void)ctx;
  if (n < kSample) return;  // the strides would be empty (the count stays 0: no contraction)
  hipLaunchKernelGGL(lc_sample_kernel<RecIn>, dim3(kSample / kThreads), dim3(kThreads), 0, st, RecIn{rec}, n, out);
}

void contraction_sample(Context* ctx, const Rec8* rec, uint64_t n, unsigned long long* out, hipStream_t st) {
  (void)ctx;
  if (n < kSample) return;  // the strides would be empty (the count stays 0: no contraction)
  hipLaunchKernelGGL(lc_sample_kernel<Rec8In>, dim3(kSample / kThreads), dim3(kThreads), 0, st, Rec8In{rec}, n, out);
}

void contraction_sample(Context* ctx, const Rec8D* rec, uint64_t n, unsigned long long* out, hipStream_t st) {
  (void)ctx;
  if (n < kSample) return;  // the strides would be empty (the count stays 0: no contraction)
  hipLaunchKernelGGL(lc_sample_kernel<Rec8DIn>, dim3(kSample / kThreads), dim3(kThreads), 0, st, Rec8DIn{rec}, n,
                     out);
}

bool contraction_sample_list(Context* ctx, const uint64_t* succ, uint64_t n, hipStream_t st) {
  if (n < kSample) return false;  // the strides would be empty
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ctx->d_counters);
  KC_HIP(hipMemsetAsync(ctx->d_counters, 0, sizeof(uint64_t), st));
  hipLaunchKernelGGL(lc_sample_kernel<ListIn>, dim3(kSample / kThreads), dim3(kThreads), 0, st, ListIn{succ, nullptr},
                     n, ctr);
  uint64_t c = 0;
  read_counters(ctx, st, 0, 1, &c);
  return contraction_sample_wins(c);
}


void rank_contracted_list(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n, const kc_rank_cfg* cfg,
                          int algorithm, const ContractOut& out, hipStream_t st) {
  contract_and_rank(ctx, ListIn{succ, w}, n, w == nullptr, true, cfg, algorithm, out, st);
}

void rank_contracted(Context* ctx, const Rec* rec, uint64_t n, bool unit, const kc_rank_cfg* cfg, int algorithm,
                     const ContractOut& out, hipStream_t st) {
  contract_and_rank(ctx, RecIn{rec}, n, unit, false, cfg, algorithm, out, st);
}

}  // namespace kc
