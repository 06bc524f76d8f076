// Device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace kc {

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Rulers: one per block of K = 2^logK consecutive ids, at a hashed offset inside
// the block.  Membership and the ruler's compact id (= its block) are computed,
// never looked up, so a chase tests "is this a ruler" without a memory access.
__device__ __forceinline__ uint64_t ruler_of_block(uint64_t b, uint64_t n, uint32_t logK,
                                                   uint32_t seed) {
  uint64_t start = b << logK;
  uint64_t full = uint64_t{1} << logK;
  uint64_t len = n - start < full ? n - start : full;
  uint32_t h = hash32(static_cast<uint32_t>(b) ^ (static_cast<uint32_t>(b >> 32) * 0x9E3779B9u) ^ seed);
  uint64_t off = (len == full) ? (h & (full - 1)) : (h % len);
  return start + off;
}
__device__ __forceinline__ bool is_ruler(uint64_t v, uint64_t n, uint32_t logK, uint32_t seed) {
  return ruler_of_block(v >> logK, n, logK, seed) == v;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Wave-aggregated atomic add of `mine` (0 or 1 per lane) on a 64-bit counter;
// returns this lane's slot.  Every lane of the wave must call it.
__device__ __forceinline__ uint64_t wave_append(unsigned long long* ctr, bool mine) {
  uint64_t mask = __ballot(mine);
  if (mask == 0) return 0;
  int leader = __ffsll(static_cast<long long>(mask)) - 1;
  unsigned long long base = 0;
  if (static_cast<int>(__lane_id()) == leader) base = atomicAdd(ctr, static_cast<unsigned long long>(__popcll(mask)));
  base = __shfl(base, leader);
  return base + __popcll(mask & lanemask_lt());
}

__device__ __forceinline__ void raise_err(uint32_t* err, uint32_t bits) { atomicOr(err, bits); }

// Child-list insertion: every `want` lane puts `child` at the front of the
// list whose head is head[slot] (sib[child] = the child inserted before it).
// Lanes of the wave that share a slot are linked among themselves first and
// make one atomic exchange for the group: a high-degree parent's children
// would otherwise serialise on one address (a 2^24-vertex star took 192 ms
// on MI355X, a random tree of the same size 2.9 ms).  Groups are formed for
// the slot of the first remaining lane while it has company (at most
// kLinkGroups groups); the remaining lanes exchange in one instruction.  A
// random tree, whose wave almost never shares a parent, thus keeps the
// plain exchange order: serialising even a few single-lane exchanges per
// wave changed the child order enough to slow the tour's chase by 7 %
// (2^26 vertices: 4.43 against 4.13 ms).  Every lane of the wave must call
// it.
constexpr int kLinkGroups = 4;
__device__ __forceinline__ void link_child(uint32_t* __restrict__ head, uint32_t* __restrict__ sib, uint64_t slot,
                                           uint32_t child, bool want) {
  const int lane = static_cast<int>(__lane_id());
  const uint64_t lt = lanemask_lt();
  uint64_t left = __ballot(want);
  for (int g = 0; g < kLinkGroups && left != 0; g++) {
    const int lead = __ffsll(static_cast<long long>(left)) - 1;
    const uint64_t s0 = __shfl(slot, lead);
    const bool mine = ((left >> lane) & 1) && slot == s0;
    const uint64_t grp = __ballot(mine);
    if (__popcll(grp) < 2) break;  // no company: everyone exchanges alone
    // the group's last lane makes the exchange; every member links to the
    // member before it, the first one to the previous head
    const int last = 63 - __clzll(static_cast<long long>(grp));
    uint32_t old = 0;
    if (lane == last) old = atomicExch(&head[s0], child);
    old = __shfl(old, last);
    const uint64_t before = grp & lt;
    const int src = before ? 63 - __clzll(static_cast<long long>(before)) : lane;
    const uint32_t prev = __shfl(child, src);
    if (mine) sib[child] = before ? prev : old;
    left &= ~grp;
  }
  if ((left >> lane) & 1) sib[child] = atomicExch(&head[slot], child);
}

// Segment of global element i: the largest s with seg_off[s] <= i (an empty
// segment shares its offset with the next one, and the search takes the last).
__device__ __forceinline__ uint64_t segment_of(const uint64_t* __restrict__ seg_off, uint64_t nseg, uint64_t i) {
  uint64_t lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) / 2;
    if (seg_off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Level-0 records of either width, seen as wide ones: decode() gives the
// wide meaning (ids with kRootFlag / kAuxFlag), encode() stores one back.
// `at` is the record's own position (Rec8D keeps unflagged successors
// relative to it; the other forms ignore it).
template <class R>
struct RecOps;
template <>
struct RecOps<Rec> {
  static constexpr bool kNarrow = false;
  static constexpr uint64_t kTag = 0;  // distinguishes the forms in launch keys
  __device__ static __forceinline__ Rec decode(Rec r, uint64_t) { return r; }
  __device__ static __forceinline__ Rec encode(uint64_t s, int64_t w, uint64_t) { return Rec{s, w}; }
  __device__ static __forceinline__ void set_s(Rec* p, uint64_t s) { p->s = s; }
  __device__ static __forceinline__ bool fits(int64_t) { return true; }
  __device__ static __forceinline__ bool fits_succ(uint64_t, uint64_t) { return true; }
};
template <>
struct RecOps<Rec8> {
  static constexpr bool kNarrow = true;
  static constexpr uint64_t kTag = 1;
  __device__ static __forceinline__ Rec decode(Rec8 r, uint64_t) {
    uint64_t s = r.s & (kAux8 - 1);
    if (r.s & kRoot8) s |= kRootFlag;
    if (r.s & kAux8) s |= kAuxFlag;
    return Rec{s, r.w};
  }
  __device__ static __forceinline__ uint32_t narrow_s(uint64_t s) {
    uint32_t t = static_cast<uint32_t>(s & (kAux8 - 1));
    if (s & kRootFlag) t |= kRoot8;
    if (s & kAuxFlag) t |= kAux8;
    return t;
  }
  __device__ static __forceinline__ Rec8 encode(uint64_t s, int64_t w, uint64_t) {
    return Rec8{narrow_s(s), static_cast<int32_t>(w)};
  }
  __device__ static __forceinline__ void set_s(Rec8* p, uint64_t s) { p->s = narrow_s(s); }
  __device__ static __forceinline__ bool fits(int64_t w) { return w >= INT32_MIN && w <= INT32_MAX; }
  __device__ static __forceinline__ bool fits_succ(uint64_t s, uint64_t) { return s < kNarrowMaxIds; }
};
// Rec8D: a flagged id (a claim, a collected or spawned id) is stored as in
// Rec8; an unflagged successor as its 30-bit two's-complement offset from
// the record's own position, so records of any position work as long as
// every successor lies within 2^29 of its element (segmented batches).
template <>
struct RecOps<Rec8D> {
  static constexpr bool kNarrow = true;
  static constexpr uint64_t kTag = 2;
  static constexpr uint32_t kOffMask = kAux8 - 1;
  __device__ static __forceinline__ Rec decode(Rec8D r, uint64_t at) {
    if (r.s & (kRoot8 | kAux8)) return RecOps<Rec8>::decode(Rec8{r.s, r.w}, at);
    const int64_t off = static_cast<int32_t>(r.s << 2) >> 2;  // sign-extend 30 bits
    return Rec{at + static_cast<uint64_t>(off), r.w};
  }
  __device__ static __forceinline__ Rec8D encode(uint64_t s, int64_t w, uint64_t at) {
    const uint32_t t = (s & (kRootFlag | kAuxFlag)) ? RecOps<Rec8>::narrow_s(s)
                                                    : static_cast<uint32_t>(s - at) & kOffMask;
    return Rec8D{t, static_cast<int32_t>(w)};
  }
  __device__ static __forceinline__ void set_s(Rec8D* p, uint64_t s) { p->s = RecOps<Rec8>::narrow_s(s); }
  __device__ static __forceinline__ bool fits(int64_t w) { return w >= INT32_MIN && w <= INT32_MAX; }
  __device__ static __forceinline__ bool fits_succ(uint64_t s, uint64_t at) {
    const int64_t off = static_cast<int64_t>(s - at);
    return off >= -(int64_t{1} << 29) && off < (int64_t{1} << 29);
  }
};

// Streaming accesses (read or written once, coalesced): non-temporal, so they
// do not evict the gather tables the same kernel keeps in L2.
__device__ __forceinline__ Rec load_stream(const Rec* p) {
  Rec r;
  r.s = __builtin_nontemporal_load(&p->s);
  r.w = __builtin_nontemporal_load(&p->w);
  return r;
}
template <typename T>
__device__ __forceinline__ void store_stream(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ Rec8 load_stream(const Rec8* p) {
  const uint64_t u = __builtin_nontemporal_load(reinterpret_cast<const uint64_t*>(p));
  return Rec8{static_cast<uint32_t>(u), static_cast<int32_t>(static_cast<uint32_t>(u >> 32))};
}
__device__ __forceinline__ Rec8D load_stream(const Rec8D* p) {
  const uint64_t u = __builtin_nontemporal_load(reinterpret_cast<const uint64_t*>(p));
  return Rec8D{static_cast<uint32_t>(u), static_cast<int32_t>(static_cast<uint32_t>(u >> 32))};
}

}  // namespace kc
