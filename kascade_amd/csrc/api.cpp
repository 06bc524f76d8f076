// C ABI of libkascade_amd: contexts, host/device pointer staging, status
// mapping.  The kernels live in the .hip sources.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <new>
#include <string>
#include <type_traits>

#include "common.hpp"

namespace kc {

// ---------------------------------------------------------------- Arena ---
namespace {

// Idle chunks per device, largest first to go when the cache is over its cap
// (a quarter of the device's memory).
struct ChunkCache {
  std::mutex mu;
  std::multimap<size_t, char*> chunks;
  size_t bytes = 0;
  size_t cap = 0;
};

ChunkCache& chunk_cache(int device) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<ChunkCache>> caches;
  std::lock_guard<std::mutex> lk(mu);
  auto& c = caches[device];
  if (!c) {
    c = std::make_unique<ChunkCache>();
    size_t free_b = 0, total_b = 0;
    c->cap = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? total_b / 4 : (size_t{16} << 30);
  }
  return *c;
}

constexpr size_t kMinChunk = size_t{64} << 20;

size_t round_chunk(size_t b) { return (b + kMinChunk - 1) / kMinChunk * kMinChunk; }

}  // namespace

Arena::~Arena() { trim(); }

void Arena::use_fixed(char* base, size_t size) {
  chunks_.assign(1, Chunk{base, size});
  blocks_.clear();
  blocks_.reserve(1024);  // a ranking's scratch stack stays far below this
  top_ = live_ = 0;
  fixed_ = true;
}

bool Arena::new_chunk(size_t want) {
  if (fixed_) return false;
  want = std::max(want, kMinChunk);
  char* p = nullptr;
  size_t size = 0;
  {
    ChunkCache& c = chunk_cache(device);
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.chunks.lower_bound(want);
    if (it != c.chunks.end()) {
      size = it->first;
      p = it->second;
      c.bytes -= size;
      c.chunks.erase(it);
    }
  }
  if (!p) {
    if (hipMalloc(reinterpret_cast<void**>(&p), want) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    size = want;
  }
  chunks_.push_back(Chunk{p, size});
  held_.fetch_add(size, std::memory_order_relaxed);
  return true;
}

// Only with no live block: the previous call is over, but its kernels may
// still run on any stream, so the device is synchronised before its chunks go.
void Arena::consolidate() {
  (void)hipDeviceSynchronize();
  for (auto& c : chunks_) (void)hipFree(c.base);
  chunks_.clear();
  blocks_.clear();
  held_.store(0, std::memory_order_relaxed);
  new_chunk(round_chunk(peak_ + peak_ / 8));
}

void* Arena::alloc(size_t bytes) {
  bytes = (bytes + 255) & ~size_t{255};
  if (bytes == 0) bytes = 256;
  if (top_ == 0 && chunks_.size() > 1) consolidate();
  // Next free position after the top block.
  size_t chunk = 0, offset = 0;
  if (top_ > 0) {
    const Block& b = blocks_[top_ - 1];
    chunk = b.chunk;
    offset = b.offset + b.size;
  }
  while (chunk < chunks_.size() && offset + bytes > chunks_[chunk].size) {
    chunk++;
    offset = 0;
  }
  if (chunk == chunks_.size() && !new_chunk(bytes)) return nullptr;
  Block nb{chunk, offset, bytes};
  if (top_ < blocks_.size()) blocks_[top_] = nb; else blocks_.push_back(nb);
  top_++;
  live_ += bytes;
  peak_ = std::max(peak_, live_);
  return chunks_[chunk].base + offset;
}

void Arena::release(size_t mark) {
  while (top_ > mark) live_ -= blocks_[--top_].size;
}

void Arena::recycle() {
  if (top_ != 0 || fixed_) return;  // never hand out live blocks or memory it does not own
  ChunkCache& c = chunk_cache(device);
  {
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto& ch : chunks_) {
      c.chunks.emplace(ch.size, ch.base);
      c.bytes += ch.size;
    }
    while (c.bytes > c.cap && !c.chunks.empty()) {
      auto it = std::prev(c.chunks.end());
      (void)hipFree(it->second);
      c.bytes -= it->first;
      c.chunks.erase(it);
    }
  }
  chunks_.clear();
  blocks_.clear();
  held_.store(0, std::memory_order_relaxed);
}

void Arena::trim() {
  if (top_ != 0 || fixed_) return;  // never free live blocks or memory it does not own
  for (auto& c : chunks_) (void)hipFree(c.base);
  chunks_.clear();
  blocks_.clear();
  held_.store(0, std::memory_order_relaxed);
}

void release_chunk_cache(int device) {
  ChunkCache& c = chunk_cache(device);
  std::lock_guard<std::mutex> lk(c.mu);
  for (auto& ch : c.chunks) (void)hipFree(ch.second);
  c.chunks.clear();
  c.bytes = 0;
}

void Arena::signature(std::vector<uint64_t>& out) const {
  out.push_back(chunks_.size());
  for (const auto& c : chunks_) {
    out.push_back(reinterpret_cast<uint64_t>(c.base));
    out.push_back(c.size);
  }
  out.push_back(top_);
  if (top_ > 0) {
    out.push_back(blocks_[top_ - 1].chunk);
    out.push_back(blocks_[top_ - 1].offset + blocks_[top_ - 1].size);
  }
}

uint64_t Arena::held() const { return held_.load(std::memory_order_relaxed); }

void Arena::reserve(size_t bytes) {
  // only an arena with nothing allocated yet: no kernel can still use its chunks
  if (fixed_ || top_ != 0 || !chunks_.empty() || bytes == 0) return;
  new_chunk(round_chunk(bytes));
}

// -------------------------------------------------------------- Profiler ---
hipEvent_t Profiler::get() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  KC_HIP(hipEventCreate(&e));
  return e;
}
void Profiler::begin(const char* name, hipStream_t st) {
  hipEvent_t a = get();
  KC_HIP(hipEventRecord(a, st));
  open_.emplace_back(name, a);
}
void Profiler::end(hipStream_t st) {
  if (open_.empty()) return;
  hipEvent_t b = get();
  (void)hipEventRecord(b, st);
  pending_.push_back(Pending{open_.back().first, open_.back().second, b});
  open_.pop_back();
}
void Profiler::flush() {
  for (auto& p : pending_) {
    float ms = 0;
    (void)hipEventSynchronize(p.b);
    if (hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess) (void)hipGetLastError();
    bool found = false;
    for (auto& t : totals)
      if (t.name == p.name) {
        t.ms += ms;
        t.count++;
        found = true;
      }
    if (!found) totals.push_back(Acc{p.name, ms, 1});
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}
void Profiler::reset() {
  flush();
  totals.clear();
}
Profiler::~Profiler() {
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : pool_) (void)hipEventDestroy(e);
  for (auto& o : open_) (void)hipEventDestroy(o.second);
}

static void raise_from_bits(Context* ctx, hipStream_t st, uint32_t e) {
  if (e == 0) return;
  KC_HIP(hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), st));
  KC_HIP(hipStreamSynchronize(st));
  if (e & kErrRange) throw KcError{KC_EINVAL, "successor id out of range (must be < n)"};
  if (e & kErrRootWeight) throw KcError{KC_EINVAL, "a root (succ[i] == i) carries a nonzero weight"};
  if (e & kErrNotList) throw KcError{KC_EINVAL, "not a list: an element has more than one predecessor"};
  if (e & kErrCycle) throw KcError{KC_ECYCLE, "cycle without a root"};
  throw KcError{KC_EDEVICE, "unknown device error"};
}

// The error word lives in the last counter slot, so one small copy and one
// stream synchronisation return both.
void check_err(Context* ctx, hipStream_t st) {
  KC_HIP(hipMemcpyAsync(ctx->h_pinned + kErrSlot, ctx->d_counters + kErrSlot, sizeof(uint64_t),
                        hipMemcpyDeviceToHost, st));
  KC_HIP(hipStreamSynchronize(st));
  raise_from_bits(ctx, st, static_cast<uint32_t>(ctx->h_pinned[kErrSlot]));
}

uint32_t check_err_overflow(Context* ctx, hipStream_t st) {
  KC_HIP(hipMemcpyAsync(ctx->h_pinned + kErrSlot, ctx->d_counters + kErrSlot, sizeof(uint64_t),
                        hipMemcpyDeviceToHost, st));
  KC_HIP(hipStreamSynchronize(st));
  const uint32_t e = static_cast<uint32_t>(ctx->h_pinned[kErrSlot]);
  const uint32_t redo = kErrOverflow | kErrNarrow;
  raise_from_bits(ctx, st, e & ~redo);
  if (!(e & redo)) return 0;
  KC_HIP(hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), st));
  KC_HIP(hipStreamSynchronize(st));
  return e & redo;
}

void read_counters(Context* ctx, hipStream_t st, int first, int count, uint64_t* out) {
  KC_HIP(hipMemcpyAsync(ctx->h_pinned + first, ctx->d_counters + first,
                        (kErrSlot + 1 - first) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  KC_HIP(hipStreamSynchronize(st));
  std::memcpy(out, ctx->h_pinned + first, count * sizeof(uint64_t));
  raise_from_bits(ctx, st, static_cast<uint32_t>(ctx->h_pinned[kErrSlot]));
}

// Host arrays are staged through device buffers; device arrays pass through.
bool is_device_ptr(const void* p) {
  if (!p) return true;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

template <typename T>
class Staged {
  using M = std::remove_const_t<T>;

 public:
  // in: copy host -> device before the call; out: device -> host after.
  Staged(Context* ctx, Scratch& scratch, T* user, uint64_t count, bool in, bool out,
         hipStream_t st)
      : user_(user), count_(count), out_(out), st_(st) {
    (void)ctx;
    check_records(count);
    if (user == nullptr || is_device_ptr(user)) {
      dev_ = const_cast<M*>(user);
      return;
    }
    dev_ = scratch.get<M>(count ? count : 1);
    staged_ = true;
    if (in && count)
      KC_HIP(hipMemcpyAsync(dev_, user, count * sizeof(T), hipMemcpyHostToDevice, st));
  }
  M* dev() const { return dev_; }
  void finish() {
    if (staged_ && out_ && count_)
      KC_HIP(hipMemcpyAsync(const_cast<M*>(user_), dev_, count_ * sizeof(T), hipMemcpyDeviceToHost, st_));
  }

 private:
  T* user_;
  M* dev_ = nullptr;
  uint64_t count_;
  bool out_;
  bool staged_ = false;
  hipStream_t st_;
};

bool load_tuning_file(Context* ctx, const std::string& path) {
  std::ifstream in(path);
  if (!in) return false;
  std::vector<TuneRow> rows, tree_rows, seg_rows;
  std::string line;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ls(line);
    int kind = KC_TUNING_LIST;
    ls >> std::ws;
    if (ls.peek() == 't' || ls.peek() == 's') {
      std::string word;
      if (!(ls >> word)) continue;
      if (word == "tree") kind = KC_TUNING_TREE;
      else if (word == "seg") kind = KC_TUNING_SEGMENTED;
      else continue;
    }
    TuneRow r{};
    if (!(ls >> r.log2n >> r.spacing_top >> r.spacing_deep >> r.chase_blocks_per_cu >> r.chase_chunk)) continue;
    auto pow2 = [](uint32_t x) { return x >= 2 && (x & (x - 1)) == 0; };
    if (!pow2(r.spacing_top) || !pow2(r.spacing_deep) || r.chase_blocks_per_cu == 0 || r.chase_chunk == 0)
      continue;
    (kind == KC_TUNING_TREE ? tree_rows : kind == KC_TUNING_SEGMENTED ? seg_rows : rows).push_back(r);
  }
  if (rows.empty() && tree_rows.empty() && seg_rows.empty()) return false;
  auto by_class = [](const TuneRow& a, const TuneRow& b) { return a.log2n < b.log2n; };
  std::sort(rows.begin(), rows.end(), by_class);
  std::sort(tree_rows.begin(), tree_rows.end(), by_class);
  std::sort(seg_rows.begin(), seg_rows.end(), by_class);
  ctx->tuning = rows;
  ctx->tuning_tree = tree_rows;
  ctx->tuning_seg = seg_rows;
  return true;
}

// tuning_table.txt next to the shared library, or $KASCADE_AMD_TUNING.
static void load_default_tuning(Context* ctx) {
  if (const char* env = std::getenv("KASCADE_AMD_TUNING")) {
    load_tuning_file(ctx, env);
    return;
  }
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&load_tuning_file), &info) && info.dli_fname) {
    std::string so(info.dli_fname);
    const size_t slash = so.rfind('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : so.substr(0, slash);
    load_tuning_file(ctx, dir + "/tuning_table.txt");
  }
}

}  // namespace kc

using kc::Context;
using kc::KcError;
using kc::Scratch;
using kc::Staged;

namespace {

template <typename F>
kc_status guarded(void* vctx, F&& body) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx) return KC_EINVAL;
  try {
    KC_HIP(hipSetDevice(ctx->device));
    body(ctx);
    if (ctx->prof.on) ctx->prof.flush();
    ctx->last_error.clear();
    return KC_OK;
  } catch (const KcError& e) {
    ctx->last_error = e.what;
    if (e.status != KC_EDEVICE) {
      // leave the context usable: drain the stream and clear the error word
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipMemset(ctx->d_err, 0, sizeof(uint32_t));
    }
    return e.status;
  } catch (const std::bad_alloc&) {
    ctx->last_error = "device allocation failed";
    return KC_ENOMEM;
  }
}

hipStream_t pick(Context* ctx, void* stream) {
  return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

}  // namespace

extern "C" {

kc_status kc_context_create(int device, uint64_t mem_budget_bytes, void** ctx_out) {
  if (!ctx_out) return KC_EINVAL;
  *ctx_out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return KC_EDEVICE;
  auto* ctx = new (std::nothrow) Context();
  if (!ctx) return KC_ENOMEM;
  ctx->device = device;
  ctx->arena.device = device;
  ctx->budget = mem_budget_bytes;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&ctx->d_counters, kc::kCounterSlots * sizeof(uint64_t)) != hipSuccess ||
      hipHostMalloc(&ctx->h_pinned, kc::kCounterSlots * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess ||
      hipMemset(ctx->d_counters, 0, kc::kCounterSlots * sizeof(uint64_t)) != hipSuccess) {
    kc_context_destroy(ctx);
    return KC_EDEVICE;
  }
  ctx->d_err = reinterpret_cast<uint32_t*>(ctx->d_counters + kc::kErrSlot);
  char* nb = nullptr;
  size_t nbytes = 0;
  if (kc::noalloc_scratch_init(device) && kc::noalloc_scratch(device, &nb, &nbytes))
    ctx->noalloc_arena.use_fixed(nb, nbytes);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  kc::load_default_tuning(ctx);
  *ctx_out = ctx;
  return KC_OK;
}

void kc_context_destroy(void* vctx) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  // the scratch may have served calls on caller streams too
  (void)hipDeviceSynchronize();
  ctx->arena.release(0);
  ctx->arena.recycle();
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->d_counters) (void)hipFree(ctx->d_counters);
  if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* kc_last_error(void* vctx) {
  Context* ctx = static_cast<Context*>(vctx);
  return ctx ? ctx->last_error.c_str() : "null context";
}

uint64_t kc_context_scratch_bytes(void* vctx) {
  Context* ctx = static_cast<Context*>(vctx);
  return ctx ? ctx->arena.held() : 0;
}

kc_status kc_context_trim(void* vctx) {
  return guarded(vctx, [&](Context* ctx) {
    // the scratch may have served calls on caller streams too
    KC_HIP(hipDeviceSynchronize());
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->graph_exec = nullptr;
    ctx->graph_key.clear();
    ctx->last_key.clear();
    ctx->arena.trim();
  });
}

kc_status kc_context_launch_stats(void* vctx, uint64_t* out3) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx || !out3) return KC_EINVAL;
  out3[0] = ctx->graph_captures;
  out3[1] = ctx->graph_replays;
  out3[2] = ctx->host_redos;
  return KC_OK;
}

kc_status kc_context_record_stats(void* vctx, uint64_t* out2) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx || !out2) return KC_EINVAL;
  out2[0] = ctx->narrow_calls;
  out2[1] = ctx->narrow_redos;
  return KC_OK;
}

kc_status kc_release_cached_scratch(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return KC_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return KC_EDEVICE;
  kc::release_chunk_cache(device);
  return KC_OK;
}

kc_status kc_rank_list(void* vctx, const uint64_t* succ, const int64_t* weight, uint64_t n,
                       uint64_t* root_out, int64_t* rank_out, const kc_rank_cfg* cfg,
                       void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && (!succ || !root_out || !rank_out)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, succ, n, true, false, st);
    Staged<const int64_t> w(ctx, scratch, weight, n, true, false, st);
    Staged<uint64_t> ro(ctx, scratch, root_out, n, false, true, st);
    Staged<int64_t> ra(ctx, scratch, rank_out, n, false, true, st);
    kc::rank_list_device(ctx, s.dev(), w.dev(), n, ro.dev(), ra.dev(), cfg, st);
    ro.finish();
    ra.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_rank_list_noalloc(void* vctx, const uint64_t* succ, const int64_t* weight, uint64_t n,
                               uint64_t* root_out, int64_t* rank_out, const kc_rank_cfg* cfg,
                               void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n > KC_NOALLOC_MAX_N)
      throw KcError{KC_EINVAL, "kc_rank_list_noalloc ranks at most 2^20 elements; use kc_rank_list"};
    if (n && (!succ || !root_out || !rank_out)) throw KcError{KC_EINVAL, "null array"};
    if (!kc::is_device_ptr(succ) || !kc::is_device_ptr(weight) || !kc::is_device_ptr(root_out) ||
        !kc::is_device_ptr(rank_out))
      throw KcError{KC_EINVAL, "kc_rank_list_noalloc needs device arrays (host arrays would be staged)"};
    if (!ctx->noalloc_arena.fixed())
      throw KcError{KC_EDEVICE, "the static scratch array's device address is not available"};
    hipStream_t st = pick(ctx, stream);
    kc_rank_cfg c{};
    if (cfg) c = *cfg;
    if (c.launch == KC_LAUNCH_AUTO) c.launch = KC_LAUNCH_DEVICE;  // a graph capture may allocate
    c.local_contraction = KC_CONTRACT_OFF;  // the scratch bound is for the uncontracted ranking
    // every allocation of the ranking comes from the static array; phase
    // timing (which creates events) is off for the call
    struct Restore {
      Context* ctx;
      bool prof;
      ~Restore() {
        ctx->active_arena = &ctx->arena;
        ctx->prof.on = prof;
      }
    } restore{ctx, ctx->prof.on};
    ctx->noalloc_arena.release(0);
    ctx->active_arena = &ctx->noalloc_arena;
    ctx->prof.on = false;
    try {
      kc::rank_list_device(ctx, succ, weight, n, root_out, rank_out, &c, st);
    } catch (const std::bad_alloc&) {
      (void)hipStreamSynchronize(st);  // kernels queued before the failed allocation still use the array
      throw KcError{KC_ENOMEM, "input needs more scratch than the static array of kc_rank_list_noalloc holds"};
    }
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_pointer_doubling(void* vctx, const kc_rank_cfg* cfg, uint64_t* succ_array,
                              int64_t* rank_array, uint64_t n, void* stream) {
  kc_rank_cfg c{};
  if (cfg) c = *cfg;
  c.algorithm = KC_ALGO_POINTER_DOUBLING;
  return kc_rank_list(vctx, succ_array, rank_array, n, succ_array, rank_array, &c, stream);
}

kc_status kc_sparse_ruling_set(void* vctx, const kc_rank_cfg* cfg, uint64_t* succ_array,
                               int64_t* rank_array, uint64_t n, void* stream) {
  kc_rank_cfg c{};
  if (cfg) c = *cfg;
  c.algorithm = KC_ALGO_SPARSE_RULING_SET;
  return kc_rank_list(vctx, succ_array, rank_array, n, succ_array, rank_array, &c, stream);
}

kc_status kc_pointer_doubling_rounds(void* vctx, const uint64_t* succ, const int64_t* weight,
                                     uint64_t n, uint32_t rounds, uint64_t* succ_out,
                                     int64_t* rank_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && (!succ || !succ_out || !rank_out)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, succ, n, true, false, st);
    Staged<const int64_t> w(ctx, scratch, weight, n, true, false, st);
    Staged<uint64_t> so(ctx, scratch, succ_out, n, false, true, st);
    Staged<int64_t> ro(ctx, scratch, rank_out, n, false, true, st);
    kc::pd_rounds_device(ctx, s.dev(), w.dev(), n, rounds, so.dev(), ro.dev(), st);
    so.finish();
    ro.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_rank_lists_segmented(void* vctx, const uint32_t* succ_local, const uint64_t* seg_off,
                                  uint64_t nseg, const int64_t* weight, uint32_t* root_local,
                                  int64_t* rank_out, const kc_rank_cfg* cfg, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    hipStream_t st = pick(ctx, stream);
    if (!seg_off) throw KcError{KC_EINVAL, "null seg_off"};
    Scratch scratch(ctx);
    Staged<const uint64_t> off(ctx, scratch, seg_off, nseg + 1, true, false, st);
    uint64_t total = 0, first = 0;
    KC_HIP(hipMemcpyAsync(&first, off.dev(), sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    KC_HIP(hipMemcpyAsync(&total, off.dev() + nseg, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    KC_HIP(hipStreamSynchronize(st));
    if (first != 0) throw KcError{KC_EINVAL, "seg_off[0] must be 0"};
    if (total && (!succ_local || !root_local || !rank_out)) throw KcError{KC_EINVAL, "null array"};
    Staged<const uint32_t> s(ctx, scratch, succ_local, total, true, false, st);
    Staged<const int64_t> w(ctx, scratch, weight, total, true, false, st);
    Staged<uint32_t> ro(ctx, scratch, root_local, total, false, true, st);
    Staged<int64_t> ra(ctx, scratch, rank_out, total, false, true, st);
    kc::rank_segmented_device(ctx, s.dev(), off.dev(), nseg, total, w.dev(), ro.dev(), ra.dev(), cfg, st);
    ro.finish();
    ra.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_root_tree(void* vctx, const uint64_t* parent, uint64_t n, uint64_t* root_out,
                       int64_t* depth_out, const kc_tree_cfg* cfg, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && (!parent || !root_out || !depth_out)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> p(ctx, scratch, parent, n, true, false, st);
    Staged<uint64_t> ro(ctx, scratch, root_out, n, false, true, st);
    Staged<int64_t> de(ctx, scratch, depth_out, n, false, true, st);
    kc::root_tree_device(ctx, p.dev(), n, ro.dev(), de.dev(), cfg, st);
    ro.finish();
    de.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_rank_via_euler_tour(void* vctx, const kc_tree_cfg* cfg, uint64_t* succ_array,
                                 int64_t* rank_array, uint64_t n, void* stream) {
  return kc_root_tree(vctx, succ_array, n, succ_array, rank_array, cfg, stream);
}

kc_status kc_compute_euler_tour(void* vctx, const uint64_t* parent, uint64_t n, uint64_t* tour_succ_out,
                                int64_t* tour_weight_out, uint64_t* tour_len_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && (!parent || !tour_succ_out)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> p(ctx, scratch, parent, n, true, false, st);
    Staged<uint64_t> so(ctx, scratch, tour_succ_out, 2 * n, false, true, st);
    Staged<int64_t> wo(ctx, scratch, tour_weight_out, 2 * n, false, true, st);
    const uint64_t len = kc::compute_euler_tour_device(ctx, p.dev(), n, so.dev(), wo.dev(), st);
    so.finish();
    wo.finish();
    KC_HIP(hipStreamSynchronize(st));
    if (tour_len_out) *tour_len_out = len;
  });
}

kc_status kc_gen_list(void* vctx, uint64_t n, uint64_t seed, int weighted, int local,
                      uint64_t* succ_out, int64_t* weight_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<uint64_t> so(ctx, scratch, succ_out, n, false, true, st);
    Staged<int64_t> wo(ctx, scratch, weight_out, n, false, true, st);
    kc::gen_list_device(n, seed, weighted, local, so.dev(), wo.dev(), st);
    so.finish();
    wo.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_gen_tree(void* vctx, uint64_t n, uint64_t seed, uint64_t* parent_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<uint64_t> po(ctx, scratch, parent_out, n, false, true, st);
    kc::gen_tree_device(n, seed, po.dev(), st);
    po.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_gen_lists_segmented(void* vctx, const uint64_t* seg_off, uint64_t nseg,
                                 const uint64_t* seeds, const uint8_t* local_flags, int weighted,
                                 uint32_t* succ_out, int64_t* weight_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    hipStream_t st = pick(ctx, stream);
    if (!seg_off || !seeds) throw KcError{KC_EINVAL, "null array"};
    Scratch scratch(ctx);
    Staged<const uint64_t> off(ctx, scratch, seg_off, nseg + 1, true, false, st);
    Staged<const uint64_t> sd(ctx, scratch, seeds, nseg, true, false, st);
    Staged<const uint8_t> lf(ctx, scratch, local_flags, nseg, true, false, st);
    uint64_t total = 0;
    KC_HIP(hipMemcpyAsync(&total, off.dev() + nseg, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    KC_HIP(hipStreamSynchronize(st));
    Staged<uint32_t> so(ctx, scratch, succ_out, total, false, true, st);
    Staged<int64_t> wo(ctx, scratch, weight_out, total, false, true, st);
    kc::gen_lists_segmented_device(off.dev(), nseg, total, sd.dev(), lf.dev(), weighted, so.dev(),
                                   wo.dev(), st);
    so.finish();
    wo.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"

namespace kc {
void initial_state_device(const uint64_t* succ, uint64_t n, uint64_t* root, int64_t* rank,
                          hipStream_t st);
void pack_device(Context* ctx, const uint64_t* succ, const int64_t* rank, uint64_t n,
                 const uint64_t* active, uint64_t na, uint64_t* psucc, int64_t* prank,
                 hipStream_t st);
void unpack_device(Context* ctx, const uint64_t* psucc, const int64_t* prank,
                   const uint64_t* active, uint64_t na, uint64_t n, uint64_t* succ, int64_t* rank,
                   hipStream_t st);
}  // namespace kc

extern "C" {

kc_status kc_set_initial_ranking_state(void* vctx, const uint64_t* succ, uint64_t n,
                                       uint64_t* root_out, int64_t* rank_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && (!succ || !root_out || !rank_out)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, succ, n, true, false, st);
    Staged<uint64_t> ro(ctx, scratch, root_out, n, false, true, st);
    Staged<int64_t> ra(ctx, scratch, rank_out, n, false, true, st);
    kc::initial_state_device(s.dev(), n, ro.dev(), ra.dev(), st);
    KC_HIP(hipGetLastError());
    ro.finish();
    ra.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_is_list(void* vctx, const uint64_t* succ, uint64_t n, int32_t* is_list_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (!is_list_out || (n && !succ)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, succ, n, true, false, st);
    *is_list_out = kc::is_list_device(ctx, s.dev(), n, st) ? 1 : 0;
  });
}

kc_status kc_leaf_info(void* vctx, const uint64_t* succ, uint64_t n, uint8_t* is_leaf_out,
                       uint64_t* num_leaves_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && !succ) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, succ, n, true, false, st);
    Staged<uint8_t> lf(ctx, scratch, is_leaf_out, n, false, true, st);
    const uint64_t count = kc::leaf_info_device(ctx, s.dev(), n, lf.dev(), st);
    lf.finish();
    KC_HIP(hipStreamSynchronize(st));
    if (num_leaves_out) *num_leaves_out = count;
  });
}

kc_status kc_reverse_list(void* vctx, const uint64_t* succ, const int64_t* dist_to_succ, uint64_t n,
                          uint64_t* pred_out, int64_t* dist_to_pred_out, uint64_t* roots_out,
                          uint64_t roots_capacity, uint64_t* num_roots_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && (!succ || !pred_out || !dist_to_pred_out)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, succ, n, true, false, st);
    Staged<const int64_t> w(ctx, scratch, dist_to_succ, n, true, false, st);
    Staged<uint64_t> p(ctx, scratch, pred_out, n, false, true, st);
    Staged<int64_t> d(ctx, scratch, dist_to_pred_out, n, false, true, st);
    const uint64_t cap = roots_out ? roots_capacity : 0;
    Staged<uint64_t> r(ctx, scratch, roots_out, std::min(cap, n), false, true, st);
    const uint64_t nroots =
        kc::reverse_list_device(ctx, s.dev(), w.dev(), n, p.dev(), d.dev(), r.dev(), cap, st);
    p.finish();
    d.finish();
    r.finish();
    KC_HIP(hipStreamSynchronize(st));
    if (num_roots_out) *num_roots_out = nroots;
  });
}

kc_status kc_compute_stats(void* vctx, int32_t level, const uint64_t* root, const int64_t* rank, uint64_t n,
                           kc_stats* out, uint64_t* tree_root_out, uint64_t* tree_size_out,
                           int64_t* tree_max_rank_out, int64_t* tree_rank_sum_out, uint64_t tree_capacity,
                           void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (!out || (n && (!root || !rank))) throw KcError{KC_EINVAL, "null array"};
    if (level < KC_STATS_NONE || level > KC_STATS_EXTENSIVE) throw KcError{KC_EINVAL, "invalid statistics level"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> ro(ctx, scratch, root, n, true, false, st);
    Staged<const int64_t> ra(ctx, scratch, rank, n, true, false, st);
    const uint64_t cap = std::min(tree_capacity, n);
    Staged<uint64_t> tr(ctx, scratch, tree_root_out, cap, false, true, st);
    Staged<uint64_t> ts(ctx, scratch, tree_size_out, cap, false, true, st);
    Staged<int64_t> tm(ctx, scratch, tree_max_rank_out, cap, false, true, st);
    Staged<int64_t> tsum(ctx, scratch, tree_rank_sum_out, cap, false, true, st);
    kc::compute_stats_device(ctx, ro.dev(), ra.dev(), n, level, out, tr.dev(), ts.dev(), tm.dev(), tsum.dev(),
                             cap, st);
    tr.finish();
    ts.finish();
    tm.finish();
    tsum.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_gen_graph(void* vctx, int32_t kind, uint64_t n, uint64_t m, uint64_t seed, uint64_t* src_out,
                       uint64_t* dst_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (kind != KC_GRAPH_GNM && kind != KC_GRAPH_RMAT) throw KcError{KC_EINVAL, "invalid graph kind"};
    if (m && (n == 0 || !src_out || !dst_out)) throw KcError{KC_EINVAL, "null array or empty vertex set"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<uint64_t> s(ctx, scratch, src_out, m, false, true, st);
    Staged<uint64_t> d(ctx, scratch, dst_out, m, false, true, st);
    kc::gen_graph_device(ctx, kind, n, m, seed, s.dev(), d.dev(), st);
    s.finish();
    d.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_gen_rgg2d(void* vctx, uint64_t n, uint64_t m, uint64_t seed, uint64_t* src_out, uint64_t* dst_out,
                       uint64_t cap, uint64_t* m_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (!m_out) throw KcError{KC_EINVAL, "null edge count"};
    hipStream_t st = pick(ctx, stream);
    if (!src_out || !dst_out) {
      *m_out = kc::gen_rgg2d_device(ctx, n, m, seed, nullptr, nullptr, 0, st);
      return;
    }
    Scratch scratch(ctx);
    Staged<uint64_t> s(ctx, scratch, src_out, cap, false, true, st);
    Staged<uint64_t> d(ctx, scratch, dst_out, cap, false, true, st);
    *m_out = kc::gen_rgg2d_device(ctx, n, m, seed, s.dev(), d.dev(), cap, st);
    s.finish();
    d.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_build_csr(void* vctx, const uint64_t* src, const uint64_t* dst, uint64_t n, uint64_t m,
                       uint64_t* xadj_out, uint64_t* adj_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (!xadj_out || (m && (!src || !dst || !adj_out))) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, src, m, true, false, st);
    Staged<const uint64_t> d(ctx, scratch, dst, m, true, false, st);
    Staged<uint64_t> x(ctx, scratch, xadj_out, n + 1, false, true, st);
    Staged<uint64_t> a(ctx, scratch, adj_out, 2 * m, false, true, st);
    kc::build_csr_device(ctx, s.dev(), d.dev(), n, m, x.dev(), a.dev(), st);
    x.finish();
    a.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_bfs_forest(void* vctx, const uint64_t* xadj, const uint64_t* adj, uint64_t n, uint64_t* parent_out,
                        void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (n && (!xadj || !parent_out)) throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    uint64_t m2 = 0;
    if (n) {
      // the adjacency length is xadj[n]
      if (kc::is_device_ptr(xadj))
        KC_HIP(hipMemcpy(&m2, xadj + n, sizeof(uint64_t), hipMemcpyDeviceToHost));
      else
        m2 = xadj[n];
    }
    if (m2 && !adj) throw KcError{KC_EINVAL, "null adjacency"};
    Staged<const uint64_t> x(ctx, scratch, xadj, n + 1, true, false, st);
    Staged<const uint64_t> a(ctx, scratch, adj, m2, true, false, st);
    Staged<uint64_t> p(ctx, scratch, parent_out, n, false, true, st);
    kc::bfs_forest_device(ctx, x.dev(), a.dev(), n, p.dev(), st);
    p.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_trace_enable(void* vctx, int on) {
  return guarded(vctx, [&](Context* ctx) {
    ctx->trace_on = on != 0;
    ctx->trace_subproblem = 0;
    ctx->trace.clear();
  });
}

uint32_t kc_trace_read(void* vctx, kc_trace_level* out, uint32_t max_entries) {
  auto* ctx = static_cast<Context*>(vctx);
  if (!ctx) return 0;
  const uint32_t k = static_cast<uint32_t>(ctx->trace.size());
  for (uint32_t i = 0; i < k && i < max_entries && out; i++) out[i] = ctx->trace[i];
  return k;
}

kc_status kc_pack(void* vctx, const uint64_t* succ, const int64_t* rank, uint64_t n,
                  const uint64_t* active, uint64_t nactive, uint64_t* psucc_out,
                  int64_t* prank_out, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (nactive && (!succ || !rank || !active || !psucc_out || !prank_out))
      throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> s(ctx, scratch, succ, n, true, false, st);
    Staged<const int64_t> r(ctx, scratch, rank, n, true, false, st);
    Staged<const uint64_t> a(ctx, scratch, active, nactive, true, false, st);
    Staged<uint64_t> ps(ctx, scratch, psucc_out, nactive, false, true, st);
    Staged<int64_t> pr(ctx, scratch, prank_out, nactive, false, true, st);
    kc::pack_device(ctx, s.dev(), r.dev(), n, a.dev(), nactive, ps.dev(), pr.dev(), st);
    ps.finish();
    pr.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

kc_status kc_unpack(void* vctx, const uint64_t* psucc, const int64_t* prank,
                    const uint64_t* active, uint64_t nactive, uint64_t* succ_inout,
                    int64_t* rank_inout, uint64_t n, void* stream) {
  return guarded(vctx, [&](Context* ctx) {
    if (nactive && (!psucc || !prank || !active || !succ_inout || !rank_inout))
      throw KcError{KC_EINVAL, "null array"};
    hipStream_t st = pick(ctx, stream);
    Scratch scratch(ctx);
    Staged<const uint64_t> ps(ctx, scratch, psucc, nactive, true, false, st);
    Staged<const int64_t> pr(ctx, scratch, prank, nactive, true, false, st);
    Staged<const uint64_t> a(ctx, scratch, active, nactive, true, false, st);
    Staged<uint64_t> s(ctx, scratch, succ_inout, n, true, true, st);
    Staged<int64_t> r(ctx, scratch, rank_inout, n, true, true, st);
    kc::unpack_device(ctx, ps.dev(), pr.dev(), a.dev(), nactive, n, s.dev(), r.dev(), st);
    s.finish();
    r.finish();
    KC_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"

extern "C" {

kc_status kc_profile_enable(void* vctx, int on) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx) return KC_EINVAL;
  ctx->prof.on = on != 0;
  return KC_OK;
}

kc_status kc_profile_reset(void* vctx) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx) return KC_EINVAL;
  ctx->prof.reset();
  return KC_OK;
}

uint32_t kc_profile_read(void* vctx, uint32_t max_entries, const char** names, double* ms,
                         uint64_t* counts) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx) return 0;
  ctx->prof.flush();
  uint32_t k = 0;
  for (auto& t : ctx->prof.totals) {
    if (k >= max_entries) break;
    if (names) names[k] = t.name.c_str();
    if (ms) ms[k] = t.ms;
    if (counts) counts[k] = t.count;
    k++;
  }
  return static_cast<uint32_t>(ctx->prof.totals.size());
}

}  // extern "C"

extern "C" {

kc_status kc_tuning_set(void* vctx, const uint32_t* rows, uint32_t nrows) {
  return kc_tuning_set_kind(vctx, KC_TUNING_LIST, rows, nrows);
}

kc_status kc_tuning_set_kind(void* vctx, int32_t kind, const uint32_t* rows, uint32_t nrows) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx || (nrows && !rows) || kind < KC_TUNING_LIST || kind > KC_TUNING_SEGMENTED) return KC_EINVAL;
  std::vector<kc::TuneRow> t;
  for (uint32_t i = 0; i < nrows; i++) {
    kc::TuneRow r{rows[5 * i], rows[5 * i + 1], rows[5 * i + 2], rows[5 * i + 3], rows[5 * i + 4]};
    auto pow2 = [](uint32_t x) { return x >= 2 && (x & (x - 1)) == 0; };
    if (!pow2(r.spacing_top) || !pow2(r.spacing_deep) || r.chase_blocks_per_cu == 0 || r.chase_chunk == 0)
      return KC_EINVAL;
    t.push_back(r);
  }
  std::sort(t.begin(), t.end(), [](const kc::TuneRow& a, const kc::TuneRow& b) { return a.log2n < b.log2n; });
  (kind == KC_TUNING_TREE ? ctx->tuning_tree : kind == KC_TUNING_SEGMENTED ? ctx->tuning_seg : ctx->tuning) = t;
  return KC_OK;
}

uint32_t kc_tuning_get(void* vctx, uint64_t n, uint32_t* out5) {
  return kc_tuning_get_kind(vctx, KC_TUNING_LIST, n, out5);
}

uint32_t kc_tuning_get_kind(void* vctx, int32_t kind, uint64_t n, uint32_t* out5) {
  Context* ctx = static_cast<Context*>(vctx);
  if (!ctx || !out5 || kind < KC_TUNING_LIST || kind > KC_TUNING_SEGMENTED) return 0;
  kc::RankTuning t = kc::tuning_for(ctx, n, nullptr, kind);
  out5[0] = t.spacing_top;
  out5[1] = t.spacing_deep;
  out5[2] = t.chase_blocks_per_cu;
  out5[3] = t.chase_chunk;
  out5[4] = t.small_n;
  return static_cast<uint32_t>(
      (kind == KC_TUNING_TREE ? ctx->tuning_tree : kind == KC_TUNING_SEGMENTED ? ctx->tuning_seg : ctx->tuning).size());
}

}  // extern "C"
