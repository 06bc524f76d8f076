// Internal declarations shared by the HIP sources of libkascade_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/kascade_amd.h"

namespace kc {

constexpr uint64_t kRootFlag = uint64_t{1} << 63;   // lib/kascade/bits.hpp:9
constexpr uint64_t kAuxFlag = uint64_t{1} << 62;    // marks "unclaimed, packed" records
constexpr uint64_t kIdMask = kAuxFlag - 1;

// One list element as the kernels see it: successor id (with flags) and the
// weight / distance that travels with it.  16 bytes so a random access is a
// single 16-byte load inside one 64-byte HBM burst.
struct alignas(16) Rec {
  uint64_t s;
  int64_t w;
};

// The narrow form of a level-0 list record, 8 bytes: ids below 2^30 with the
// root and aux flags in bits 31 and 30, and weights and claim distances that
// fit int32.  A dependent random walk that claims every record it reads runs
// 10 % faster over 8-byte records than over 16-byte ones (2^27 records on
// MI355X: 5.61 against 6.27 ms), because twice as many of them stay in the
// caches.  Used for the top level of single-list rankings from 2^22
// elements; anything that does not fit is redone with wide records.
struct alignas(8) Rec8 {
  uint32_t s;
  int32_t w;
};
// The same 8 bytes with the successor stored relative to the record's own
// position (flagged ids stay absolute): for segmented batches of any total
// size, whose successors lie inside their segment.
struct alignas(8) Rec8D {
  uint32_t s;
  int32_t w;
};
constexpr uint32_t kRoot8 = 1u << 31;
constexpr uint32_t kAux8 = 1u << 30;
constexpr uint64_t kNarrowMaxIds = uint64_t{1} << 30;

// Device-side error word: kernels OR these bits in; the host maps them to kc_status.
enum ErrBits : uint32_t {
  kErrRange = 1u,      // successor id >= n
  kErrRootWeight = 2u, // nonzero weight on a tail
  kErrNotList = 4u,    // an element reached twice (in-degree > 1)
  kErrCycle = 8u,      // a cycle without a tail
  kErrOverflow = 16u,  // a device-sized level outgrew a capacity (redo host-sized; never reported)
  kErrNarrow = 32u,    // a claim distance outgrew a narrow record (redo wide; never reported)
};

// Scratch pool: chunks reused across calls, handed out as a stack.  When a
// call needed more than one chunk, the next call starts by replacing them with
// one chunk of the largest size any call has used, so a steady workload runs
// from a single chunk and allocates nothing.  Chunks of a destroyed context go
// to a per-device cache that later contexts take from before they allocate.
class Arena {
 public:
  ~Arena();
  int device = 0;
  // Returns nullptr on allocation failure.
  void* alloc(size_t bytes);
  size_t mark() const { return top_; }
  void release(size_t mark);
  // Hands every chunk to the device's cache (no live blocks; device idle).
  void recycle();
  // Frees every chunk (no live blocks).
  void trim();
  uint64_t held() const;
  // A fresh arena's first chunk, sized up front (a no-op once it holds any):
  // an arena that grows later does so by consolidating, which waits for the
  // whole device.
  void reserve(size_t bytes);
  uint64_t peak() const { return peak_; }
  // Appends what decides the addresses of the next allocations (chunks and
  // the stack position), so equal signatures hand out equal addresses.
  void signature(std::vector<uint64_t>& out) const;
  // Serves every allocation from [base, base + size), memory the arena does
  // not own: it never allocates, frees or caches a chunk, and an allocation
  // that does not fit fails (alloc returns nullptr).
  void use_fixed(char* base, size_t size);
  bool fixed() const { return fixed_; }

 private:
  struct Chunk {
    char* base;
    size_t size;
  };
  struct Block {
    size_t chunk;
    size_t offset;
    size_t size;
  };
  bool new_chunk(size_t want);
  void consolidate();
  std::vector<Chunk> chunks_;
  std::vector<Block> blocks_;
  size_t top_ = 0;   // number of live blocks
  size_t live_ = 0;  // bytes in live blocks
  size_t peak_ = 0;  // most bytes live at once, over the arena's life
  bool fixed_ = false;
  std::atomic<uint64_t> held_{0};  // readable from other threads (server statistics)
};
// Frees the chunks cached for `device` (kc_release_cached_scratch).
void release_chunk_cache(int device);
// The static scratch array of kc_rank_list_noalloc (noalloc.hip).
// noalloc_scratch_init looks up this device's copy once per device (the
// device must be current); noalloc_scratch returns it.  Both return false
// when the address is not available.
bool noalloc_scratch_init(int device);
bool noalloc_scratch(int device, char** base, size_t* bytes);

// Per-phase device timing with HIP events on the launching stream (the
// library's counterpart of the reference's kamping phase timers).
class Profiler {
 public:
  bool on = false;
  void begin(const char* name, hipStream_t st);
  void end(hipStream_t st);
  // Resolves recorded events (the stream must be synchronised).
  void flush();
  void reset();
  struct Acc {
    std::string name;
    double ms = 0;
    uint64_t count = 0;
  };
  std::vector<Acc> totals;
  ~Profiler();

 private:
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  hipEvent_t get();
  std::vector<hipEvent_t> pool_;
  std::vector<Pending> pending_;
  std::vector<std::pair<std::string, hipEvent_t>> open_;
};

// Device counter block: slots [0, kErrSlot) are scratch counters of the
// current operation, slot kErrSlot holds the error word.
constexpr int kCounterSlots = 64;
constexpr int kErrSlot = kCounterSlots - 1;

struct TuneRow;
struct Context {
  Profiler prof;
  std::vector<TuneRow> tuning;       // lists; sorted by log2n; empty = built-in table
  std::vector<TuneRow> tuning_tree;  // tree rooting, by vertices; the same
  std::vector<TuneRow> tuning_seg;   // segmented batches, by total elements; the same
  int device = 0;
  uint64_t budget = 0;
  hipStream_t stream = nullptr;
  Arena arena;
  // The static scratch of kc_rank_list_noalloc (fixed over this device's
  // copy of the library's static array; unset when its address was not
  // available), and the arena Scratch allocates from: `arena` except inside
  // kc_rank_list_noalloc.
  Arena noalloc_arena;
  Arena* active_arena = &arena;
  uint32_t* d_err = nullptr;       // error word (inside the counter block)
  uint64_t* d_counters = nullptr;  // kCounterSlots device counters
  uint64_t* h_pinned = nullptr;    // pinned host mirror of counters
  std::string last_error;
  int num_cus = 256;
  // per-level trace (kc_trace_*)
  bool trace_on = false;
  int trace_subproblem = 0;
  std::vector<kc_trace_level> trace;
  // HIP graph of the last repeated device-sized ranking, and the keys (every
  // argument of the launch sequence) of that graph and of the last call
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint64_t> graph_key;
  std::vector<uint64_t> last_key;
  uint64_t graph_replays = 0;
  uint64_t graph_captures = 0;
  uint64_t host_redos = 0;  // device-sized rankings redone host-sized after an overflow
  uint64_t narrow_calls = 0;  // list rankings whose top level used narrow records
  uint64_t narrow_redos = 0;  // narrow-record rankings redone with wide records
};

// Scoped stack allocation from the arena.
class Scratch {
 public:
  explicit Scratch(Context* ctx) : arena_(ctx->active_arena), mark_(arena_->mark()) {}
  ~Scratch() { arena_->release(mark_); }
  template <typename T>
  T* get(uint64_t count) {
    void* p = arena_->alloc(count * sizeof(T) + 256);
    if (!p) throw std::bad_alloc();
    return static_cast<T*>(p);
  }

 private:
  Arena* arena_;
  size_t mark_;
};

struct KcError {
  kc_status status;
  std::string what;
};

#define KC_HIP(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw ::kc::KcError{KC_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// RAII phase timer: no-op unless the context's profiler is on.
class Phase {
 public:
  Phase(Context* ctx, const char* name, hipStream_t st) : ctx_(ctx), st_(st) {
    if (ctx_->prof.on) ctx_->prof.begin(name, st_);
  }
  ~Phase() {
    if (ctx_->prof.on) ctx_->prof.end(st_);
  }

 private:
  Context* ctx_;
  hipStream_t st_;
};

// Kernel grids count work-items in 32 bits (blocks x threads < 2^32), and the
// one-thread-per-record passes launch ceil(records / 256) blocks of 256, so a
// ranking holds at most kMaxRecords records (a tree's tour: two per vertex).
constexpr uint64_t kMaxRecords = (uint64_t{1} << 32) - 256;
inline void check_records(uint64_t records) {
  if (records > kMaxRecords) throw KcError{KC_EINVAL, "more than 2^32 - 256 records (elements, or twice the vertices)"};
}

inline unsigned blocks_for(uint64_t n, unsigned threads) {
  uint64_t b = (n + threads - 1) / threads;
  return static_cast<unsigned>(b ? b : 1);
}

// Reads the device error word (synchronising `stream`) and throws on error.
void check_err(Context* ctx, hipStream_t stream);
// After a device-sized ranking (synchronising `stream`): raises input errors;
// returns the redo bits, 0 when the call is done: kErrOverflow (a capacity
// overflowed: redo host-sized), kErrNarrow (a claim distance outgrew a narrow
// record: redo with wide records).
uint32_t check_err_overflow(Context* ctx, hipStream_t stream);
// Copies `count` counters starting at slot `first` to the host (synchronising).
void read_counters(Context* ctx, hipStream_t stream, int first, int count, uint64_t* out);

// ---- tuning ----------------------------------------------------------------
struct RankTuning {
  uint32_t spacing_top;   // elements per ruler at the top level
  uint32_t spacing_deep;  // elements per ruler below
  uint32_t max_levels;
  uint32_t small_n;       // sizes at or below this go to the one-workgroup solver
  uint32_t chase_blocks_per_cu;
  uint32_t chase_chunk;  // rulers a wave takes per work-counter atomic
  // spacings set by the caller are used as given; otherwise each level caps
  // its spacing so that the lists' head prefixes (about K elements per list,
  // ranked as a sub-problem) stay within an eighth of the level
  bool spacing_top_fixed;
  bool spacing_deep_fixed;
  uint32_t launch;  // kc_launch
  uint32_t spawn;   // kc_spawn
  uint32_t records;  // kc_records
  // level 0 of a forest caps its spacing by the cache window (level_spacing);
  // off for rows measured on such forests (the segmented table)
  bool window_cap;
};
// Per-size-class tuning table row: log2(n) class and the knobs picked for it
// (written by kascade_amd/autotune.py into tuning_table.txt, which contexts
// load at creation; a built-in table is used otherwise).
struct TuneRow {
  uint32_t log2n;
  uint32_t spacing_top;
  uint32_t spacing_deep;
  uint32_t chase_blocks_per_cu;
  uint32_t chase_chunk;
};
// The knobs for ranking n list elements (KC_TUNING_LIST), rooting a tree of
// n vertices (KC_TUNING_TREE; the tour has 2n records) or a segmented batch
// of n elements in total (KC_TUNING_SEGMENTED; below its first class, the
// list table): each kind has its own table.
RankTuning tuning_for(const Context* ctx, uint64_t n, const kc_rank_cfg* cfg, int kind = KC_TUNING_LIST);
// Loads a tuning table file (one row per line, '#' comments; rows prefixed
// with "tree" are tree-rooting rows, "seg" segmented-batch rows); false if
// absent.
bool load_tuning_file(Context* ctx, const std::string& path);

// ---- internal entry points (device pointers only) -----------------------
// Level-0 ranking of separate succ / weight arrays into root / rank arrays.
void rank_list_device(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n,
                      uint64_t* root, int64_t* rank, const kc_rank_cfg* cfg, hipStream_t st);
// Where a ranking writes its outputs.  Element e produces output only if its
// low `shift` bits are zero and e >> shift < n_out; it goes to index
// (e >> shift) * stride, and root ids are reported >> shift.  shift = 0,
// stride = 1 is the plain layout; stride 2 writes into the fields of an
// array of records; shift 1 keeps only the even elements of an interleaved
// Euler tour (element 2v = the upward edge of v).
struct OutMap {
  uint32_t stride;
  uint32_t shift;
};
__host__ __device__ inline bool out_selected(const OutMap& om, uint64_t e, uint64_t n_out) {
  return (e & ((uint64_t{1} << om.shift) - 1)) == 0 && (e >> om.shift) < n_out;
}
__host__ __device__ inline uint64_t out_pos(const OutMap& om, uint64_t e) {
  return (e >> om.shift) * om.stride;
}
// One job of a mixed batch (device-side table entry).
struct BatchJob {
  uint64_t n;
  const uint64_t* succ;  // list successors / tree parents, local ids (device)
  const int64_t* w;      // list weights (device) or null
  uint64_t* root_out;    // local root ids (device)
  int64_t* rank_out;     // ranks / depths (device)
  uint64_t base;         // set by rank_batch_device: offset in its kind's id space
  uint32_t kind;         // kc_job_kind
  uint32_t pad;
};

// Level-0 outputs of a mixed batch written straight into each job's arrays:
// list element e (< nl) of the job whose base is the largest <= e, and tree
// vertex t of tour element nl + 2t (downward edges nl + 2t + 1 have none).
// list_block / tree_block hold the job of the first item of every block of
// 256 items, so a lookup advances over a few jobs at most.  lists == nullptr:
// outputs go through the OutMap instead.
struct JobOut {
  const BatchJob* lists = nullptr;
  const BatchJob* trees = nullptr;
  const uint32_t* list_block = nullptr;
  const uint32_t* tree_block = nullptr;
  uint32_t nlists = 0;
  uint32_t ntrees = 0;
  uint64_t nl = 0;
  // Segmented batch: roots are written as u32 ids local to their segment
  // (segment s owns [seg_off[s], seg_off[s + 1])), ranks at the element.
  uint32_t* seg_root = nullptr;
  const uint64_t* seg_off = nullptr;
  uint64_t nseg = 0;
};

// Euler tours inside a ranking: elements [lo, n) are interleaved tour pairs
// (lo + 2t = upward edge / dummy, lo + 2t + 1 = downward edge).  Only the
// upward edges' results are ever read, and the head of every tour with a
// root is known from the build (`heads`, nheads device ids).  The sparse
// ruling set then starts a chase at every head (no unreached prefixes in the
// tours) and writes no claim into downward edges; an unreached upward edge
// is either an isolated root (a self loop: its own root, rank 0) or lies on a
// cycle (an error).
struct TourRegion {
  uint64_t lo;
  const uint64_t* heads;
  uint64_t nheads;
};
// Ranks packed records (ids < n, consumed) into root (ids) / rank; outputs as
// `out_shift` selects (see OutMap; n_out outputs).  max_tails bounds the
// number of lists.  Device-sized launches (tune.launch) synchronise at the
// end and return false when a capacity overflowed: the caller then rebuilds
// the records and calls again with tune.launch = KC_LAUNCH_HOST.
[[nodiscard]] bool rank_records(Context* ctx, Rec* rec, uint64_t n, uint64_t n_out, uint64_t* root,
                                int64_t* rank, uint64_t max_tails, const RankTuning& tune, int algorithm,
                                hipStream_t st, uint32_t out_shift = 0, const TourRegion* tours = nullptr,
                                const JobOut* jobs_out = nullptr);
// KC_ALGO_AUTO: pointer doubling up to these sizes, the sparse ruling set
// above.  Measured on MI355X (whole calls, random weighted lists: PD 0.28 ms
// against SRS 0.31 at 2^19, 0.51 against 0.34 at 2^20; random trees: PD 0.23
// against 0.26 at 2^17 vertices, 0.29 against 0.27 at 2^18).
constexpr uint64_t kAutoPdMaxList = uint64_t{1} << 19;   // list elements
constexpr uint64_t kAutoPdMaxTree = uint64_t{1} << 17;   // tree vertices
// Validates cfg->algorithm and resolves KC_ALGO_AUTO for a ranking of n list
// elements (tree = false) or n tree vertices (tree = true).
int resolve_algorithm(const kc_rank_cfg* cfg, uint64_t n, bool tree);
// The narrow form of rank_records (the sparse ruling set, device-sized
// levels, no per-job outputs): returns the redo bits of check_err_overflow
// (0: done).
uint32_t rank_records_narrow(Context* ctx, Rec8* rec, uint64_t n, uint64_t n_out, uint64_t* root, int64_t* rank,
                             uint64_t max_tails, const RankTuning& tune, hipStream_t st, uint32_t out_shift = 0,
                             const TourRegion* tours = nullptr, const JobOut* jobs_out = nullptr);
// The same with position-relative narrow records (segmented batches: no
// tours).  narrow_level0_fits: every id the top level's claims and
// unreached prefixes can write into such records (sublist and prefix
// capacities) is below 2^30, for a ranking of n records with `tails` lists.
uint32_t rank_records_narrow(Context* ctx, Rec8D* rec, uint64_t n, uint64_t n_out, uint64_t* root, int64_t* rank,
                             uint64_t max_tails, const RankTuning& tune, hipStream_t st, const JobOut* jobs_out);
bool narrow_level0_fits(const Context* ctx, uint64_t n, uint64_t tails, const RankTuning& tune);
// Narrow records apply to a top level ranked by the ruling set with
// device-sized levels, below max_records (2^30 for absolute ids, flags in
// bits 31 and 30; position-relative records have no such limit) and from
// 2^22 (below, the records fit the caches either way: A/B on MI355X, lists
// of 2^20 0.353 against 0.338 ms wide, 2^22 0.524 both, 2^24 1.29 against
// 1.44, 2^26 4.31 against 4.97).
constexpr uint64_t kNarrowMin = uint64_t{1} << 22;
bool narrow_wanted(const Context* ctx, uint64_t records, int algo, const RankTuning& tune,
                   uint64_t max_records = kNarrowMaxIds);
// The level-0 ids (rulers, tails, spawned rulers) must stay below 2^30 too:
// rulers are at most records / 2, spawned ones at most a quarter.
bool narrow_ids_fit(uint64_t records, uint64_t tails);
// True when rank_records ranks n records with the sparse ruling set at the top
// level (the only top level that writes through a JobOut).
bool top_level_is_ruling_set(uint64_t n, const RankTuning& tune, int algorithm);
// Local contraction of a list ranking's level 0 (contract.hip).
// contraction_mode gives KC_CONTRACT_ON or OFF, or AUTO when a sample must
// decide: contraction_sample then adds to *out (zeroed before) how many of
// 65536 evenly spread packed records lie inside runs of consecutive
// successors, and
// contraction_sample_wins turns the count into the decision.  The ranking
// itself contracts, ranks the contracted list and writes every element's
// result (root element ids to root, or u32 ids local to their segment to
// seg_root): rank_contracted_list from the caller's arrays, with the checks
// pack_records makes, rank_contracted from packed records (unit_weights: the
// caller gave no weights).
struct ContractOut {
  uint64_t* root = nullptr;
  int64_t* rank = nullptr;
  uint32_t* seg_root = nullptr;
  const uint64_t* seg_off = nullptr;
  uint64_t nseg = 0;
};
uint32_t contraction_mode(const kc_rank_cfg* cfg, uint64_t n);
void contraction_sample(Context* ctx, const Rec* rec, uint64_t n, unsigned long long* out, hipStream_t st);
void contraction_sample(Context* ctx, const Rec8* rec, uint64_t n, unsigned long long* out, hipStream_t st);
void contraction_sample(Context* ctx, const Rec8D* rec, uint64_t n, unsigned long long* out, hipStream_t st);
bool contraction_sample_wins(uint64_t sampled_interior);
// The same decision from the caller's successor array, before any pack
// (synchronises `st`): for lists of kContractEarlySample elements or more,
// where the synchronisation costs about 1 % and a contracting ranking saves
// the pack.
bool contraction_sample_list(Context* ctx, const uint64_t* succ, uint64_t n, hipStream_t st);
constexpr uint64_t kContractEarlySample = uint64_t{1} << 24;
void rank_contracted_list(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n, const kc_rank_cfg* cfg,
                          int algorithm, const ContractOut& out, hipStream_t st);
void rank_contracted(Context* ctx, const Rec* rec, uint64_t n, bool unit_weights, const kc_rank_cfg* cfg,
                     int algorithm, const ContractOut& out, hipStream_t st);
void pd_rounds_device(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n,
                      uint32_t rounds, uint64_t* succ_out, int64_t* rank_out, hipStream_t st);
// Packs succ/weight into records with validation (range, tail weight).
void pack_records(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n, Rec* rec,
                  hipStream_t st);
void validate_list(Context* ctx, const uint64_t* succ, uint64_t n, hipStream_t st);
// graph inputs (graph.hip)
void gen_graph_device(Context* ctx, int kind, uint64_t n, uint64_t m, uint64_t seed, uint64_t* src, uint64_t* dst,
                      hipStream_t st);
// Random geometric graph: returns the edge count; writes the edges when src
// and dst are given (at most cap).  rgg2d_radius: the radius in 2^-31 units.
uint64_t gen_rgg2d_device(Context* ctx, uint64_t n, uint64_t m, uint64_t seed, uint64_t* src, uint64_t* dst,
                          uint64_t cap, hipStream_t st);
uint64_t rgg2d_radius(uint64_t n, uint64_t m);
void build_csr_device(Context* ctx, const uint64_t* src, const uint64_t* dst, uint64_t n, uint64_t m,
                      uint64_t* xadj, uint64_t* adj, hipStream_t st);
void bfs_forest_device(Context* ctx, const uint64_t* xadj, const uint64_t* adj, uint64_t n, uint64_t* parent,
                       hipStream_t st);
// successor utilities and statistics (listutil.hip)
bool is_list_device(Context* ctx, const uint64_t* succ, uint64_t n, hipStream_t st);
uint64_t leaf_info_device(Context* ctx, const uint64_t* succ, uint64_t n, uint8_t* is_leaf,
                          hipStream_t st);
uint64_t reverse_list_device(Context* ctx, const uint64_t* succ, const int64_t* w, uint64_t n,
                             uint64_t* pred, int64_t* dpred, uint64_t* roots, uint64_t roots_cap,
                             hipStream_t st);
void compute_stats_device(Context* ctx, const uint64_t* root, const int64_t* rank, uint64_t n, int level,
                          kc_stats* out, uint64_t* tree_root, uint64_t* tree_size, int64_t* tree_max,
                          int64_t* tree_sum, uint64_t tree_cap, hipStream_t st);
void root_tree_device(Context* ctx, const uint64_t* parent, uint64_t n, uint64_t* root,
                      int64_t* depth, const kc_tree_cfg* cfg, hipStream_t st);
// compute_euler_tour in the reference's element order: tour_succ (and
// tour_w, may be null) receive 2(n - roots) + roots entries; returns that
// length.  The arrays need capacity 2n.
uint64_t compute_euler_tour_device(Context* ctx, const uint64_t* parent, uint64_t n, uint64_t* tour_succ,
                                   int64_t* tour_w, hipStream_t st);
void rank_segmented_device(Context* ctx, const uint32_t* succ, const uint64_t* seg_off,
                           uint64_t nseg, uint64_t total, const int64_t* w, uint32_t* root,
                           int64_t* rank, const kc_rank_cfg* cfg, hipStream_t st);

// generators
void gen_list_device(uint64_t n, uint64_t seed, int weighted, int local, uint64_t* succ,
                     int64_t* w, hipStream_t st);
void gen_tree_device(uint64_t n, uint64_t seed, uint64_t* parent, hipStream_t st);
void gen_lists_segmented_device(const uint64_t* seg_off, uint64_t nseg, uint64_t total,
                                const uint64_t* seeds, const uint8_t* local_flags, int weighted,
                                uint32_t* succ, int64_t* w, hipStream_t st);

// Device radix sort of (key, value) pairs by key, stable, keys < 2^key_bits.
void radix_sort_pairs_u32(Context* ctx, uint32_t* keys, uint32_t* vals, uint64_t n,
                          unsigned key_bits, hipStream_t st);
void radix_sort_pairs_u64(Context* ctx, uint64_t* keys, uint64_t* vals, uint64_t n,
                          unsigned key_bits, hipStream_t st);
// Exclusive scan of u64 on the device (enqueued on `st`; returns 0).
uint64_t exclusive_scan_u64(Context* ctx, const uint64_t* in, uint64_t* out, uint64_t n,
                            hipStream_t st);

}  // namespace kc
