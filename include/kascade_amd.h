/*
 * kascade_amd.h -- C ABI of the MI355X list-ranking / tree-rooting library.
 *
 * Every entry point is a drop-in for one of KAscade's lib/kascade entry points
 * (cited per declaration; paths are relative to the niklas-uhl/kascade tree).
 * The MPI communicator and Distribution arguments of the reference disappear:
 * one call ranks the whole array on one GPU.
 *
 * Conventions (lib/kascade/types.hpp:8-9):
 *   ids are uint64_t (idx_t), ranks are int64_t (rank_t); ids must be < 2^63
 *   because bit 63 is the root flag (lib/kascade/bits.hpp:9-21).
 *   One call ranks at most 2^32 - 256 records: list elements, or twice the
 *   vertices of a tree (its Euler tour); kernel grids count work-items in 32
 *   bits.  Larger calls return KC_EINVAL.
 *
 * Pointers may be device memory (hipMalloc / torch tensors) or host memory;
 * host arrays are staged through device buffers by the library.  All entry
 * points are synchronous with respect to the host: they enqueue on `stream`
 * and return when the results are written, like the reference's collective
 * calls.  A context is externally synchronised, like a communicator: one host
 * thread at a time.
 *
 * Stream ordering.  `stream` is a hipStream_t passed as void*:
 *   - a stream the caller created: the call runs on it, after the work
 *     already queued there;
 *   - NULL: the context's stream, which is non-blocking and NOT ordered with
 *     any other stream, HIP's null stream included: device inputs written on
 *     other streams must be complete (the caller synchronises that stream)
 *     before the call.
 * Because every call returns only when its results are written, work queued
 * after it on any stream sees them.  Server jobs (kc_submit) run on the
 * server's own streams: their device arrays must be complete before
 * kc_submit and must not be touched until the ticket completes.
 */
#ifndef KASCADE_AMD_H
#define KASCADE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes replace the reference's std::runtime_error / KASSERT aborts
 * (pointer_doubling_generic.hpp:446-447, sparse_ruling_set.cpp:374,
 * eulertour.cpp:318). */
typedef enum kc_status {
  KC_OK = 0,
  KC_EINVAL = 1,    /* bad argument, id >= n, in-degree > 1 where a list is required,
                       or nonzero weight on a root (pointer_doubling_generic.hpp:431) */
  KC_ECYCLE = 2,    /* a cycle without a root (the reference would not terminate) */
  KC_ENOMEM = 3,    /* device allocation failed */
  KC_EDEVICE = 4,   /* HIP runtime error */
  KC_EBUDGET = 5,   /* job larger than the server's memory budget */
  KC_EAGAIN = 6,    /* result not ready yet (kc_poll) */
  KC_ENOTFOUND = 7  /* unknown ticket */
} kc_status;

/* Mirrors kascade::Algorithm (lib/kascade/configuration.hpp:9-18) for the
 * members that exist on one GPU. */
typedef enum kc_algorithm {
  KC_ALGO_AUTO = 0,             /* pointer doubling up to 2^19 list elements or 2^17 tree
                                   vertices (measured faster there), the sparse ruling set above
                                   or when the config sets a ruling-set knob (ruler_spacing,
                                   ruler_spacing_deep, max_levels, spawn ON); batched serving
                                   always ranks with the sparse ruling set */
  KC_ALGO_POINTER_DOUBLING = 1, /* sync Wyllie doubling, pointer_doubling_generic.hpp:608-700 */
  KC_ALGO_SPARSE_RULING_SET = 2 /* ruler chase + recursion, sparse_ruling_set.cpp:45-393 */
} kc_algorithm;

/* How the levels of a ranking are sized and launched (kc_rank_cfg.launch).
 * No reference counterpart: the reference sizes each round with MPI
 * collectives.  Device-sized levels read their element counts from device
 * memory inside capacities fixed from (n, number of lists), so a ranking
 * needs no host round trip until it ends and can be captured as a HIP
 * graph.  An input whose unreached list prefixes exceed the capacity (far
 * beyond what random inputs reach) is redone with host-sized levels. */
typedef enum kc_launch {
  KC_LAUNCH_AUTO = 0,   /* device-sized; a call repeated with the same arrays and
                           sizes is captured as a HIP graph (once per context)
                           and replayed */
  KC_LAUNCH_DEVICE = 1, /* device-sized, kernels launched one by one */
  KC_LAUNCH_HOST = 2    /* each level's counts read on the host to size the next */
} kc_launch;

/* Knobs of kascade::SparseRulingSetConfig / PointerDoublingConfig
 * (configuration.hpp:39-47, :65-95) that mean something on one GPU.  Zero
 * means "choose" (the per-size autotuned table). */
typedef struct kc_rank_cfg {
  int32_t algorithm;        /* kc_algorithm */
  uint32_t ruler_spacing;   /* elements per ruler at the top level (power of two), 0 = auto */
  uint32_t ruler_spacing_deep; /* elements per ruler below the top level, 0 = auto */
  uint32_t max_levels;      /* recursion depth cap (sparse_ruling_set_rounds), 0 = auto */
  uint32_t validate;        /* 1: reject in-degree > 1 before ranking (is_list, successor_utils.cpp:31-63) */
  uint32_t chase_blocks_per_cu; /* resident 256-lane chase workgroups per CU, 0 = auto */
  uint32_t chase_chunk;     /* rulers a wavefront takes per work-counter atomic, 0 = auto */
  uint32_t launch;          /* kc_launch */
  uint32_t spawn;           /* kc_spawn */
  uint32_t local_contraction; /* kc_contract */
  uint32_t records;         /* kc_records */
  uint32_t reserved;        /* zero */
} kc_rank_cfg;

/* Local contraction of list rankings (kc_rank_cfg.local_contraction), the
 * counterpart of the reference's use_local_contraction
 * (pointer_doubling_generic.hpp:471-543, sparse_ruling_set_detail/
 * local_contraction.hpp:28-104): runs of consecutive successors
 * (succ[v] = v + 1) are collapsed into one element each before ranking and
 * expanded afterwards.  Results are identical either way. */
typedef enum kc_contract {
  KC_CONTRACT_AUTO = 0, /* when a sample finds at least half of the elements inside runs */
  KC_CONTRACT_OFF = 1,
  KC_CONTRACT_ON = 2
} kc_contract;

/* Ruler spawn in the sparse ruling set's chase (kc_rank_cfg.spawn), the
 * counterpart of the reference's `spawn` (sparse_ruling_set.cpp:186-231,
 * :250-256): once most chases have ended, the remaining ones are suspended,
 * new rulers are started at unreached elements (candidates at a hashed place
 * in every block of ids, finer than the rulers) and everything resumes, so
 * the longest remaining gaps are split instead of walked to the end by a few
 * lanes.  In an Euler tour the candidates are upward edges (the tour
 * elements that carry claims); AUTO leaves spawn off for tours (2^26-vertex
 * tree on MI355X: 8.58 ms with spawn against 8.38 without at 128
 * elements per ruler).  Results are identical either way. */
typedef enum kc_spawn {
  KC_SPAWN_AUTO = 0, /* on for list levels (no tours) of at least 2^20 elements and 256 or more per ruler */
  KC_SPAWN_OFF = 1,
  KC_SPAWN_ON = 2    /* on every level with more than 8 elements per ruler, tours included */
} kc_spawn;

/* Record width of the top level (kc_rank_cfg.records).  AUTO: 8-byte records
 * (32-bit successor, 32-bit weight and claim distance; in segmented batches
 * the successor relative to its element) for sparse-ruling-set rankings from
 * 2^22 records whose ids, weights and distances fit, 16-byte records
 * otherwise; WIDE: always 16-byte records.  Results are identical either
 * way. */
typedef enum kc_records {
  KC_RECORDS_AUTO = 0,
  KC_RECORDS_WIDE = 1
} kc_records;

/* ---- context ------------------------------------------------------------ */
kc_status kc_context_create(int device, uint64_t mem_budget_bytes, void **ctx_out);
void kc_context_destroy(void *ctx);
/* Human-readable text of the last error on this context. */
const char *kc_last_error(void *ctx);
/* Device memory currently held by the context's scratch pool, in bytes. */
uint64_t kc_context_scratch_bytes(void *ctx);
/* Releases the scratch pool (frees its device memory). */
kc_status kc_context_trim(void *ctx);
/* kc_context_destroy hands the context's scratch chunks to a per-device
 * cache (at most a quarter of the device's memory) that later contexts take
 * from before they allocate, so servers and contexts that come and go do not
 * allocate and free device memory each time.  This frees that cache.  No
 * reference counterpart (the reference allocates per call with std::vector). */
kc_status kc_release_cached_scratch(int device);
/* Launch statistics of the context since its creation (see kc_launch):
 * out3 = {HIP graphs captured, graph replays, device-sized rankings redone
 * host-sized after a capacity overflow}.  No reference counterpart. */
kc_status kc_context_launch_stats(void *ctx, uint64_t *out3);
/* Level-0 record width statistics since the context's creation: out2 =
 * {list rankings whose top level used 8-byte records, such rankings redone
 * with 16-byte records (a weight or distance outside int32)}.  No reference
 * counterpart. */
kc_status kc_context_record_stats(void *ctx, uint64_t *out2);

/* Per-phase device timing (HIP events on the launching stream), the
 * counterpart of the reference's kamping::measurements::timer phases
 * (e.g. "chase_ruler", "ruler_propagation", sparse_ruling_set.cpp:147,315).
 * kc_profile_read returns the number of phases and fills up to max_entries
 * (name, total ms, launches); names stay valid until the next reset. */
kc_status kc_profile_enable(void *ctx, int on);
kc_status kc_profile_reset(void *ctx);
uint32_t kc_profile_read(void *ctx, uint32_t max_entries, const char **names, double *ms,
                         uint64_t *counts);

/* Per-level trace of the rankings run while tracing is on (the counterpart of
 * RulerTrace, lib/kascade/sparse_ruling_set_detail/trace.hpp:14-166):
 * levels in the order they ran; subproblem 0 = the ruler sublist recursion,
 * 1 = the unclaimed head prefixes.  For an SRS level, chains are the
 * sublists the rulers chase (the ruler included, the stopping element
 * excluded); the longest chain is the GPU's counterpart of the reference's
 * ruler_chasing_rounds.  kc_trace_enable(ctx, 1) clears and starts;
 * kc_trace_read returns the number of entries and fills up to max_entries. */
typedef enum kc_trace_kind { KC_TRACE_SRS = 0, KC_TRACE_PD = 1, KC_TRACE_SMALL = 2 } kc_trace_kind;
typedef struct kc_trace_level {
  int32_t level;
  int32_t kind;       /* kc_trace_kind */
  int32_t subproblem; /* 0 ruler sublist, 1 unclaimed prefixes */
  uint32_t rounds;    /* pointer doubling rounds (KC_TRACE_PD) */
  uint64_t n;         /* elements at this level (local_subproblem_size) */
  uint64_t spacing;   /* elements per ruler (KC_TRACE_SRS) */
  uint64_t rulers;    /* rulers picked (local_num_ruler) */
  uint64_t tail_rulers; /* list tails reached by chases, ranked as sublists of their own */
  uint64_t unreached; /* head-prefix elements no ruler reached (num_unreached) */
  uint64_t chain_length_min;
  uint64_t chain_length_max;
  double chain_length_avg;
  uint64_t spawned;   /* rulers spawned during the chase (spawned_rulers) */
} kc_trace_level;
kc_status kc_trace_enable(void *ctx, int on);
uint32_t kc_trace_read(void *ctx, kc_trace_level *out, uint32_t max_entries);

/* Per-size-class tuning (the autotuner's table): rows of 5 uint32
 * {log2(n) class, ruler spacing top, ruler spacing deep, chase workgroups per
 * CU, rulers per work-counter atomic}; a size uses the row of the largest
 * class not above log2(n).  Contexts load tuning_table.txt next to the library
 * (or $KASCADE_AMD_TUNING) at creation.  kc_tuning_get reports the knobs a
 * size n would use (out5: spacing top, spacing deep, workgroups per CU,
 * chunk, one-workgroup threshold) and returns the number of loaded rows. */
kc_status kc_tuning_set(void *ctx, const uint32_t *rows, uint32_t nrows);
uint32_t kc_tuning_get(void *ctx, uint64_t n, uint32_t *out5);
/* The same per workload kind: KC_TUNING_LIST is the table above;
 * KC_TUNING_TREE is tree rooting, whose classes count vertices (the tour has
 * 2n records) and whose rows the table file prefixes with "tree";
 * KC_TUNING_SEGMENTED is kc_rank_lists_segmented, whose classes count the
 * batch's elements and whose rows the file prefixes with "seg" (a batch below
 * its first class uses the list table).  Mixed serving batches use the list
 * table.  No reference counterpart (the reference picks ruler counts by
 * formula, ruler_selection.hpp:18-79). */
typedef enum kc_tuning_kind {
  KC_TUNING_LIST = 0,  /* = KC_JOB_LIST */
  KC_TUNING_TREE = 1,  /* = KC_JOB_TREE */
  KC_TUNING_SEGMENTED = 2
} kc_tuning_kind;
kc_status kc_tuning_set_kind(void *ctx, int32_t kind, const uint32_t *rows, uint32_t nrows);
uint32_t kc_tuning_get_kind(void *ctx, int32_t kind, uint64_t n, uint32_t *out5);

/* ---- list ranking ------------------------------------------------------- */

/* Replaces kascade::set_initial_ranking_state (list_ranking.hpp:47-50,
 * list_ranking.cpp:60-80): root := succ, rank := 1, rank := 0 where succ[i]==i. */
kc_status kc_set_initial_ranking_state(void *ctx, const uint64_t *succ, uint64_t n,
                                       uint64_t *root_out, int64_t *rank_out, void *stream);

/* Ranks a forest of lists.  succ[i] is the successor of i, a tail points to
 * itself.  weight == NULL means unit weights as set_initial_ranking_state
 * (rank 1, 0 on tails); otherwise weight[i] is the distance from i to succ[i]
 * and must be 0 on tails.  Outputs: root_out[i] = the tail of i's list,
 * rank_out[i] = sum of weights from i (inclusive) to the tail (exclusive).
 * root_out/rank_out may alias succ/weight (the reference's in-place contract,
 * pointer_doubling.hpp:14-17).  Replaces kascade::pointer_doubling
 * (pointer_doubling.hpp:18-23) and kascade::sparse_ruling_set
 * (sparse_ruling_set.hpp:13-18); cfg->algorithm selects which. */
kc_status kc_rank_list(void *ctx, const uint64_t *succ, const int64_t *weight, uint64_t n,
                       uint64_t *root_out, int64_t *rank_out, const kc_rank_cfg *cfg,
                       void *stream);

/* kc_rank_list without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch is a static device array of
 * KC_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
 * the process ends; the call itself calls no hipMalloc or hipFree and
 * captures no HIP graph (cfg->launch AUTO runs as KC_LAUNCH_DEVICE).
 *   - n must be at most KC_NOALLOC_MAX_N (else KC_EINVAL); with the built-in
 *     tuning every list or forest of that size fits.  A caller-fixed ruler
 *     spacing or an adversarial input that needs more scratch gets
 *     KC_ENOMEM before any kernel writes past the array.
 *   - succ, weight (may be NULL), root_out and rank_out must be device
 *     memory (else KC_EINVAL): host arrays would need staging buffers.
 *   - Every call on one GPU shares that GPU's array, whatever the context:
 *     calls that overlap on one GPU (two host threads, or two streams of
 *     one caller) race and must be serialised by the caller.  The array
 *     keeps the last call's contents; every call clears what it needs.
 * Results are identical to kc_rank_list's. */
#define KC_NOALLOC_MAX_N ((uint64_t)1 << 20)
#define KC_NOALLOC_SCRATCH_BYTES ((size_t)192 << 20)
kc_status kc_rank_list_noalloc(void *ctx, const uint64_t *succ, const int64_t *weight, uint64_t n,
                               uint64_t *root_out, int64_t *rank_out, const kc_rank_cfg *cfg,
                               void *stream);

/* In-place forms with the reference signatures: succ_array in = successors,
 * out = roots; rank_array in = distance to successor (0 on roots), out =
 * distance to root.  kascade::pointer_doubling (pointer_doubling.hpp:18-23). */
kc_status kc_pointer_doubling(void *ctx, const kc_rank_cfg *cfg, uint64_t *succ_array,
                              int64_t *rank_array, uint64_t n, void *stream);
/* kascade::sparse_ruling_set (sparse_ruling_set.hpp:13-18). */
kc_status kc_sparse_ruling_set(void *ctx, const kc_rank_cfg *cfg, uint64_t *succ_array,
                               int64_t *rank_array, uint64_t n, void *stream);

/* State after exactly `rounds` synchronous doubling rounds (each a Jacobi step
 * of DoublingWithoutAggregation::execute_doubling_step,
 * pointer_doubling_generic.hpp:187-243).  succ_out keeps the root flag
 * (bit 63) on entries that point at a root, as the reference's arrays do
 * inside the doubling loop (:418-438, :677-688).  Exposed so the per-round
 * state can be checked; not a ranking entry point. */
kc_status kc_pointer_doubling_rounds(void *ctx, const uint64_t *succ, const int64_t *weight,
                                     uint64_t n, uint32_t rounds, uint64_t *succ_out,
                                     int64_t *rank_out, void *stream);

/* Order-preserving compaction of the active elements (pack.hpp:23-80): the
 * packed id of active[k] is k; psucc_out[k] = packed id of succ[active[k]],
 * prank_out[k] = rank[active[k]].  `active` must be ascending and closed under
 * succ.  kc_unpack writes packed results back (pack.hpp:81-175). */
kc_status kc_pack(void *ctx, const uint64_t *succ, const int64_t *rank, uint64_t n,
                  const uint64_t *active, uint64_t nactive, uint64_t *psucc_out,
                  int64_t *prank_out, void *stream);
kc_status kc_unpack(void *ctx, const uint64_t *psucc, const int64_t *prank,
                    const uint64_t *active, uint64_t nactive, uint64_t *succ_inout,
                    int64_t *rank_inout, uint64_t n, void *stream);

/* ---- batched segmented ranking ----------------------------------------- */
/* nseg independent segments; segment s owns elements [seg_off[s], seg_off[s+1])
 * and its successors are local ids in [0, size_s).  Equivalent to ranking the
 * forest with global ids seg_off[s] + local.  root_local receives local ids.
 * weight == NULL: unit weights as set_initial_ranking_state.  The whole batch
 * is one ranking with one ruler spacing, capped at the mean segment length /
 * 16 (a chase ends at every segment's tail, so short segments need no denser
 * rulers of their own).  The knobs come from the segmented tuning table
 * (KC_TUNING_SEGMENTED) by the batch's total size: from 2^27 elements 256
 * per ruler at the top level, 16 below, 2 chase workgroups per CU (on 20,000
 * lists of 2^8..2^20 elements, 2.5e9 in all: 94.6 ms on MI355X; 105.2 with
 * the list table's knobs).  Smaller batches use the list table with the
 * top-level spacing capped, when the batch holds more than 64 lists, so that
 * the chases in flight (taken in id order, K ids apart) stay inside a window
 * of records that fits the GPU's 256 MB Infinity Cache.  Batches from 2^22
 * elements pack 8-byte records (the successor stored relative to its
 * element, the weight in 32 bits) when every segment is below 2^29 elements
 * and the weights fit; otherwise, or when a claim distance outgrows int32,
 * 16-byte records.  cfg->local_contraction applies to the batch as a
 * whole. */
kc_status kc_rank_lists_segmented(void *ctx, const uint32_t *succ_local, const uint64_t *seg_off,
                                  uint64_t nseg, const int64_t *weight, uint32_t *root_local,
                                  int64_t *rank_out, const kc_rank_cfg *cfg, void *stream);

/* ---- tree rooting ------------------------------------------------------- */
/* How the children of every vertex are grouped before tour linking. */
typedef enum kc_tour_build {
  KC_TOUR_AUTO = 0,   /* = KC_TOUR_LINKED */
  KC_TOUR_SORTED = 1, /* stable device radix sort of (parent, child): the tour is reproducible */
  KC_TOUR_LINKED = 2  /* child lists by atomic exchange: one pass, child order may vary per run */
} kc_tour_build;

typedef struct kc_tree_cfg {
  kc_rank_cfg rank;   /* how the Euler tour is ranked (eulertour.cpp:282-321) */
  int32_t tour_build; /* kc_tour_build; (root, depth) are identical either way */
  uint32_t reserved[3];
} kc_tree_cfg;

/* parent[v] = parent of v, roots are their own parent.  Outputs root_out[v]
 * (the tree root) and depth_out[v] (edges to the root).  Replaces
 * kascade::rank_via_euler_tour (eulertour.hpp:22-27, eulertour.cpp:337-377)
 * after set_initial_ranking_state, as the harness calls it
 * (bench/detail/algorithm_impl.hpp:191-222).  May alias (in-place). */
kc_status kc_root_tree(void *ctx, const uint64_t *parent, uint64_t n, uint64_t *root_out,
                       int64_t *depth_out, const kc_tree_cfg *cfg, void *stream);
/* In-place form with the reference signature: succ_array in = parents, out =
 * roots; rank_array out = depth (input values ignored, eulertour.cpp:45-53). */
kc_status kc_rank_via_euler_tour(void *ctx, const kc_tree_cfg *cfg, uint64_t *succ_array,
                                 int64_t *rank_array, uint64_t n, void *stream);
/* The tour successor array alone, in the reference's element order
 * (kascade::compute_euler_tour, eulertour.cpp:379-392, as one PE builds it):
 * one element per directed tree edge, grouped by source vertex in ascending
 * order and, within a vertex, in the order of make_graph's CSR
 * (successor_utils.cpp:236-260): children above the vertex in descending
 * order, its parent, children below it in descending order; then one dummy
 * per root that has children, in vertex order, then one self-looping slot
 * per childless root.  succ[(u_prev, v)] = (v, u_next) in that cyclic order;
 * the wrap to a root's first neighbour goes to its dummy
 * (compute_euler_tour_impl, eulertour.cpp:33-127).  *tour_len_out =
 * 2(n - roots) + roots; tour_succ_out (and tour_weight_out, may be NULL:
 * +1 upward, -1 downward, 0 dummies) need capacity 2n. */
kc_status kc_compute_euler_tour(void *ctx, const uint64_t *parent, uint64_t n,
                                uint64_t *tour_succ_out, int64_t *tour_weight_out,
                                uint64_t *tour_len_out, void *stream);

/* ---- list utilities and statistics --------------------------------------- */
/* *is_list_out = 1 iff no element has two predecessors (self loops ignored).
 * Replaces kascade::is_list (successor_utils.cpp:23-63).  KC_EINVAL for ids
 * >= n. */
kc_status kc_is_list(void *ctx, const uint64_t *succ, uint64_t n, int32_t *is_list_out, void *stream);
/* is_leaf_out[v] = 1 iff v has no predecessor (may be NULL), and the count.
 * Replaces kascade::LeafInfo (successor_utils.cpp:72-123): as there, a root
 * without predecessor counts as a leaf. */
kc_status kc_leaf_info(void *ctx, const uint64_t *succ, uint64_t n, uint8_t *is_leaf_out,
                       uint64_t *num_leaves_out, void *stream);
/* pred_out[succ[u]] = u and dist_to_pred_out[succ[u]] = dist_to_succ[u] (1 when
 * dist_to_succ is NULL); elements without predecessor get (self, 0).  The
 * original roots (succ[i] == i) are written in ascending order to roots_out
 * (up to roots_capacity; may be NULL) and counted in *num_roots_out.  Replaces
 * kascade::reverse_list (successor_utils.cpp:159-232).  KC_EINVAL if an
 * element has two predecessors. */
kc_status kc_reverse_list(void *ctx, const uint64_t *succ, const int64_t *dist_to_succ, uint64_t n,
                          uint64_t *pred_out, int64_t *dist_to_pred_out, uint64_t *roots_out,
                          uint64_t roots_capacity, uint64_t *num_roots_out, void *stream);

/* Statistics levels, as StatsLevel (bench/detail/stats.hpp:9-15). */
typedef enum kc_stats_level {
  KC_STATS_NONE = 1,
  KC_STATS_BASIC = 2,
  KC_STATS_REDUCED_EXTENSIVE = 3, /* aggregates without the per-tree list */
  KC_STATS_EXTENSIVE = 4
} kc_stats_level;

/* BasicStats and ExtensiveStats (bench/detail/stats.hpp:17-39). */
typedef struct kc_stats {
  uint64_t num_trees;
  int64_t max_rank;
  double avg_rank;
  uint64_t num_nontrivial_trees; /* trees of more than one element */
  uint64_t nontrivial_size_sum;
  int64_t nontrivial_rank_sum;
  uint64_t max_size;
  double avg_size;
  double nontrivial_avg_size;
  double nontrivial_avg_rank;
} kc_stats;

/* compute_stats (bench/detail/stats.cpp:147-162) over a ranking's outputs.
 * With KC_STATS_EXTENSIVE the nontrivial trees' (root, size, max rank, rank
 * sum) are written in ascending root order, up to tree_capacity entries (the
 * arrays may be NULL); their number is out->num_nontrivial_trees. */
kc_status kc_compute_stats(void *ctx, int32_t level, const uint64_t *root, const int64_t *rank, uint64_t n,
                           kc_stats *out, uint64_t *tree_root_out, uint64_t *tree_size_out,
                           int64_t *tree_max_rank_out, int64_t *tree_rank_sum_out, uint64_t tree_capacity,
                           void *stream);

/* ---- synthetic inputs (seeded; identical to the CPU generators) ---------- */
/* Random list: order pi = seeded Feistel permutation, succ[pi(i)] = pi(i+1);
 * weighted: weights uniform in [1,100], 0 on the tail; else 1 / 0.
 * local != 0: blocks of 10 consecutive ids visited in permuted block order. */
kc_status kc_gen_list(void *ctx, uint64_t n, uint64_t seed, int weighted, int local,
                      uint64_t *succ_out, int64_t *weight_out, void *stream);
/* Random recursive tree (parent of generation index i uniform in [0,i)),
 * relabelled by a seeded permutation. */
kc_status kc_gen_tree(void *ctx, uint64_t n, uint64_t seed, uint64_t *parent_out, void *stream);
/* Segment-local variant for batches: segment s is a list of size
 * seg_off[s+1]-seg_off[s] generated with seed seeds[s] (local ids, u32). */
kc_status kc_gen_lists_segmented(void *ctx, const uint64_t *seg_off, uint64_t nseg,
                                 const uint64_t *seeds, const uint8_t *local_flags, int weighted,
                                 uint32_t *succ_out, int64_t *weight_out, void *stream);

/* ---- graph inputs (the reference's "bfs" / "eulertour" input processing) -- */
typedef enum kc_graph_kind {
  KC_GRAPH_GNM = 0, /* m edges, both endpoints uniform (KaGen gnm-undirected) */
  KC_GRAPH_RMAT = 1 /* m edges, Graph500 R-MAT probabilities (0.57, 0.19, 0.19, 0.05) */
} kc_graph_kind;
/* m seeded random undirected edges on n vertices (identical to the CPU
 * generator); a self loop's head moves to the next vertex. */
kc_status kc_gen_graph(void *ctx, int32_t kind, uint64_t n, uint64_t m, uint64_t seed, uint64_t *src_out,
                       uint64_t *dst_out, void *stream);
/* Random geometric graph in the unit square, the role of KaGen's rgg2d in the
 * reference's Euler-tour suites (experiments/submission/
 * sparse-ruling-set-eulertour.suite.yaml:19-28): n seeded points with 31-bit
 * fixed-point coordinates and an undirected edge (i, j), i < j, for every pair
 * within radius r = sqrt(2m / (pi n^2)), so that about m edges are expected
 * (fewer near the border).  With src_out or dst_out NULL only *m_out (the
 * edge count) is computed; otherwise cap must be at least that count and the
 * edges are written grouped by i ascending (identical to the CPU generator). */
kc_status kc_gen_rgg2d(void *ctx, uint64_t n, uint64_t m, uint64_t seed, uint64_t *src_out, uint64_t *dst_out,
                       uint64_t cap, uint64_t *m_out, void *stream);
/* CSR of the undirected graph: xadj_out (n+1 entries), adj_out (2m entries,
 * both directions of every edge; the order within a vertex is unspecified). */
kc_status kc_build_csr(void *ctx, const uint64_t *src, const uint64_t *dst, uint64_t n, uint64_t m,
                       uint64_t *xadj_out, uint64_t *adj_out, void *stream);
/* BFS forest of a symmetric CSR graph (generate_bfs_tree,
 * bench/input/bfs_forest_generation.cpp:52-76): isolated vertices are their
 * own parents; every other component is a BFS tree rooted at its smallest
 * vertex id (parent_out[root] = root).  Roots and depths are unique; a
 * vertex's parent is one of its neighbours one level up.  KC_EINVAL for a
 * non-symmetric adjacency. */
kc_status kc_bfs_forest(void *ctx, const uint64_t *xadj, const uint64_t *adj, uint64_t n, uint64_t *parent_out,
                        void *stream);

/* ---- diagnostics ------------------------------------------------------- */
/* Random-gather ceiling: dst[i] = src[idx[i]] over n uint64 with idx a seeded
 * random permutation (n = 2^27 exceeds every cache).  Reports the effective
 * rate credited 24 bytes per element (idx 8 + src 8 + dst 8), in GB/s, and the
 * kernel time.  The reference point for the list-ranking kernels' gathers. */
kc_status kc_probe_random_gather(void *ctx, uint64_t n, uint32_t reps, double *gbps_out,
                                 double *ms_out);
/* The same with a memory mode: 0 plain, 1 non-temporal loads, 2 an uncached
 * source allocation, 3 a fine-grained source allocation (2 and 3 are kept
 * allocated until the process ends). */
kc_status kc_probe_random_gather_mode(void *ctx, uint64_t n, uint32_t reps, int mode,
                                      double *gbps_out, double *ms_out);
/* n random gathers from a table of m uint64 (m <= n): gathers per second as a
 * function of the table size (which cache level serves a random lookup). */
kc_status kc_probe_table_gather(void *ctx, uint64_t n, uint64_t m, uint32_t reps,
                                double *gathers_per_s, double *ms_out);
/* Access patterns of 16-byte records over one random cycle of n records (n a
 * power of two, >= 256): 0 independent random reads, 1 independent random
 * writes, 2 independent read + 16-byte write of the same record, 3 dependent
 * walks (n/256 walkers of 256 steps) reading only, 4 the same walks writing a
 * 16-byte claim into every record (the ruler chase's pattern); 5 and 6 the
 * walks of 3 and 4 over 8-byte records (the narrow level-0 form; n at most
 * 2^30).  Records per second and the kernel time: the cost of a random write
 * against a random read. */
kc_status kc_probe_record_access(void *ctx, uint64_t n, int pattern, uint32_t reps,
                                 double *records_per_s, double *ms_out);

/* ---- serving ------------------------------------------------------------ */
typedef enum kc_job_kind { KC_JOB_LIST = 0, KC_JOB_TREE = 1 } kc_job_kind;

/* A job's arrays (host or device memory) are owned by the caller and must
 * stay valid until the ticket completes; device arrays are ranked in place,
 * host arrays are staged.  LIST: succ/weight in, root/rank out.  TREE: succ =
 * parent in, root/rank(depth) out, weight ignored. */
typedef struct kc_job {
  int32_t kind; /* kc_job_kind */
  uint32_t reserved;
  uint64_t n;
  const uint64_t *succ;
  const int64_t *weight; /* may be NULL */
  uint64_t *root_out;
  int64_t *rank_out;
} kc_job;

typedef struct kc_server_cfg {
  uint64_t mem_budget_bytes;   /* device-memory admission budget */
  uint32_t max_batch_jobs;     /* jobs coalesced into one segmented launch, 0 = auto */
  uint32_t num_slots;          /* batches in flight on separate streams; 0 = 3, or 1 with one_job_per_launch */
  uint64_t max_batch_elements; /* elements per batch, 0 = 2^26 */
  uint32_t one_job_per_launch; /* 1: no coalescing (the baseline mode) */
  uint32_t tour_build;         /* kc_tour_build for tree jobs */
  uint32_t reserved[2];
} kc_server_cfg;

typedef struct kc_server_stats {
  uint64_t jobs_completed;
  uint64_t elements_completed;
  uint64_t batches_launched;
  double p50_latency_ms;
  double p99_latency_ms;
  double mean_latency_ms;
  double jobs_per_s;       /* completions / (last completion - first submit) */
  double elements_per_s;
  uint64_t peak_device_bytes; /* peak admitted footprint (the budget's accounting) */
  uint64_t batches_rerun;     /* coalesced passes that failed and were re-run job by job */
  uint64_t jobs_failed;       /* jobs that completed with an error status */
  uint64_t scratch_bytes;     /* device memory the slots' scratch pools hold now */
} kc_server_stats;

kc_status kc_server_create(void *ctx, const kc_server_cfg *cfg, void **server_out);
/* Stops admitting, drains in-flight work, frees everything. */
void kc_server_destroy(void *server);
/* Enqueue a job; returns at once.  KC_EBUDGET if the job can never fit. */
kc_status kc_submit(void *server, const kc_job *job, uint64_t *ticket_out);
/* KC_OK once the job's outputs are written, KC_EAGAIN before, or its error. */
kc_status kc_poll(void *server, uint64_t ticket);
kc_status kc_wait(void *server, uint64_t ticket);
kc_status kc_server_stats_get(void *server, kc_server_stats *out);
kc_status kc_server_reset_stats(void *server);

#ifdef __cplusplus
}
#endif
#endif /* KASCADE_AMD_H */
